// Device-side types and helpers of the gfx950 mesh stage.
//
// All device code is compiled with -ffp-contract=off and without fast math:
// every f64 and f32 expression keeps the reference's operation order (Rust
// evaluates left to right and never fuses), so values equal the CPU's bit
// for bit. Wavefronts are 64 lanes (CDNA); cooperative loops stride by 64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smg {

constexpr int WAVE = 64;
constexpr double EPS = 1e-9;            // geom/arrangement.rs:26
constexpr double CLASSIFY_EPS = 2e-3;   // field/lattice.rs:20
constexpr float INSIDE_TOL = 2e-5f;     // field/kernel.rs:42
constexpr int64_t KEY_BIAS = int64_t(1) << 20;  // field/lattice.rs:15

enum : uint8_t { ST_NONE = 0, ST_OUT = 1, ST_IN = 2 };  // sub-block / brick states
enum : uint32_t { S_EXPOSED = 0, S_BURIED = 1, S_CONTAINED = 2 };
enum : uint32_t { F_SOLVENT = 0, F_BEYOND = 1, F_SPHERE = 2, F_CIRCLE = 3, F_VERTEX = 4 };

// A cell grid (reference geom/cells.rs:9-130) in HBM.
struct DGrid {
    const uint32_t* starts;
    const uint32_t* items;
    double ox, oy, oz, cell;
    uint32_t dx, dy, dz, pad;
};

// The resident f64 arrangement, SoA.
struct DArr {
    const double *cx, *cy, *cz, *R, *r;
    const uint8_t* status;
    const uint32_t* orig;
    const uint32_t *nb_off, *nb_items;
    const double *mx, *my, *mz, *ax, *ay, *az, *e1x, *e1y, *e1z, *crho;
    const uint32_t *ci, *cj;
    const uint32_t *cb_off, *cb_items;
    const double *vx, *vy, *vz;
    const uint32_t* vatoms;
    DGrid gs, gc, gv;
    uint32_t n_spheres, n_circles, n_vertices, pad;
    double probe, max_expanded, rho_max;
};

// Frame-local f32 copies (reference field/gpu.rs:253-318, kernel.rs:604-747).
struct DLoc {
    float *px, *py, *pz, *pr, *pr2, *pr2tol;
    float *mx, *my, *mz, *ux, *uy, *uz, *ex, *ey, *ez, *rho;
    float *vx, *vy, *vz;
};

// Lattice and stage parameters of one mesh run.
struct DRun {
    double h, band, probe, t_min;
    int64_t frame[3];
    int64_t blo[3];       // dense brick box: lowest key per axis
    uint32_t bdim[3];     // dense brick box size per axis
    uint32_t b, sub, side, spb;  // brick edge, sub-block edge, b+1, (b+1)^3
    float h32, fband, fprobe, fcutoff2;
};

__device__ inline double ddist(double ax, double ay, double az, double bx, double by, double bz) {
    double x = ax - bx, y = ay - by, z = az - bz;
    return sqrt(x * x + y * y + z * z);
}

// geom/cells.rs:91-105 for one axis.
__device__ inline uint32_t gcoord(double p, double o, double cell, uint32_t dim) {
    double f = floor((p - o) / cell);
    if (!(f > 0.0)) return 0;
    return f >= (double)(dim - 1) ? dim - 1 : (uint32_t)f;
}

struct Box {
    uint32_t x0, x1, y0, y1, z0, z1;
};
__device__ inline Box grid_box(const DGrid& g, double x, double y, double z, double r) {
    Box b;
    b.x0 = gcoord(x - r, g.ox, g.cell, g.dx);
    b.x1 = gcoord(x + r, g.ox, g.cell, g.dx);
    b.y0 = gcoord(y - r, g.oy, g.cell, g.dy);
    b.y1 = gcoord(y + r, g.oy, g.cell, g.dy);
    b.z0 = gcoord(z - r, g.oz, g.cell, g.dz);
    b.z1 = gcoord(z + r, g.oz, g.cell, g.dz);
    return b;
}

// Items of the cells overlapping the cube of half-size r (CellGrid::
// for_each_near), `lane` of `lanes` threads taking every lanes-th item of
// each row.
template <class F>
__device__ inline void for_near(const DGrid& g, double x, double y, double z, double r, uint32_t lane,
                                uint32_t lanes, F&& f) {
    Box b = grid_box(g, x, y, z, r);
    for (uint32_t zc = b.z0; zc <= b.z1; zc++)
        for (uint32_t yc = b.y0; yc <= b.y1; yc++) {
            uint32_t row = (zc * g.dy + yc) * g.dx;
            uint32_t a = g.starts[row + b.x0], e = g.starts[row + b.x1 + 1];
            for (uint32_t i = a + lane; i < e; i += lanes) f(g.items[i]);
        }
}

__device__ inline double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ inline float wave_minf(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }

__device__ inline uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }

// f64::total_cmp as an order-preserving integer key.
__device__ inline uint64_t total_key(double d) {
    int64_t i = __double_as_longlong(d);
    i ^= (int64_t)(((uint64_t)(i >> 63)) >> 1);
    return (uint64_t)i ^ 0x8000000000000000ull;
}

}  // namespace smg
