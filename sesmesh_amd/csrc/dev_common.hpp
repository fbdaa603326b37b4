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
// Workgroup size of the passes over the large LDS tiers: with one or two
// workgroups per CU, 1024 threads give each SIMD four wavefronts to hide
// LDS and memory latency with, where 256 would give it one.
constexpr int WIDE = 1024;
constexpr double EPS = 1e-9;            // geom/arrangement.rs:26
constexpr double CLASSIFY_EPS = 2e-3;   // field/lattice.rs:20
constexpr float INSIDE_TOL = 2e-5f;     // field/kernel.rs:42
constexpr int64_t KEY_BIAS = int64_t(1) << 20;  // field/lattice.rs:15

enum : uint8_t { ST_NONE = 0, ST_OUT = 1, ST_IN = 2 };  // sub-block / brick states
enum : uint32_t { S_EXPOSED = 0, S_BURIED = 1, S_CONTAINED = 2 };
enum : uint32_t { F_SOLVENT = 0, F_BEYOND = 1, F_SPHERE = 2, F_CIRCLE = 3, F_VERTEX = 4 };
// Surface kinds (field/mod.rs:33-41). SAS and vdW are unions of spheres:
// their arrangements have no circles or SAS vertices, and the vdW one is
// built with probe 0.
enum : uint32_t { K_SES = 0, K_SAS = 1, K_VDW = 2 };

// Pointers into device global memory (address space 1). The arrangement's
// pointers are read from the molecule table in memory; typed like this,
// loads through them compile to global loads, not flat ones.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
inline gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}

// A cell grid (reference geom/cells.rs:9-130) in HBM.
struct DGrid {
    gptr<const uint32_t> starts;
    gptr<const uint32_t> items;
    double ox, oy, oz, cell;
    uint32_t dx, dy, dz, pad;
};

// Packed records read by the hot gathers: one 32- or 64-byte load per
// candidate instead of one load per field.
struct alignas(32) SphereRec {
    double x, y, z, R;  // centre, expanded radius
};
struct alignas(64) CircleRec {
    double mx, my, mz, ax, ay, az, rho, pad;  // centre, axis, radius
};
struct alignas(32) VertexRec {
    double x, y, z, pad;
};
// A sphere's squared thresholds of the refinement's distance tests,
// sqrt_threshold(R - EPS) (blocking) and sqrt_threshold(R) (inside), with R
// the expanded radius: computed once when the arrangement is uploaded.
struct alignas(16) SphereThr {
    double blk, nin;
};

// The resident f64 arrangement, SoA, plus packed records of the fields the
// gathers read together.
struct DArr {
    gptr<const double> cx, cy, cz, R, r;
    gptr<const uint8_t> status;
    gptr<const uint32_t> orig;
    gptr<const uint32_t> nb_off, nb_items;
    gptr<const double> mx, my, mz, ax, ay, az, e1x, e1y, e1z, crho;
    gptr<const uint32_t> ci, cj;
    gptr<const uint32_t> cb_off, cb_items;
    gptr<const double> vx, vy, vz;
    gptr<const uint32_t> vatoms;
    gptr<const SphereRec> sph;
    gptr<const CircleRec> cir;
    gptr<const VertexRec> ver;
    gptr<const SphereThr> thr;
    DGrid gs, gc, gv;
    DGrid fs, fc, fv;  // the same items on cells 0.4 as wide, for gather3
    uint32_t n_spheres, n_circles, n_vertices, kind;
    double probe, max_expanded, rho_max;
};

// Frame-local f32 copies (reference field/gpu.rs:253-318, kernel.rs:604-747),
// one float4 per record: spheres (x, y, z, expanded radius), circles
// (centre, radius), (axis, 0), (e1, 0), SAS vertices (x, y, z, 0).
struct DLoc {
    gptr<float4> p4, m4, u4, e4, v4;
};

// Loads of the packed records through global pointers (as vector loads).
typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
__device__ inline SphereRec load_rec(gptr<const SphereRec> p, uint32_t k) {
    dvec4 v = ((gptr<const dvec4>)p)[k];
    return SphereRec{v.x, v.y, v.z, v.w};
}
__device__ inline CircleRec load_rec(gptr<const CircleRec> p, uint32_t k) {
    gptr<const dvec4> q = (gptr<const dvec4>)p + 2 * (size_t)k;
    dvec4 a = q[0], b = q[1];
    return CircleRec{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}
__device__ inline SphereThr load_rec(gptr<const SphereThr> p, uint32_t k) {
    typedef double dvec2 __attribute__((ext_vector_type(2)));
    dvec2 v = ((gptr<const dvec2>)p)[k];
    return SphereThr{v.x, v.y};
}
__device__ inline VertexRec load_rec(gptr<const VertexRec> p, uint32_t k) {
    dvec4 v = ((gptr<const dvec4>)p)[k];
    return VertexRec{v.x, v.y, v.z, v.w};
}
__device__ inline float4 load_f4(gptr<float4> p, uint32_t k) {
    fvec4 v = ((gptr<const fvec4>)p)[k];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline void store_f4(gptr<float4> p, uint32_t k, float x, float y, float z, float w) {
    fvec4 v = {x, y, z, w};
    ((gptr<fvec4>)p)[k] = v;
}

// Squared f32 radius and squared radius less the inside tolerance of a
// sphere record (kernel.rs:627-637).
__device__ inline float rec_r2(float r) { return r * r; }
__device__ inline float rec_r2tol(float r) {
    float rt = fmaxf(r - INSIDE_TOL, 0.0f);
    return rt * rt;
}

// Lattice and stage parameters of one mesh run.
struct DRun {
    double h, band, probe, t_min;
    int64_t frame[3];
    int64_t blo[3];       // dense brick box: lowest key per axis
    uint32_t bdim[3];     // dense brick box size per axis
    uint32_t b, sub, side, spb;  // brick edge, sub-block edge, b+1, (b+1)^3
    float h32, fband, fprobe, fcutoff2;
};

// One molecule of a run: its resident arrangement, its f32 copies and its
// lattice, plus where its per-molecule arrays start in the run's
// concatenated buffers. A single mesh is a run of one molecule.
struct Mol {
    DArr A;
    DLoc L;
    DRun P;
    uint64_t word_off;   // first word of its touched-brick bitmap
    uint64_t dense_off;  // first entry of its dense brick-index grid
};

// Work queue of a persistent launch: items 0 .. *count - 1 (entries of
// `list`, or the indices themselves without one), handed out one per
// workgroup from the ticket counter `next`, which is zero before the launch.
// Every workgroup loops until the queue is empty, so the grid is sized by
// residency, not by the count, and the count may be known only on the
// device (a stream-captured pipeline never reads it on the host).
struct Queue {
    const uint32_t* list;
    const uint32_t* count;
    uint32_t* next;
};

// Cancellation of the device's current run. Every run has a generation
// number, which its first kernel stores in g_run; the host cancels run g by
// writing g to a word of coherent pinned host memory (g_cancel_host, its
// device address), in either order. Work-item kernels stop taking items once
// g_cancel equals g_run. The host word is read over the bus, so only every
// 64th item taken polls it (every launch polls at its item 0) and copies a
// match into g_cancel, which every other take reads from device memory. A
// copy into g_cancel on the host's own stream raises it too; such a copy may
// wait for free compute units behind kernels that fill the device, the
// polled word does not.
__device__ uint32_t g_cancel = 0u;
__device__ uint32_t g_run = 0xffffffffu;
__device__ const uint32_t* g_cancel_host = nullptr;
__device__ __forceinline__ bool cancel_raised() {
    return __hip_atomic_load(&g_cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
           __hip_atomic_load(&g_run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool poll_host_cancel() {
    const uint32_t* h = g_cancel_host;
    if (!h) return false;
    const uint32_t g = __hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (g != __hip_atomic_load(&g_run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __hip_atomic_store(&g_cancel, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}
__global__ void k_run_begin(const uint32_t* gen) {
    __hip_atomic_store(&g_run, *gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Takes the workgroup's next item; false once the queue is empty. Every
// thread of the workgroup calls it (it holds two barriers).
__device__ __forceinline__ bool queue_take(const Queue& q, uint32_t* ticket, uint32_t& item) {
    if (threadIdx.x == 0) {
        uint32_t t = cancel_raised() ? 0xffffffffu : atomicAdd(q.next, 1u);
        if ((t & 63u) == 0u && poll_host_cancel()) t = 0xffffffffu;
        *ticket = t;
    }
    __syncthreads();
    // The ticket and item are the workgroup's: reading them as wave-uniform
    // keeps everything derived from them (molecule, lattice, arrangement
    // pointers) in scalar registers.
    uint32_t t = __builtin_amdgcn_readfirstlane(*ticket);
    __syncthreads();
    if (t >= *q.count) return false;
    item = __builtin_amdgcn_readfirstlane(q.list ? q.list[t] : t);
    return true;
}

// Runs body(item) for the workgroup's items. PERSIST: loops over the queue
// (a residency-sized grid). Otherwise one item per workgroup, blockIdx.x
// of the queue's items (a grid of at least the count; workgroups past it do
// nothing): cheaper for short items and for the first, largest pass. A
// dispatch holds fewer than 2^32 work-items: the host launches a pass with
// more workgroups than that persistent.
template <bool PERSIST, class F>
__device__ __forceinline__ void for_each_item(const Queue& q, uint32_t* ticket, F&& body) {
    if (PERSIST) {
        uint32_t j;
        while (queue_take(q, ticket, j)) {
            body(j);
            __syncthreads();
        }
    } else {
        uint32_t t = blockIdx.x;
        if (t >= *q.count) return;
        // one reading of the flag for the whole workgroup (the body holds barriers)
        __shared__ uint32_t stop;
        if (threadIdx.x == 0) stop = cancel_raised() || ((t & 63u) == 0u && poll_host_cancel());
        __syncthreads();
        if (!stop) body(q.list ? q.list[t] : t);
    }
}

// The molecule owning item i of a concatenation: the last m < nm with
// off[m] <= i (off ascending, off[0] = 0).
template <class T>
__device__ inline uint32_t find_mol(const T* off, uint32_t nm, T i) {
    uint32_t lo = 0, hi = nm - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

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

// A wavefront-uniform double moved to scalar registers: values the whole
// workgroup shares (a ball centre, a radius) then stay out of the vector
// registers, which the list passes need.
__device__ inline double uniform_f64(double v) {
    const unsigned long long x = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }

// f64::total_cmp as an order-preserving integer key.
__device__ inline uint64_t total_key(double d) {
    int64_t i = __double_as_longlong(d);
    i ^= (int64_t)(((uint64_t)(i >> 63)) >> 1);
    return (uint64_t)i ^ 0x8000000000000000ull;
}
// Inverse of total_key for keys of non-negative values.
__device__ inline double key_value(uint64_t k) { return __longlong_as_double((int64_t)(k ^ 0x8000000000000000ull)); }
// Inverse of total_key for keys of any sign.
__device__ inline double key_value_any(uint64_t k) {
    int64_t i = (int64_t)(k ^ 0x8000000000000000ull);
    i ^= (int64_t)(((uint64_t)(i >> 63)) >> 1);
    return __longlong_as_double(i);
}

// Smallest s >= 0 with sqrt(s) >= m. A correctly rounded sqrt is
// monotonic, so for every s >= 0: sqrt(s) >= m  <=>  s >= threshold. The
// distance tests `ddist(p, c) < R` and `!(ddist(p, c) >= R - EPS)` thus
// become `s < threshold` on the same sum of squares, without the sqrt.
// Host and device give the same value: both square roots are correctly
// rounded.
__host__ __device__ inline double sqrt_threshold(double m) {
    if (!(m > 0.0)) return 0.0;
    double s = m * m;
    if (sqrt(s) >= m) {
        for (;;) {
            double p = __builtin_bit_cast(double, __builtin_bit_cast(int64_t, s) - 1);
            if (p >= 0.0 && sqrt(p) >= m)
                s = p;
            else
                break;
        }
    } else {
        do {
            s = __builtin_bit_cast(double, __builtin_bit_cast(int64_t, s) + 1);
        } while (!(sqrt(s) >= m));
    }
    return s;
}

// Block-wide exclusive scan of one value per thread (blockDim.x a multiple
// of 64, at most 256); `tmp` holds 4 words of LDS.
__device__ inline uint32_t block_scan_256(uint32_t v, uint32_t* tmp, uint32_t& total) {
    uint32_t lane = lane_id(), wv = threadIdx.x / WAVE;
    uint32_t inc = v;
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, WAVE);
        if ((int)lane >= o) inc += y;
    }
    if (lane == WAVE - 1) tmp[wv] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t q = 0; q < blockDim.x / WAVE; q++) {
        if (q < wv) before += tmp[q];
        all += tmp[q];
    }
    __syncthreads();
    total = all;
    return before + inc - v;
}

// Open-addressing map global sphere id -> near-list slot in LDS.
__device__ inline void hash_put(uint32_t* hk, uint32_t* hv, uint32_t mask, uint32_t key, uint32_t val) {
    uint32_t h = (key * 2654435761u) & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
        uint32_t prev = atomicCAS(&hk[h], 0u, key + 1u);
        if (prev == 0u) {
            hv[h] = val;
            return;
        }
        h = (h + 1u) & mask;
    }
}
__device__ inline uint32_t hash_get(const uint32_t* hk, const uint32_t* hv, uint32_t mask, uint32_t key) {
    uint32_t h = (key * 2654435761u) & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
        uint32_t k = hk[h];
        if (k == key + 1u) return hv[h];
        if (k == 0u) return 0xffffffffu;
        h = (h + 1u) & mask;
    }
    return 0xffffffffu;
}

constexpr uint32_t KIND_SHIFT = 30, ID_MASK = (1u << KIND_SHIFT) - 1u;

// The candidates the sampling gather dropped as hidden, per sampled
// sub-block j: cnt[j] entries (candidate id | class << KIND_SHIFT, class 0
// for S, 1 for C) at ids + off[j]; cnt[j] is NO_LIST where there is no list
// (a sub-block not sampled, or the pool of `cap` entries full). `used`
// counts the entries taken. The refinement of the sub-block reuses them:
// the sampling ball plus COVER_PAD holds the refinement's ball, its near
// spheres and candidates are among the refinement's, and the cover limit
// is the same, so each of them is hidden there too.
// The refinement's candidate superset of each sampled sub-block j: ids (a
// sphere id, or a circle / SAS vertex id | class << KIND_SHIFT) of every
// sphere, circle and vertex the refinement's gather of that sub-block
// tests positively, found by the sampling's grid walk; the refinement
// tests just these instead of walking the grids again. cnt[j] NO_LIST: no
// list (the refinement walks). ids null: none are kept.
struct SupList {
    uint32_t *off, *cnt, *ids, *used;
    uint32_t cap;
};
// Fixed-size slots of the blocker records: sampled sub-block j (dense
// index slot[j] among the sampled ones, below `slots`) has `words` words at
// ids + slot[j] * words; cnt[j] is its record's length, NO_LIST where it
// has none (not sampled, or the record did not fit). ids null: none are
// kept.
struct BlkPool {
    const uint32_t* slot;
    uint32_t *cnt, *ids;
    uint32_t words, slots;
};
struct HidList {
    uint32_t *off, *cnt, *ids, *used;
    uint32_t cap;
    SupList sup;   // handed over the same way
    SupList bsup;  // per surface brick: the classification's superset of its sampling walks (see gather_refine)
    BlkPool blk;   // per sampled sub-block: its blocker record (see BlkRecord)
};

// The blocker record a sampling workgroup leaves for the refinement of the
// same sub-block, in the pool of HidList::blk (u32 words): n_near, n_s,
// n_c, nw (64-bit words per bitset), n_v and `same`; then the n_near
// near-sphere ids by near-list slot; then per kept candidate, classes S, C
// and V, in the refinement's list order (class, f64 total_cmp key at the
// ball centre, index): its id (class in the top bits, as the refinement's
// lists hold it), its hint slot, its key (low half first) and its blocker
// bitset over near-list slots (nw words, low half first; none for V). The
// blockers are those the sampling's reach test keeps, a superset of those
// any refinement point can meet (the test's ball and limits are larger).
// `same` is 1 when no item lies between the sampling's reach and the
// refinement's slightly larger one: the near spheres and candidates are
// then the refinement's own, so the refinement may take its lists from the
// record; otherwise it gathers and takes only the bitsets, once its near
// list and candidates are found to be the record's.
constexpr uint32_t BLK_HEAD = 6;
__host__ __device__ inline uint32_t blk_entry_words(uint32_t nw) { return 4u + 2u * nw; }
constexpr uint32_t NO_LIST = 0xffffffffu;
constexpr uint32_t KEY_BUCKETS = 16;  // key buckets per candidate class in the gathers

// Blocker bitsets of candidate rows p < rows: row p lists sphere indices
// (nb_items for p < ns, else cb_items, from off[p], len[p] of them); bit
// `slot` of row p is set for each listed sphere the near hash holds at
// `slot`. Row p's bits are words bits[w * cc + p]. 8 lanes per row, each
// loading 4 of its items before it probes any, so 4 loads are in flight per
// lane. keep(r, p, slot) may leave out a blocker that can never block row
// p's candidate where it is evaluated; r = prep(p) is computed once per row
// by each of its lanes, and finish(r, p) runs once the row's items are done,
// by all 8 of its lanes together (len[p] is no longer read then). By the
// whole workgroup; the caller synchronises afterwards.
template <class Prep, class Keep, class Finish>
__device__ inline void blocker_bits(gptr<const uint32_t> nb_items, gptr<const uint32_t> cb_items, uint32_t ns,
                                    uint32_t rows, const uint32_t* off, const uint32_t* len, const uint32_t* hk,
                                    const uint32_t* hv, uint32_t hmask, unsigned long long* bits, uint32_t cc,
                                    Prep&& prep, Keep&& keep, Finish&& finish) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t p = tid / 8; p < rows; p += nt / 8) {
        gptr<const uint32_t> row = (p < ns ? nb_items : cb_items) + off[p];
        const uint32_t n = len[p];
        auto r = prep(p);
        for (uint32_t q0 = tid % 8; q0 < n; q0 += 32) {
            uint32_t it[4];
            for (int u = 0; u < 4; u++) it[u] = q0 + 8 * u < n ? row[q0 + 8 * u] : 0xffffffffu;
            for (int u = 0; u < 4; u++) {
                if (it[u] == 0xffffffffu) continue;
                const uint32_t slot = hash_get(hk, hv, hmask, it[u]);
                if (slot != 0xffffffffu && keep(r, p, slot)) atomicOr(&bits[(slot >> 6) * cc + p], 1ull << (slot & 63));
            }
        }
        finish(r, p);
    }
}

// Candidates hidden everywhere. A class S or C candidate is tested, at a
// point x of a sub-block, by projecting x onto it and checking whether the
// projection lies inside one of its blockers; if it does the candidate
// cannot give the value. When one blocker contains, with a margin, every
// projection of every point of the sub-block's ball, the candidate can
// never give the value there, and dropping it from the lists changes no
// result: the value is a minimum over the unblocked candidates (sampling),
// and the refinement's least unblocked candidate, ties broken by list order,
// is the same once blocked candidates leave the sorted list.
//
// The projections of the ball (b, rb) onto sphere (c, R) fill the cap of
// directions within alpha of b - c, sin(alpha) = rb / |b - c| (when the ball
// holds c, every direction). The farthest cap point from blocker centre ck
// is at angle min(theta + alpha, pi) from ck - c (theta the angle between
// b - c and ck - c), at squared distance R^2 + d^2 - 2 R d cos(that angle).
// For a circle (centre m, unit axis u, radius r) the projections fill the
// arc within beta of the in-plane direction of b - m, sin(beta) = rb / |w|
// (the whole circle when |w| <= rb, which also holds the degenerate axis
// case); with L the in-plane distance of ck from the axis the farthest arc
// point is at r^2 + |ck - m|^2 - 2 r L cos(min(delta + beta, pi)).
//
// COVER_PAD widens the ball for the f32 rounding of the sample points, and
// COVER_MARGIN (squared distance below (Rk - COVER_MARGIN)^2) exceeds the
// rounding of the f32 sampling arithmetic at coordinates up to 2000 (about
// 1e-3) by a factor of four, and the f64 refinement's and the blocker
// tolerances (2e-5, 1e-9) by far more, so every such projection tests
// blocked in either precision.
constexpr double COVER_PAD = 1e-3, COVER_MARGIN = 4e-3;

// Both tests are multiplied through so that they need no division and at
// most two square roots. Cap: with v = b - c, e = ck - c, D2 = |v|^2,
// h2 = D2 - rb^2, d = |e|, cos(theta + alpha) d = (v.e sqrt(h2) - |v x e| rb)
// / D2; theta + alpha >= pi exactly when v.e <= -d sqrt(h2), and then the
// farthest point is at R + d. Arc: the same with w (b - m in the plane), p
// (ck - m in the plane), L = |p|, W2 = |w|^2, h2 = W2 - rb^2; the full
// circle (h2 <= 0) or an arc reaching the far side gives r^2 + |ck - m|^2 +
// 2 r L.
__device__ inline bool cap_covered(const double b[3], double rb, const dvec4& c, const dvec4& k) {
    const double lim = k.w - COVER_MARGIN;
    if (!(lim > 0.0)) return false;
    const double vx = b[0] - c.x, vy = b[1] - c.y, vz = b[2] - c.z;
    const double h2 = (vx * vx + vy * vy + vz * vz) - rb * rb;
    if (!(h2 > 0.0)) return false;  // the ball reaches the centre: every direction
    const double ex = k.x - c.x, ey = k.y - c.y, ez = k.z - c.z;
    const double d2 = ex * ex + ey * ey + ez * ez, dot = vx * ex + vy * ey + vz * ez;
    const double R = c.w;
    if (dot <= 0.0 && dot * dot >= d2 * h2) {
        const double m = lim - R;  // R + d < lim
        return m > 0.0 && d2 < m * m;
    }
    const double cx = vy * ez - vz * ey, cy = vz * ex - vx * ez, cz = vx * ey - vy * ex;
    const double D2 = h2 + rb * rb;
    return (R * R + d2 - lim * lim) * D2 < 2.0 * R * (dot * sqrt(h2) - sqrt(cx * cx + cy * cy + cz * cz) * rb);
}

// Circle record: a = (mx, my, mz, ax), e = (ay, az, rho, -).
__device__ inline bool arc_covered(const double b[3], double rb, const dvec4& a, const dvec4& e, const dvec4& k) {
    const double lim = k.w - COVER_MARGIN;
    if (!(lim > 0.0)) return false;
    const double ux = a.w, uy = e.x, uz = e.y, r = e.z;
    const double yx = b[0] - a.x, yy = b[1] - a.y, yz = b[2] - a.z;
    const double al = yx * ux + yy * uy + yz * uz;
    const double wx = yx - ux * al, wy = yy - uy * al, wz = yz - uz * al;
    const double W2 = wx * wx + wy * wy + wz * wz, h2 = W2 - rb * rb;
    const double fx = k.x - a.x, fy = k.y - a.y, fz = k.z - a.z;
    const double fa = fx * ux + fy * uy + fz * uz;
    const double px = fx - ux * fa, py = fy - uy * fa, pz = fz - uz * fa;
    const double L2 = px * px + py * py + pz * pz, wp = wx * px + wy * py + wz * pz;
    const double A0 = r * r + (fx * fx + fy * fy + fz * fz) - lim * lim;
    if (!(h2 > 0.0) || (wp <= 0.0 && wp * wp >= L2 * h2)) return A0 < 0.0 && A0 * A0 > 4.0 * r * r * L2;
    const double cx = wy * pz - wz * py, cy = wz * px - wx * pz, cz = wx * py - wy * px;
    return A0 * W2 < 2.0 * r * (wp * sqrt(h2) - sqrt(cx * cx + cy * cy + cz * cz) * rb);
}

// The converse tests: false only if no projection of the ball (b, rb) onto
// the candidate comes within t of the blocker centre k (t its blocking
// radius plus COVER_MARGIN), so the blocker can never block there and
// leaving it out of the candidate's bitset changes no test. The nearest cap
// point to k is at |d - R| when ck - c lies within the cap's directions
// (theta <= alpha), else at angle theta - alpha: R^2 + d^2 - 2 R d cos(theta
// - alpha), with cos(theta - alpha) d = (v.e sqrt(h2) + |v x e| rb) / D2.
// For the arc likewise with the in-plane angle, the whole circle (h2 <= 0)
// reaching to r^2 + |ck - m|^2 - 2 r L.
//
// Each test is split in two: what depends only on the ball and the
// candidate (its row of the bitsets) is computed once per row, with the
// operations and order of the whole test, so every decision is unchanged.
// T is double, or float for callers whose limits carry REACH_SLACK on both
// the ball radius and the reach: f32 inputs within 1.2e-4 of the values
// they stand for (coordinates below 2048, or lengths of a few A relative to
// the ball centre) and the test's own f32 arithmetic on lengths of a few A
// (about 1e-5) stay far inside it, so a blocker the f32 test leaves out
// cannot come within its blocking radius of any projection either.
constexpr float REACH_SLACK = 1e-3f;
template <class T>
struct ReachRow {
    T ox, oy, oz;  // sphere centre c, or circle centre m
    T ux, uy, uz;  // circle axis
    T vx, vy, vz;  // b - c, or w (b - m in the circle's plane)
    T R;           // sphere radius, or circle radius r
    T D2, h2;      // |v|^2 (W2 for the arc) and that less rb^2
    T sh2;         // sqrt(h2) where h2 > 0
    T rb;
};

// Cap row: sphere (cx, cy, cz) of radius R.
template <class T>
__device__ inline ReachRow<T> cap_row(const T b[3], T rb, T cx, T cy, T cz, T R) {
    ReachRow<T> w;
    w.ox = cx;
    w.oy = cy;
    w.oz = cz;
    w.ux = w.uy = w.uz = T(0);
    w.vx = b[0] - cx;
    w.vy = b[1] - cy;
    w.vz = b[2] - cz;
    w.R = R;
    w.D2 = w.vx * w.vx + w.vy * w.vy + w.vz * w.vz;
    w.h2 = w.D2 - rb * rb;
    w.sh2 = w.h2 > T(0) ? sqrt(w.h2) : T(0);
    w.rb = rb;
    return w;
}

template <class T>
__device__ inline bool cap_reaches(const ReachRow<T>& w, T kx, T ky, T kz, T t) {
    const T ex = kx - w.ox, ey = ky - w.oy, ez = kz - w.oz;
    const T d2 = ex * ex + ey * ey + ez * ez, dot = w.vx * ex + w.vy * ey + w.vz * ez;
    const T R = w.R;
    if (!(w.h2 > T(0)) || (dot >= T(0) && dot * dot >= d2 * w.h2)) {  // nearest point |d - R| away
        const T lo = R - t, hi = R + t;
        return d2 < hi * hi && !(lo > T(0) && d2 <= lo * lo);
    }
    const T cx = w.vy * ez - w.vz * ey, cy = w.vz * ex - w.vx * ez, cz = w.vx * ey - w.vy * ex;
    return (R * R + d2 - t * t) * w.D2 < T(2) * R * (dot * w.sh2 + sqrt(cx * cx + cy * cy + cz * cz) * w.rb);
}

// Arc row: circle centre (mx, my, mz), unit axis (ax, ay, az), radius r.
template <class T>
__device__ inline ReachRow<T> arc_row(const T b[3], T rb, T mx, T my, T mz, T ax, T ay, T az, T r) {
    ReachRow<T> w;
    w.ox = mx;
    w.oy = my;
    w.oz = mz;
    w.ux = ax;
    w.uy = ay;
    w.uz = az;
    w.R = r;
    const T yx = b[0] - mx, yy = b[1] - my, yz = b[2] - mz;
    const T al = yx * w.ux + yy * w.uy + yz * w.uz;
    w.vx = yx - w.ux * al;
    w.vy = yy - w.uy * al;
    w.vz = yz - w.uz * al;
    w.D2 = w.vx * w.vx + w.vy * w.vy + w.vz * w.vz;
    w.h2 = w.D2 - rb * rb;
    w.sh2 = w.h2 > T(0) ? sqrt(w.h2) : T(0);
    w.rb = rb;
    return w;
}

template <class T>
__device__ inline bool arc_reaches(const ReachRow<T>& w, T kx, T ky, T kz, T t) {
    const T r = w.R;
    const T fx = kx - w.ox, fy = ky - w.oy, fz = kz - w.oz;
    const T fa = fx * w.ux + fy * w.uy + fz * w.uz;
    const T px = fx - w.ux * fa, py = fy - w.uy * fa, pz = fz - w.uz * fa;
    const T L2 = px * px + py * py + pz * pz, wp = w.vx * px + w.vy * py + w.vz * pz;
    const T A0 = r * r + (fx * fx + fy * fy + fz * fz) - t * t;
    if (!(w.h2 > T(0)) || (wp >= T(0) && wp * wp >= L2 * w.h2))  // the arc holds the nearest circle point
        return A0 < T(0) || A0 * A0 < T(4) * r * r * L2;
    const T cx = w.vy * pz - w.vz * py, cy = w.vz * px - w.vx * pz, cz = w.vx * py - w.vy * px;
    return A0 * w.D2 < T(2) * r * (wp * w.sh2 + sqrt(cx * cx + cy * cy + cz * cz) * w.rb);
}

// Marks the candidates q < total hidden everywhere in the ball (b, rb). A
// blocker that covers a candidate's cap or arc holds points of its sphere
// (or circle) well inside, so it overlaps that sphere (both circle spheres)
// by more than the rows' EPS: neighbour rows hold every such sphere
// (geom/arrangement.rs:519-550) and circle-blocker rows every common
// neighbour cutting the circle (:672-751). So the near list itself is
// searched, without the rows. kind(q) is 0 for class S (sphere id(q)), 1
// for class C (circle id(q)), anything else for candidates without
// blockers; geom(kind, id, g0, g1) gives a candidate's record in the frame
// of b (S: centre, radius; C: centre, axis x | axis y, axis z, radius);
// blocker(slot) the near sphere at a near-list slot as (centre, t) with t
// the square of its blocking radius; mark(q) is called once per hidden
// candidate, by one lane. A covering blocker holds the projection of b
// itself, the cap's (arc's) central point, strictly within its blocking
// radius: that test, without a square root, comes first, as pre(pp, slot)
// with pp = prep(x, y, z) of the projection. It may pass more blockers than
// the exact test (their cover tests then fail), never fewer. 4 lanes per
// candidate. By the whole workgroup; the caller synchronises afterwards.
constexpr uint32_t HIDE_STEPS = 8;  // prefilter steps per round of mark_hidden

template <class Kind, class Id, class Geom, class Blocker, class Prep, class Pre, class Mark>
__device__ inline void mark_hidden(const double b[3], double rb, uint32_t total, uint32_t n_near, Kind&& kind,
                                   Id&& id, Geom&& geom, Blocker&& blocker, Prep&& prep, Pre&& pre, Mark&& mark) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = lane_id();
    for (uint32_t q = tid / 4; q < total; q += nt / 4) {
        const uint32_t kd = kind(q);
        bool covered = false, live = false;
        dvec4 g0 = dvec4{0.0, 0.0, 0.0, 0.0}, g1 = g0;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (kd <= 1u) {
            geom(kd, id(q), g0, g1);
            // The projection of b: c + R (b - c) / |b - c|, or m + r w / |w|.
            if (kd == 0u) {
                px = b[0] - g0.x;
                py = b[1] - g0.y;
                pz = b[2] - g0.z;
            } else {
                const double yx = b[0] - g0.x, yy = b[1] - g0.y, yz = b[2] - g0.z;
                const double al = yx * g0.w + yy * g1.x + yz * g1.y;
                px = yx - g0.w * al;
                py = yy - g1.x * al;
                pz = yz - g1.y * al;
            }
            const double len2 = px * px + py * py + pz * pz;
            // The cap needs the ball off the centre; a circle point needs w != 0.
            if (kd == 0u ? len2 > rb * rb : len2 > 0.0) {
                const double sc = (kd == 0u ? g0.w : g1.z) / sqrt(len2);
                px = g0.x + px * sc;
                py = g0.y + py * sc;
                pz = g0.z + pz * sc;
                live = true;
            }
        }
        // Rounds: each lane moves on, by at most HIDE_STEPS blockers, to its
        // next blocker holding the projected centre, then every lane that
        // found one runs its cover test in the same round, so a wavefront
        // does not pay a masked f64 test for every step of its slowest lane.
        // The step bound keeps a round short when the near list is long and
        // blockers pass often (coarse lattices). A candidate's four lanes
        // stop once one of them has found a cover.
        uint32_t k = tid % 4;
        const auto pp = prep(px, py, pz);
        while (__ballot(live && k < n_near) != 0ull) {
            bool pass = false;
            if (live) {
                for (uint32_t st = 0; st < HIDE_STEPS && k < n_near; st++) {
                    if (pre(pp, k)) {
                        pass = true;
                        break;
                    }
                    k += 4;
                }
            }
            if (pass) {
                const dvec4 kt = blocker(k);
                const dvec4 ks = dvec4{kt.x, kt.y, kt.z, sqrt(kt.w) - COVER_MARGIN};
                covered = kd == 0u ? cap_covered(b, rb, g0, ks) : arc_covered(b, rb, g0, g1, ks);
                k += 4;
            }
            if (((__ballot(covered) >> (lane & ~3u)) & 0xfull) != 0ull) live = false;
        }
        const unsigned long long bal = __ballot(covered);
        if ((lane & 3u) == 0u && ((bal >> lane) & 0xfull) != 0ull) mark(q);
    }
}

// Candidate order: (class, key, index), the class in the top bits of id.
__device__ inline bool cand_less(unsigned long long ka, uint32_t ia, unsigned long long kb, uint32_t ib) {
    uint32_t ca = ia >> KIND_SHIFT, cb = ib >> KIND_SHIFT;
    return ca < cb || (ca == cb && (ka < kb || (ka == kb && ia < ib)));
}

constexpr uint32_t ROWS_MAX = 256;

// The items of the cells of the sphere, circle and vertex grids that
// overlap cubes of half-sizes r[0], r[1], r[2] about (x, y, z) (three
// CellGrid::for_each_near), visited by the whole workgroup as one flat
// range: the rows' item ranges are read in parallel, one row per thread,
// then each thread takes every blockDim.x-th item, so a gather costs three
// dependent memory round trips instead of three per row. f(grid, item);
// the visiting order is unspecified. Every thread of the block must call it.
template <class F>
__device__ inline void gather3(const DArr& A, double x, double y, double z, const double r[3], uint32_t* tmp4,
                               F&& f) {
    __shared__ uint32_t rs[ROWS_MAX], rp[ROWS_MAX];
    // The fine grids visit fewer items outside the ball; the coarse ones
    // take balls whose fine rows would not fit.
    Box b0 = grid_box(A.fs, x, y, z, r[0]);
    Box b1 = grid_box(A.fc, x, y, z, r[1]);
    Box b2 = grid_box(A.fv, x, y, z, r[2]);
    uint32_t n0 = (b0.z1 - b0.z0 + 1) * (b0.y1 - b0.y0 + 1);
    uint32_t n1 = (b1.z1 - b1.z0 + 1) * (b1.y1 - b1.y0 + 1);
    uint32_t n2 = (b2.z1 - b2.z0 + 1) * (b2.y1 - b2.y0 + 1);
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const DGrid *g0 = &A.fs, *g1 = &A.fc, *g2 = &A.fv;
    if (n0 + n1 + n2 > ROWS_MAX || n0 + n1 + n2 > nt) {
        g0 = &A.gs;
        g1 = &A.gc;
        g2 = &A.gv;
        b0 = grid_box(A.gs, x, y, z, r[0]);
        b1 = grid_box(A.gc, x, y, z, r[1]);
        b2 = grid_box(A.gv, x, y, z, r[2]);
        n0 = (b0.z1 - b0.z0 + 1) * (b0.y1 - b0.y0 + 1);
        n1 = (b1.z1 - b1.z0 + 1) * (b1.y1 - b1.y0 + 1);
        n2 = (b2.z1 - b2.z0 + 1) * (b2.y1 - b2.y0 + 1);
    }
    const uint32_t nrows = n0 + n1 + n2;
    if (nrows > ROWS_MAX || nrows > nt) {
        for_near(A.gs, x, y, z, r[0], tid, nt, [&](uint32_t k) { f(0u, k); });
        for_near(A.gc, x, y, z, r[1], tid, nt, [&](uint32_t k) { f(1u, k); });
        for_near(A.gv, x, y, z, r[2], tid, nt, [&](uint32_t k) { f(2u, k); });
        __syncthreads();  // the caller reads what f gathered
        return;
    }
    uint32_t len = 0, start = 0;
    if (tid < nrows) {
        uint32_t k = tid < n0 ? 0u : (tid < n0 + n1 ? 1u : 2u);
        uint32_t q = tid - (k == 0 ? 0u : (k == 1 ? n0 : n0 + n1));
        const DGrid& G = k == 0 ? *g0 : (k == 1 ? *g1 : *g2);
        const Box& B = k == 0 ? b0 : (k == 1 ? b1 : b2);
        uint32_t ny = B.y1 - B.y0 + 1;
        uint32_t row = ((B.z0 + q / ny) * G.dy + (B.y0 + q % ny)) * G.dx;
        start = G.starts[row + B.x0];
        len = G.starts[row + B.x1 + 1] - start;
    }
    uint32_t total;
    uint32_t pre = block_scan_256(len, tmp4, total);
    if (tid < nrows) {
        rs[tid] = start;
        rp[tid] = pre;
    }
    __syncthreads();
    for (uint32_t i = tid; i < total; i += nt) {
        uint32_t lo = 0, hi = nrows - 1;  // last row whose range starts at or before i
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (rp[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        uint32_t k = lo < n0 ? 0u : (lo < n0 + n1 ? 1u : 2u);
        gptr<const uint32_t> items = k == 0 ? g0->items : (k == 1 ? g1->items : g2->items);
        f(k, items[rs[lo] + (i - rp[lo])]);
    }
    __syncthreads();
}

}  // namespace smg
