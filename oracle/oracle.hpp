// CPU oracle of the sesmesh mesh stage (test infrastructure only).
//
// A scalar C++ restatement of the reference's CPU back end, used by the tests
// and by the CPU baseline of bench.py as the checker of the HIP product path.
// Nothing in the product (sesmesh_amd/) links or calls this code.
//
// Every function cites the reference file and lines it restates (paths are
// relative to the reference repository root, crates/sesmesh/src/ unless
// noted). Arithmetic follows the reference's operation order with no FMA
// contraction (built with -ffp-contract=off) so results are bit-identical to
// the reference algorithm executed in IEEE f64/f32.
//
// Scope: the SES kind only (the workloads this project measures are SES).
//
// How each part is pinned (tests/test_oracle_golden.py, test_oracle_math.py):
// * the whole path (arrangement, classification, sampling, refinement, MC
//   table, assembly, area/volume) reproduces the reference's recorded
//   validation runs (docs/validation/compare/runs.csv, campaign "acc"):
//   vertex/triangle/component/cavity counts and every printed digit of area
//   and volume for 1CRN, 1EJG, 1UBQ, 3FX2, 2PTN at 0.5 and 0.25 A, 1AON at
//   0.5 A;
// * the MC table hashes equal to the reference's table data;
// * ses_eval and the arrangement agree with a brute-force evaluator over all
//   spheres, pairs and triples within 1e-9;
// * lattice samples of a single sphere equal the closed form;
// * refined vertices lie within 1e-4 h + 2e-6 of the surface, meshes are
//   closed and oriented, and results do not depend on threads or brick size.
// No part is left unpinned.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

namespace orc {

using u32 = uint32_t;
using u64 = uint64_t;

// ---------------------------------------------------------------- vectors
// geom/vec.rs:9-124
struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    constexpr V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    V3 operator/(double s) const { return {x / s, y / s, z / s}; }
    V3 operator-() const { return {-x, -y, -z}; }
    bool operator==(const V3& o) const { return x == o.x && y == o.y && z == o.z; }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const {
        return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
    }
    double norm2() const { return dot(*this); }
    double norm() const { return std::sqrt(norm2()); }
    double dist(const V3& o) const { return (*this - o).norm(); }
    V3 normalized() const { return *this / norm(); }
    // f64::min / f64::max ignore a NaN operand, like fmin / fmax.
    V3 vmin(const V3& o) const { return {std::fmin(x, o.x), std::fmin(y, o.y), std::fmin(z, o.z)}; }
    V3 vmax(const V3& o) const { return {std::fmax(x, o.x), std::fmax(y, o.y), std::fmax(z, o.z)}; }
    bool finite() const { return std::isfinite(x) && std::isfinite(y) && std::isfinite(z); }
    double operator[](int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};

// geom/vec.rs:112-124
void orthonormal_basis(const V3& u, V3& e1, V3& e2);

// Rust `f64::total_cmp` as a strict "less than" (sort keys).
inline bool total_lt(double a, double b) {
    int64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    ia ^= (int64_t)(((uint64_t)(ia >> 63)) >> 1);
    ib ^= (int64_t)(((uint64_t)(ib >> 63)) >> 1);
    return ia < ib;
}

// ---------------------------------------------------------------- cells
// geom/cells.rs:9-130
struct CellGrid {
    V3 origin;
    double cell = 1;
    size_t dims[3] = {1, 1, 1};
    std::vector<u32> starts, items;

    void build(const std::vector<V3>& points, double cell);
    void coord(const V3& p, size_t c[3]) const;
    size_t flat(const size_t c[3]) const { return (c[2] * dims[1] + c[1]) * dims[0] + c[0]; }
    template <class F>
    void for_each_near(const V3& center, double radius, F&& f) const {
        V3 r(radius, radius, radius);
        size_t lo[3], hi[3];
        coord(center - r, lo);
        coord(center + r, hi);
        for (size_t z = lo[2]; z <= hi[2]; z++)
            for (size_t y = lo[1]; y <= hi[1]; y++) {
                size_t c0[3] = {0, y, z};
                size_t row = flat(c0);
                size_t a = starts[row + lo[0]], b = starts[row + hi[0] + 1];
                for (size_t i = a; i < b; i++) f(items[i]);
            }
    }
};

// geom/cells.rs:132-146
u64 morton(const u32 c[3]);

// ---------------------------------------------------------------- arrangement
// geom/arrangement.rs:43-135 (rows only; the builder pushes rows in order)
struct Csr {
    std::vector<u32> off{0};
    std::vector<u32> items;
    size_t rows() const { return off.size() - 1; }
    const u32* row(size_t i, size_t& n) const {
        n = off[i + 1] - off[i];
        return items.data() + off[i];
    }
    void close_row() { off.push_back((u32)items.size()); }
};

enum Status : uint8_t { EXPOSED = 0, BURIED = 1, CONTAINED = 2 };

// geom/arrangement.rs:137-152
struct Circle {
    u32 i, j;
    V3 center, axis, e1;
    double radius;
};

// geom/arrangement.rs:184-190
struct SasVertex {
    V3 position;
    u32 atoms[3];
};

struct Sphere {
    double c[3];
    double r;
};

// geom/arrangement.rs:214-235 (without arcs: the field never reads them)
struct Arrangement {
    double probe = 0;
    std::vector<V3> centers;
    std::vector<double> radii, expanded;
    std::vector<u32> orig;
    std::vector<uint8_t> status;
    Csr neighbors;
    std::vector<Circle> circles;
    Csr circle_blockers;
    std::vector<SasVertex> vertices;
    CellGrid cells;
    double max_expanded = 0;
    size_t n_contained = 0, n_exposed = 0;
};

// geom/arrangement.rs:437-645 with with_arcs = false (field/point.rs:
// 229-237): with_circles for the SES field; without circles every sphere
// that is not contained is exposed. Returns an error message or "".
std::string build_arrangement(const std::vector<Sphere>& spheres, double probe, Arrangement& a,
                              bool with_circles = true);

// field/mod.rs:33-41
enum SurfaceKind : int { K_SES = 0, K_SAS = 1, K_VDW = 2 };

// ---------------------------------------------------------------- field
enum FeatureKind : u32 { F_SOLVENT = 0, F_BEYOND = 1, F_SPHERE = 2, F_CIRCLE = 3, F_VERTEX = 4 };
struct Feature {
    u32 kind = F_BEYOND;
    u32 id = 0;
    bool operator==(const Feature& o) const { return kind == o.kind && id == o.id; }
};

// field/point.rs:67-73
struct PointEval {
    double distance;
    V3 closest;
    Feature feature;
};

// field/point.rs:75-95
struct LocalCandidates {
    std::vector<u32> inside, near, spheres, s_off, s_blk, circles, c_off, c_blk, vertices;
    std::vector<std::pair<double, u32>> order;
    void clear();
};

// field/point.rs:164-171
struct NearEval {
    double value;
    Feature feature;
    V3 closest;
};

// field/point.rs:189-274
struct SesField {
    Arrangement arr;
    int kind = K_SES;
    double probe = 0;
    CellGrid circle_cells, vertex_cells;
    double rho_max = 0;
    double extent = 0;

    // SES with probe 0 is the vdW kind (point.rs:224-228).
    std::string build(const std::vector<Sphere>& spheres, double probe, int kind = K_SES);
    bool inside_sas(const V3& x) const;
    PointEval ses_eval(const V3& x, double cutoff) const;
    void try_sphere_with(const V3& x, u32 i, const u32* blk, size_t nb, PointEval& best) const;
    void try_circle_with(const V3& x, u32 ci, const u32* blk, size_t nb, PointEval& best) const;
    void try_vertex(const V3& x, u32 vi, PointEval& best) const;
    void gather_local(const V3& b, double rho, double band, LocalCandidates& out) const;
    NearEval eval_local(const V3& x, double band, const LocalCandidates& l) const;
    // (value, index or -1)
    std::pair<double, int64_t> sphere_eval(const V3& x, double cutoff) const;
    // SesField::sample: value, closest, feature (field/point.rs:292-369)
    double sample(const V3& x, V3& closest, Feature& f) const;
    // SesField::clamped_value (field/point.rs:760-765)
    double clamped_value(const V3& x, double band) const;
};

// Algorithmic operation counts (one per IEEE add, sub, mul, div, sqrt, min
// or compare; abs not counted) of the f32 lane kernel and of the f64
// refinement's eval_local, for roofline accounting. Counting is per thread
// and off unless `on` is set.
struct OpCounts {
    bool on = false;
    uint64_t f32_ops = 0, samples = 0, f64_ops = 0, evals = 0;
};
extern thread_local OpCounts g_ops;
extern thread_local bool g_in_eval_local;

// ---------------------------------------------------------------- lattice
constexpr int64_t KEY_BIAS = int64_t(1) << 20;  // field/lattice.rs:15
constexpr double CLASSIFY_EPS = 2e-3;           // field/lattice.rs:20

// field/lattice.rs:25-91
struct Lattice {
    double h = 0;
    int64_t b = 0;
    int64_t frame[3] = {0, 0, 0};
    float h32 = 0;
    void init(double h, size_t b, const V3& frame_center);
    V3 point(const int64_t idx[3]) const {
        return V3((double)idx[0] * h, (double)idx[1] * h, (double)idx[2] * h);
    }
    V3 frame_origin() const { return point(frame); }
    void local(const int64_t idx[3], float out[3]) const {
        for (int a = 0; a < 3; a++) out[a] = (float)(idx[a] - frame[a]) * h32;
    }
    V3 brick_center(const int32_t key[3]) const {
        auto c = [&](int32_t k) { return ((double)k + 0.5) * (double)b * h; };
        return V3(c(key[0]), c(key[1]), c(key[2]));
    }
    double brick_radius() const { return 0.5 * std::sqrt(3.0) * (double)b * h; }
    void brick_origin(const int32_t key[3], int64_t o[3]) const {
        for (int a = 0; a < 3; a++) o[a] = (int64_t)key[a] * b;
    }
    int64_t brick_of(double v) const { return (int64_t)std::floor(v / ((double)b * h)); }
};

u64 pack(const int32_t key[3]);
void unpack(u64 p, int32_t key[3]);

// ---------------------------------------------------------------- engine
constexpr float INSIDE_TOL = 2e-5f;     // field/kernel.rs:42
constexpr double F32_MAX_COORD = 2000;  // field/kernel.rs:46

// field/kernel.rs:49-84
struct Gathered {
    size_t n_inside = 0;
    std::vector<float> nx, ny, nz, nr, nr2, nr2tol;
    std::vector<float> sx, sy, sz, sr;
    std::vector<u32> s_off, s_blk;
    std::vector<float> cx, cy, cz, ux, uy, uz, ex, ey, ez, rho;
    std::vector<u32> c_off, c_blk;
    std::vector<float> vx, vy, vz;
    void clear();
};

// Per sub-block state: 0 = surface may pass, 1 = uniform outside,
// 2 = uniform inside (field/kernel.rs:125-127 `Option<bool>`).
enum SubState : uint8_t { SS_NONE = 0, SS_OUT = 1, SS_IN = 2 };

struct BrickScratch {
    Gathered g;
    std::vector<float> xs, ys, zs, out, values;
    std::vector<u32> groups[8];
    std::vector<u32> local_of, touched;
    std::vector<std::pair<double, u32>> order;
};

// field/kernel.rs:188-261
struct Engine {
    SesField* field = nullptr;
    Lattice lattice;
    double band = 0;
    bool f32_ok = true;
    double coord_max = 0;

    std::string init(SesField* f, double spacing, size_t brick);
    size_t samples_per_brick() const {
        size_t s = (size_t)lattice.b + 1;
        return s * s * s;
    }
    void sub_layout(int64_t& sub, int64_t& nsub) const;
    size_t sub_blocks() const {
        int64_t sub, nsub;
        sub_layout(sub, nsub);
        return (size_t)(nsub * nsub * nsub);
    }
    BrickScratch scratch() const;
    // (touched count, surface bricks sorted by packed key)
    std::string surface_bricks(size_t& touched, std::vector<std::array<int32_t, 3>>& bricks) const;
    void sub_states(const int32_t key[3], uint8_t st[8]) const;
    void assign(const int32_t key[3], const uint8_t st[8], BrickScratch& s) const;
    bool prepare_sub(const int32_t key[3], size_t sb, BrickScratch& s, size_t& padded) const;
    void gather_near(const V3& b, double rho, double reach, BrickScratch& s) const;
    void gather_ses(const V3& b, double rho, BrickScratch& s) const;
    void gather_spheres(const V3& b, double rho, double band, BrickScratch& s) const;
    void evaluate(const int32_t key[3], BrickScratch& s) const;
};

// field/lattice.rs:105-163
std::string touched_bricks(const SesField& f, const Lattice& l, std::vector<u64>& keys);
// field/lattice.rs:190-227: SS_NONE / SS_OUT / SS_IN
uint8_t uniform_sign(const SesField& f, const V3& center, double rho);
// field/kernel.rs:837-952, one lane (scalar)
float ses_lane(const Gathered& g, float x, float y, float z, float probe, float cutoff2);
// field/kernel.rs:953-981, one lane (SAS and vdW)
float sphere_lane(const Gathered& g, float x, float y, float z, float band);

// ---------------------------------------------------------------- mesh
// mesh/extract.rs:28-51
constexpr int MAX_ITER = 8;
constexpr int MAX_BISECT = 64;
constexpr double TOL = 1e-5;
constexpr double T_MIN = 1e-4;
constexpr u32 CODE_SHIFT = 28;
constexpr u32 ID_MASK = (1u << CODE_SHIFT) - 1;

enum PatchKind : uint8_t { CONVEX = 0, SADDLE = 1, CONCAVE = 2 };

// mesh/tables.rs: corner offsets, EDGE_LOWER, generated triangle table.
extern const uint8_t EDGE_LOWER[3][4];
struct McTable {
    uint16_t offset[257];
    std::vector<std::array<uint8_t, 3>> tris;
};
const McTable& mc_table();
void generate_mc_table(McTable& t);  // restates mesh/tables.rs:47-274

// mesh/extract.rs:56-72 + per-brick ends (ChunkMesh, :77-84)
struct BrickMesh {
    std::vector<u32> edges;
    std::vector<std::array<float, 3>> positions, normals;
    std::vector<u32> atoms;
    std::vector<uint8_t> patches;
    std::vector<std::array<u32, 3>> tris;
    std::vector<u32> vertex_end, tri_end;
    uint64_t evals = 0;
    // Per vertex: the refinement's edge parameter (unclamped, f64) and the
    // number of evaluations; kept for the device comparison tests.
    std::vector<double> t;
    std::vector<u32> vevals;
};

double edge_clamp(const Engine& e);                                            // extract.rs:47-51
void mesh_brick(const Engine& e, const int32_t key[3], const float* values, BrickMesh& out);  // :146-223
void cells(size_t b, const float* values, BrickMesh& out);                     // :226-265

// mesh/mod.rs:47-66
struct Mesh {
    std::vector<std::array<float, 3>> positions, normals;
    std::vector<std::array<u32, 3>> triangles;
    std::vector<u32> vertex_atom;
    std::vector<uint8_t> vertex_patch;
    uint64_t evals = 0;
    std::vector<double> t;
    std::vector<u32> vevals;
};

// mesh/extract.rs:606-691
std::string assemble(const std::vector<std::array<int32_t, 3>>& bricks, std::vector<BrickMesh>& parts,
                     size_t chunk, Mesh& mesh);

// mesh/mod.rs:181-300 (component ids, per-component sums in the reference order)
struct ComponentStats {
    size_t components = 0, cavities = 0;
    double area = 0, volume = 0;
};
ComponentStats measure(const Mesh& m);

// The measure stage (mesh/mod.rs validate, components; post.rs grouping).
struct Validation {
    u64 boundary_edges = 0, non_manifold_edges = 0, misoriented_edges = 0, degenerate_triangles = 0,
        duplicate_triangles = 0, invalid_indices = 0, bad_vertices = 0;
};
struct Component {
    u64 vertices = 0, triangles = 0;
    double area = 0, volume = 0;
    int64_t euler = 0;
    u64 first_triangle = 0;
};
struct Measured {
    Validation validation;
    bool indices_valid = false;  // components and grouping are measured only then
    std::vector<u32> vertex_component;
    std::vector<Component> components;
    std::vector<std::array<u32, 3>> grouped;
    double area = 0, volume = 0;
    int64_t euler = 0;
    size_t cavities = 0;
};
Measured measure_full(const std::vector<std::array<float, 3>>& pos, const std::vector<std::array<float, 3>>& nrm,
                      const std::vector<std::array<u32, 3>>& tris);

// Writers (export/fmt.rs, obj.rs, ply.rs, msms.rs). The version string is
// the crate version the files name (Cargo.toml:23).
constexpr const char* SESMESH_VERSION = "1.3.2";
struct ExportMesh {
    std::vector<std::array<float, 3>> positions, normals;
    std::vector<std::array<u32, 3>> triangles;
    std::vector<u32> vertex_atom, vertex_component;
    std::vector<uint8_t> vertex_patch;
};
struct ObjOptions {
    int precision = 4;
    bool normals = true;
    std::vector<std::string> comments;
    std::string object_name = "ses";
    std::vector<std::array<float, 3>> vertex_colors;
    std::vector<std::pair<size_t, std::string>> groups;
};
struct PlyOptions {
    std::vector<std::string> comments;
    bool normals = true;
    std::vector<std::array<float, 3>> vertex_colors;
};
struct MsmsHeader {
    std::string description;
    size_t spheres = 0;
    double density = 0, probe_radius = 0;
};
void push_fixed(std::string& b, float x, int prec);
std::string write_obj(const ExportMesh& m, const ObjOptions& o);
std::string write_ply(const ExportMesh& m, const PlyOptions& o);
void write_msms(const ExportMesh& m, const MsmsHeader& h, std::string& vert, std::string& face);

// The whole pipeline of surface.rs:1001-1079 for the CPU back end, with
// `threads` worker threads over chunks of par::chunk_len(bricks) bricks.
struct Surface {
    SesField field;
    Engine engine;
    size_t touched = 0;
    std::vector<std::array<int32_t, 3>> bricks;
    Mesh mesh;
    double t_geometry = 0, t_classify = 0, t_mesh = 0, t_assemble = 0;
};
std::string compute_surface(const std::vector<Sphere>& spheres, double probe, double spacing, size_t brick,
                            int threads, Surface& out, int kind = K_SES);

size_t chunk_len(size_t n);  // par.rs:66-68
std::string classify_bricks(Surface& S, int threads);

}  // namespace orc
