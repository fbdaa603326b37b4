// The MI355X (gfx950) mesh stage of sesmesh behind the C ABI of
// include/sesmesh_gpu.h: device open and self-test, resident arrangement,
// and the whole mesh stage as a sequence of kernels on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sesmesh_gpu.h"
#include "dev_classify.hpp"
#include "dev_common.hpp"
#include "dev_sample.hpp"
#include "dev_refine.hpp"
#include "dev_cells.hpp"
#include "host.hpp"

namespace smg {
void make_mc_table(uint16_t offset[257], std::vector<uint8_t>& edges);
}

using namespace smg;

// ------------------------------------------------------------------ errors
static thread_local std::string g_error;
static int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}
#define HIP_TRY(call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(SMG_E_CALL, std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

extern "C" const char* smg_last_error(void) { return g_error.c_str(); }

// ------------------------------------------------------------------ scans
namespace {

constexpr uint32_t SCAN_BLOCK = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Exclusive scan of tiles: out[i] = sum(in[tile start..i)), sums[tile].
__global__ void __launch_bounds__(256) k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t* sums, size_t n) {
    __shared__ uint32_t tmp[4];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], acc = 0;
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        acc += v[k];
    }
    uint32_t total;
    uint32_t before = block_scan_256(acc, tmp, total);
    for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = before;
        before += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// In-place exclusive scan of the tile sums by one block; total at sums[n].
__global__ void __launch_bounds__(256) k_scan_sums(uint32_t* sums, size_t n) {
    __shared__ uint32_t tmp[4];
    uint32_t carry = 0;
    for (size_t c0 = 0; c0 < n; c0 += SCAN_TILE) {
        size_t base = c0 + (size_t)threadIdx.x * SCAN_ITEMS;
        uint32_t v[SCAN_ITEMS], acc = 0;
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
            v[k] = base + k < n ? sums[base + k] : 0u;
            acc += v[k];
        }
        uint32_t total;
        uint32_t before = carry + block_scan_256(acc, tmp, total);
        for (uint32_t k = 0; k < SCAN_ITEMS; k++) {
            if (base + k < n) sums[base + k] = before;
            before += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) sums[n] = carry;
}

__global__ void k_scan_add(uint32_t* out, const uint32_t* sums, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / SCAN_TILE];
}

__global__ void k_popcount(const uint32_t* bits, uint32_t* cnt, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = (uint32_t)__popc(bits[i]);
}

// Set bits of the touched bitmap, in ascending order, as linear indices.
__global__ void k_bits_to_list(const uint32_t* bits, const uint32_t* off, size_t n, uint32_t* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w = bits[i], at = off[i];
    while (w) {
        uint32_t bit = __ffs(w) - 1;
        out[at++] = (uint32_t)(i * 32 + bit);
        w &= w - 1;
    }
}

// Surface bricks: keys in order, and the dense index grid.
__global__ void k_compact_bricks(DRun P, const uint32_t* touched, const uint32_t* keep, const uint32_t* off,
                                 uint32_t n, int32_t* keys, uint32_t* dense) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    uint32_t at = off[i];
    int32_t key[3];
    key_of_lin(P, touched[i], key);
    keys[3 * at] = key[0];
    keys[3 * at + 1] = key[1];
    keys[3 * at + 2] = key[2];
    dense[touched[i]] = at;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Flags of sub-blocks to sample (state NONE) or to refine (edges > 0).
__global__ void k_flag_states(const uint8_t* st, uint32_t n, uint32_t* flag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = st[i] == ST_NONE ? 1u : 0u;
}
__global__ void k_flag_nonzero(const uint32_t* c, uint32_t n, uint32_t* flag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = c[i] ? 1u : 0u;
}
__global__ void k_compact_index(const uint32_t* flag, const uint32_t* off, uint32_t n, uint32_t* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[off[i]] = i;
}

// Frame-local f32 copies (field/gpu.rs:253-318, kernel.rs:627-637, 717-746).
__global__ void k_localize(DArr A, DLoc L, double fx, double fy, double fz) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n_spheres) {
        L.px[i] = (float)(A.cx[i] - fx);
        L.py[i] = (float)(A.cy[i] - fy);
        L.pz[i] = (float)(A.cz[i] - fz);
        float r32 = (float)A.R[i];
        L.pr[i] = r32;
        L.pr2[i] = r32 * r32;
        float rt = fmaxf(r32 - INSIDE_TOL, 0.0f);
        L.pr2tol[i] = rt * rt;
    }
    if (i < A.n_circles) {
        L.mx[i] = (float)(A.mx[i] - fx);
        L.my[i] = (float)(A.my[i] - fy);
        L.mz[i] = (float)(A.mz[i] - fz);
        L.ux[i] = (float)A.ax[i];
        L.uy[i] = (float)A.ay[i];
        L.uz[i] = (float)A.az[i];
        L.ex[i] = (float)A.e1x[i];
        L.ey[i] = (float)A.e1y[i];
        L.ez[i] = (float)A.e1z[i];
        L.rho[i] = (float)A.crho[i];
    }
    if (i < A.n_vertices) {
        L.vx[i] = (float)(A.vx[i] - fx);
        L.vy[i] = (float)(A.vy[i] - fy);
        L.vz[i] = (float)(A.vz[i] - fz);
    }
}

// Self-test operations (reference crates/sesmesh-gpu/src/selftest.rs:25-38
// for f32; the same expressions in f64).
template <class T>
__host__ __device__ inline void selftest_ops(T x, T y, T r[8]) {
    T dx = x - y, dy = y * (T)0.5 - x, dz = x * y - (T)1.0;
    T d = sqrt(dx * dx + dy * dy + dz * dz);
    r[0] = x + y;
    r[1] = x * y;
    r[2] = x / y;
    r[3] = sqrt(fabs(x));
    r[4] = d;
    r[5] = (d - y) / d;
    r[6] = fmin(x, y);
    r[7] = fabs(d - fabs(y));
}
__global__ void k_selftest(const float* a32, const float* b32, float* o32, const double* a64, const double* b64,
                           double* o64, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    selftest_ops<float>(a32[i], b32[i], o32 + 8 * (size_t)i);
    selftest_ops<double>(a64[i], b64[i], o64 + 8 * (size_t)i);
}

inline uint32_t blocks_for(size_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

// ------------------------------------------------------------------ objects
struct smg_device {
    int ordinal = 0;
    hipDeviceProp_t prop{};
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    uint32_t* host_u32 = nullptr;  // pinned staging for counts
};

struct smg_arrangement {
    Arrangement a;
};

struct DevBuf {
    std::vector<void*> ptrs;
    size_t bytes = 0;
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
        bytes = 0;
    }
};

struct smg_resident {
    smg_device* dev = nullptr;
    DArr A{};
    DevBuf buf;
    // Host copies needed to set up a run: centre bounding box.
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    size_t n_spheres = 0;
};

struct smg_mesh {
    smg_device* dev = nullptr;
    DRun P{};
    size_t touched = 0, bricks = 0, vertices = 0, triangles = 0, sampled = 0, overflow = 0;
    uint64_t evals = 0, samples = 0;
    uint32_t words = 0;
    // Device arrays owned by the mesh.
    int32_t* keys = nullptr;
    float* values = nullptr;
    float* pos = nullptr;
    float* nrm = nullptr;
    uint32_t* atom = nullptr;
    uint8_t* patch = nullptr;
    double* t = nullptr;
    uint32_t* vevals = nullptr;
    uint32_t* tris = nullptr;
    ~smg_mesh() {
        void* ps[] = {keys, values, pos, nrm, atom, patch, t, vevals, tris};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
};

// ------------------------------------------------------------------ device
static std::once_flag g_table_once;
static uint16_t g_tri_offset[257];
static std::vector<uint8_t> g_tri_edges;

extern "C" int smg_device_open(int ordinal, smg_device** out, smg_device_info* info) {
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return fail(SMG_E_NO_RUNTIME, "hipGetDeviceCount failed");
    if (ordinal < 0 || ordinal >= count)
        return fail(SMG_E_NO_DEVICE, "no HIP device " + std::to_string(ordinal) + " (" + std::to_string(count) +
                                         " present)");
    smg_device* d = new smg_device();
    d->ordinal = ordinal;
    if (hipSetDevice(ordinal) != hipSuccess || hipGetDeviceProperties(&d->prop, ordinal) != hipSuccess) {
        delete d;
        return fail(SMG_E_CALL, "cannot open device");
    }
    std::string arch = d->prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0) {
        delete d;
        return fail(SMG_E_UNSUPPORTED, "device architecture " + arch + " is not gfx950 (built for gfx950 only)");
    }
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        delete d;
        return fail(SMG_E_CALL, "hipStreamCreate failed");
    }
    for (auto& e : d->ev) (void)hipEventCreate(&e);
    {  // keep stream-ordered allocations cached between mesh stages
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, ordinal) == hipSuccess) {
            uint64_t keep = ~0ull;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    (void)hipHostMalloc((void**)&d->host_u32, 64 * sizeof(uint32_t), hipHostMallocDefault);
    std::call_once(g_table_once, [] { make_mc_table(g_tri_offset, g_tri_edges); });
    uint8_t lower[12] = {0, 2, 4, 6, 0, 1, 4, 5, 0, 1, 2, 3};
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_tri_offset), g_tri_offset, sizeof(g_tri_offset)) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(c_tri_edges), g_tri_edges.data(), g_tri_edges.size()) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(c_edge_lower), lower, sizeof(lower)) != hipSuccess) {
        smg_device_close(d);
        return fail(SMG_E_CALL, "cannot upload the Marching Cubes table");
    }
    uint64_t m32 = 0, m64 = 0;
    int rc = smg_selftest(d, 4096, 1, &m32, &m64);
    if (rc != SMG_OK) {
        smg_device_close(d);
        return rc;
    }
    if (m32 || m64) {
        smg_device_close(d);
        return fail(SMG_E_SELFTEST, "device arithmetic differs from the host's: " + std::to_string(m32) +
                                        " f32 and " + std::to_string(m64) + " f64 results");
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        std::snprintf(info->name, sizeof(info->name), "%s", d->prop.name);
        std::snprintf(info->arch, sizeof(info->arch), "%s", d->prop.gcnArchName);
        info->compute_units = d->prop.multiProcessorCount;
        info->clock_khz = d->prop.clockRate;
        info->total_memory = d->prop.totalGlobalMem;
        info->lds_per_block = d->prop.sharedMemPerBlock;
        info->wavefront = d->prop.warpSize;
    }
    *out = d;
    return SMG_OK;
}

extern "C" void smg_device_close(smg_device* d) {
    if (!d) return;
    (void)hipSetDevice(d->ordinal);
    for (auto& e : d->ev)
        if (e) (void)hipEventDestroy(e);
    if (d->host_u32) (void)hipHostFree(d->host_u32);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

// Inputs as crates/sesmesh-gpu/src/selftest.rs:40-80 (xorshift64*; one in 16
// subnormal, one in 16 over the whole exponent range, else within 2^+-30),
// and the f64 analogue (2^+-60).
namespace {
struct XorShift {
    uint64_t s;
    explicit XorShift(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull | 1ull) {}
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1Dull;
    }
    uint64_t exponent(uint64_t lo, uint64_t hi) { return lo + next() % (hi - lo + 1); }
};
float input_f32(XorShift& g, size_t i) {
    uint64_t r = g.next();
    uint64_t sign = (r >> 63) << 31, mant = (r >> 8) & 0x7FFFFF;
    uint64_t ex = i % 16 == 0 ? 0 : (i % 16 == 1 ? g.exponent(1, 254) : g.exponent(127 - 30, 127 + 30));
    if (ex == 0) mant |= 1;
    uint32_t bits = (uint32_t)(sign | (ex << 23) | mant);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
double input_f64(XorShift& g, size_t i) {
    uint64_t r = g.next();
    uint64_t sign = r & 0x8000000000000000ull, mant = g.next() & 0xFFFFFFFFFFFFFull;
    uint64_t ex = i % 16 == 0 ? 0 : (i % 16 == 1 ? g.exponent(1, 2046) : g.exponent(1023 - 60, 1023 + 60));
    if (ex == 0) mant |= 1;
    uint64_t bits = sign | (ex << 52) | mant;
    double f;
    std::memcpy(&f, &bits, 8);
    return f;
}
template <class T, class U>
bool same_bits(T a, T b) {
    U x, y;
    std::memcpy(&x, &a, sizeof(T));
    std::memcpy(&y, &b, sizeof(T));
    return x == y || (a != a && b != b);
}
}  // namespace

extern "C" int smg_selftest(smg_device* d, size_t n, uint64_t seed, uint64_t* m32, uint64_t* m64) {
    if (!d || n == 0) return fail(SMG_E_INVALID, "selftest: no device or no values");
    HIP_TRY(hipSetDevice(d->ordinal));
    XorShift g(seed);
    std::vector<float> a32(n), b32(n), o32(8 * n);
    std::vector<double> a64(n), b64(n), o64(8 * n);
    for (size_t i = 0; i < n; i++) {
        a32[i] = input_f32(g, i);
        b32[i] = input_f32(g, i + 7);
        a64[i] = input_f64(g, i);
        b64[i] = input_f64(g, i + 7);
    }
    float *da32, *db32, *do32;
    double *da64, *db64, *do64;
    HIP_TRY(hipMalloc(&da32, n * 4));
    HIP_TRY(hipMalloc(&db32, n * 4));
    HIP_TRY(hipMalloc(&do32, 8 * n * 4));
    HIP_TRY(hipMalloc(&da64, n * 8));
    HIP_TRY(hipMalloc(&db64, n * 8));
    HIP_TRY(hipMalloc(&do64, 8 * n * 8));
    HIP_TRY(hipMemcpy(da32, a32.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db32, b32.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(da64, a64.data(), n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db64, b64.data(), n * 8, hipMemcpyHostToDevice));
    k_selftest<<<blocks_for(n, 256), 256, 0, d->stream>>>(da32, db32, do32, da64, db64, do64, (uint32_t)n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d->stream));
    HIP_TRY(hipMemcpy(o32.data(), do32, 8 * n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(o64.data(), do64, 8 * n * 8, hipMemcpyDeviceToHost));
    for (void* p : {(void*)da32, (void*)db32, (void*)do32, (void*)da64, (void*)db64, (void*)do64}) (void)hipFree(p);
    uint64_t c32 = 0, c64 = 0;
    for (size_t i = 0; i < n; i++) {
        float r32[8];
        double r64[8];
        selftest_ops<float>(a32[i], b32[i], r32);
        selftest_ops<double>(a64[i], b64[i], r64);
        for (int k = 0; k < 8; k++) {
            c32 += !same_bits<float, uint32_t>(r32[k], o32[8 * i + k]);
            c64 += !same_bits<double, uint64_t>(r64[k], o64[8 * i + k]);
        }
    }
    if (m32) *m32 = c32;
    if (m64) *m64 = c64;
    return SMG_OK;
}

// ------------------------------------------------------------- arrangement
extern "C" int smg_arrangement_build(const double* xyzr, size_t n, double probe, int threads, smg_arrangement** out) {
    *out = nullptr;
    if (!xyzr && n) return fail(SMG_E_INVALID, "no spheres");
    smg_arrangement* a = new smg_arrangement();
    std::string e = build_arrangement(xyzr, n, probe, threads, a->a);
    if (!e.empty()) {
        delete a;
        return fail(SMG_E_INVALID, e);
    }
    *out = a;
    return SMG_OK;
}
extern "C" void smg_arrangement_free(smg_arrangement* a) { delete a; }

static void grid_ref(const Grid& g, smg_grid_ref& r) {
    r.origin = g.origin;
    r.cell = g.cell;
    for (int a = 0; a < 3; a++) r.dims[a] = g.dims[a];
    r.starts = g.starts.data();
    r.items = g.items.data();
    r.n_items = g.items.size();
}

extern "C" int smg_arrangement_view(const smg_arrangement* p, smg_arrangement_ref* r) {
    if (!p || !r) return fail(SMG_E_INVALID, "null arrangement");
    const Arrangement& a = p->a;
    std::memset(r, 0, sizeof(*r));
    r->n_spheres = a.radius.size();
    r->center = a.center.data();
    r->radius = a.radius.data();
    r->expanded = a.expanded.data();
    r->status = a.status.data();
    r->orig = a.orig.data();
    r->nb_off = a.nb_off.data();
    r->nb_items = a.nb_items.data();
    r->n_circles = a.circle.size() / 10;
    r->circle_ij = a.circle_ij.data();
    r->circle = a.circle.data();
    r->cb_off = a.cb_off.data();
    r->cb_items = a.cb_items.data();
    r->n_vertices = a.vertex.size() / 3;
    r->vertex = a.vertex.data();
    r->vertex_atoms = a.vertex_atoms.data();
    grid_ref(a.sphere_grid, r->sphere_grid);
    grid_ref(a.circle_grid, r->circle_grid);
    grid_ref(a.vertex_grid, r->vertex_grid);
    r->probe = a.probe;
    r->max_expanded = a.max_expanded;
    r->rho_max = a.rho_max;
    return SMG_OK;
}

// ---------------------------------------------------------------- resident
template <class T>
static int upload(DevBuf& b, const T* src, size_t count, const T** dst) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    HIP_TRY(hipMalloc(&p, bytes));
    b.ptrs.push_back(p);
    b.bytes += bytes;
    if (count) HIP_TRY(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T*)p;
    return SMG_OK;
}

#define UP(src, count, dst)                                   \
    do {                                                      \
        int rc_ = upload(r->buf, (src), (count), &(dst));     \
        if (rc_ != SMG_OK) {                                  \
            smg_resident_free(r);                             \
            return rc_;                                       \
        }                                                     \
    } while (0)

static int upload_grid(smg_resident* r, const smg_grid_ref& g, DGrid& d) {
    size_t cells = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    int rc = upload(r->buf, g.starts, cells + 1, &d.starts);
    if (rc == SMG_OK) rc = upload(r->buf, g.items, g.n_items, &d.items);
    d.ox = g.origin[0];
    d.oy = g.origin[1];
    d.oz = g.origin[2];
    d.cell = g.cell;
    d.dx = g.dims[0];
    d.dy = g.dims[1];
    d.dz = g.dims[2];
    return rc;
}

extern "C" int smg_resident_new(smg_device* dev, const smg_arrangement_ref* a, smg_resident** out) {
    *out = nullptr;
    if (!dev || !a) return fail(SMG_E_INVALID, "null device or arrangement");
    if (a->n_spheres >= (1u << 30) || a->n_circles >= (1u << 30) || a->n_vertices >= (1u << 30))
        return fail(SMG_E_LENGTH, "arrangement too large for 30-bit candidate indices");
    if (!(a->probe > 0.0)) return fail(SMG_E_UNSUPPORTED, "the device path computes the SES (probe > 0)");
    HIP_TRY(hipSetDevice(dev->ordinal));
    smg_resident* r = new smg_resident();
    r->dev = dev;
    size_t n = a->n_spheres, nc = a->n_circles, nv = a->n_vertices;
    r->n_spheres = n;
    std::vector<double> x(n), y(n), z(n);
    for (size_t i = 0; i < n; i++) {
        x[i] = a->center[3 * i];
        y[i] = a->center[3 * i + 1];
        z[i] = a->center[3 * i + 2];
    }
    if (n) {
        for (int q = 0; q < 3; q++) r->lo[q] = r->hi[q] = a->center[q];
        for (size_t i = 0; i < n; i++)
            for (int q = 0; q < 3; q++) {
                r->lo[q] = std::fmin(r->lo[q], a->center[3 * i + q]);
                r->hi[q] = std::fmax(r->hi[q], a->center[3 * i + q]);
            }
    }
    DArr& A = r->A;
    UP(x.data(), n, A.cx);
    UP(y.data(), n, A.cy);
    UP(z.data(), n, A.cz);
    UP(a->expanded, n, A.R);
    UP(a->radius, n, A.r);
    UP(a->status, n, A.status);
    UP(a->orig, n, A.orig);
    UP(a->nb_off, n + 1, A.nb_off);
    UP(a->nb_items, (size_t)a->nb_off[n], A.nb_items);
    std::vector<double> cf[10];
    std::vector<uint32_t> ci(nc), cj(nc);
    for (int q = 0; q < 10; q++) cf[q].resize(nc);
    for (size_t c = 0; c < nc; c++) {
        for (int q = 0; q < 10; q++) cf[q][c] = a->circle[10 * c + q];
        ci[c] = a->circle_ij[2 * c];
        cj[c] = a->circle_ij[2 * c + 1];
    }
    UP(cf[0].data(), nc, A.mx);
    UP(cf[1].data(), nc, A.my);
    UP(cf[2].data(), nc, A.mz);
    UP(cf[3].data(), nc, A.ax);
    UP(cf[4].data(), nc, A.ay);
    UP(cf[5].data(), nc, A.az);
    UP(cf[6].data(), nc, A.e1x);
    UP(cf[7].data(), nc, A.e1y);
    UP(cf[8].data(), nc, A.e1z);
    UP(cf[9].data(), nc, A.crho);
    UP(ci.data(), nc, A.ci);
    UP(cj.data(), nc, A.cj);
    UP(a->cb_off, nc + 1, A.cb_off);
    UP(a->cb_items, (size_t)a->cb_off[nc], A.cb_items);
    std::vector<double> vxyz[3];
    for (int q = 0; q < 3; q++) vxyz[q].resize(nv);
    for (size_t v = 0; v < nv; v++)
        for (int q = 0; q < 3; q++) vxyz[q][v] = a->vertex[3 * v + q];
    UP(vxyz[0].data(), nv, A.vx);
    UP(vxyz[1].data(), nv, A.vy);
    UP(vxyz[2].data(), nv, A.vz);
    UP(a->vertex_atoms, 3 * nv, A.vatoms);
    int rc = upload_grid(r, a->sphere_grid, A.gs);
    if (rc == SMG_OK) rc = upload_grid(r, a->circle_grid, A.gc);
    if (rc == SMG_OK) rc = upload_grid(r, a->vertex_grid, A.gv);
    if (rc != SMG_OK) {
        smg_resident_free(r);
        return rc;
    }
    A.n_spheres = (uint32_t)n;
    A.n_circles = (uint32_t)nc;
    A.n_vertices = (uint32_t)nv;
    A.probe = a->probe;
    A.max_expanded = a->max_expanded;
    A.rho_max = a->rho_max;
    *out = r;
    return SMG_OK;
}

extern "C" void smg_resident_free(smg_resident* r) {
    if (!r) return;
    (void)hipSetDevice(r->dev->ordinal);
    r->buf.release();
    delete r;
}

extern "C" size_t smg_resident_bytes(const smg_resident* r) { return r ? r->buf.bytes : 0; }

// ------------------------------------------------------------- mesh stage
namespace {

// Temporary device buffers of one run, freed at the end.
struct Temps {
    std::vector<void*> ptrs;
    hipStream_t s;
    explicit Temps(hipStream_t st) : s(st) {}
    ~Temps() {
        for (void* p : ptrs) (void)hipFreeAsync(p, s);
    }
    template <class T>
    hipError_t get(T** p, size_t count) {
        hipError_t e = hipMallocAsync((void**)p, std::max<size_t>(count, 1) * sizeof(T), s);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Exclusive scan of n u32 values into `out`; the total is written to
// host[slot] after a synchronisation by the caller.
int scan_u32(hipStream_t s, Temps& T, const uint32_t* in, uint32_t* out, size_t n, uint32_t** total_dev) {
    size_t tiles = std::max<size_t>((n + SCAN_TILE - 1) / SCAN_TILE, 1);
    uint32_t* sums;
    HIP_TRY(T.get(&sums, tiles + 1));
    if (n) {
        k_scan_tiles<<<(uint32_t)tiles, SCAN_BLOCK, 0, s>>>(in, out, sums, n);
        k_scan_sums<<<1, SCAN_BLOCK, 0, s>>>(sums, tiles);
        k_scan_add<<<blocks_for(n, 256), 256, 0, s>>>(out, sums, n);
    } else {
        k_fill_u32<<<1, 64, 0, s>>>(sums + tiles, 0u, 1);
    }
    HIP_TRY(hipGetLastError());
    *total_dev = sums + tiles;
    return SMG_OK;
}

int read_u32(smg_device* d, const uint32_t* dev_ptr, uint32_t& v) {
    HIP_TRY(hipMemcpyAsync(d->host_u32, dev_ptr, 4, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    v = d->host_u32[0];
    return SMG_OK;
}

#define TRY(x)                  \
    do {                        \
        int rc_ = (x);          \
        if (rc_ != SMG_OK) {    \
            return rc_;         \
        }                       \
    } while (0)

// List capacities, tried in turn: LDS variants sized for common and for
// large sub-blocks, then global scratch for the rest. Each pass takes the
// previous pass's overflow list.
constexpr Caps SAMPLE_CAPS[3] = {{128, 64, 256, 64, 256}, {256, 128, 512, 128, 512}, {1024, 512, 4096, 2048, 4096}};
constexpr RCaps REFINE_CAPS[3] = {{128, 64, 256, 256, 0}, {256, 128, 512, 512, 0}, {1024, 512, 4096, 4096, 0}};
constexpr uint32_t LDS_EDGES = 512;      // crossing-edge buffer of the LDS passes
constexpr uint32_t SCRATCH_CHUNK = 256;   // workgroups per scratch launch
constexpr uint32_t STATIC_LDS = 6144;     // static LDS of k_sample / k_refine, bound

}  // namespace

static int mesh_stage(smg_resident* r, const smg_mesh_params* prm, smg_mesh* M, smg_mesh_times* times) {
    smg_device* d = r->dev;
    hipStream_t s = d->stream;
    const DArr& A = r->A;
    HIP_TRY(hipSetDevice(d->ordinal));
    // Engine::new (field/kernel.rs:204-261).
    double spacing = prm->spacing;
    uint32_t b = prm->brick;
    if (!(std::isfinite(spacing) && spacing >= 0.01)) return fail(SMG_E_INVALID, "grid spacing must be at least 0.01 A");
    if (b < 4 || b > 32 || b % 2) return fail(SMG_E_UNSUPPORTED, "the device path needs an even brick size in 4..=32");
    if (prm->refine != SMG_REFINE_EXACT) return fail(SMG_E_UNSUPPORTED, "only exact refinement is implemented");
    double center[3] = {0, 0, 0}, half = 0.0;
    if (r->n_spheres) {
        double dd[3];
        for (int q = 0; q < 3; q++) {
            dd[q] = r->hi[q] - r->lo[q];
            center[q] = (r->lo[q] + r->hi[q]) * 0.5;
        }
        half = 0.5 * std::fmax(std::fmax(dd[0], dd[1]), dd[2]);
    }
    DRun& P = M->P;
    P.h = spacing;
    P.b = b;
    P.sub = b / 2;
    P.side = b + 1;
    P.spb = P.side * P.side * P.side;
    P.band = std::sqrt(3.0) * spacing * 1.01;
    P.probe = A.probe;
    for (int q = 0; q < 3; q++) P.frame[q] = (int64_t)std::round(center[q] / spacing);
    P.h32 = (float)spacing;
    P.fband = (float)P.band;
    P.fprobe = (float)A.probe;
    float cut = (float)(A.probe + P.band);
    P.fcutoff2 = cut * cut;
    double reach = half + 2.0 * A.max_expanded + 2.0 * A.probe + 4.0 * (double)b * spacing;
    if (!(reach < 2000.0 && reach / spacing < (double)(1u << 23)))
        return fail(SMG_E_UNSUPPORTED, "coordinates too far from the frame for the f32 sampling kernels");
    double coord_max = std::fmax(std::fmax(std::fabs(center[0]), std::fabs(center[1])), std::fabs(center[2])) + half +
                       A.max_expanded + A.probe + 2.0 * P.band;
    {  // mesh/extract.rs:47-51
        double cm = std::fmin(std::fmax(coord_max, 1.0), (double)3.4028234663852886e38 / 4.0);
        float m = (float)cm;
        uint32_t bits;
        std::memcpy(&bits, &m, 4);
        bits += 1;
        float nx;
        std::memcpy(&nx, &bits, 4);
        P.t_min = std::fmin(std::fmax(4.0 * (double)(nx - m) / spacing, 1e-4), 0.25);
    }
    Temps T(s);
    hipEvent_t* ev = d->ev;
    HIP_TRY(hipEventRecord(ev[0], s));

    // Frame-local f32 copies.
    DLoc L{};
    size_t nmax = std::max<size_t>({A.n_spheres, A.n_circles, A.n_vertices, 1});
    float* lbuf;
    HIP_TRY(T.get(&lbuf, 19 * nmax));
    float** fields[19] = {&L.px, &L.py, &L.pz, &L.pr, &L.pr2, &L.pr2tol, &L.mx, &L.my, &L.mz, &L.ux,
                          &L.uy, &L.uz, &L.ex, &L.ey, &L.ez, &L.rho,   &L.vx, &L.vy, &L.vz};
    for (int q = 0; q < 19; q++) *fields[q] = lbuf + q * nmax;
    double fo[3] = {(double)P.frame[0] * spacing, (double)P.frame[1] * spacing, (double)P.frame[2] * spacing};
    k_localize<<<blocks_for(nmax, 256), 256, 0, s>>>(A, L, fo[0], fo[1], fo[2]);
    HIP_TRY(hipGetLastError());

    // Touched bricks (lattice.rs:105-163) in a dense bitmap over their box.
    double bh = (double)b * spacing;
    unsigned long long* lohi;
    HIP_TRY(T.get(&lohi, 6));
    unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    HIP_TRY(hipMemcpyAsync(lohi, init, sizeof(init), hipMemcpyHostToDevice, s));
    if (A.n_spheres) k_brick_range<<<blocks_for(A.n_spheres, 256), 256, 0, s>>>(A, bh, lohi);
    unsigned long long hl[6];
    HIP_TRY(hipMemcpyAsync(hl, lohi, sizeof(hl), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    M->touched = 0;
    if (hl[0] == ~0ull) {  // no non-contained sphere: no surface
        M->bricks = M->vertices = M->triangles = 0;
        return SMG_OK;
    }
    uint64_t total_cells = 1;
    for (int q = 0; q < 3; q++) {
        int64_t lo = (int64_t)hl[q] - (int64_t(1) << 41), hi = (int64_t)hl[3 + q] - (int64_t(1) << 41);
        if (std::llabs(lo) >= KEY_BIAS - 1 || std::llabs(hi) >= KEY_BIAS - 1)
            return fail(SMG_E_INVALID, "coordinates too far from the origin for grid spacing " + std::to_string(spacing) + " A");
        P.blo[q] = lo;
        P.bdim[q] = (uint32_t)(hi - lo + 1);
        total_cells *= P.bdim[q];
    }
    if (total_cells >= (1ull << 32))
        return fail(SMG_E_UNSUPPORTED, "brick box too large for the dense touched-brick map");
    size_t nwords = (size_t)((total_cells + 31) / 32);
    uint32_t *bits, *wcount, *woff, *dense;
    HIP_TRY(T.get(&bits, nwords));
    HIP_TRY(T.get(&wcount, nwords));
    HIP_TRY(T.get(&woff, nwords));
    HIP_TRY(T.get(&dense, total_cells));
    HIP_TRY(hipMemsetAsync(bits, 0, nwords * 4, s));
    k_mark_touched<<<blocks_for(A.n_spheres, 128), 128, 0, s>>>(A, P, bh, bits);
    k_popcount<<<blocks_for(nwords, 256), 256, 0, s>>>(bits, wcount, nwords);
    uint32_t* ntouched_d;
    TRY(scan_u32(s, T, wcount, woff, nwords, &ntouched_d));
    uint32_t ntouched;
    TRY(read_u32(d, ntouched_d, ntouched));
    M->touched = ntouched;
    uint32_t *touched, *keep, *koff;
    HIP_TRY(T.get(&touched, ntouched));
    HIP_TRY(T.get(&keep, ntouched));
    HIP_TRY(T.get(&koff, ntouched));
    k_bits_to_list<<<blocks_for(nwords, 256), 256, 0, s>>>(bits, woff, nwords, touched);
    // Brick classification (lattice.rs:168-183), one wavefront per brick.
    HIP_TRY(hipEventRecord(ev[6], s));
    if (ntouched) k_classify_bricks<<<ntouched, WAVE, 0, s>>>(A, P, touched, ntouched, keep);
    HIP_TRY(hipEventRecord(ev[7], s));
    uint32_t* nb_d;
    TRY(scan_u32(s, T, keep, koff, ntouched, &nb_d));
    uint32_t nb;
    TRY(read_u32(d, nb_d, nb));
    M->bricks = nb;
    HIP_TRY(hipMalloc(&M->keys, std::max<size_t>(nb, 1) * 3 * sizeof(int32_t)));
    k_fill_u32<<<blocks_for(total_cells, 256), 256, 0, s>>>(dense, 0xffffffffu, total_cells);
    if (ntouched) k_compact_bricks<<<blocks_for(ntouched, 256), 256, 0, s>>>(P, touched, keep, koff, ntouched, M->keys, dense);
    uint8_t* st;
    HIP_TRY(T.get(&st, 8 * (size_t)nb));
    HIP_TRY(hipEventRecord(ev[8], s));
    if (nb) k_sub_states<<<8 * nb, WAVE, 0, s>>>(A, P, M->keys, nb, st);
    HIP_TRY(hipEventRecord(ev[9], s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[1], s));

    // Samples (field/kernel.rs:415-449).
    HIP_TRY(hipMalloc(&M->values, std::max<size_t>((size_t)nb * P.spb, 1) * sizeof(float)));
    if (nb) k_assign_uniform<<<nb, 256, 0, s>>>(P, st, nb, M->values);
    uint32_t nsub = 8 * nb;
    uint32_t *sflag, *soff, *swork;
    HIP_TRY(T.get(&sflag, nsub));
    HIP_TRY(T.get(&soff, nsub));
    HIP_TRY(T.get(&swork, nsub));
    if (nsub) k_flag_states<<<blocks_for(nsub, 256), 256, 0, s>>>(st, nsub, sflag);
    uint32_t* nsw_d;
    TRY(scan_u32(s, T, sflag, soff, nsub, &nsw_d));
    if (nsub) k_compact_index<<<blocks_for(nsub, 256), 256, 0, s>>>(sflag, soff, nsub, swork);
    uint32_t nsw;
    TRY(read_u32(d, nsw_d, nsw));
    M->sampled = nsw;
    unsigned long long* nsamp;
    HIP_TRY(T.get(&nsamp, 1));
    HIP_TRY(hipMemsetAsync(nsamp, 0, 8, s));
    const size_t lds_max = d->prop.sharedMemPerBlock;
    M->overflow = 0;
    {
        const uint32_t* list = swork;
        uint32_t count = nsw;
        HIP_TRY(hipEventRecord(ev[10], s));
        for (int pass = 0; pass < 3 && count; pass++) {
            const Caps& cap = SAMPLE_CAPS[pass];
            Layout lay = make_layout(cap);
            bool lds = pass < 2;
            if (lds && lay.bytes + STATIC_LDS > lds_max) continue;
            uint32_t *next, *nnext;
            HIP_TRY(T.get(&next, count));
            HIP_TRY(T.get(&nnext, 1));
            HIP_TRY(hipMemsetAsync(nnext, 0, 4, s));
            if (lds) {
                k_sample<false><<<count, 256, lay.bytes, s>>>(A, L, P, cap, lay, M->keys, st, list, count, M->values, nullptr,
                                                       0, next, nnext, nsamp);
            } else {
                uint32_t chunk = std::min(count, SCRATCH_CHUNK);
                char* scratch;
                HIP_TRY(T.get(&scratch, (size_t)chunk * lay.bytes));
                for (uint32_t c0 = 0; c0 < count; c0 += chunk)
                    k_sample<true><<<std::min(chunk, count - c0), 256, 0, s>>>(A, L, P, cap, lay, M->keys, st, list + c0,
                                                                         std::min(chunk, count - c0), M->values,
                                                                         scratch, lay.bytes, next, nnext, nsamp);
            }
            if (pass == 0) HIP_TRY(hipEventRecord(ev[11], s));
            HIP_TRY(hipGetLastError());
            TRY(read_u32(d, nnext, count));
            if (pass == 0) M->overflow = count;
            list = next;
        }
        if (count) return fail(SMG_E_INTERNAL, std::to_string(count) + " sub-blocks exceed the largest candidate lists");
    }
    HIP_TRY(hipEventRecord(ev[2], s));

    // Crossing edges and vertex numbering (mesh/extract.rs:176-216, 606-627).
    uint32_t words = (3 * b * b * b + 63) / 64;
    M->words = words;
    unsigned long long* mask;
    uint32_t *wpre, *bcount, *scount, *vbase;
    HIP_TRY(T.get(&mask, (size_t)nb * words));
    HIP_TRY(T.get(&wpre, (size_t)nb * words));
    HIP_TRY(T.get(&bcount, nb));
    HIP_TRY(T.get(&scount, nsub));
    HIP_TRY(T.get(&vbase, nb));
    if (nb) k_edges<<<nb, 256, 0, s>>>(P, M->values, nb, words, mask, wpre, bcount, scount);
    uint32_t* nv_d;
    TRY(scan_u32(s, T, bcount, vbase, nb, &nv_d));
    uint32_t *rflag, *roff, *rwork;
    HIP_TRY(T.get(&rflag, nsub));
    HIP_TRY(T.get(&roff, nsub));
    HIP_TRY(T.get(&rwork, nsub));
    if (nsub) k_flag_nonzero<<<blocks_for(nsub, 256), 256, 0, s>>>(scount, nsub, rflag);
    uint32_t* nrw_d;
    TRY(scan_u32(s, T, rflag, roff, nsub, &nrw_d));
    if (nsub) k_compact_index<<<blocks_for(nsub, 256), 256, 0, s>>>(rflag, roff, nsub, rwork);
    uint32_t nverts, nrw;
    TRY(read_u32(d, nv_d, nverts));
    TRY(read_u32(d, nrw_d, nrw));
    M->vertices = nverts;
    HIP_TRY(hipEventRecord(ev[3], s));

    // Exact refinement and attributes (mesh/extract.rs:425-597).
    size_t nvv = std::max<size_t>(nverts, 1);
    HIP_TRY(hipMalloc(&M->pos, 3 * nvv * sizeof(float)));
    HIP_TRY(hipMalloc(&M->nrm, 3 * nvv * sizeof(float)));
    HIP_TRY(hipMalloc(&M->atom, nvv * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&M->patch, nvv));
    HIP_TRY(hipMalloc(&M->t, nvv * sizeof(double)));
    HIP_TRY(hipMalloc(&M->vevals, nvv * sizeof(uint32_t)));
    VOut vo{M->pos, M->nrm, M->atom, M->patch, M->t, M->vevals};
    unsigned long long* evd;
    HIP_TRY(T.get(&evd, 1));
    HIP_TRY(hipMemsetAsync(evd, 0, 8, s));
    {
        const uint32_t* list = rwork;
        uint32_t count = nrw, sub_edges = 3 * P.sub * P.sub * P.sub;
        HIP_TRY(hipEventRecord(ev[12], s));
        for (int pass = 0; pass < 3 && count; pass++) {
            RCaps cap = REFINE_CAPS[pass];
            bool lds = pass < 2;
            cap.edges = lds ? std::min(sub_edges, LDS_EDGES) : sub_edges;
            RLayout lay = make_rlayout(cap);
            if (lds && lay.bytes + STATIC_LDS > lds_max) continue;
            uint32_t *next, *nnext;
            HIP_TRY(T.get(&next, count));
            HIP_TRY(T.get(&nnext, 1));
            HIP_TRY(hipMemsetAsync(nnext, 0, 4, s));
            if (lds) {
                k_refine<false><<<count, 256, lay.bytes, s>>>(A, P, cap, lay, M->keys, M->values, list, count, mask, wpre,
                                                       words, vbase, vo, nullptr, 0, next, nnext, evd);
            } else {
                uint32_t chunk = std::min(count, SCRATCH_CHUNK);
                char* scratch;
                HIP_TRY(T.get(&scratch, (size_t)chunk * lay.bytes));
                for (uint32_t c0 = 0; c0 < count; c0 += chunk)
                    k_refine<true><<<std::min(chunk, count - c0), 256, 0, s>>>(A, P, cap, lay, M->keys, M->values, list + c0,
                                                                         std::min(chunk, count - c0), mask, wpre,
                                                                         words, vbase, vo, scratch, lay.bytes, next,
                                                                         nnext, evd);
            }
            if (pass == 0) HIP_TRY(hipEventRecord(ev[13], s));
            HIP_TRY(hipGetLastError());
            TRY(read_u32(d, nnext, count));
            list = next;
        }
        if (count) return fail(SMG_E_INTERNAL, std::to_string(count) + " sub-blocks exceed the largest refinement lists");
    }
    HIP_TRY(hipEventRecord(ev[4], s));

    // Cell pass and assembly (mesh/extract.rs:226-265, 606-691).
    uint32_t *tcount, *tbase, *err;
    HIP_TRY(T.get(&tcount, nb));
    HIP_TRY(T.get(&tbase, nb));
    HIP_TRY(T.get(&err, 1));
    HIP_TRY(hipMemsetAsync(err, 0, 4, s));
    if (nb) k_cell_count<<<nb, 256, 0, s>>>(P, M->values, nb, tcount);
    uint32_t* nt_d;
    TRY(scan_u32(s, T, tcount, tbase, nb, &nt_d));
    uint32_t ntris;
    TRY(read_u32(d, nt_d, ntris));
    M->triangles = ntris;
    HIP_TRY(hipMalloc(&M->tris, std::max<size_t>(ntris, 1) * 3 * sizeof(uint32_t)));
    if (nb)
        k_cell_emit<<<nb, 256, 0, s>>>(P, M->values, M->keys, nb, tbase, dense, vbase, mask, wpre, words, M->tris, err);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[5], s));
    uint32_t bad;
    TRY(read_u32(d, err, bad));
    if (bad) return fail(SMG_E_INTERNAL, "a triangle references an edge without a vertex");
    unsigned long long evh = 0, sh = 0;
    HIP_TRY(hipMemcpy(&evh, evd, 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&sh, nsamp, 8, hipMemcpyDeviceToHost));
    M->evals = evh;
    M->samples = sh;
    if (times) {
        float ms[5];
        for (int q = 0; q < 5; q++) HIP_TRY(hipEventElapsedTime(&ms[q], ev[q], ev[q + 1]));
        times->classify_ms = ms[0];
        times->sample_ms = ms[1];
        times->edges_ms = ms[2];
        times->refine_ms = ms[3];
        times->cells_ms = ms[4];
        HIP_TRY(hipEventElapsedTime(&ms[0], ev[0], ev[5]));
        times->total_ms = ms[0];
        float k[4];
        for (int q = 0; q < 4; q++) HIP_TRY(hipEventElapsedTime(&k[q], ev[6 + 2 * q], ev[7 + 2 * q]));
        times->k_classify_ms = k[0];
        times->k_substates_ms = k[1];
        times->k_sample_ms = k[2];
        times->k_refine_ms = k[3];
    }
    return SMG_OK;
}

extern "C" int smg_mesh_new(smg_resident* r, const smg_mesh_params* p, smg_mesh** out, smg_mesh_times* times) {
    *out = nullptr;
    if (!r || !p) return fail(SMG_E_INVALID, "null resident or parameters");
    smg_mesh* M = new smg_mesh();
    M->dev = r->dev;
    if (times) std::memset(times, 0, sizeof(*times));
    int rc = mesh_stage(r, p, M, times);
    if (rc == SMG_OK) {
        hipError_t e = hipStreamSynchronize(r->dev->stream);
        if (e != hipSuccess) rc = fail(SMG_E_CALL, std::string("mesh stage: ") + hipGetErrorString(e));
    }
    if (rc != SMG_OK) {
        (void)hipStreamSynchronize(r->dev->stream);
        delete M;
        return rc;
    }
    *out = M;
    return SMG_OK;
}

extern "C" void smg_mesh_free(smg_mesh* m) {
    if (!m) return;
    (void)hipSetDevice(m->dev->ordinal);
    delete m;
}

extern "C" int smg_mesh_info_get(const smg_mesh* m, smg_mesh_info* info) {
    if (!m || !info) return fail(SMG_E_INVALID, "null mesh");
    std::memset(info, 0, sizeof(*info));
    info->touched_bricks = m->touched;
    info->surface_bricks = m->bricks;
    info->vertices = m->vertices;
    info->triangles = m->triangles;
    info->refine_evaluations = m->evals;
    info->sampled_subblocks = m->sampled;
    info->overflow_subblocks = m->overflow;
    info->evaluated_samples = m->samples;
    for (int q = 0; q < 3; q++) info->frame[q] = m->P.frame[q];
    info->band = m->P.band;
    info->edge_clamp = m->P.t_min;
    return SMG_OK;
}

extern "C" int smg_mesh_copy(const smg_mesh* m, float* positions, float* normals, uint32_t* atoms, uint8_t* patches,
                             uint32_t* triangles, int32_t* bricks, float* values, double* t, uint32_t* evals) {
    if (!m) return fail(SMG_E_INVALID, "null mesh");
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    size_t nv = m->vertices, nt = m->triangles, nb = m->bricks;
    if (positions && nv) HIP_TRY(hipMemcpy(positions, m->pos, 3 * nv * 4, hipMemcpyDeviceToHost));
    if (normals && nv) HIP_TRY(hipMemcpy(normals, m->nrm, 3 * nv * 4, hipMemcpyDeviceToHost));
    if (atoms && nv) HIP_TRY(hipMemcpy(atoms, m->atom, nv * 4, hipMemcpyDeviceToHost));
    if (patches && nv) HIP_TRY(hipMemcpy(patches, m->patch, nv, hipMemcpyDeviceToHost));
    if (t && nv) HIP_TRY(hipMemcpy(t, m->t, nv * 8, hipMemcpyDeviceToHost));
    if (evals && nv) HIP_TRY(hipMemcpy(evals, m->vevals, nv * 4, hipMemcpyDeviceToHost));
    if (triangles && nt) HIP_TRY(hipMemcpy(triangles, m->tris, 3 * nt * 4, hipMemcpyDeviceToHost));
    if (bricks && nb) HIP_TRY(hipMemcpy(bricks, m->keys, 3 * nb * 4, hipMemcpyDeviceToHost));
    if (values && nb) HIP_TRY(hipMemcpy(values, m->values, (size_t)nb * m->P.spb * 4, hipMemcpyDeviceToHost));
    return SMG_OK;
}

// The generated Marching Cubes table (for the tests).
extern "C" int smg_mc_table(uint16_t offset[257], uint8_t* edges, size_t cap) {
    std::call_once(g_table_once, [] { make_mc_table(g_tri_offset, g_tri_edges); });
    std::memcpy(offset, g_tri_offset, sizeof(g_tri_offset));
    std::memcpy(edges, g_tri_edges.data(), std::min(cap, g_tri_edges.size()));
    return SMG_OK;
}
