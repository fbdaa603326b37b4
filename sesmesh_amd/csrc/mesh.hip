// The MI355X (gfx950) mesh stage of sesmesh behind the C ABI of
// include/sesmesh_gpu.h: device open and self-test, resident arrangement,
// and the whole mesh stage as a sequence of kernels on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sesmesh_gpu.h"
#include "dev_classify.hpp"
#include "dev_common.hpp"
#include "dev_sample.hpp"
#include "dev_refine.hpp"
#include "dev_cells.hpp"
#include "dev_uniform.hpp"
#include "dev_fast.hpp"
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

// Surface bricks: keys (molecule lattice) in order, their molecules, and
// each molecule's dense index grid.
__global__ void k_compact_bricks(const Mol* __restrict__ mols, const uint64_t* wofs, uint32_t nm,
                                 const uint32_t* touched, const uint32_t* keep, const uint32_t* off, uint32_t n,
                                 int32_t* keys, uint32_t* bmol, uint32_t* dense) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    uint32_t at = off[i], bit = touched[i];
    uint32_t m = find_mol(wofs, nm, (uint64_t)(bit >> 5));
    uint64_t lin = bit - 32 * wofs[m];
    int32_t key[3];
    key_of_lin(mols[m].P, lin, key);
    keys[3 * at] = key[0];
    keys[3 * at + 1] = key[1];
    keys[3 * at + 2] = key[2];
    bmol[at] = m;
    dense[mols[m].dense_off + lin] = at;
}

// First touched brick and first surface brick of every molecule
// (rows[m] for m <= nm; row nm holds the totals).
__global__ void k_mol_ranges(const uint64_t* wofs, uint32_t nm, const uint32_t* woff, uint64_t nwords,
                             uint32_t ntouched, const uint32_t* koff, uint32_t nb, uint32_t* tfirst,
                             uint32_t* fbrick) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > nm) return;
    uint32_t tf = (m < nm && wofs[m] < nwords) ? woff[wofs[m]] : ntouched;
    tfirst[m] = tf;
    fbrick[m] = tf < ntouched ? koff[tf] : nb;
}

// Where each molecule's sub-blocks to sample, vertices and triangles start.
__global__ void k_mol_summary(uint32_t nm, const uint32_t* fbrick, uint32_t nb, const uint32_t* soff, uint32_t nsw,
                              const uint32_t* vbase, uint32_t nverts, const uint32_t* tbase, uint32_t ntris,
                              uint32_t* sfirst, uint32_t* vfirst, uint32_t* trfirst) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > nm) return;
    uint32_t fb = fbrick[m];
    sfirst[m] = fb < nb ? soff[8 * fb] : nsw;
    vfirst[m] = fb < nb ? vbase[fb] : nverts;
    trfirst[m] = fb < nb ? tbase[fb] : ntris;
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
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

// Frame-local f32 copies (field/gpu.rs:253-318, kernel.rs:627-637, 717-746);
// one thread per index below max(spheres, circles, vertices) of each
// molecule (loc_off: first index of each molecule).
__global__ void k_localize(const Mol* __restrict__ mols, const uint64_t* loc_off, uint32_t nm) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= loc_off[nm]) return;
    uint32_t m = find_mol(loc_off, nm, g);
    const DArr& A = mols[m].A;
    const DLoc& L = mols[m].L;
    const DRun& P = mols[m].P;
    const double fx = (double)P.frame[0] * P.h, fy = (double)P.frame[1] * P.h, fz = (double)P.frame[2] * P.h;
    uint32_t i = (uint32_t)(g - loc_off[m]);
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

// The meshes of one run: one part per molecule, concatenated in the
// device arrays; part i owns bricks [bfirst[i], bfirst[i+1]), vertices
// [vfirst[i], vfirst[i+1]) and triangles [tfirst[i], tfirst[i+1]).
struct smg_mesh {
    smg_device* dev = nullptr;
    uint32_t spb = 0;  // samples per brick
    std::vector<smg_mesh_info> info;
    std::vector<uint32_t> bfirst, vfirst, tfirst;
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
namespace {
// Packs the arrangement's arrays into one host buffer (256-byte aligned
// pieces) so a resident costs one allocation and one copy.
struct Packer {
    std::vector<char> bytes;
    std::vector<std::pair<size_t, void*>> fix;  // offset, address of the pointer field
    template <class T>
    void add(const T* src, size_t count, gptr<const T>* field) {
        size_t at = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(at + std::max<size_t>(count, 1) * sizeof(T));
        if (count) std::memcpy(bytes.data() + at, src, count * sizeof(T));
        fix.push_back({at, (void*)field});
    }
    void grid(const smg_grid_ref& g, DGrid& d) {
        size_t cells = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
        add(g.starts, cells + 1, &d.starts);
        add(g.items, g.n_items, &d.items);
        d.ox = g.origin[0];
        d.oy = g.origin[1];
        d.oz = g.origin[2];
        d.cell = g.cell;
        d.dx = g.dims[0];
        d.dy = g.dims[1];
        d.dz = g.dims[2];
    }
};
}  // namespace

extern "C" int smg_resident_new(smg_device* dev, const smg_arrangement_ref* a, smg_resident** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!dev || !a) return fail(SMG_E_INVALID, "null device or arrangement");
    if (a->n_spheres >= (1u << 30) || a->n_circles >= (1u << 30) || a->n_vertices >= (1u << 30))
        return fail(SMG_E_LENGTH, "arrangement too large for 30-bit candidate indices");
    if (!(a->probe > 0.0)) return fail(SMG_E_UNSUPPORTED, "the device path computes the SES (probe > 0)");
    HIP_TRY(hipSetDevice(dev->ordinal));
    size_t n = a->n_spheres, nc = a->n_circles, nv = a->n_vertices;
    std::unique_ptr<smg_resident> r(new smg_resident());
    r->dev = dev;
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
    std::vector<double> cf[10];
    std::vector<uint32_t> ci(nc), cj(nc);
    for (int q = 0; q < 10; q++) cf[q].resize(nc);
    for (size_t c = 0; c < nc; c++) {
        for (int q = 0; q < 10; q++) cf[q][c] = a->circle[10 * c + q];
        ci[c] = a->circle_ij[2 * c];
        cj[c] = a->circle_ij[2 * c + 1];
    }
    std::vector<double> vxyz[3];
    for (int q = 0; q < 3; q++) vxyz[q].resize(nv);
    for (size_t v = 0; v < nv; v++)
        for (int q = 0; q < 3; q++) vxyz[q][v] = a->vertex[3 * v + q];
    DArr& A = r->A;
    Packer pk;
    pk.add(x.data(), n, &A.cx);
    pk.add(y.data(), n, &A.cy);
    pk.add(z.data(), n, &A.cz);
    pk.add(a->expanded, n, &A.R);
    pk.add(a->radius, n, &A.r);
    pk.add(a->status, n, &A.status);
    pk.add(a->orig, n, &A.orig);
    pk.add(a->nb_off, n + 1, &A.nb_off);
    pk.add(a->nb_items, (size_t)a->nb_off[n], &A.nb_items);
    gptr<const double>* cfield[10] = {&A.mx, &A.my, &A.mz, &A.ax, &A.ay, &A.az, &A.e1x, &A.e1y, &A.e1z, &A.crho};
    for (int q = 0; q < 10; q++) pk.add(cf[q].data(), nc, cfield[q]);
    pk.add(ci.data(), nc, &A.ci);
    pk.add(cj.data(), nc, &A.cj);
    pk.add(a->cb_off, nc + 1, &A.cb_off);
    pk.add(a->cb_items, (size_t)a->cb_off[nc], &A.cb_items);
    pk.add(vxyz[0].data(), nv, &A.vx);
    pk.add(vxyz[1].data(), nv, &A.vy);
    pk.add(vxyz[2].data(), nv, &A.vz);
    pk.add(a->vertex_atoms, 3 * nv, &A.vatoms);
    pk.grid(a->sphere_grid, A.gs);
    pk.grid(a->circle_grid, A.gc);
    pk.grid(a->vertex_grid, A.gv);
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, pk.bytes.size()));
    r->buf.ptrs.push_back(p);
    r->buf.bytes = pk.bytes.size();
    HIP_TRY(hipMemcpy(p, pk.bytes.data(), pk.bytes.size(), hipMemcpyHostToDevice));
    for (auto& f : pk.fix) {
        uintptr_t v = (uintptr_t)p + f.first;
        std::memcpy(f.second, &v, sizeof(v));
    }
    A.n_spheres = (uint32_t)n;
    A.n_circles = (uint32_t)nc;
    A.n_vertices = (uint32_t)nv;
    A.probe = a->probe;
    A.max_expanded = a->max_expanded;
    A.rho_max = a->rho_max;
    *out = r.release();
    return SMG_OK;
}

extern "C" void smg_resident_free(smg_resident* r) {
    if (!r) return;
    (void)hipSetDevice(r->dev->ordinal);
    r->buf.release();
    delete r;
}

extern "C" size_t smg_resident_bytes(const smg_resident* r) { return r ? r->buf.bytes : 0; }

// ------------------------------------------------------------ field values
namespace {
// SesField::clamped_value (field/point.rs:759-765) at points[w], one
// wavefront per point: probe - ses_eval(x, probe + band).distance, where
// the distance is 0 outside the SAS (inside_sas, point.rs:372-381).
__global__ void k_field_clamped(DArr A, const double* points, uint32_t n, double band, double* out) {
    uint32_t w = blockIdx.x;
    if (w >= n) return;
    double x = points[3 * w], y = points[3 * w + 1], z = points[3 * w + 2];
    bool inside = false;
    for_near(A.gs, x, y, z, A.max_expanded, lane_id(), WAVE, [&](uint32_t k) {
        if (!inside && A.status[k] != S_CONTAINED && ddist(x, y, z, A.cx[k], A.cy[k], A.cz[k]) < A.R[k]) inside = true;
    });
    double dist = wave_any(inside) ? ses_distance_wave(A, x, y, z, A.probe + band) : 0.0;
    if (lane_id() == 0) out[w] = A.probe - dist;
}
}  // namespace

extern "C" int smg_field_clamped(smg_resident* r, const double* points, size_t n, double band, double* values) {
    if (!r || (n && (!points || !values))) return fail(SMG_E_INVALID, "null resident, points or values");
    if (n >= (1u << 31)) return fail(SMG_E_LENGTH, "too many points");
    if (!(band >= 0.0)) return fail(SMG_E_INVALID, "band must be non-negative");
    if (n == 0) return SMG_OK;
    smg_device* d = r->dev;
    HIP_TRY(hipSetDevice(d->ordinal));
    hipStream_t s = d->stream;
    double *pd, *vd;
    HIP_TRY(hipMallocAsync((void**)&pd, n * 24, s));
    HIP_TRY(hipMallocAsync((void**)&vd, n * 8, s));
    HIP_TRY(hipMemcpyAsync(pd, points, n * 24, hipMemcpyHostToDevice, s));
    k_field_clamped<<<(uint32_t)n, WAVE, 0, s>>>(r->A, pd, (uint32_t)n, band, vd);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(values, vd, n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipFreeAsync(pd, s));
    HIP_TRY(hipFreeAsync(vd, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SMG_OK;
}

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

// Exclusive scan of n u32 values into `out`; *total_dev receives the device
// address of the total.
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

// List capacities, tried in turn: LDS variants sized for common sub-blocks
// (several workgroups per CU) and for those of coarse lattices (up to the
// 160 KB of LDS one gfx950 workgroup may hold), then global scratch for the
// rest. Each pass takes the previous pass's overflow list.
constexpr int TIERS = 4;  // the last one in global scratch
constexpr Caps SAMPLE_CAPS[TIERS] = {{128, 64, 256, 64, 256},
                                     {256, 128, 384, 128, 512},
                                     {640, 256, 768, 256, 2048},
                                     {1024, 512, 4096, 2048, 4096}};
constexpr RCaps REFINE_CAPS[TIERS] = {
    {128, 64, 256, 256, 0}, {256, 128, 384, 512, 0}, {640, 256, 704, 1024, 0}, {1024, 512, 4096, 4096, 0}};
// Sub-block classification: one LDS tier; bricks beyond it take one
// wavefront per ball, which beats larger tiers at one or two workgroups per CU.
constexpr RCaps UNIFORM_CAPS[1] = {{192, 128, 256, 512, 0, 1}};
constexpr uint32_t LDS_EDGES = 512;      // crossing-edge buffer of the LDS passes
constexpr uint32_t SCRATCH_CHUNK = 256;   // workgroups per scratch launch
constexpr uint32_t STATIC_LDS = 6144;     // static LDS of k_sample / k_refine, bound

// Uniform-sign classification of `n` units (touched bricks or surface
// bricks): shared-gather workgroups with the LDS capacities in turn, then
// one wavefront per ball for units whose lists still overflow. Records
// `done` after the first pass (the kernel's time).
template <uint32_t MODE>
int run_uniform(smg_device* d, Temps& T, const Mol* mols, const uint64_t* wofs, uint32_t nm, const uint32_t* touched,
                const uint32_t* bmol, const int32_t* keys, uint32_t n, uint32_t* keep, uint8_t* st, hipEvent_t done) {
    hipStream_t s = d->stream;
    const uint32_t* list = nullptr;
    uint32_t count = n;
    for (int pass = 0; pass < 1 && count; pass++) {
        const RCaps& cap = UNIFORM_CAPS[pass];
        RLayout lay = make_rlayout(cap);
        if (lay.bytes + STATIC_LDS > d->prop.sharedMemPerBlock) continue;
        uint32_t *next, *nnext;
        HIP_TRY(T.get(&next, count));
        HIP_TRY(T.get(&nnext, 1));
        HIP_TRY(hipMemsetAsync(nnext, 0, 4, s));
        k_uniform<MODE, false><<<count, 256, lay.bytes, s>>>(mols, wofs, nm, touched, bmol, keys, cap, lay, list, count,
                                                             keep, st, nullptr, 0, next, nnext);
        HIP_TRY(hipGetLastError());
        if (pass == 0) HIP_TRY(hipEventRecord(done, s));
        TRY(read_u32(d, nnext, count));
        list = next;
    }
    if (count) {
        if (!list) {  // no LDS pass fits: every unit
            uint32_t* all;
            HIP_TRY(T.get(&all, count));
            k_iota<<<blocks_for(count, 256), 256, 0, s>>>(all, count);
            list = all;
            HIP_TRY(hipEventRecord(done, s));
        }
        if (MODE == UNIFORM_TOUCHED)
            k_classify_bricks<<<count, WAVE, 0, s>>>(mols, wofs, nm, touched, list, count, keep);
        else
            k_sub_states<<<8 * count, WAVE, 0, s>>>(mols, bmol, keys, list, count, st);
        HIP_TRY(hipGetLastError());
    }
    if (n == 0) HIP_TRY(hipEventRecord(done, s));
    return SMG_OK;
}

// Engine::new's lattice for one molecule (field/kernel.rs:204-261) and a
// brick box holding every touched brick: the centres' bounding box grown by
// the largest expanded radius, with the operations k_mark_touched uses, so
// the box bounds hold for each sphere (rounding is monotonic). Lattice
// order inside any such box is the packed-key order.
int molecule_run(const smg_resident* r, const smg_mesh_params* prm, uint32_t b, DRun& P, bool& empty) {
    const DArr& A = r->A;
    double spacing = prm->spacing;
    if (!(std::isfinite(spacing) && spacing >= 0.01)) return fail(SMG_E_INVALID, "grid spacing must be at least 0.01 A");
    if (prm->brick != b) return fail(SMG_E_UNSUPPORTED, "one brick size per run");
    if (prm->refine != SMG_REFINE_EXACT && prm->refine != SMG_REFINE_FAST)
        return fail(SMG_E_INVALID, "unknown refinement mode " + std::to_string(prm->refine));
    double center[3] = {0, 0, 0}, half = 0.0;
    if (r->n_spheres) {
        double dd[3];
        for (int q = 0; q < 3; q++) {
            dd[q] = r->hi[q] - r->lo[q];
            center[q] = (r->lo[q] + r->hi[q]) * 0.5;
        }
        half = 0.5 * std::fmax(std::fmax(dd[0], dd[1]), dd[2]);
    }
    P = DRun{};
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
        float mf = (float)cm;
        uint32_t bits;
        std::memcpy(&bits, &mf, 4);
        bits += 1;
        float nx;
        std::memcpy(&nx, &bits, 4);
        P.t_min = std::fmin(std::fmax(4.0 * (double)(nx - mf) / spacing, 1e-4), 0.25);
    }
    empty = r->n_spheres == 0;
    if (empty) return SMG_OK;
    double bh = (double)b * spacing, rr = A.max_expanded + CLASSIFY_EPS;
    uint64_t cells = 1;
    for (int q = 0; q < 3; q++) {
        double flo = std::floor((r->lo[q] - rr) / bh), fhi = std::floor((r->hi[q] + rr) / bh);
        if (!(std::fabs(flo) < (double)(KEY_BIAS - 1) && std::fabs(fhi) < (double)(KEY_BIAS - 1)))
            return fail(SMG_E_INVALID, "coordinates too far from the origin for grid spacing " + std::to_string(spacing) + " A");
        P.blo[q] = (int64_t)flo;
        P.bdim[q] = (uint32_t)((int64_t)fhi - (int64_t)flo + 1);
        cells *= P.bdim[q];
    }
    if (cells >= (1ull << 31)) return fail(SMG_E_UNSUPPORTED, "brick box too large for the dense touched-brick map");
    return SMG_OK;
}

}  // namespace

// The mesh stage of `nm` molecules in one pass: every kernel runs once over
// the concatenated work of all molecules, each work item finding its
// molecule's arrangement, f32 copies and lattice in the molecule table.
static int mesh_many(smg_resident* const* R, const smg_mesh_params* prm, size_t nm, smg_mesh* M,
                     smg_mesh_times* times) {
    if (nm == 0) return fail(SMG_E_INVALID, "no molecules");
    if (nm >= (1u << 24)) return fail(SMG_E_LENGTH, "too many molecules in one run");
    smg_device* d = R[0]->dev;
    for (size_t m = 0; m < nm; m++)
        if (!R[m] || R[m]->dev != d) return fail(SMG_E_INVALID, "all molecules of a run must be resident on one device");
    hipStream_t s = d->stream;
    HIP_TRY(hipSetDevice(d->ordinal));
    const uint32_t b = prm[0].brick;
    if (b < 4 || b > 32 || b % 2) return fail(SMG_E_UNSUPPORTED, "the device path needs an even brick size in 4..=32");
    const uint32_t spb = (b + 1) * (b + 1) * (b + 1);
    M->spb = spb;
    const bool fast = prm[0].refine == SMG_REFINE_FAST;
    for (size_t m = 0; m < nm; m++)
        if ((prm[m].refine == SMG_REFINE_FAST) != fast) return fail(SMG_E_UNSUPPORTED, "one refinement mode per run");
    // Molecule table, offsets of the concatenated per-molecule arrays.
    std::vector<Mol> H(nm);
    std::vector<uint64_t> sph_off(nm + 1, 0), loc_off(nm + 1, 0), wofs(nm + 1, 0);
    uint64_t dense_total = 0;
    for (size_t m = 0; m < nm; m++) {
        Mol& X = H[m];
        X.A = R[m]->A;
        bool empty = false;
        TRY(molecule_run(R[m], &prm[m], b, X.P, empty));
        uint64_t cells = empty ? 0 : (uint64_t)X.P.bdim[0] * X.P.bdim[1] * X.P.bdim[2];
        uint64_t nmax = std::max<uint64_t>({X.A.n_spheres, X.A.n_circles, X.A.n_vertices});
        sph_off[m + 1] = sph_off[m] + X.A.n_spheres;
        loc_off[m + 1] = loc_off[m] + nmax;
        wofs[m + 1] = wofs[m] + (cells + 31) / 32;
        X.word_off = wofs[m];
        X.dense_off = dense_total;
        dense_total += cells;
    }
    const uint64_t nwords = wofs[nm];
    if (nwords >= (1ull << 27)) return fail(SMG_E_UNSUPPORTED, "brick boxes of the run too large for 32-bit bit indices");
    Temps T(s);
    hipEvent_t* ev = d->ev;
    HIP_TRY(hipEventRecord(ev[0], s));
    float* lbuf;
    HIP_TRY(T.get(&lbuf, 19 * loc_off[nm]));
    for (size_t m = 0; m < nm; m++) {
        DLoc& L = H[m].L;
        size_t nmax = loc_off[m + 1] - loc_off[m];
        float* base = lbuf + 19 * loc_off[m];
        gptr<float>* fields[19] = {&L.px, &L.py, &L.pz, &L.pr, &L.pr2, &L.pr2tol, &L.mx, &L.my, &L.mz, &L.ux,
                                   &L.uy, &L.uz, &L.ex, &L.ey, &L.ez, &L.rho,   &L.vx, &L.vy, &L.vz};
        for (int q = 0; q < 19; q++) *fields[q] = to_global(base + q * nmax);
    }
    Mol* mols;
    uint64_t *sph_d, *loc_d, *wofs_d;
    HIP_TRY(T.get(&mols, nm));
    HIP_TRY(T.get(&sph_d, nm + 1));
    HIP_TRY(T.get(&loc_d, nm + 1));
    HIP_TRY(T.get(&wofs_d, nm + 1));
    HIP_TRY(hipMemcpyAsync(mols, H.data(), nm * sizeof(Mol), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sph_d, sph_off.data(), (nm + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(loc_d, loc_off.data(), (nm + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(wofs_d, wofs.data(), (nm + 1) * 8, hipMemcpyHostToDevice, s));
    const uint32_t nmu = (uint32_t)nm;
    if (loc_off[nm]) k_localize<<<blocks_for(loc_off[nm], 256), 256, 0, s>>>(mols, loc_d, nmu);
    HIP_TRY(hipGetLastError());

    // Touched bricks (lattice.rs:105-163) in per-molecule dense bitmaps.
    uint32_t *bits, *wcount, *woff, *dense;
    HIP_TRY(T.get(&bits, nwords));
    HIP_TRY(T.get(&wcount, nwords));
    HIP_TRY(T.get(&woff, nwords));
    HIP_TRY(T.get(&dense, dense_total));
    if (nwords) HIP_TRY(hipMemsetAsync(bits, 0, nwords * 4, s));
    if (sph_off[nm] && nwords) k_mark_touched<<<blocks_for(sph_off[nm], 128), 128, 0, s>>>(mols, sph_d, nmu, bits);
    if (nwords) k_popcount<<<blocks_for(nwords, 256), 256, 0, s>>>(bits, wcount, nwords);
    uint32_t* ntouched_d;
    TRY(scan_u32(s, T, wcount, woff, nwords, &ntouched_d));
    uint32_t ntouched;
    TRY(read_u32(d, ntouched_d, ntouched));
    uint32_t *touched, *keep, *koff;
    HIP_TRY(T.get(&touched, ntouched));
    HIP_TRY(T.get(&keep, ntouched));
    HIP_TRY(T.get(&koff, ntouched));
    if (nwords) k_bits_to_list<<<blocks_for(nwords, 256), 256, 0, s>>>(bits, woff, nwords, touched);
    // Brick classification (lattice.rs:168-183), one wavefront per brick.
    HIP_TRY(hipEventRecord(ev[6], s));
    if (ntouched) k_classify_bricks<<<ntouched, WAVE, 0, s>>>(mols, wofs_d, nmu, touched, nullptr, ntouched, keep);
    HIP_TRY(hipEventRecord(ev[7], s));
    uint32_t* nb_d;
    TRY(scan_u32(s, T, keep, koff, ntouched, &nb_d));
    uint32_t nb;
    TRY(read_u32(d, nb_d, nb));
    HIP_TRY(hipMalloc(&M->keys, std::max<size_t>(nb, 1) * 3 * sizeof(int32_t)));
    uint32_t *bmol, *tfirst, *fbrick;
    HIP_TRY(T.get(&bmol, nb));
    HIP_TRY(T.get(&tfirst, nm + 1));
    HIP_TRY(T.get(&fbrick, nm + 1));
    if (dense_total) k_fill_u32<<<blocks_for(dense_total, 256), 256, 0, s>>>(dense, 0xffffffffu, dense_total);
    if (ntouched)
        k_compact_bricks<<<blocks_for(ntouched, 256), 256, 0, s>>>(mols, wofs_d, nmu, touched, keep, koff, ntouched,
                                                                    M->keys, bmol, dense);
    k_mol_ranges<<<blocks_for(nm + 1, 256), 256, 0, s>>>(wofs_d, nmu, woff, nwords, ntouched, koff, nb, tfirst, fbrick);
    uint8_t* st;
    HIP_TRY(T.get(&st, 8 * (size_t)nb));
    HIP_TRY(hipEventRecord(ev[8], s));
    TRY(run_uniform<UNIFORM_SUBBLOCKS>(d, T, mols, wofs_d, nmu, nullptr, bmol, M->keys, nb, nullptr, st, ev[9]));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[1], s));

    // Samples (field/kernel.rs:415-449).
    HIP_TRY(hipMalloc(&M->values, std::max<size_t>((size_t)nb * spb, 1) * sizeof(float)));
    if (nb) k_assign_uniform<<<nb, 256, 0, s>>>(mols, bmol, st, nb, M->values);
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
    unsigned long long *nsamp, *evd;
    uint32_t* ovcount;
    HIP_TRY(T.get(&nsamp, nm));
    HIP_TRY(T.get(&evd, nm));
    HIP_TRY(T.get(&ovcount, 1));
    HIP_TRY(hipMemsetAsync(nsamp, 0, nm * 8, s));
    HIP_TRY(hipMemsetAsync(evd, 0, nm * 8, s));
    const size_t lds_max = d->prop.sharedMemPerBlock;
    uint32_t overflow0 = 0;
    {
        const uint32_t* list = swork;
        uint32_t count = nsw;
        HIP_TRY(hipEventRecord(ev[10], s));
        for (int pass = 0; pass < TIERS && count; pass++) {
            const Caps& cap = SAMPLE_CAPS[pass];
            Layout lay = make_layout(cap);
            bool lds = pass < TIERS - 1;
            if (lds && lay.bytes + STATIC_LDS > lds_max) continue;
            uint32_t *next, *nnext;
            HIP_TRY(T.get(&next, count));
            HIP_TRY(T.get(&nnext, 1));
            HIP_TRY(hipMemsetAsync(nnext, 0, 4, s));
            if (lds) {
                k_sample<false><<<count, 256, lay.bytes, s>>>(mols, bmol, cap, lay, M->keys, st, list, count,
                                                              M->values, nullptr, 0, next, nnext, nsamp);
            } else {
                uint32_t chunk = std::min(count, SCRATCH_CHUNK);
                char* scratch;
                HIP_TRY(T.get(&scratch, (size_t)chunk * lay.bytes));
                for (uint32_t c0 = 0; c0 < count; c0 += chunk)
                    k_sample<true><<<std::min(chunk, count - c0), 256, 0, s>>>(
                        mols, bmol, cap, lay, M->keys, st, list + c0, std::min(chunk, count - c0), M->values, scratch,
                        lay.bytes, next, nnext, nsamp);
            }
            if (pass == 0) HIP_TRY(hipEventRecord(ev[11], s));
            HIP_TRY(hipGetLastError());
            TRY(read_u32(d, nnext, count));
            if (pass == 0) overflow0 = count;
            list = next;
        }
        if (count) return fail(SMG_E_INTERNAL, std::to_string(count) + " sub-blocks exceed the largest candidate lists");
    }
    HIP_TRY(hipEventRecord(ev[2], s));

    // Crossing edges and vertex numbering (mesh/extract.rs:176-216, 606-627).
    uint32_t words = (3 * b * b * b + 63) / 64;
    unsigned long long* mask;
    uint32_t *wpre, *bcount, *scount, *vbase;
    HIP_TRY(T.get(&mask, (size_t)nb * words));
    HIP_TRY(T.get(&wpre, (size_t)nb * words));
    HIP_TRY(T.get(&bcount, nb));
    HIP_TRY(T.get(&scount, nsub));
    HIP_TRY(T.get(&vbase, nb + 1));
    if (nb) k_edges<<<nb, 256, 0, s>>>(b, M->values, nb, words, mask, wpre, bcount, scount);
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
    {
        const uint32_t* list = rwork;
        uint32_t count = nrw, sub_edges = 3 * (b / 2) * (b / 2) * (b / 2);
        HIP_TRY(hipEventRecord(ev[12], s));
        for (int pass = 0; pass < TIERS && count; pass++) {
            RCaps cap = REFINE_CAPS[pass];
            bool lds = pass < TIERS - 1;
            cap.edges = lds ? std::min(sub_edges, LDS_EDGES) : sub_edges;
            RLayout lay = make_rlayout(cap);
            FLayout flay = make_flayout(cap, lay.bytes);
            const uint32_t bytes = fast ? flay.bytes : lay.bytes;
            if (lds && bytes + STATIC_LDS > lds_max) continue;
            uint32_t *next, *nnext;
            HIP_TRY(T.get(&next, count));
            HIP_TRY(T.get(&nnext, 1));
            HIP_TRY(hipMemsetAsync(nnext, 0, 4, s));
            if (lds && fast) {
                k_refine_fast<false><<<count, 256, bytes, s>>>(mols, bmol, cap, lay, flay, M->keys, M->values, list,
                                                               count, mask, wpre, words, vbase, vo, nullptr, 0, next,
                                                               nnext, evd);
            } else if (lds) {
                k_refine<false><<<count, 256, lay.bytes, s>>>(mols, bmol, cap, lay, M->keys, M->values, list, count, mask,
                                                              wpre, words, vbase, vo, nullptr, 0, next, nnext, evd);
            } else {
                uint32_t chunk = std::min(count, SCRATCH_CHUNK);
                char* scratch;
                HIP_TRY(T.get(&scratch, (size_t)chunk * bytes));
                for (uint32_t c0 = 0; c0 < count; c0 += chunk) {
                    uint32_t cnt = std::min(chunk, count - c0);
                    if (fast)
                        k_refine_fast<true><<<cnt, 256, 0, s>>>(mols, bmol, cap, lay, flay, M->keys, M->values,
                                                                 list + c0, cnt, mask, wpre, words, vbase, vo, scratch,
                                                                 bytes, next, nnext, evd);
                    else
                        k_refine<true><<<cnt, 256, 0, s>>>(mols, bmol, cap, lay, M->keys, M->values, list + c0, cnt,
                                                            mask, wpre, words, vbase, vo, scratch, lay.bytes, next,
                                                            nnext, evd);
                }
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
    HIP_TRY(T.get(&tbase, nb + 1));
    HIP_TRY(T.get(&err, 1));
    HIP_TRY(hipMemsetAsync(err, 0, 4, s));
    if (nb) k_cell_count<<<nb, 256, 0, s>>>(b, M->values, nb, tcount);
    uint32_t* nt_d;
    TRY(scan_u32(s, T, tcount, tbase, nb, &nt_d));
    uint32_t ntris;
    TRY(read_u32(d, nt_d, ntris));
    HIP_TRY(hipMalloc(&M->tris, std::max<size_t>(ntris, 1) * 3 * sizeof(uint32_t)));
    if (nb)
        k_cell_emit<<<nb, 256, 0, s>>>(mols, bmol, fbrick, M->values, M->keys, nb, tbase, dense, vbase, mask, wpre,
                                       words, M->tris, err);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[5], s));
    // Per-molecule ranges.
    uint32_t *sfirst, *vfirst, *trfirst;
    HIP_TRY(T.get(&sfirst, nm + 1));
    HIP_TRY(T.get(&vfirst, nm + 1));
    HIP_TRY(T.get(&trfirst, nm + 1));
    k_mol_summary<<<blocks_for(nm + 1, 256), 256, 0, s>>>(nmu, fbrick, nb, soff, nsw, vbase, nverts, tbase, ntris,
                                                          sfirst, vfirst, trfirst);
    HIP_TRY(hipGetLastError());
    std::vector<uint32_t> h_tf(nm + 1), h_fb(nm + 1), h_sf(nm + 1), h_vf(nm + 1), h_trf(nm + 1);
    std::vector<unsigned long long> h_ns(nm), h_ev(nm);
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(h_tf.data(), tfirst, (nm + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_fb.data(), fbrick, (nm + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_sf.data(), sfirst, (nm + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_vf.data(), vfirst, (nm + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_trf.data(), trfirst, (nm + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_ns.data(), nsamp, nm * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h_ev.data(), evd, nm * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&bad, err, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (bad) return fail(SMG_E_INTERNAL, "a triangle references an edge without a vertex");
    M->info.assign(nm, smg_mesh_info{});
    M->bfirst = h_fb;
    M->vfirst = h_vf;
    M->tfirst = h_trf;
    for (size_t m = 0; m < nm; m++) {
        smg_mesh_info& I = M->info[m];
        I.touched_bricks = h_tf[m + 1] - h_tf[m];
        I.surface_bricks = h_fb[m + 1] - h_fb[m];
        I.vertices = h_vf[m + 1] - h_vf[m];
        I.triangles = h_trf[m + 1] - h_trf[m];
        I.refine_evaluations = h_ev[m];
        I.sampled_subblocks = h_sf[m + 1] - h_sf[m];
        I.overflow_subblocks = nm == 1 ? overflow0 : 0;
        I.evaluated_samples = h_ns[m];
        for (int q = 0; q < 3; q++) I.frame[q] = H[m].P.frame[q];
        I.band = H[m].P.band;
        I.edge_clamp = H[m].P.t_min;
    }
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

extern "C" int smg_mesh_many(smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                             smg_mesh** out, smg_mesh_times* times) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!residents || !params || count == 0) return fail(SMG_E_INVALID, "no molecules");
    for (size_t m = 0; m < count; m++)
        if (!residents[m]) return fail(SMG_E_INVALID, "null resident");
    smg_mesh* M = new smg_mesh();
    M->dev = residents[0]->dev;
    if (times) std::memset(times, 0, sizeof(*times));
    int rc = mesh_many(residents, params, count, M, times);
    if (rc == SMG_OK) {
        hipError_t e = hipStreamSynchronize(M->dev->stream);
        if (e != hipSuccess) rc = fail(SMG_E_CALL, std::string("mesh stage: ") + hipGetErrorString(e));
    }
    if (rc != SMG_OK) {
        (void)hipStreamSynchronize(M->dev->stream);
        delete M;
        return rc;
    }
    *out = M;
    return SMG_OK;
}

extern "C" int smg_mesh_new(smg_resident* r, const smg_mesh_params* p, smg_mesh** out, smg_mesh_times* times) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!r || !p) return fail(SMG_E_INVALID, "null resident or parameters");
    return smg_mesh_many(&r, p, 1, out, times);
}

extern "C" void smg_mesh_free(smg_mesh* m) {
    if (!m) return;
    (void)hipSetDevice(m->dev->ordinal);
    delete m;
}

extern "C" size_t smg_mesh_count(const smg_mesh* m) { return m ? m->info.size() : 0; }

extern "C" int smg_mesh_part_info(const smg_mesh* m, size_t part, smg_mesh_info* info) {
    if (!m || !info) return fail(SMG_E_INVALID, "null mesh");
    if (part >= m->info.size()) return fail(SMG_E_INVALID, "no part " + std::to_string(part));
    *info = m->info[part];
    return SMG_OK;
}

extern "C" int smg_mesh_info_get(const smg_mesh* m, smg_mesh_info* info) { return smg_mesh_part_info(m, 0, info); }

extern "C" int smg_mesh_part_copy(const smg_mesh* m, size_t part, float* positions, float* normals, uint32_t* atoms,
                                  uint8_t* patches, uint32_t* triangles, int32_t* bricks, float* values, double* t,
                                  uint32_t* evals) {
    if (!m) return fail(SMG_E_INVALID, "null mesh");
    if (part >= m->info.size()) return fail(SMG_E_INVALID, "no part " + std::to_string(part));
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    size_t v0 = m->vfirst[part], nv = m->vfirst[part + 1] - v0;
    size_t t0 = m->tfirst[part], nt = m->tfirst[part + 1] - t0;
    size_t b0 = m->bfirst[part], nb = m->bfirst[part + 1] - b0;
    if (positions && nv) HIP_TRY(hipMemcpy(positions, m->pos + 3 * v0, 3 * nv * 4, hipMemcpyDeviceToHost));
    if (normals && nv) HIP_TRY(hipMemcpy(normals, m->nrm + 3 * v0, 3 * nv * 4, hipMemcpyDeviceToHost));
    if (atoms && nv) HIP_TRY(hipMemcpy(atoms, m->atom + v0, nv * 4, hipMemcpyDeviceToHost));
    if (patches && nv) HIP_TRY(hipMemcpy(patches, m->patch + v0, nv, hipMemcpyDeviceToHost));
    if (t && nv) HIP_TRY(hipMemcpy(t, m->t + v0, nv * 8, hipMemcpyDeviceToHost));
    if (evals && nv) HIP_TRY(hipMemcpy(evals, m->vevals + v0, nv * 4, hipMemcpyDeviceToHost));
    if (triangles && nt) HIP_TRY(hipMemcpy(triangles, m->tris + 3 * t0, 3 * nt * 4, hipMemcpyDeviceToHost));
    if (bricks && nb) HIP_TRY(hipMemcpy(bricks, m->keys + 3 * b0, 3 * nb * 4, hipMemcpyDeviceToHost));
    if (values && nb)
        HIP_TRY(hipMemcpy(values, m->values + b0 * m->spb, nb * m->spb * 4, hipMemcpyDeviceToHost));
    return SMG_OK;
}

extern "C" int smg_mesh_copy(const smg_mesh* m, float* positions, float* normals, uint32_t* atoms, uint8_t* patches,
                             uint32_t* triangles, int32_t* bricks, float* values, double* t, uint32_t* evals) {
    return smg_mesh_part_copy(m, 0, positions, normals, atoms, patches, triangles, bricks, values, t, evals);
}

// The generated Marching Cubes table (for the tests).
extern "C" int smg_mc_table(uint16_t offset[257], uint8_t* edges, size_t cap) {
    std::call_once(g_table_once, [] { make_mc_table(g_tri_offset, g_tri_edges); });
    std::memcpy(offset, g_tri_offset, sizeof(g_tri_offset));
    std::memcpy(edges, g_tri_edges.data(), std::min(cap, g_tri_edges.size()));
    return SMG_OK;
}
