// The MI355X (gfx950) mesh stage of sesmesh behind the C ABI of
// include/sesmesh_gpu.h: device open and self-test, resident arrangement,
// and the whole mesh stage as a sequence of kernels on one stream.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sesmesh_gpu.h"
#include "dev_classify.hpp"
#include "dev_common.hpp"
#include "dev_sample.hpp"
#include "dev_refine.hpp"
#include "dev_cells.hpp"
#include "dev_uniform.hpp"
#include "dev_fast.hpp"
#include "dev_mlp.hpp"
#include "dev_measure.hpp"
#include "dev_export.hpp"
#include "dev_arrange.hpp"
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

#define TRY(x)                  \
    do {                        \
        int rc_ = (x);          \
        if (rc_ != SMG_OK) {    \
            return rc_;         \
        }                       \
    } while (0)

extern "C" const char* smg_last_error(void) { return g_error.c_str(); }

// ------------------------------------------------------------------ scans
namespace {

constexpr uint32_t SCAN_BLOCK = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Counts live in device memory (a captured run never reads them on the
// host); kernels over a counted range stride over it with whatever grid
// they were given.
#define GRID_STRIDE(i, n) \
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)(n); i += (size_t)gridDim.x * blockDim.x)

// Exclusive scan of in[0 .. *n), tile pass: out[i] = sum(in[tile start..i)),
// sums[tile]. Tiles at or past the count do nothing.
__global__ void __launch_bounds__(256) k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t* sums,
                                                    const uint32_t* n_dev) {
    __shared__ uint32_t tmp[4];
    const size_t n = *n_dev;
    if ((size_t)blockIdx.x * SCAN_TILE >= n) return;
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

// In-place exclusive scan of the tile sums of the count *n by one block;
// the grand total goes to *total.
__global__ void __launch_bounds__(256) k_scan_sums(uint32_t* sums, const uint32_t* n_dev, uint32_t* total_out) {
    __shared__ uint32_t tmp[4];
    const size_t n = ((size_t)*n_dev + SCAN_TILE - 1) / SCAN_TILE;
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
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(uint32_t* out, const uint32_t* sums, const uint32_t* n_dev) {
    GRID_STRIDE(i, *n_dev) out[i] += sums[i / SCAN_TILE];
}

__global__ void k_popcount(const uint32_t* bits, uint32_t* cnt, const uint32_t* n) {
    GRID_STRIDE(i, *n) cnt[i] = (uint32_t)__popc(bits[i]);
}

// Set bits of the touched bitmap, in ascending order, as linear indices.
__global__ void k_bits_to_list(const uint32_t* bits, const uint32_t* off, const uint32_t* n, uint32_t* out) {
    GRID_STRIDE(i, *n) {
        uint32_t w = bits[i], at = off[i];
        while (w) {
            uint32_t bit = __ffs(w) - 1;
            out[at++] = (uint32_t)(i * 32 + bit);
            w &= w - 1;
        }
    }
}

// Surface bricks: keys (molecule lattice) in order, their molecules, and
// each molecule's dense index grid. Bricks at or past `cap` are dropped
// (the caller has clamped the count and flagged the run).
__global__ void k_compact_bricks(const Mol* __restrict__ mols, const uint64_t* wofs, uint32_t nm,
                                 const uint32_t* touched, const uint32_t* keep, const uint32_t* off,
                                 const uint32_t* n, uint32_t cap, int32_t* keys, uint32_t* bmol, uint32_t* dense) {
    GRID_STRIDE(i, *n) {
        if (!keep[i] || off[i] >= cap) continue;
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
}

// First touched brick and first surface brick of every molecule
// (rows[m] for m <= nm; row nm holds the totals).
__global__ void k_mol_ranges(const uint64_t* wofs, uint32_t nm, const uint32_t* woff, const uint32_t* nwords,
                             const uint32_t* ntouched, const uint32_t* koff, const uint32_t* nb, uint32_t* tfirst,
                             uint32_t* fbrick) {
    GRID_STRIDE(m, nm + 1) {
        uint32_t nt = *ntouched;
        uint32_t tf = (m < nm && wofs[m] < *nwords) ? woff[wofs[m]] : nt;
        tfirst[m] = tf;
        fbrick[m] = tf < nt ? min(koff[tf], *nb) : *nb;
    }
}

// Where each molecule's sub-blocks to sample, vertices and triangles start.
__global__ void k_mol_summary(uint32_t nm, const uint32_t* fbrick, const uint32_t* nb, const uint32_t* soff,
                              const uint32_t* nsw, const uint32_t* vbase, const uint32_t* nverts,
                              const uint32_t* tbase, const uint32_t* ntris, uint32_t* sfirst, uint32_t* vfirst,
                              uint32_t* trfirst) {
    GRID_STRIDE(m, nm + 1) {
        uint32_t fb = fbrick[m];
        bool in = fb < *nb;
        sfirst[m] = in ? soff[8 * fb] : *nsw;
        vfirst[m] = in ? min(vbase[fb], *nverts) : *nverts;
        trfirst[m] = in ? min(tbase[fb], *ntris) : *ntris;
    }
}

// p[i] = v for i below n, or below *n_dev when given.
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n, const uint32_t* n_dev) {
    GRID_STRIDE(i, n_dev ? (size_t)*n_dev : n) p[i] = v;
}

// *n = min(*n, cap), setting `bit` in *flags when it was larger; out2, when
// given, receives k * the clamped count.
__global__ void k_clamp(uint32_t* n, uint32_t cap, uint32_t* flags, uint32_t bit, uint32_t* out2, uint32_t k) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t v = *n;
    if (v > cap) {
        v = cap;
        *n = cap;
        atomicOr(flags, bit);
    }
    if (out2) *out2 = k * v;
}

// Flags of sub-blocks to sample (state NONE) or to refine (edges > 0).
__global__ void k_flag_states(const uint8_t* st, const uint32_t* n, uint32_t* flag) {
    GRID_STRIDE(i, *n) flag[i] = st[i] == ST_NONE ? 1u : 0u;
}
__global__ void k_flag_nonzero(const uint32_t* c, const uint32_t* n, uint32_t* flag) {
    GRID_STRIDE(i, *n) flag[i] = c[i] ? 1u : 0u;
}
__global__ void k_compact_index(const uint32_t* flag, const uint32_t* off, const uint32_t* n, uint32_t* out) {
    GRID_STRIDE(i, *n) if (flag[i]) out[off[i]] = (uint32_t)i;
}

// Frame-local f32 copies (field/gpu.rs:253-318, kernel.rs:627-637, 717-746);
// one thread per index below max(spheres, circles, vertices) of each
// molecule (loc_off: first index of each molecule).
__device__ inline void localize_one(const Mol* __restrict__ mols, const uint64_t* loc_off, uint32_t nm, uint64_t g) {
    uint32_t m = find_mol(loc_off, nm, g);
    const DArr& A = mols[m].A;
    const DLoc& L = mols[m].L;
    const DRun& P = mols[m].P;
    const double fx = (double)P.frame[0] * P.h, fy = (double)P.frame[1] * P.h, fz = (double)P.frame[2] * P.h;
    uint32_t i = (uint32_t)(g - loc_off[m]);
    if (i < A.n_spheres) {
        SphereRec S = load_rec(A.sph, i);
        store_f4(L.p4, i, (float)(S.x - fx), (float)(S.y - fy), (float)(S.z - fz), (float)S.R);
    }
    if (i < A.n_circles) {
        CircleRec C = load_rec(A.cir, i);
        store_f4(L.m4, i, (float)(C.mx - fx), (float)(C.my - fy), (float)(C.mz - fz), (float)C.rho);
        store_f4(L.u4, i, (float)C.ax, (float)C.ay, (float)C.az, 0.0f);
        store_f4(L.e4, i, (float)A.e1x[i], (float)A.e1y[i], (float)A.e1z[i], 0.0f);
    }
    if (i < A.n_vertices) {
        VertexRec V = load_rec(A.ver, i);
        store_f4(L.v4, i, (float)(V.x - fx), (float)(V.y - fy), (float)(V.z - fz), 0.0f);
    }
}

__global__ void k_localize(const Mol* __restrict__ mols, const uint64_t* loc_off, uint32_t nm) {
    GRID_STRIDE(g, loc_off[nm]) localize_one(mols, loc_off, nm, g);
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

// The scratch memory of smg_mesh_into: every intermediate buffer of its
// runs is carved from this array, so that entry point allocates nothing.
// Each device holds its own copy for the life of the process; host code
// uses only the device address hipGetSymbolAddress gives for it.
static __device__ __align__(256) char noalloc_scratch[SMG_NOALLOC_SCRATCH_BYTES];

struct smg_device;
static int noalloc_setup(smg_device* d);

// ------------------------------------------------------------------ objects
struct smg_device {
    int ordinal = 0;
    hipDeviceProp_t prop{};
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    uint32_t* host_u32 = nullptr;  // pinned staging for counts
    // Workgroups of 256 (or 64) threads resident on the whole device, per
    // kernel and dynamic LDS size: the grid of a persistent launch.
    std::map<std::pair<const void*, size_t>, uint32_t> resident;
    std::mutex mu;
    smg_tuning tune{0, 0, 64, 1, 0};  // launch shapes of the next runs (smg_tune)
    smg_model* model = nullptr;       // sizing model of runs with sizing 1
    // Static LDS (the __shared__ arrays) of the kernels whose dynamic LDS
    // depends on the brick or the list tier, read once (hipFuncGetAttributes).
    uint32_t static_lds[8] = {};
    bool static_ready = false;
    // Holders: the handle itself and every resident, mesh, plan and model
    // made on the device. smg_device_close drops the handle's reference;
    // the device is destroyed when the last holder is freed, so objects may
    // be freed in any order relative to the device.
    std::atomic<int> refs{1};
    // smg_mesh_into: this device's copy of the fixed scratch array (its
    // address from hipGetSymbolAddress at open), pinned images of a
    // one-molecule table and summary, and the lock that serialises its calls.
    char* scratch_ptr = nullptr;
    std::string scratch_error;
    char* noalloc_stage_h = nullptr;
    char* noalloc_sum_h = nullptr;
    std::mutex noalloc_mu;
    // Cancellation and progress of the device's runs (smg_device_set_control):
    // the caller's flag, polled by the host. Raising a run's cancellation
    // writes its generation into cancel_host (the ordinal's word, which the
    // work-item kernels poll) and copies it into g_cancel on cstream.
    smg_control ctl{};
    uint32_t* cancel_dev = nullptr;
    uint32_t* cancel_host = nullptr;
    hipStream_t cstream = nullptr;
    bool raised = false;
    uint32_t gen = 0;  // generation of the current run
};

static void device_destroy(smg_device* d);

static void device_retain(smg_device* d) {
    if (d) d->refs.fetch_add(1, std::memory_order_relaxed);
}
static void device_release(smg_device* d) {
    if (d && d->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) device_destroy(d);
}

// A counted reference to a device, the first member of every object made
// on one (so it is released after the object's own destructor has freed
// its device memory).
struct DevRef {
    smg_device* p = nullptr;
    DevRef() = default;
    DevRef(const DevRef& o) : p(o.p) { device_retain(p); }
    DevRef& operator=(const DevRef& o) { return *this = o.p; }
    DevRef& operator=(smg_device* d) {
        device_retain(d);
        device_release(p);
        p = d;
        return *this;
    }
    ~DevRef() { device_release(p); }
    operator smg_device*() const { return p; }
    smg_device* operator->() const { return p; }
};

// The sizing model: FP32 master parameters, Adam moments, BF16 copies of
// W1 and W2 for the matrix cores, and its labelled training set (features
// of touched bricks, occupancy, log candidate load).
struct smg_model {
    DevRef dev;
    uint64_t id = 0;  // unique per model of the process (plans recorded with it check it)
    float *P = nullptr, *Mo = nullptr, *Ve = nullptr;
    uint16_t *W1b = nullptr, *W2b = nullptr;
    uint64_t step = 0;
    float* X = nullptr;
    uint8_t* occ = nullptr;
    float* tgt = nullptr;
    size_t n = 0, cap = 0;
    // Placement: a larger tier runs a sub-block at lower occupancy, so a
    // prediction may only skip the first tier, and only for loads well
    // past it (measured: never slower than no sizing on 0.5-1.0 A).
    float scale = 0.3f;
    uint32_t max_tier = 1;
    // Ratios of the collected runs for capacity predictions, per spacing
    // bucket (a third of an octave from 1/8 A): vertices per surface brick
    // per b^2 (crossing edges grow with a brick's face area, and with the
    // surface's roughness on the lattice), triangles per vertex.
    static constexpr int SIZING_BUCKETS = 16;
    static int spacing_bucket(double h) {
        double f = std::floor(3.0 * std::log2(h / 0.125));
        return (int)std::min<double>(std::max<double>(f, 0.0), SIZING_BUCKETS - 1);
    }
    double sum_vertices[SIZING_BUCKETS] = {}, sum_brick_b2[SIZING_BUCKETS] = {}, sum_triangles[SIZING_BUCKETS] = {};
    ~smg_model() {
        void* ps[] = {P, Mo, Ve, W1b, W2b, X, occ, tgt};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
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
    DevRef dev;
    DArr A{};
    DevBuf buf;
    // Host copies needed to set up a run: centre bounding box.
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    size_t n_spheres = 0;
    bool on_device = false;  // arrangement built by smg_resident_build on the device
    double build_ms = 0.0;
};

// The meshes of one run: one part per molecule, concatenated in the
// device arrays; part i owns bricks [bfirst[i], bfirst[i+1]), vertices
// [vfirst[i], vfirst[i+1]) and triangles [tfirst[i], tfirst[i+1]).
struct smg_mesh {
    DevRef dev;
    uint32_t spb = 0;  // samples per brick
    std::vector<smg_mesh_info> info;
    std::vector<uint32_t> bfirst, vfirst, tfirst;
    // Device arrays, owned by the mesh unless it views a plan's buffers
    // (then `plan` holds the view; a view whose plan was freed is detached
    // and has no arrays).
    smg_plan* plan = nullptr;
    bool detached = false;
    int32_t* keys = nullptr;
    float* values = nullptr;
    float* pos = nullptr;
    float* nrm = nullptr;
    uint32_t* atom = nullptr;
    uint8_t* patch = nullptr;
    double* t = nullptr;
    uint32_t* vevals = nullptr;
    uint32_t* tris = nullptr;
    // The measure stage (smg_mesh_measure): per-part results, components of
    // all parts in order (part p's from cfirst_part[p]), vertex components
    // (device, owned by the mesh); triangles are then grouped by component.
    bool measured = false;
    std::vector<smg_measure_info> minfo;
    std::vector<smg_component> comps;
    std::vector<uint32_t> cfirst_part;
    uint32_t* vcomp = nullptr;
    ~smg_mesh();
};

// ------------------------------------------------------------------ device
static std::once_flag g_table_once;

// The cancellation word of each device ordinal: coherent pinned host memory
// that the device's work-item kernels poll (g_cancel_host holds its device
// address). Every handle opened on the ordinal shares it and it is never
// freed, so no handle's close can free memory another handle's kernels read;
// run generations are drawn from one counter per ordinal, so a raised
// generation names one run of one handle.
static std::mutex g_cancel_mu;
static uint32_t* g_cancel_word[64] = {};
static std::atomic<uint32_t> g_generation[64];
static uint32_t* cancel_word(int ordinal) {
    if (ordinal < 0 || ordinal >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(g_cancel_mu);
    if (!g_cancel_word[ordinal]) {
        uint32_t* w = nullptr;
        if (hipHostMalloc((void**)&w, 64 * sizeof(uint32_t), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
            return nullptr;
        w[0] = 0u;  // no generation is 0
        g_cancel_word[ordinal] = w;
    }
    return g_cancel_word[ordinal];
}
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
    d->cancel_host = cancel_word(ordinal);
    const uint32_t* cancel_host_dev = nullptr;
    void* cancel_host_sym = nullptr;
    if (!d->host_u32 || !d->cancel_host ||
        hipHostGetDevicePointer((void**)&cancel_host_dev, d->cancel_host, 0) != hipSuccess ||
        hipStreamCreateWithFlags(&d->cstream, hipStreamNonBlocking) != hipSuccess ||
        hipGetSymbolAddress((void**)&d->cancel_dev, HIP_SYMBOL(g_cancel)) != hipSuccess ||
        hipGetSymbolAddress(&cancel_host_sym, HIP_SYMBOL(g_cancel_host)) != hipSuccess ||
        hipMemcpy(cancel_host_sym, &cancel_host_dev, sizeof(cancel_host_dev), hipMemcpyHostToDevice) != hipSuccess) {
        smg_device_close(d);
        return fail(SMG_E_CALL, "cannot set up cancellation on the device");
    }
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
    rc = noalloc_setup(d);
    if (rc != SMG_OK) {
        smg_device_close(d);
        return rc;
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        std::snprintf(info->name, sizeof(info->name), "%s", d->prop.name);
        if (!info->name[0]) {  // some runtimes leave the marketing name empty
            char nm[256] = {};
            if (hipDeviceGetName(nm, sizeof(nm) - 1, ordinal) == hipSuccess && nm[0])
                std::snprintf(info->name, sizeof(info->name), "%s", nm);
            else
                std::snprintf(info->name, sizeof(info->name), "AMD %s, %d CUs, %.0f GiB", d->prop.gcnArchName,
                              d->prop.multiProcessorCount, (double)d->prop.totalGlobalMem / (1u << 30));
        }
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

extern "C" void smg_device_close(smg_device* d) { device_release(d); }

static void device_destroy(smg_device* d) {
    (void)hipSetDevice(d->ordinal);
    for (auto& e : d->ev)
        if (e) (void)hipEventDestroy(e);
    if (d->host_u32) (void)hipHostFree(d->host_u32);
    if (d->noalloc_stage_h) (void)hipHostFree(d->noalloc_stage_h);
    if (d->noalloc_sum_h) (void)hipHostFree(d->noalloc_sum_h);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    if (d->cstream) (void)hipStreamDestroy(d->cstream);
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
extern "C" int smg_arrangement_build_kind(const double* xyzr, size_t n, double probe, int kind, int threads,
                                          smg_arrangement** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!xyzr && n) return fail(SMG_E_INVALID, "no spheres");
    smg_arrangement* a = new smg_arrangement();
    std::string e = build_arrangement(xyzr, n, probe, kind, threads, a->a);
    if (!e.empty()) {
        delete a;
        return fail(SMG_E_INVALID, e);
    }
    *out = a;
    return SMG_OK;
}

extern "C" int smg_arrangement_build(const double* xyzr, size_t n, double probe, int threads, smg_arrangement** out) {
    return smg_arrangement_build_kind(xyzr, n, probe, SMG_KIND_SES, threads, out);
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
    r->kind = a.kind;
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
    if (a->kind < SMG_KIND_SES || a->kind > SMG_KIND_VDW)
        return fail(SMG_E_INVALID, "unknown surface kind " + std::to_string(a->kind));
    if (!(std::isfinite(a->probe) && a->probe >= 0.0)) return fail(SMG_E_INVALID, "probe radius must be finite and >= 0");
    if (a->kind == SMG_KIND_SES && !(a->probe > 0.0))
        return fail(SMG_E_INVALID, "an SES arrangement has probe > 0 (probe 0 is the vdW kind)");
    if (a->kind != SMG_KIND_SES && (a->n_circles || a->n_vertices))
        return fail(SMG_E_INVALID, "SAS and vdW arrangements have no circles or SAS vertices");
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
    std::vector<SphereRec> srec(n);
    for (size_t i = 0; i < n; i++) srec[i] = SphereRec{x[i], y[i], z[i], a->expanded[i]};
    std::vector<CircleRec> crec(nc);
    for (size_t c = 0; c < nc; c++)
        crec[c] = CircleRec{cf[0][c], cf[1][c], cf[2][c], cf[3][c], cf[4][c], cf[5][c], cf[9][c], 0.0};
    std::vector<VertexRec> vrec(nv);
    for (size_t v = 0; v < nv; v++) vrec[v] = VertexRec{vxyz[0][v], vxyz[1][v], vxyz[2][v], 0.0};
    std::vector<SphereThr> thr(n);
    for (size_t i = 0; i < n; i++)
        thr[i] = SphereThr{sqrt_threshold(a->expanded[i] - EPS), sqrt_threshold(a->expanded[i])};
    pk.add(srec.data(), n, &A.sph);
    pk.add(thr.data(), n, &A.thr);
    pk.add(crec.data(), nc, &A.cir);
    pk.add(vrec.data(), nv, &A.ver);
    pk.grid(a->sphere_grid, A.gs);
    pk.grid(a->circle_grid, A.gc);
    pk.grid(a->vertex_grid, A.gv);
    // Finer grids over the same points (0.4 of the cell edge; a third and a
    // quarter measured no faster) for the sub-block gathers, which may
    // visit items in any order.
    Grid fine[3];
    {
        const double fcell = std::fmax(0.8 * a->max_expanded, 1e-6);
        std::vector<double> cpts(3 * nc);
        for (size_t c = 0; c < nc; c++)
            for (int q = 0; q < 3; q++) cpts[3 * c + q] = cf[q][c];
        build_grid(a->center, n, fcell, fine[0]);
        build_grid(cpts.data(), nc, fcell, fine[1]);
        build_grid(a->vertex, nv, fcell, fine[2]);
        smg_grid_ref fr[3];
        for (int q = 0; q < 3; q++) grid_ref(fine[q], fr[q]);
        pk.grid(fr[0], A.fs);
        pk.grid(fr[1], A.fc);
        pk.grid(fr[2], A.fv);
    }
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
    A.kind = (uint32_t)a->kind;
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
// the distance is 0 outside the SAS (inside_sas, point.rs:372-381); for the
// SAS and vdW kinds sphere_eval(x, band) (point.rs:742-756), the least
// |x - c| - R below band.
__global__ void k_field_clamped(DArr A, const double* points, uint32_t n, double band, double* out) {
    for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
        double x = points[3 * (size_t)w], y = points[3 * (size_t)w + 1], z = points[3 * (size_t)w + 2];
        if (A.kind != K_SES) {
            double best = band;
            for_near(A.gs, x, y, z, band + A.max_expanded, lane_id(), WAVE, [&](uint32_t k) {
                if (A.status[k] != S_CONTAINED) best = fmin(best, ddist(x, y, z, A.cx[k], A.cy[k], A.cz[k]) - A.R[k]);
            });
            best = wave_min(best);
            if (lane_id() == 0) out[w] = best;
            continue;
        }
        bool inside = false;
        for_near(A.gs, x, y, z, A.max_expanded, lane_id(), WAVE, [&](uint32_t k) {
            if (!inside && A.status[k] != S_CONTAINED && ddist(x, y, z, A.cx[k], A.cy[k], A.cz[k]) < A.R[k])
                inside = true;
        });
        double dist = wave_any(inside) ? ses_distance_wave(A, x, y, z, A.probe + band) : 0.0;
        if (lane_id() == 0) out[w] = A.probe - dist;
    }
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
    k_field_clamped<<<std::min<uint32_t>((uint32_t)n, 1u << 24), WAVE, 0, s>>>(r->A, pd, (uint32_t)n, band, vd);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(values, vd, n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipFreeAsync(pd, s));
    HIP_TRY(hipFreeAsync(vd, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SMG_OK;
}

// ------------------------------------------------------------- mesh stage
namespace {

// Temporary device buffers of one run, freed at the end (stream-ordered).
struct Temps {
    std::vector<void*> ptrs;
    hipStream_t s;
    explicit Temps(hipStream_t st) : s(st) {}
    ~Temps() {
        for (void* p : ptrs) (void)hipFreeAsync(p, s);
    }
    hipError_t get(void** p, size_t bytes) {
        hipError_t e = hipMallocAsync(p, std::max<size_t>(bytes, 1), s);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Cancellation (control.rs:104-118): the caller's flag, and the device flag
// raised once per run while kernels run.
bool cancel_requested(const smg_device* d) { return d->ctl.cancel && *d->ctl.cancel != 0; }
void raise_cancel(smg_device* d) {
    if (d->raised) return;
    __atomic_store_n(d->cancel_host, d->gen, __ATOMIC_SEQ_CST);
    d->host_u32[63] = d->gen;
    (void)hipMemcpyAsync(d->cancel_dev, d->host_u32 + 63, 4, hipMemcpyHostToDevice, d->cstream);
    d->raised = true;
}
int check_cancel(smg_device* d) {
    if (!cancel_requested(d)) return SMG_OK;
    raise_cancel(d);
    return fail(SMG_E_CANCELLED, "cancelled");
}
// Before a run is enqueued: the caller's flag is clear; the run takes the
// next generation (its table image carries it to the device).
int prepare_run(smg_device* d) {
    TRY(check_cancel(d));
    uint32_t g;  // never 0 (the word's initial value) nor g_run's initial value
    do g = g_generation[d->ordinal].fetch_add(1u, std::memory_order_relaxed) + 1u;
    while (g == 0u || g == 0xffffffffu);
    d->gen = g;
    d->raised = false;
    return SMG_OK;
}
void report(const smg_device* d, int32_t stage, uint64_t done, uint64_t total) {
    if (d->ctl.progress) d->ctl.progress(d->ctl.user, stage, done, total);
}
// Waits for the stream; with a cancellation flag, polls it meanwhile and
// raises the device flag once it is set (the wait then ends as soon as the
// kernels stop), returning SMG_E_CANCELLED.
int wait_stream(smg_device* d, hipStream_t s) {
    if (!d->ctl.cancel) {
        HIP_TRY(hipStreamSynchronize(s));
        return SMG_OK;
    }
    for (;;) {
        hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) return fail(SMG_E_CALL, std::string("hipStreamQuery: ") + hipGetErrorString(e));
        if (cancel_requested(d)) raise_cancel(d);
        // A short sleep, not a spin: back-to-back stream queries slowed the
        // kernels they waited for several times over (a 1024-molecule
        // batch's sampling 0.23 -> 0.7-1.4 s with a control installed).
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    if (d->raised) return fail(SMG_E_CANCELLED, "cancelled");
    return check_cancel(d);
}

int read_u32(smg_device* d, const uint32_t* dev_ptr, uint32_t& v) {
    HIP_TRY(hipMemcpyAsync(d->host_u32, dev_ptr, 4, hipMemcpyDeviceToHost, d->stream));
    TRY(wait_stream(d, d->stream));
    v = d->host_u32[0];
    return SMG_OK;
}


// List capacities, tried in turn: LDS variants sized for common sub-blocks
// (several workgroups per CU) and for those of coarse lattices (up to the
// 160 KB of LDS one gfx950 workgroup may hold), then global scratch for the
// rest. Each pass takes the previous pass's overflow list.
constexpr int TIERS = 5;  // the last one in global scratch
// The first tiers hold 192 candidates: 6 sampling and 5 refinement
// workgroups fit a CU's 160 KB of LDS (256 held 5 and 4), and 0.25 A
// sub-blocks need at most ~170. The second holds what most 0.5 A
// sub-blocks need (near spheres to ~190, candidates to ~250) at 4 and 3
// workgroups per CU.
constexpr Caps SAMPLE_CAPS[TIERS] = {{128, 64, 192, 64, 256},
                                     {192, 96, 256, 64, 512},
                                     {256, 128, 384, 128, 512},
                                     {640, 256, 768, 256, 2048},
                                     {1024, 512, 4096, 2048, 4096}};
constexpr RCaps REFINE_CAPS[TIERS] = {{128, 64, 192, 256, 0, 0, 0},
                                      {192, 96, 256, 512, 0, 0, 0},
                                      {256, 128, 384, 512, 0, 0, 0},
                                      {640, 256, 704, 1024, 0, 0, 0},
                                      {1024, 512, 4096, 4096, 0, 0, 0}};
constexpr int WIDE_FROM = 2;  // refinement passes over tiers from this one run WIDE threads
static_assert(TIERS == SPLIT_TIERS, "the learned-sizing split covers every tier");
// Sub-block classification: three LDS tiers (5, 3 and 1 workgroups per
// CU); bricks beyond them take one wavefront per brick.
constexpr int UNIFORM_TIERS = 3;
constexpr RCaps UNIFORM_CAPS[UNIFORM_TIERS] = {
    {128, 64, 192, 256, 0, 1, 0}, {192, 128, 256, 512, 0, 1, 0}, {384, 256, 640, 1024, 0, 1, 0}};
constexpr uint32_t LDS_EDGES = 512;      // crossing-edge buffer of the LDS passes
constexpr uint32_t SCRATCH_CHUNK = 1024;  // default workgroups of a scratch pass (one list slice each)
// Workgroups of the wavefront passes that take what every list tier left: few
// bricks reach them, and a residency-sized grid (8,192 one-wavefront groups)
// costs tens of microseconds to launch and drain even when none does.
constexpr uint32_t WAVE_PASS_WGS = 2048;
constexpr uint32_t HID_PER_SUB = 64;      // hidden-candidate pool entries per sampled sub-block
constexpr uint32_t SUP_PER_SUB = 160;     // refinement-superset pool entries per sampled sub-block
constexpr uint32_t BLK_PER_SUB = 320;     // blocker-record slot words per sampled sub-block
constexpr uint32_t BSUP_PER_BRICK = 320;  // brick-superset pool entries per touched brick
constexpr uint64_t BSUP_MAX = 1ull << 28; // brick-superset pool entries at most
constexpr uint32_t NO_CAP = 0xffffffffu;

// Device counters of a run (u32 slots, zero at the start of the run).
enum : uint32_t {
    K_TOUCHED, K_NB, K_NSUB, K_NSW, K_NV, K_NRW, K_NT, K_FLAGS, K_ERR,
    // Per sub-block classification tier: overflow count, queue ticket.
    K_UOVER, K_Q_UNIFORM0 = K_UOVER + UNIFORM_TIERS,
    // Per pass: items in its list (predicted there, then appended by the
    // passes before it), items predicted there, and the last pass's overflow.
    K_SLIST0 = K_Q_UNIFORM0 + UNIFORM_TIERS, K_SPRED0 = K_SLIST0 + TIERS, K_SLAST = K_SPRED0 + TIERS,
    K_RLIST0, K_RPRED0 = K_RLIST0 + TIERS, K_RLAST = K_RPRED0 + TIERS,
    K_Q_SUBST, K_Q_SAMPLE0, K_Q_REFINE0 = K_Q_SAMPLE0 + TIERS,
    K_HID = K_Q_REFINE0 + TIERS,  // entries taken from the hidden-candidate pool
    K_SUP,                        // entries taken from the refinement-superset pool
    K_BSUP,                       // entries taken from the brick-superset pool
    K_N
};
// Capacity of each tier in candidate-load units (max(S + C, 2 near, 4 SAS
// vertices, 4 inside) of the sampling lists; the refinement lists alike).
constexpr float SAMPLE_LOAD_CAP[TIERS] = {192.0f, 256.0f, 384.0f, 768.0f, 2048.0f};
constexpr float REFINE_LOAD_CAP[TIERS] = {192.0f, 256.0f, 384.0f, 704.0f, 2048.0f};
// K_FLAGS bits: a device count exceeded its capacity.
enum : uint32_t { OVER_BRICKS = 1, OVER_VERTICES = 2, OVER_TRIANGLES = 4, OVER_TABLE = 8 };

// Kernels whose LDS request depends on the brick or the list tier; a
// launch needs its dynamic bytes plus the kernel's static bytes within the
// device's per-workgroup LDS.
enum : int { LK_EDGES, LK_CELLS, LK_UNIFORM, LK_SAMPLE, LK_REFINE, LK_REFINE_WIDE, LK_REFINE_FAST, LK_REFINE_FAST_WIDE, LK_N };
static_assert(LK_N == 8, "smg_device::static_lds holds every kind");

// Which LDS variants a run at brick b takes: the edge and cell kernels
// stage a brick in LDS when it fits, and each LDS list tier runs only when
// it fits (the global-scratch tier always runs).
struct LdsPlan {
    uint32_t edges_dyn, cells_dyn, edges_staged, cells_staged;
    uint32_t uniform_dyn[UNIFORM_TIERS], uniform_fits[UNIFORM_TIERS];
    uint32_t sample_dyn[TIERS], sample_fits[TIERS];
    uint32_t refine_dyn[TIERS], refine_fits[TIERS];
};

LdsPlan lds_plan(uint32_t b, bool fast, uint64_t lds_max, const uint32_t st[LK_N]) {
    LdsPlan L{};
    const uint32_t spb = (b + 1) * (b + 1) * (b + 1);
    const uint32_t words = (3 * b * b * b + 63) / 64;
    auto fits = [&](uint64_t dyn, int k) { return (uint32_t)(dyn + st[k] <= lds_max); };
    const uint32_t signs = 8 * ((spb + 63) / 64 + 1);  // sign bits of the samples, one zero word past
    L.edges_dyn = signs + 8 * words;
    L.edges_staged = fits(L.edges_dyn, LK_EDGES);
    L.cells_dyn = signs + 8 * 12 * words + 4 * TRI_SLOTS;
    L.cells_staged = fits(L.cells_dyn, LK_CELLS);
    for (int u = 0; u < UNIFORM_TIERS; u++) {
        L.uniform_dyn[u] = make_rlayout(UNIFORM_CAPS[u]).bytes;
        L.uniform_fits[u] = fits(L.uniform_dyn[u], LK_UNIFORM);
    }
    const uint32_t sub_edges = 3 * (b / 2) * (b / 2) * (b / 2);
    for (int p = 0; p < TIERS; p++) {
        const bool lds = p < TIERS - 1;
        L.sample_dyn[p] = lds ? make_layout(SAMPLE_CAPS[p]).bytes : 0;
        L.sample_fits[p] = lds ? fits(L.sample_dyn[p], LK_SAMPLE) : 1;
        RCaps cap = REFINE_CAPS[p];
        cap.edges = lds ? std::min(sub_edges, LDS_EDGES) : sub_edges;
        RLayout rl = make_rlayout(cap);
        L.refine_dyn[p] = lds ? (fast ? make_flayout(cap, rl.bytes).bytes : rl.bytes) : 0;
        const int kind = fast ? (p >= WIDE_FROM ? LK_REFINE_FAST_WIDE : LK_REFINE_FAST)
                              : (p >= WIDE_FROM ? LK_REFINE_WIDE : LK_REFINE);
        L.refine_fits[p] = lds ? fits(L.refine_dyn[p], kind) : 1;
    }
    return L;
}

// The static LDS of each kind on this device (the largest of the kind's
// launched variants).
int device_statics(smg_device* d, const uint32_t** out) {
    std::lock_guard<std::mutex> lk(d->mu);
    if (!d->static_ready) {
        const void* kinds[LK_N][4] = {
            {(const void*)k_edges, (const void*)k_edges},
            {(const void*)k_cell_emit, (const void*)k_cell_emit},
            {(const void*)k_uniform<false>, (const void*)k_uniform<true>, (const void*)k_classify_touched<false>,
             (const void*)k_classify_touched<true>},
            {(const void*)k_sample<false, false>, (const void*)k_sample<false, true>},
            {(const void*)k_refine<false, false>, (const void*)k_refine<false, true>},
            {(const void*)k_refine_wide<false, false>, (const void*)k_refine_wide<false, true>},
            {(const void*)k_refine_fast<false, false>, (const void*)k_refine_fast<false, true>},
            {(const void*)k_refine_fast_wide<false, false>, (const void*)k_refine_fast_wide<false, true>}};
        for (int k = 0; k < LK_N; k++) {
            uint32_t m = 0;
            for (const void* f : kinds[k]) {
                if (!f) continue;
                hipFuncAttributes a{};
                if (hipFuncGetAttributes(&a, f) != hipSuccess)
                    return fail(SMG_E_CALL, "hipFuncGetAttributes failed for a mesh-stage kernel");
                m = std::max(m, (uint32_t)a.sharedSizeBytes);
            }
            d->static_lds[k] = m;
        }
        d->static_ready = true;
    }
    *out = d->static_lds;
    return SMG_OK;
}

// Run-wide extents the kernels read from the uploaded table.
struct RunHdr {
    uint32_t nwords, dense, locals, spheres;
    uint32_t gen;  // the run's generation (cancellation)
};

inline uint32_t g1(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n, 1u << 30)); }
// Workgroups of `block` threads one dispatch may hold (fewer than 2^32
// work-items): a pass with more items than that runs persistent.
inline uint64_t dispatch_max(uint32_t block) { return std::min<uint64_t>(1u << 30, 0xffffffffull / block); }
inline bool fits_dispatch(uint64_t items, uint32_t block) { return items <= dispatch_max(block); }
inline uint32_t gridn(uint64_t n, uint32_t block) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(n, dispatch_max(block))); }
inline uint32_t stride_grid(uint64_t n, uint32_t bs = 256) { return g1(std::min<uint64_t>(blocks_for(n, bs), 8192)); }

// Workgroups of `block` threads of `kernel` with `lds` dynamic LDS bytes that
// the whole device holds at once.
uint32_t resident(smg_device* d, const void* kernel, uint32_t block, size_t lds) {
    std::lock_guard<std::mutex> lk(d->mu);
    auto key = std::make_pair(kernel, lds * 1024 + block);
    auto it = d->resident.find(key);
    if (it != d->resident.end()) return it->second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    uint32_t n = (uint32_t)per_cu * (uint32_t)std::max(d->prop.multiProcessorCount, 1);
    d->resident[key] = n;
    return n;
}

// Where a run's buffers come from and how it learns its sizes.
// Exact: a size is read back from the device when a buffer depends on it
// (three synchronisations), temporaries come from the stream-ordered pool
// and the mesh owns its outputs. Planned: sizes are a plan's capacities and
// nothing is read back before the end, so the whole run can be captured
// into a graph; every buffer is a plan slot, allocated on the plan's first
// run and reused, in the same order, by every later run and by the capture.
struct Run {
    smg_device* d = nullptr;
    hipStream_t s = nullptr;
    Temps* T = nullptr;                       // exact
    std::vector<void*>* slots = nullptr;      // planned
    std::vector<size_t>* slot_bytes = nullptr;
    size_t next_slot = 0;
    // Allocation-free (smg_mesh_into): sizes are capacities as in a planned
    // run, and every buffer is carved from a fixed device array.
    char* arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
    uint32_t scratch_wgs = 0;  // workgroups of the global-scratch passes (0: the tuning's)
    float* lbuf = nullptr;  // the run's f32-copy buffer (its table points into it)
    bool planned() const { return slots != nullptr || arena != nullptr; }
    int slot(void** p, size_t bytes) {
        if (arena) {
            size_t at = (arena_used + 255) & ~(size_t)255;
            if (at > arena_size || bytes > arena_size - at)
                return fail(SMG_E_LENGTH, "the run needs more than the " + std::to_string(arena_size >> 20) +
                                              " MiB of scratch smg_mesh_into has (a smaller molecule, or smg_mesh_new)");
            *p = arena + at;
            arena_used = at + bytes;
            return SMG_OK;
        }
        if (next_slot < slots->size()) {
            if ((*slot_bytes)[next_slot] != bytes) return fail(SMG_E_INTERNAL, "plan buffers changed between runs");
            *p = (*slots)[next_slot++];
            return SMG_OK;
        }
        void* q = nullptr;
        HIP_TRY(hipMalloc(&q, bytes));
        slots->push_back(q);
        slot_bytes->push_back(bytes);
        next_slot++;
        *p = q;
        return SMG_OK;
    }
    // Allocations of an unplanned run check the caller's cancellation flag
    // first: a large allocation holds the host for a while, and the flag is
    // not polled meanwhile.
    template <class X>
    int temp(X** p, size_t count) {
        size_t bytes = std::max<size_t>(count, 1) * sizeof(X);
        if (planned()) return slot((void**)p, bytes);
        TRY(check_cancel(d));
        HIP_TRY(T->get((void**)p, bytes));
        return SMG_OK;
    }
    template <class X>
    int output(X** p, size_t count) {
        size_t bytes = std::max<size_t>(count, 1) * sizeof(X);
        if (arena && *p) return SMG_OK;  // the caller's buffer, sized for the run's capacity
        if (planned()) return slot((void**)p, bytes);
        TRY(check_cancel(d));
        HIP_TRY(hipMalloc((void**)p, bytes));
        return SMG_OK;
    }
    // A device count: read back (exact) or the capacity (planned).
    int size(const uint32_t* dev, uint32_t cap, uint32_t& v) {
        if (planned()) {
            v = cap;
            return SMG_OK;
        }
        return read_u32(d, dev, v);
    }
};

// Exclusive scan of in[0 .. *n) into out, at most `cap` items; the total
// goes to *total (device).
int scan_u32(Run& X, const uint32_t* in, uint32_t* out, const uint32_t* n, uint64_t cap, uint32_t* total) {
    uint64_t tiles = std::max<uint64_t>((cap + SCAN_TILE - 1) / SCAN_TILE, 1);
    uint32_t* sums;
    TRY(X.temp(&sums, tiles));
    k_scan_tiles<<<g1(tiles), SCAN_BLOCK, 0, X.s>>>(in, out, sums, n);
    k_scan_sums<<<1, SCAN_BLOCK, 0, X.s>>>(sums, n, total);
    k_scan_add<<<stride_grid(cap), 256, 0, X.s>>>(out, sums, n);
    HIP_TRY(hipGetLastError());
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

// A run's molecule table: one row per molecule (rows past `nm` are empty
// molecules, so a plan's table keeps its size), with the offsets of each
// molecule's spheres, f32 copies and bitmap words in the concatenations.
struct Table {
    size_t nm = 0, rows = 0;
    uint32_t b = 0;
    bool fast = false;
    std::vector<Mol> H;
    std::vector<uint64_t> sph, loc, wofs;  // rows + 1 each
    uint64_t dense = 0;
    uint64_t words() const { return wofs[rows]; }
    uint64_t locals() const { return loc[rows]; }
    uint64_t spheres() const { return sph[rows]; }
};

int build_table(smg_resident* const* R, const smg_mesh_params* prm, size_t nm, size_t rows, Table& t) {
    if (nm == 0) return fail(SMG_E_INVALID, "no molecules");
    if (nm >= (1u << 24)) return fail(SMG_E_LENGTH, "too many molecules in one run");
    smg_device* d = R[0]->dev;
    for (size_t m = 0; m < nm; m++)
        if (!R[m] || R[m]->dev != d) return fail(SMG_E_INVALID, "all molecules of a run must be resident on one device");
    const uint32_t b = prm[0].brick;
    if (b < 4 || b > 32 || b % 2) return fail(SMG_E_UNSUPPORTED, "the device path needs an even brick size in 4..=32");
    t.fast = prm[0].refine == SMG_REFINE_FAST;
    for (size_t m = 0; m < nm; m++)
        if ((prm[m].refine == SMG_REFINE_FAST) != t.fast) return fail(SMG_E_UNSUPPORTED, "one refinement mode per run");
    t.nm = nm;
    t.rows = std::max(rows, nm);
    t.b = b;
    t.H.assign(t.rows, Mol{});
    t.sph.assign(t.rows + 1, 0);
    t.loc.assign(t.rows + 1, 0);
    t.wofs.assign(t.rows + 1, 0);
    t.dense = 0;
    for (size_t m = 0; m < t.rows; m++) {
        Mol& X = t.H[m];
        uint64_t cells = 0, nmax = 0;
        if (m < nm) {
            X.A = R[m]->A;
            bool empty = false;
            TRY(molecule_run(R[m], &prm[m], b, X.P, empty));
            cells = empty ? 0 : (uint64_t)X.P.bdim[0] * X.P.bdim[1] * X.P.bdim[2];
            nmax = std::max<uint64_t>({X.A.n_spheres, X.A.n_circles, X.A.n_vertices});
        }
        t.sph[m + 1] = t.sph[m] + X.A.n_spheres;
        t.loc[m + 1] = t.loc[m] + nmax;
        t.wofs[m + 1] = t.wofs[m] + (cells + 31) / 32;
        X.word_off = t.wofs[m];
        X.dense_off = t.dense;
        t.dense += cells;
    }
    if (t.words() >= (1ull << 27)) return fail(SMG_E_UNSUPPORTED, "brick boxes of the run too large for 32-bit bit indices");
    if (t.locals() >= (1ull << 32) || t.spheres() >= (1ull << 32))
        return fail(SMG_E_UNSUPPORTED, "too many spheres in one run");
    return SMG_OK;
}

// The uploaded image of a table: header, rows, three offset arrays.
struct StageLayout {
    size_t mols, sph, loc, wofs, bytes;
};
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
StageLayout stage_layout(size_t rows) {
    StageLayout L;
    L.mols = 256;
    L.sph = align256(L.mols + rows * sizeof(Mol));
    L.loc = align256(L.sph + (rows + 1) * 8);
    L.wofs = align256(L.loc + (rows + 1) * 8);
    L.bytes = align256(L.wofs + (rows + 1) * 8);
    return L;
}

// f32 copies per molecule: LOC_ARRAYS float4 arrays of max(spheres,
// circles, SAS vertices) entries.
constexpr int LOC_ARRAYS = 5;

// Writes the table into `host` (laid out as stage_layout), pointing each
// molecule's f32 copies into `lbuf`.
void write_stage(Table& t, float* lbuf, char* host, uint32_t gen) {
    StageLayout L = stage_layout(t.rows);
    for (size_t m = 0; m < t.rows; m++) {
        DLoc& Lc = t.H[m].L;
        size_t nmax = t.loc[m + 1] - t.loc[m];
        float4* base = (float4*)lbuf + LOC_ARRAYS * t.loc[m];
        gptr<float4>* fields[LOC_ARRAYS] = {&Lc.p4, &Lc.m4, &Lc.u4, &Lc.e4, &Lc.v4};
        for (int q = 0; q < LOC_ARRAYS; q++) *fields[q] = to_global(base + q * nmax);
    }
    RunHdr h{(uint32_t)t.words(), (uint32_t)t.dense, (uint32_t)t.locals(), (uint32_t)t.spheres(), gen};
    std::memset(host, 0, L.bytes);
    std::memcpy(host, &h, sizeof(h));
    std::memcpy(host + L.mols, t.H.data(), t.rows * sizeof(Mol));
    std::memcpy(host + L.sph, t.sph.data(), (t.rows + 1) * 8);
    std::memcpy(host + L.loc, t.loc.data(), (t.rows + 1) * 8);
    std::memcpy(host + L.wofs, t.wofs.data(), (t.rows + 1) * 8);
}

// Host-side extents a run's buffers are sized for, and the capacities of
// the device counts (NO_CAP: none, the exact path).
struct Extents {
    uint64_t rows, locals, spheres, words, dense;
    uint32_t nb, nv, nt;
};

// The per-run summary read back at the end, one image on the device and
// on the host: samples and evaluations per molecule (u64), the counters,
// then per-molecule first touched brick, surface brick, sampled sub-block,
// vertex and triangle (rows + 1 each, the last holding the totals).
struct SumLayout {
    size_t nsamp, evals, cnt, tfirst, fbrick, sfirst, vfirst, trfirst, zero_words, bytes;
};
SumLayout sum_layout(size_t rows) {
    SumLayout L;
    L.nsamp = 0;
    L.evals = 8 * rows;
    L.cnt = 16 * rows;
    L.tfirst = L.cnt + 4 * K_N;
    L.fbrick = L.tfirst + 4 * (rows + 1);
    L.sfirst = L.fbrick + 4 * (rows + 1);
    L.vfirst = L.sfirst + 4 * (rows + 1);
    L.trfirst = L.vfirst + 4 * (rows + 1);
    L.bytes = L.trfirst + 4 * (rows + 1);
    L.zero_words = L.tfirst / 4;
    return L;
}

// Learned sizing of a run: `predict` sends each sub-block first to the LDS
// tier its brick's predicted load fits; `collect` appends the run's touched
// bricks, labelled by the exact classifier and the exact gathers, to the
// model's training set.
struct Sizing {
    smg_model* predict = nullptr;
    smg_model* collect = nullptr;
};

int grow_dataset(smg_model* m, size_t need);

// The mesh stage of a table's molecules in one pass: every kernel runs once
// over the concatenated work of all of them, each work item finding its
// molecule's arrangement, f32 copies and lattice in the table. Sets M's
// device arrays and copies the summary into `sum_host` (asynchronously).
int pipeline(Run& X, Table& t, const Extents& E, char* stage_host, char* sum_host, smg_mesh* M, bool timed,
             const Sizing& sz) {
    smg_device* d = X.d;
    hipStream_t s = X.s;
    const uint32_t b = t.b, spb = (b + 1) * (b + 1) * (b + 1);
    const uint32_t rows = (uint32_t)E.rows;
    const bool fast = t.fast;
    const smg_tuning tune = d->tune;
    const uint32_t scratch_wgs = X.scratch_wgs                    ? X.scratch_wgs
                                 : tune.scratch_workgroups > 0 ? (uint32_t)tune.scratch_workgroups
                                                               : SCRATCH_CHUNK;
    M->spb = spb;
    const uint32_t* statics;
    TRY(device_statics(d, &statics));
    const LdsPlan LP = lds_plan(b, fast, d->prop.sharedMemPerBlock, statics);
    auto rec = [&](int i) -> int {
        if (timed) HIP_TRY(hipEventRecord(d->ev[i], s));
        return SMG_OK;
    };
    // Table upload and run counters.
    StageLayout SL = stage_layout(rows);
    SumLayout UL = sum_layout(rows);
    char *stage, *sum;
    float* lbuf;
    TRY(X.temp(&stage, SL.bytes));
    TRY(X.temp(&sum, UL.bytes));
    TRY(X.temp(&lbuf, 4 * LOC_ARRAYS * E.locals));
    X.lbuf = lbuf;
    write_stage(t, lbuf, stage_host, d->gen);
    TRY(rec(0));
    HIP_TRY(hipMemcpyAsync(stage, stage_host, SL.bytes, hipMemcpyHostToDevice, s));
    const RunHdr* hdr = (const RunHdr*)stage;
    k_run_begin<<<1, 1, 0, s>>>(&hdr->gen);
    const Mol* mols = (const Mol*)(stage + SL.mols);
    const uint64_t* sph_d = (const uint64_t*)(stage + SL.sph);
    const uint64_t* loc_d = (const uint64_t*)(stage + SL.loc);
    const uint64_t* wofs_d = (const uint64_t*)(stage + SL.wofs);
    unsigned long long* nsamp = (unsigned long long*)(sum + UL.nsamp);
    unsigned long long* evd = (unsigned long long*)(sum + UL.evals);
    uint32_t* cnt = (uint32_t*)(sum + UL.cnt);
    uint32_t* tfirst = (uint32_t*)(sum + UL.tfirst);
    uint32_t* fbrick = (uint32_t*)(sum + UL.fbrick);
    uint32_t* sfirst = (uint32_t*)(sum + UL.sfirst);
    uint32_t* vfirst = (uint32_t*)(sum + UL.vfirst);
    uint32_t* trfirst = (uint32_t*)(sum + UL.trfirst);
    const bool report_phases = !X.planned();  // a planned run reports from smg_plan_run
    if (report_phases) {
        report(d, SMG_STAGE_CLASSIFY, 0, 1);
        TRY(check_cancel(d));
    }
    k_fill_u32<<<stride_grid(UL.zero_words), 256, 0, s>>>((uint32_t*)sum, 0u, UL.zero_words, nullptr);
    k_localize<<<stride_grid(E.locals), 256, 0, s>>>(mols, loc_d, rows);
    HIP_TRY(hipGetLastError());

    // Touched bricks (lattice.rs:105-163) in per-molecule dense bitmaps.
    uint32_t *bits, *wcount, *woff, *dense, *touched, *keep, *koff;
    TRY(X.temp(&bits, E.words));
    TRY(X.temp(&wcount, E.words));
    TRY(X.temp(&woff, E.words));
    TRY(X.temp(&dense, E.dense));
    TRY(X.temp(&touched, E.dense));
    TRY(X.temp(&keep, E.dense));
    TRY(X.temp(&koff, E.dense));
    k_fill_u32<<<stride_grid(E.words), 256, 0, s>>>(bits, 0u, 0, &hdr->nwords);
    k_mark_touched<<<stride_grid(E.spheres, 128), 128, 0, s>>>(mols, sph_d, rows, bits);
    k_popcount<<<stride_grid(E.words), 256, 0, s>>>(bits, wcount, &hdr->nwords);
    TRY(scan_u32(X, wcount, woff, &hdr->nwords, E.words, &cnt[K_TOUCHED]));
    k_bits_to_list<<<stride_grid(E.words), 256, 0, s>>>(bits, woff, &hdr->nwords, touched);
    // Brick classification (lattice.rs:168-183). With shared gathers, one
    // workgroup per touched brick also finds a surface brick's sub-block
    // states (in touched order, stt), in the LDS tiers in turn, one
    // wavefront per brick for those whose lists overflow them all;
    // otherwise one wavefront per brick classifies it alone.
    TRY(rec(6));
    // Batches of many molecules keep the two passes: their molecules are
    // small, with a larger share of touched bricks that are not surface
    // bricks (a full gather each in the fused pass), and their planned runs
    // would launch a workgroup per brick-box cell (measured: 4096 molecules,
    // 0.26-0.56 s fused against 0.24 s).
    const bool fused = tune.uniform_shared && LP.uniform_fits[0] && rows == 1;
    uint8_t* stt = nullptr;
    SupList bsup_t{};  // per touched brick: the superset of its sampling walks (not in allocation-free runs)
    if (fused) {
        TRY(X.temp(&stt, 8 * (size_t)E.dense));
        uint32_t ntc;  // touched bricks (their capacity in a planned run)
        TRY(X.size(&cnt[K_TOUCHED], (uint32_t)E.dense, ntc));
        if (!X.arena) {
            TRY(X.temp(&bsup_t.off, E.dense));
            TRY(X.temp(&bsup_t.cnt, E.dense));
            const uint64_t bcap = std::min<uint64_t>((uint64_t)BSUP_PER_BRICK * std::max<uint32_t>(ntc, 1u), BSUP_MAX);
            TRY(X.temp(&bsup_t.ids, bcap));
            bsup_t.used = &cnt[K_BSUP];
            bsup_t.cap = (uint32_t)bcap;
            k_fill_u32<<<stride_grid(E.dense), 256, 0, s>>>(bsup_t.cnt, NO_LIST, 0, &cnt[K_TOUCHED]);
            HIP_TRY(hipGetLastError());
        }
        const uint32_t* list = nullptr;  // the first pass takes every touched brick
        const uint32_t* count = &cnt[K_TOUCHED];
        bool first = true;
        for (int u = 0; u < UNIFORM_TIERS; u++) {
            if (!LP.uniform_fits[u]) continue;
            RLayout lay = make_rlayout(UNIFORM_CAPS[u]);
            uint32_t* over;
            TRY(X.temp(&over, E.dense));
            Queue q{list, count, &cnt[K_Q_UNIFORM0 + u]};
            // The first pass takes one workgroup per touched brick: in an exact
            // run a grid of their count; in a planned run, whose count is known
            // only on the device, a grid of the brick boxes' cells (few enough
            // here that the idle workgroups cost little). The other passes
            // loop over their queues.
            const uint64_t first_grid = X.planned() ? E.dense : ntc;
            if (first && fits_dispatch(first_grid, 256))
                k_classify_touched<false><<<gridn(first_grid, 256), 256, lay.bytes, s>>>(
                    mols, wofs_d, rows, touched, UNIFORM_CAPS[u], lay, q, keep, stt, over, &cnt[K_UOVER + u], bsup_t);
            else
                k_classify_touched<true><<<g1(std::min<uint64_t>(E.dense, resident(d, (const void*)k_classify_touched<true>,
                                                                                    256, lay.bytes))),
                                           256, lay.bytes, s>>>(mols, wofs_d, rows, touched, UNIFORM_CAPS[u], lay, q,
                                                                keep, stt, over, &cnt[K_UOVER + u], bsup_t);
            HIP_TRY(hipGetLastError());
            first = false;
            list = over;
            count = &cnt[K_UOVER + u];
        }
        Queue rest{list, count, &cnt[K_Q_SUBST]};
        uint32_t grid = g1(std::min<uint64_t>({E.dense, resident(d, (const void*)k_classify_touched_wave, WAVE, 0),
                                                 (uint64_t)WAVE_PASS_WGS}));
        k_classify_touched_wave<<<grid, WAVE, 0, s>>>(mols, wofs_d, rows, touched, rest, keep, stt);
        HIP_TRY(hipGetLastError());
    } else {
        Queue q{nullptr, &cnt[K_TOUCHED], nullptr};
        if (!fits_dispatch(E.dense, WAVE)) return fail(SMG_E_UNSUPPORTED, "brick boxes of the run too large for one dispatch");
        k_classify_bricks<<<gridn(E.dense, WAVE), WAVE, 0, s>>>(mols, wofs_d, rows, touched, q, keep);
    }
    TRY(rec(7));
    TRY(scan_u32(X, keep, koff, &cnt[K_TOUCHED], E.dense, &cnt[K_NB]));
    k_clamp<<<1, 64, 0, s>>>(&cnt[K_NB], E.nb, &cnt[K_FLAGS], OVER_BRICKS, &cnt[K_NSUB], 8u);
    HIP_TRY(hipGetLastError());
    uint32_t nb;
    TRY(X.size(&cnt[K_NB], E.nb, nb));
    if (report_phases) {
        report(d, SMG_STAGE_CLASSIFY, 1, 1);
        report(d, SMG_STAGE_MESH, 0, nb);
        TRY(check_cancel(d));
    }
    uint32_t* bmol;
    uint8_t* st;
    TRY(X.output(&M->keys, (size_t)nb * 3));
    TRY(X.temp(&bmol, nb));
    TRY(X.temp(&st, 8 * (size_t)nb));
    k_fill_u32<<<stride_grid(E.dense), 256, 0, s>>>(dense, 0xffffffffu, 0, &hdr->dense);
    k_compact_bricks<<<stride_grid(E.dense), 256, 0, s>>>(mols, wofs_d, rows, touched, keep, koff, &cnt[K_TOUCHED],
                                                            nb, M->keys, bmol, dense);
    k_mol_ranges<<<stride_grid(rows + 1), 256, 0, s>>>(wofs_d, rows, woff, &hdr->nwords, &cnt[K_TOUCHED], koff,
                                                         &cnt[K_NB], tfirst, fbrick);
    HIP_TRY(hipGetLastError());
    // Predicted candidate loads of the surface bricks (learned sizing).
    float* pred = nullptr;
    if (sz.predict) {
        float* feat;
        TRY(X.temp(&feat, (size_t)nb * MLP_IN));
        TRY(X.temp(&pred, (size_t)nb * MLP_OUT));
        k_features<<<gridn(nb, WAVE), WAVE, 0, s>>>(mols, wofs_d, rows, nullptr, M->keys, bmol, Queue{nullptr, &cnt[K_NB], nullptr},
                                           feat);
        k_mlp_forward<<<g1(std::min<uint64_t>((nb + MLP_ROWS - 1) / MLP_ROWS, 4096)), 256, 0, s>>>(
            feat, &cnt[K_NB], sz.predict->P, sz.predict->W1b, sz.predict->W2b, pred);
        HIP_TRY(hipGetLastError());
    }
    // Training labels: features of every touched brick, and per surface
    // brick the largest candidate load of its sampled sub-blocks.
    float *cfeat = nullptr, *ctgt = nullptr;
    uint8_t* cocc = nullptr;
    uint32_t* cload = nullptr;
    uint32_t ntouched = 0;
    if (sz.collect) {
        TRY(X.size(&cnt[K_TOUCHED], (uint32_t)E.dense, ntouched));
        TRY(X.temp(&cfeat, (size_t)ntouched * MLP_IN));
        TRY(X.temp(&cocc, ntouched));
        TRY(X.temp(&ctgt, ntouched));
        TRY(X.temp(&cload, nb));
        k_features<<<gridn(ntouched, WAVE), WAVE, 0, s>>>(mols, wofs_d, rows, touched, nullptr, nullptr,
                                                 Queue{nullptr, &cnt[K_TOUCHED], nullptr}, cfeat);
        k_fill_u32<<<stride_grid(nb), 256, 0, s>>>(cload, 0u, 0, &cnt[K_NB]);
        HIP_TRY(hipGetLastError());
    }
    // Sub-block states (kernel.rs:467-482): from the touched-order states of
    // the fused classification, or one shared gather per brick in the LDS
    // tiers in turn, one wavefront per brick for bricks whose lists overflow
    // them all.
    TRY(rec(8));
    SupList bsup_s{};  // the brick supersets in surface-brick order, for the sampling
    if (fused) {
        if (bsup_t.ids) {
            bsup_s = bsup_t;
            TRY(X.temp(&bsup_s.off, nb));
            TRY(X.temp(&bsup_s.cnt, nb));
            k_fill_u32<<<stride_grid(nb), 256, 0, s>>>(bsup_s.cnt, NO_LIST, 0, &cnt[K_NB]);
        }
        k_gather_states<<<stride_grid(E.dense), 256, 0, s>>>(keep, koff, &cnt[K_TOUCHED], nb, stt, st, bsup_t, bsup_s);
        HIP_TRY(hipGetLastError());
        TRY(rec(9));
    } else {
        const uint32_t* list = nullptr;  // the first pass takes every surface brick
        const uint32_t* count = &cnt[K_NB];
        bool first = true;
        for (int u = 0; u < UNIFORM_TIERS && tune.uniform_shared; u++) {
            RLayout lay = make_rlayout(UNIFORM_CAPS[u]);
            if (!LP.uniform_fits[u]) continue;
            uint32_t* over;
            TRY(X.temp(&over, nb));
            Queue q{list, count, &cnt[K_Q_UNIFORM0 + u]};
            uint32_t items = nb;
            if (!first && !X.planned()) TRY(X.size(count, 0, items));
            if ((first || !X.planned()) && fits_dispatch(items, 256))
                k_uniform<false><<<gridn(items, 256), 256, lay.bytes, s>>>(mols, bmol, M->keys, UNIFORM_CAPS[u], lay, q, st,
                                                                   over, &cnt[K_UOVER + u]);
            else
                k_uniform<true><<<g1(std::min<uint64_t>(nb, resident(d, (const void*)k_uniform<true>, 256, lay.bytes))),
                                  256, lay.bytes, s>>>(mols, bmol, M->keys, UNIFORM_CAPS[u], lay, q, st, over,
                                                       &cnt[K_UOVER + u]);
            HIP_TRY(hipGetLastError());
            if (first) TRY(rec(9));
            first = false;
            list = over;
            count = &cnt[K_UOVER + u];
        }
        if (first) TRY(rec(9));
        Queue rest{list, count, &cnt[K_Q_SUBST]};
        uint32_t grid = g1(std::min<uint64_t>({nb, resident(d, (const void*)k_sub_states, WAVE, 0), (uint64_t)WAVE_PASS_WGS}));
        k_sub_states<<<grid, WAVE, 0, s>>>(mols, bmol, M->keys, rest, st);
        HIP_TRY(hipGetLastError());
    }
    TRY(rec(1));

    // Samples (field/kernel.rs:415-449).
    TRY(X.output(&M->values, (size_t)nb * spb));
    k_assign_uniform<<<stride_grid(nb, 1), 256, 0, s>>>(mols, bmol, st, &cnt[K_NB], b, M->values);
    const uint64_t nsub = 8 * (uint64_t)nb;
    uint32_t *sflag, *soff, *swork;
    TRY(X.temp(&sflag, nsub));
    TRY(X.temp(&soff, nsub));
    TRY(X.temp(&swork, nsub));
    k_flag_states<<<stride_grid(nsub), 256, 0, s>>>(st, &cnt[K_NSUB], sflag);
    TRY(scan_u32(X, sflag, soff, &cnt[K_NSUB], nsub, &cnt[K_NSW]));
    k_compact_index<<<stride_grid(nsub), 256, 0, s>>>(sflag, soff, &cnt[K_NSUB], swork);
    HIP_TRY(hipGetLastError());
    // Passes over the LDS tiers from the tuning's first one that fit, then
    // global scratch; each takes a list: sub-blocks predicted for it (all
    // of them in the first pass without learned sizing) followed by those
    // the passes before it could not hold. The first pass takes one
    // sub-block per workgroup (a grid of the count, or of its capacity when
    // the run is planned); later passes are persistent over their queues.
    uint32_t nsw;
    TRY(X.size(&cnt[K_NSW], (uint32_t)nsub, nsw));
    // The hidden candidates the sampling finds, per sub-block, for the
    // refinement (HidList): one pool of HID_PER_SUB entries per sampled
    // sub-block; a sub-block that finds it full gets no list.
    HidList hid{};
    hid.bsup = bsup_s;
    TRY(X.temp(&hid.off, nsub));
    TRY(X.temp(&hid.cnt, nsub));
    const uint64_t hcap = std::min<uint64_t>((uint64_t)HID_PER_SUB * std::max<uint32_t>(nsw, 1u), 0xfffffff0ull);
    TRY(X.temp(&hid.ids, hcap));
    hid.used = &cnt[K_HID];
    hid.cap = (uint32_t)hcap;
    k_fill_u32<<<stride_grid(nsub), 256, 0, s>>>(hid.cnt, NO_LIST, nsub, &cnt[K_NSUB]);
    HIP_TRY(hipGetLastError());
    // The refinement's candidate supersets (SupList), SUP_PER_SUB entries per
    // sampled sub-block; not kept by allocation-free runs, whose fixed
    // scratch they would crowd.
    if (!X.arena) {
        TRY(X.temp(&hid.sup.off, nsub));
        TRY(X.temp(&hid.sup.cnt, nsub));
        const uint64_t scap = std::min<uint64_t>((uint64_t)SUP_PER_SUB * std::max<uint32_t>(nsw, 1u), 0xfffffff0ull);
        TRY(X.temp(&hid.sup.ids, scap));
        hid.sup.used = &cnt[K_SUP];
        hid.sup.cap = (uint32_t)scap;
        k_fill_u32<<<stride_grid(nsub), 256, 0, s>>>(hid.sup.cnt, NO_LIST, nsub, &cnt[K_NSUB]);
        HIP_TRY(hipGetLastError());
        // The sampling's blocker records for the refinement (BlkRecord): a
        // slot of BLK_PER_SUB words per sampled sub-block, at its dense
        // index; a record that does not fit is not kept and that
        // sub-block's refinement builds its own bitsets.
        hid.blk.slot = soff;
        hid.blk.words = BLK_PER_SUB;
        hid.blk.slots = std::max<uint32_t>(nsw, 1u);
        TRY(X.temp(&hid.blk.cnt, nsub));
        TRY(X.temp(&hid.blk.ids, (size_t)BLK_PER_SUB * hid.blk.slots));
        k_fill_u32<<<stride_grid(nsub), 256, 0, s>>>(hid.blk.cnt, NO_LIST, nsub, &cnt[K_NSUB]);
        HIP_TRY(hipGetLastError());
    }
    {
        bool runs[TIERS];
        for (int p = 0; p < TIERS; p++) {
            runs[p] = p == TIERS - 1 || (p >= tune.sample_first_tier && LP.sample_fits[p]);
        }
        int first = 0;
        while (!runs[first]) first++;
        uint32_t* lists[TIERS + 1] = {};
        for (int p = 0; p < TIERS; p++)
            if (runs[p] && (p != first || pred)) TRY(X.temp(&lists[p], nsub));
        TRY(X.temp(&lists[TIERS], nsub));  // overflow of the last pass (an error)
        const uint32_t* count0 = &cnt[K_NSW];
        if (pred) {
            TierSplit T{};
            for (int t = 0; t < TIERS; t++) {
                int p = std::max(t, first);
                while (!runs[p]) p++;
                T.list[t] = lists[t];
                T.pass[t] = (uint32_t)p;
                T.cap[t] = SAMPLE_LOAD_CAP[t];
            }
            T.scale = sz.predict->scale;
            T.max_tier = sz.predict->max_tier;
            k_split_tiers<<<stride_grid(nsub), 256, 0, s>>>(swork, &cnt[K_NSW], pred, T, &cnt[K_SLIST0], &cnt[K_SPRED0]);
            HIP_TRY(hipGetLastError());
            count0 = &cnt[K_SLIST0 + first];
        } else {
            lists[first] = swork;
        }
        TRY(rec(10));
        for (int pass = first; pass < TIERS; pass++) {
            if (!runs[pass]) continue;
            const Caps& cap = SAMPLE_CAPS[pass];
            Layout lay = make_layout(cap);
            int nx = pass + 1;
            while (nx < TIERS && !runs[nx]) nx++;
            uint32_t* next = lists[nx];
            uint32_t* nnext = nx < TIERS ? &cnt[K_SLIST0 + nx] : &cnt[K_SLAST];
            Queue q{lists[pass], pass == first ? count0 : &cnt[K_SLIST0 + pass], &cnt[K_Q_SAMPLE0 + pass]};
            // One sub-block per workgroup where the grid can be the count:
            // every LDS pass of an exact run (its count read back), the
            // first pass of a planned one (its capacity).
            uint32_t items = (uint32_t)nsub;
            if (!X.planned()) {
                TRY(X.size(q.count, 0, items));
                if (items == 0 && pass != first) continue;
            } else if (pass == first) {
                items = nsw;
            }
            // 256 threads in every pass: the lane kernel takes one sample per
            // thread (at most (sub + 1)^3 per sub-block), so wider
            // workgroups would idle (measured: no gain on the large tiers).
            if (pass < TIERS - 1 && (!X.planned() || pass == first) && fits_dispatch(items, 256)) {
                k_sample<false, false><<<gridn(items, 256), 256, lay.bytes, s>>>(mols, bmol, cap, lay, M->keys, st, q,
                                                                         M->values, nullptr, 0, next, nnext, nsamp,
                                                                         cload, hid);
            } else if (pass < TIERS - 1) {
                uint32_t grid =
                    g1(std::min<uint64_t>(nsub, resident(d, (const void*)k_sample<false, true>, 256, lay.bytes)));
                k_sample<false, true><<<grid, 256, lay.bytes, s>>>(mols, bmol, cap, lay, M->keys, st, q, M->values,
                                                                   nullptr, 0, next, nnext, nsamp, cload, hid);
            } else {
                uint32_t grid = g1(std::min<uint64_t>(items, scratch_wgs));
                char* scratch;
                TRY(X.temp(&scratch, (size_t)grid * lay.bytes));
                k_sample<true, true><<<grid, 256, 0, s>>>(mols, bmol, cap, lay, M->keys, st, q, M->values, scratch,
                                                          lay.bytes, next, nnext, nsamp, cload, hid);
            }
            HIP_TRY(hipGetLastError());
            if (pass == first) TRY(rec(11));
        }
    }
    if (sz.collect) {
        k_labels<<<stride_grid(ntouched), 256, 0, s>>>(keep, koff, &cnt[K_TOUCHED], cload, &cnt[K_NB], cocc, ctgt);
        HIP_TRY(hipGetLastError());
    }
    TRY(rec(2));

    // Crossing edges and vertex numbering (mesh/extract.rs:176-216, 606-627).
    const uint32_t words = (3 * b * b * b + 63) / 64;
    unsigned long long* mask;
    uint32_t *wpre, *bcount, *scount, *vbase, *tcount;
    TRY(X.temp(&mask, (size_t)nb * words));
    TRY(X.temp(&wpre, (size_t)nb * words));
    TRY(X.temp(&bcount, nb));
    TRY(X.temp(&scount, nsub));
    TRY(X.temp(&vbase, (size_t)nb + 1));
    TRY(X.temp(&tcount, nb));  // triangles per brick, for the cell pass
    if (!LP.edges_staged) return fail(SMG_E_INTERNAL, "the edge pass's sign bits do not fit in LDS");
    k_edges<<<stride_grid(nb, 1), 256, LP.edges_dyn, s>>>(b, M->values, &cnt[K_NB], words, mask, wpre, bcount, scount,
                                                          tcount);
    TRY(scan_u32(X, bcount, vbase, &cnt[K_NB], nb, &cnt[K_NV]));
    k_clamp<<<1, 64, 0, s>>>(&cnt[K_NV], E.nv, &cnt[K_FLAGS], OVER_VERTICES, nullptr, 0u);
    uint32_t *rflag, *roff, *rwork;
    TRY(X.temp(&rflag, nsub));
    TRY(X.temp(&roff, nsub));
    TRY(X.temp(&rwork, nsub));
    k_flag_nonzero<<<stride_grid(nsub), 256, 0, s>>>(scount, &cnt[K_NSUB], rflag);
    TRY(scan_u32(X, rflag, roff, &cnt[K_NSUB], nsub, &cnt[K_NRW]));
    k_compact_index<<<stride_grid(nsub), 256, 0, s>>>(rflag, roff, &cnt[K_NSUB], rwork);
    HIP_TRY(hipGetLastError());
    uint32_t nverts, nrw;
    TRY(X.size(&cnt[K_NV], E.nv, nverts));
    TRY(X.size(&cnt[K_NRW], (uint32_t)nsub, nrw));
    if (report_phases) {
        report(d, SMG_STAGE_MESH, nb / 2, nb);  // sampled; edges numbered
        TRY(check_cancel(d));
    }
    TRY(rec(3));

    // Refinement and attributes (mesh/extract.rs:425-597).
    TRY(X.output(&M->pos, 3 * (size_t)nverts));
    TRY(X.output(&M->nrm, 3 * (size_t)nverts));
    TRY(X.output(&M->atom, nverts));
    TRY(X.output(&M->patch, nverts));
    TRY(X.output(&M->t, nverts));
    TRY(X.output(&M->vevals, nverts));
    VOut vo{M->pos, M->nrm, M->atom, M->patch, M->t, M->vevals, nverts};
    {
        const uint32_t sub_edges = 3 * (b / 2) * (b / 2) * (b / 2);
        bool runs[TIERS];
        uint32_t bytes_of[TIERS];
        RCaps caps[TIERS];
        RLayout lays[TIERS];
        FLayout flays[TIERS];
        for (int p = 0; p < TIERS; p++) {
            RCaps cap = REFINE_CAPS[p];
            bool lds = p < TIERS - 1;
            cap.edges = lds ? std::min(sub_edges, LDS_EDGES) : sub_edges;
            cap.gmax = (uint32_t)tune.refine_group_max;
            caps[p] = cap;
            lays[p] = make_rlayout(cap);
            flays[p] = make_flayout(cap, lays[p].bytes);
            bytes_of[p] = fast ? flays[p].bytes : lays[p].bytes;
            runs[p] = !lds || (p >= tune.refine_first_tier && LP.refine_fits[p]);
        }
        int first = 0;
        while (!runs[first]) first++;
        uint32_t* lists[TIERS + 1] = {};
        for (int p = 0; p < TIERS; p++)
            if (runs[p] && (p != first || pred)) TRY(X.temp(&lists[p], nsub));
        TRY(X.temp(&lists[TIERS], nsub));
        const uint32_t* count0 = &cnt[K_NRW];
        if (pred) {
            TierSplit T{};
            for (int t = 0; t < TIERS; t++) {
                int p = std::max(t, first);
                while (!runs[p]) p++;
                T.list[t] = lists[t];
                T.pass[t] = (uint32_t)p;
                T.cap[t] = REFINE_LOAD_CAP[t];
            }
            T.scale = sz.predict->scale;
            T.max_tier = sz.predict->max_tier;
            k_split_tiers<<<stride_grid(nsub), 256, 0, s>>>(rwork, &cnt[K_NRW], pred, T, &cnt[K_RLIST0], &cnt[K_RPRED0]);
            HIP_TRY(hipGetLastError());
            count0 = &cnt[K_RLIST0 + first];
        } else {
            lists[first] = rwork;
        }
        TRY(rec(12));
        for (int pass = first; pass < TIERS; pass++) {
            if (!runs[pass]) continue;
            const RCaps& cap = caps[pass];
            const RLayout& lay = lays[pass];
            const FLayout& flay = flays[pass];
            const uint32_t bytes = bytes_of[pass];
            int nx = pass + 1;
            while (nx < TIERS && !runs[nx]) nx++;
            uint32_t* next = lists[nx];
            uint32_t* nnext = nx < TIERS ? &cnt[K_RLIST0 + nx] : &cnt[K_RLAST];
            Queue q{lists[pass], pass == first ? count0 : &cnt[K_RLIST0 + pass], &cnt[K_Q_REFINE0 + pass]};
            uint32_t items = (uint32_t)nsub;
            if (!X.planned()) {
                TRY(X.size(q.count, 0, items));
                if (items == 0 && pass != first) continue;
            } else if (pass == first) {
                items = nrw;
            }
            // The large tiers (one or two workgroups per CU) run WIDE
            // threads: more wavefronts per SIMD, more edges in flight.
            const bool wide = pass >= WIDE_FROM;
            if (pass < TIERS - 1 && (!X.planned() || pass == first) && fits_dispatch(items, wide ? WIDE : 256)) {
                if (fast && wide)
                    k_refine_fast_wide<false, false><<<gridn(items, WIDE), WIDE, bytes, s>>>(
                        mols, bmol, cap, lay, flay, M->keys, M->values, q, hid, mask, wpre, words, vbase, vo, nullptr, 0,
                        next, nnext, evd);
                else if (fast)
                    k_refine_fast<false, false><<<gridn(items, 256), 256, bytes, s>>>(mols, bmol, cap, lay, flay, M->keys,
                                                                              M->values, q, hid, mask, wpre, words, vbase,
                                                                              vo, nullptr, 0, next, nnext, evd);
                else if (wide)
                    k_refine_wide<false, false><<<gridn(items, WIDE), WIDE, bytes, s>>>(mols, bmol, cap, lay, M->keys,
                                                                               M->values, q, hid, mask, wpre, words, vbase,
                                                                               vo, nullptr, 0, next, nnext, evd);
                else
                    k_refine<false, false><<<gridn(items, 256), 256, bytes, s>>>(mols, bmol, cap, lay, M->keys, M->values, q, hid,
                                                                         mask, wpre, words, vbase, vo, nullptr, 0,
                                                                         next, nnext, evd);
            } else if (pass < TIERS - 1) {
                const void* kern = fast ? (wide ? (const void*)k_refine_fast_wide<false, true>
                                                : (const void*)k_refine_fast<false, true>)
                                        : (wide ? (const void*)k_refine_wide<false, true>
                                                : (const void*)k_refine<false, true>);
                const uint32_t block = wide ? WIDE : 256;
                uint32_t grid = g1(std::min<uint64_t>(nsub, resident(d, kern, block, bytes)));
                if (fast && wide)
                    k_refine_fast_wide<false, true><<<grid, WIDE, bytes, s>>>(mols, bmol, cap, lay, flay, M->keys,
                                                                              M->values, q, hid, mask, wpre, words, vbase,
                                                                              vo, nullptr, 0, next, nnext, evd);
                else if (fast)
                    k_refine_fast<false, true><<<grid, 256, bytes, s>>>(mols, bmol, cap, lay, flay, M->keys,
                                                                        M->values, q, hid, mask, wpre, words, vbase, vo,
                                                                        nullptr, 0, next, nnext, evd);
                else if (wide)
                    k_refine_wide<false, true><<<grid, WIDE, bytes, s>>>(mols, bmol, cap, lay, M->keys, M->values, q, hid,
                                                                         mask, wpre, words, vbase, vo, nullptr, 0,
                                                                         next, nnext, evd);
                else
                    k_refine<false, true><<<grid, 256, bytes, s>>>(mols, bmol, cap, lay, M->keys, M->values, q, hid, mask,
                                                                   wpre, words, vbase, vo, nullptr, 0, next, nnext,
                                                                   evd);
            } else {
                uint32_t grid = g1(std::min<uint64_t>(items, scratch_wgs));
                char* scratch;
                TRY(X.temp(&scratch, (size_t)grid * bytes));
                if (fast)
                    k_refine_fast_wide<true, true><<<grid, WIDE, 0, s>>>(mols, bmol, cap, lay, flay, M->keys,
                                                                         M->values, q, hid, mask, wpre, words, vbase, vo,
                                                                         scratch, bytes, next, nnext, evd);
                else
                    k_refine_wide<true, true><<<grid, WIDE, 0, s>>>(mols, bmol, cap, lay, M->keys, M->values, q, hid, mask,
                                                                    wpre, words, vbase, vo, scratch, bytes, next,
                                                                    nnext, evd);
            }
            HIP_TRY(hipGetLastError());
            if (pass == first) TRY(rec(13));
        }
    }
    TRY(rec(4));

    // Cell pass and assembly (mesh/extract.rs:226-265, 606-691).
    uint32_t* tbase;
    TRY(X.temp(&tbase, (size_t)nb + 1));
    // The emit kernel stages a brick's samples and the edge masks it
    // references in LDS when they fit.
    const uint32_t cell_lds = LP.cells_dyn;
    const uint32_t staged = LP.cells_staged;
    TRY(scan_u32(X, tcount, tbase, &cnt[K_NB], nb, &cnt[K_NT]));
    k_clamp<<<1, 64, 0, s>>>(&cnt[K_NT], E.nt, &cnt[K_FLAGS], OVER_TRIANGLES, nullptr, 0u);
    HIP_TRY(hipGetLastError());
    uint32_t ntris;
    TRY(X.size(&cnt[K_NT], E.nt, ntris));
    if (report_phases) {
        report(d, SMG_STAGE_MESH, nb - nb / 10, nb);  // refined
        TRY(check_cancel(d));
    }
    TRY(X.output(&M->tris, 3 * (size_t)ntris));
    int32_t* tgt;
    TRY(X.temp(&tgt, 8 * (size_t)nb));
    k_cell_targets<<<stride_grid(8 * (uint64_t)nb), 256, 0, s>>>(mols, bmol, fbrick, M->keys, dense, &cnt[K_NB], tgt);
    k_cell_emit<<<stride_grid(nb, 1), 256, staged ? cell_lds : 0, s>>>(b, tgt, M->values, &cnt[K_NB], tbase, vbase,
                                                                        mask, wpre, words, M->tris, ntris, &cnt[K_ERR],
                                                                        staged);
    HIP_TRY(hipGetLastError());
    TRY(rec(5));
    k_mol_summary<<<stride_grid(rows + 1), 256, 0, s>>>(rows, fbrick, &cnt[K_NB], soff, &cnt[K_NSW], vbase,
                                                          &cnt[K_NV], tbase, &cnt[K_NT], sfirst, vfirst, trfirst);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(sum_host, sum, UL.bytes, hipMemcpyDeviceToHost, s));
    if (sz.collect && ntouched) {
        smg_model* mo = sz.collect;
        TRY(grow_dataset(mo, mo->n + ntouched));
        HIP_TRY(hipMemcpyAsync(mo->X + mo->n * MLP_IN, cfeat, (size_t)ntouched * MLP_IN * 4, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(mo->occ + mo->n, cocc, ntouched, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(mo->tgt + mo->n, ctgt, (size_t)ntouched * 4, hipMemcpyDeviceToDevice, s));
        mo->n += ntouched;
    }
    return SMG_OK;
}

// Reads a finished run's summary into the mesh: per-molecule parts and
// counts. `flags` receives the capacity-overflow bits (the parts are then
// not set).
int finish(smg_mesh* M, const Table& t, const char* sum_host, bool timed, smg_mesh_times* times, uint32_t& flags) {
    SumLayout UL = sum_layout(t.rows);
    const unsigned long long* nsamp = (const unsigned long long*)(sum_host + UL.nsamp);
    const unsigned long long* evd = (const unsigned long long*)(sum_host + UL.evals);
    const uint32_t* cnt = (const uint32_t*)(sum_host + UL.cnt);
    const uint32_t* tf = (const uint32_t*)(sum_host + UL.tfirst);
    const uint32_t* fb = (const uint32_t*)(sum_host + UL.fbrick);
    const uint32_t* sf = (const uint32_t*)(sum_host + UL.sfirst);
    const uint32_t* vf = (const uint32_t*)(sum_host + UL.vfirst);
    const uint32_t* trf = (const uint32_t*)(sum_host + UL.trfirst);
    flags = cnt[K_FLAGS];
    if (flags) return SMG_OK;
    if (cnt[K_SLAST])
        return fail(SMG_E_INTERNAL, std::to_string(cnt[K_SLAST]) + " sub-blocks exceed the largest candidate lists");
    if (cnt[K_RLAST])
        return fail(SMG_E_INTERNAL, std::to_string(cnt[K_RLAST]) + " sub-blocks exceed the largest refinement lists");
    // Retries: sub-blocks a pass took after an earlier pass could not hold them.
    uint64_t sretry = 0, rretry = 0;
    for (int p = 0; p < TIERS; p++) {
        sretry += cnt[K_SLIST0 + p] - cnt[K_SPRED0 + p];
        rretry += cnt[K_RLIST0 + p] - cnt[K_RPRED0 + p];
    }
    if (cnt[K_ERR]) return fail(SMG_E_INTERNAL, "a triangle references an edge without a vertex");
    const size_t nm = t.nm;
    M->info.assign(nm, smg_mesh_info{});
    M->bfirst.assign(fb, fb + nm + 1);
    M->vfirst.assign(vf, vf + nm + 1);
    M->tfirst.assign(trf, trf + nm + 1);
    for (size_t m = 0; m < nm; m++) {
        smg_mesh_info& I = M->info[m];
        I.touched_bricks = tf[m + 1] - tf[m];
        I.surface_bricks = M->bfirst[m + 1] - M->bfirst[m];
        I.vertices = M->vfirst[m + 1] - M->vfirst[m];
        I.triangles = M->tfirst[m + 1] - M->tfirst[m];
        I.refine_evaluations = evd[m];
        I.sampled_subblocks = sf[m + 1] - sf[m];
        I.overflow_subblocks = nm == 1 ? sretry : 0;
        I.sample_retries = sretry;
        I.refine_retries = rretry;
        I.evaluated_samples = nsamp[m];
        for (int q = 0; q < 3; q++) I.frame[q] = t.H[m].P.frame[q];
        I.band = t.H[m].P.band;
        I.edge_clamp = t.H[m].P.t_min;
    }
    if (timed && times) {
        hipEvent_t* ev = M->dev->ev;
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

}  // namespace

// The stream-ordered pool keeps what runs free for the next run. After a
// run whose temporaries held more than an eighth of the device's memory
// (a batch of thousands of molecules), all but 8 GiB of it goes back to
// the device, so a plan or another resident can take it.
static void trim_pool(smg_device* d) {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, d->ordinal) != hipSuccess) return;
    uint64_t reserved = 0;
    if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) != hipSuccess) return;
    const uint64_t keep = 8ull << 30;
    if (reserved > d->prop.totalGlobalMem / 8 && reserved > keep) (void)hipMemPoolTrimTo(pool, keep);
    (void)hipGetLastError();
}

// The learned sizing a run asks for (params[0].sizing: 0 none, 1 the
// device's sizing model).
static int run_sizing(smg_device* d, const smg_mesh_params* prm, size_t nm, Sizing& sz) {
    for (size_t m = 0; m < nm; m++)
        if (prm[m].sizing != prm[0].sizing) return fail(SMG_E_UNSUPPORTED, "one sizing mode per run");
    if (prm[0].sizing == 1) {
        if (!d->model) return fail(SMG_E_INVALID, "sizing 1 needs a sizing model on the device (smg_device_set_model)");
        sz.predict = d->model;
    } else if (prm[0].sizing != 0) {
        return fail(SMG_E_INVALID, "unknown sizing mode " + std::to_string(prm[0].sizing));
    }
    return SMG_OK;
}

// The exact path: sizes read back from the device as the run goes. With
// `collect`, the run's labelled bricks go to that model's training set.
static int mesh_many(smg_resident* const* R, const smg_mesh_params* prm, size_t nm, smg_mesh* M,
                     smg_mesh_times* times, smg_model* collect = nullptr) {
    Table t;
    TRY(build_table(R, prm, nm, nm, t));
    smg_device* d = R[0]->dev;
    Sizing sz;
    TRY(run_sizing(d, prm, nm, sz));
    sz.collect = collect;
    HIP_TRY(hipSetDevice(d->ordinal));
    Temps T(d->stream);
    Run X;
    X.d = d;
    X.s = d->stream;
    X.T = &T;
    Extents E{t.rows, t.locals(), t.spheres(), t.words(), t.dense, NO_CAP, NO_CAP, NO_CAP};
    std::vector<char> stage(stage_layout(t.rows).bytes), sum(sum_layout(t.rows).bytes);
    TRY(prepare_run(d));
    TRY(pipeline(X, t, E, stage.data(), sum.data(), M, times != nullptr, sz));
    TRY(wait_stream(d, d->stream));
    trim_pool(d);
    uint32_t flags = 0;
    TRY(finish(M, t, sum.data(), times != nullptr, times, flags));
    if (flags) return fail(SMG_E_INTERNAL, "device counts exceeded the sizes read back");
    uint64_t nb = M->bfirst.empty() ? 0 : M->bfirst.back();
    report(d, SMG_STAGE_MESH, nb, nb);
    report(d, SMG_STAGE_ASSEMBLE, 0, 1);
    report(d, SMG_STAGE_ASSEMBLE, 1, 1);
    return SMG_OK;
}

// ------------------------------------------------- arrangement on the device
namespace {

template <class X>
int own(DevBuf& b, X** p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(X);
    HIP_TRY(hipMalloc((void**)p, bytes));
    b.ptrs.push_back(*p);
    b.bytes += bytes;
    return SMG_OK;
}

template <class K>
int sort_pairs(Run& X, const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, size_t n, uint32_t bits) {
    if (n == 0) return SMG_OK;
    size_t tmp_bytes = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, vout, n, 0u, bits, X.s));
    char* tmp;
    TRY(X.temp(&tmp, tmp_bytes));
    HIP_TRY(rocprim::radix_sort_pairs((void*)tmp, tmp_bytes, kin, kout, vin, vout, n, 0u, bits, X.s));
    return SMG_OK;
}

// A cell grid of n interleaved device points with the given shape (CellGrid
// ::new, geom/cells.rs:27-77): cells in (z, y, x) order, each listing its
// points in index order. Its starts and items are owned by `keep`.
int device_grid(Run& X, DevBuf& keep, const double* p3, size_t n, const Grid& shape, DGrid& g) {
    const uint64_t cells = (uint64_t)shape.dims[0] * shape.dims[1] * shape.dims[2];
    if (cells >= (1ull << 31)) return fail(SMG_E_LENGTH, "cell grid too large");
    uint32_t *starts, *items;
    TRY(own(keep, &starts, cells + 1));
    TRY(own(keep, &items, n));
    g.starts = (gptr<const uint32_t>)starts;
    g.items = (gptr<const uint32_t>)items;
    g.ox = shape.origin[0];
    g.oy = shape.origin[1];
    g.oz = shape.origin[2];
    g.cell = shape.cell;
    g.dx = shape.dims[0];
    g.dy = shape.dims[1];
    g.dz = shape.dims[2];
    if (n == 0) {
        HIP_TRY(hipMemsetAsync(starts, 0, (cells + 1) * 4, X.s));
        return SMG_OK;
    }
    uint32_t *count, *key, *idx, *skey, *cnt;
    TRY(X.temp(&count, cells));
    TRY(X.temp(&key, n));
    TRY(X.temp(&idx, n));
    TRY(X.temp(&skey, n));
    TRY(X.temp(&cnt, 1));
    HIP_TRY(hipMemsetAsync(count, 0, cells * 4, X.s));
    k_arr_cells<<<stride_grid(n), 256, 0, X.s>>>(p3, (uint32_t)n, g, key, idx, count);
    HIP_TRY(hipGetLastError());
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < cells) bits++;
    TRY(sort_pairs(X, key, skey, idx, items, n, bits));
    const uint32_t c32 = (uint32_t)cells;
    HIP_TRY(hipMemcpyAsync(cnt, &c32, 4, hipMemcpyHostToDevice, X.s));
    TRY(scan_u32(X, count, starts, cnt, cells, starts + cells));
    return SMG_OK;
}

// Bounds of n interleaved device points (n > 0) and, with `vals`, the
// largest of them.
int device_bounds(Run& X, const double* p3, size_t n, double lo[3], double hi[3], const double* vals, double* vmax) {
    const uint32_t blocks = std::min<uint32_t>(blocks_for(n, 256), 512);
    double* part;
    unsigned long long* vm;
    TRY(X.temp(&part, 6 * (size_t)blocks));
    TRY(X.temp(&vm, 1));
    HIP_TRY(hipMemsetAsync(vm, 0, 8, X.s));
    k_arr_bounds<<<blocks, 256, 0, X.s>>>(p3, (uint32_t)n, part, vals, vm);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(6 * (size_t)blocks);
    unsigned long long hv = 0;
    HIP_TRY(hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, X.s));
    HIP_TRY(hipMemcpyAsync(&hv, vm, 8, hipMemcpyDeviceToHost, X.s));
    TRY(wait_stream(X.d, X.s));
    for (int a = 0; a < 3; a++) {
        lo[a] = h[a];
        hi[a] = h[3 + a];
    }
    for (uint32_t b = 1; b < blocks; b++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], h[6 * b + a]);
            hi[a] = std::fmax(hi[a], h[6 * b + 3 + a]);
        }
    if (vmax) std::memcpy(vmax, &hv, 8);
    return SMG_OK;
}

// Arrangement::build_with and the field's grids (arrangement.rs:437-645,
// point.rs:213-274) on the device, into R. `undecided` is set, and nothing
// of R is usable, when a circle's exposed-arc test was left to the host.
int arrange_device(smg_device* d, const double* xyzr, size_t n_in, const std::vector<uint32_t>& active, double probe,
                   int kind, double maxe, const double lo[3], const double hi[3], smg_resident& R, bool& undecided) {
    undecided = false;
    Temps T(d->stream);
    Run X;
    X.d = d;
    X.s = d->stream;
    X.T = &T;
    const hipStream_t s = X.s;
    const uint32_t n = (uint32_t)active.size();
    DevBuf& keep = R.buf;
    DArr& A = R.A;
    const double cell = std::fmax(2.0 * maxe, 1e-6), fcell = std::fmax(0.8 * maxe, 1e-6);
    double* dxyzr;
    uint32_t *dact, *cnt;
    TRY(X.temp(&dxyzr, 4 * n_in));
    TRY(X.temp(&dact, n));
    TRY(X.temp(&cnt, 8));
    if (n_in) HIP_TRY(hipMemcpyAsync(dxyzr, xyzr, 32 * n_in, hipMemcpyHostToDevice, s));
    if (n) HIP_TRY(hipMemcpyAsync(dact, active.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(cnt, &n, 4, hipMemcpyHostToDevice, s));
    // Spheres in Morton order.
    uint64_t *mkey, *skey;
    uint32_t *mval, *orig;
    TRY(X.temp(&mkey, n));
    TRY(X.temp(&skey, n));
    TRY(X.temp(&mval, n));
    TRY(own(keep, &orig, n));
    k_arr_morton<<<stride_grid(n), 256, 0, s>>>(dxyzr, dact, n, lo[0], lo[1], lo[2], cell, mkey, mval);
    HIP_TRY(hipGetLastError());
    TRY(sort_pairs(X, mkey, skey, mval, orig, n, 63));
    double *cx, *cy, *cz, *c3, *r, *Rx;
    SphereRec* srec;
    TRY(own(keep, &cx, n));
    TRY(own(keep, &cy, n));
    TRY(own(keep, &cz, n));
    TRY(own(keep, &r, n));
    TRY(own(keep, &Rx, n));
    TRY(own(keep, &srec, n));
    TRY(X.temp(&c3, 3 * (size_t)n));
    k_arr_gather<<<stride_grid(n), 256, 0, s>>>(dxyzr, orig, n, probe, cx, cy, cz, c3, r, Rx, srec);
    HIP_TRY(hipGetLastError());
    Grid shape;
    shape.cell = cell;
    if (n) grid_shape(lo, hi, n, cell, shape);
    TRY(device_grid(X, keep, c3, n, shape, A.gs));
    // Containment and the full neighbour rows.
    uint8_t* contained;
    uint32_t *nbc, *nboff, *nbitems;
    TRY(X.temp(&contained, n));
    TRY(X.temp(&nbc, n));
    TRY(X.temp(&nboff, (size_t)n + 1));
    ABuild B{cx, cy, cz, Rx, orig, contained, nboff, nullptr, n, maxe};
    k_arr_contain<<<stride_grid(n, 64), 64, 0, s>>>(B, A.gs, contained);
    k_arr_neighbours<false><<<stride_grid(n, 64), 64, 0, s>>>(B, A.gs, nbc, nullptr, nullptr);
    HIP_TRY(hipGetLastError());
    TRY(scan_u32(X, nbc, nboff, cnt, n, nboff + n));
    uint32_t nbt = 0;
    TRY(read_u32(d, nboff + n, nbt));
    TRY(X.temp(&nbitems, nbt));
    k_arr_neighbours<true><<<stride_grid(n, 64), 64, 0, s>>>(B, A.gs, nullptr, nboff, nbitems);
    HIP_TRY(hipGetLastError());
    B.nb_items = nbitems;
    uint32_t* exposed;
    TRY(X.temp(&exposed, n));
    HIP_TRY(hipMemsetAsync(exposed, 0, 4 * (size_t)n, s));
    // Circles, blockers and triple points (SES only).
    uint32_t nc = 0, ncb = 0, nv = 0;
    CircleOut o{};
    double* m3 = nullptr;
    double* v3 = nullptr;
    uint32_t npairs = 0;
    uint32_t *pc = nullptr, *poff = nullptr, *pi = nullptr, *pj = nullptr;
    if (kind == SMG_KIND_SES && nbt) {
        TRY(X.temp(&pc, n));
        TRY(X.temp(&poff, (size_t)n + 1));
        k_arr_pairs<false><<<stride_grid(n), 256, 0, s>>>(B, pc, nullptr, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
        TRY(scan_u32(X, pc, poff, cnt, n, poff + n));
        TRY(read_u32(d, poff + n, npairs));
    }
    if (npairs) {
        TRY(X.temp(&pi, npairs));
        TRY(X.temp(&pj, npairs));
        k_arr_pairs<true><<<stride_grid(n), 256, 0, s>>>(B, nullptr, poff, pi, pj);
        uint32_t *kp, *nb, *nvt, *coff, *boff, *voff, *flags;
        TRY(X.temp(&kp, npairs));
        TRY(X.temp(&nb, npairs));
        TRY(X.temp(&nvt, npairs));
        TRY(X.temp(&coff, (size_t)npairs + 1));
        TRY(X.temp(&boff, (size_t)npairs + 1));
        TRY(X.temp(&voff, (size_t)npairs + 1));
        TRY(X.temp(&flags, 1));
        HIP_TRY(hipMemsetAsync(flags, 0, 4, s));
        const uint32_t cgrid = g1(std::min<uint64_t>(blocks_for(npairs, ARR_WAVES), 65536));
        k_arr_circles<false><<<cgrid, ARR_WAVES * WAVE, 0, s>>>(B, pi, pj, npairs, kp, nb, nvt, nullptr, nullptr,
                                                                 nullptr, o, flags);
        HIP_TRY(hipGetLastError());
        uint32_t fl = 0;
        TRY(read_u32(d, flags, fl));
        if (fl) {
            undecided = true;
            return SMG_OK;
        }
        HIP_TRY(hipMemcpyAsync(cnt + 1, &npairs, 4, hipMemcpyHostToDevice, s));
        TRY(scan_u32(X, kp, coff, cnt + 1, npairs, coff + npairs));
        TRY(scan_u32(X, nb, boff, cnt + 1, npairs, boff + npairs));
        TRY(scan_u32(X, nvt, voff, cnt + 1, npairs, voff + npairs));
        TRY(read_u32(d, coff + npairs, nc));
        TRY(read_u32(d, boff + npairs, ncb));
        TRY(read_u32(d, voff + npairs, nv));
        double** soa[10] = {&o.mx, &o.my, &o.mz, &o.ax, &o.ay, &o.az, &o.e1x, &o.e1y, &o.e1z, &o.rho};
        for (double** f : soa) TRY(own(keep, f, nc));
        TRY(X.temp(&m3, 3 * (size_t)nc));
        o.m3 = m3;
        TRY(own(keep, &o.rec, nc));
        TRY(own(keep, &o.ci, nc));
        TRY(own(keep, &o.cj, nc));
        TRY(own(keep, &o.cb_off, (size_t)nc + 1));
        TRY(own(keep, &o.cb_items, ncb));
        TRY(own(keep, &o.vx, nv));
        TRY(own(keep, &o.vy, nv));
        TRY(own(keep, &o.vz, nv));
        TRY(X.temp(&v3, 3 * (size_t)nv));
        o.v3 = v3;
        TRY(own(keep, &o.vrec, nv));
        TRY(own(keep, &o.vatoms, 3 * (size_t)nv));
        o.exposed = exposed;
        k_arr_circles<true><<<cgrid, ARR_WAVES * WAVE, 0, s>>>(B, pi, pj, npairs, kp, nb, nvt, coff, boff, voff, o,
                                                                flags);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(o.cb_off + nc, &ncb, 4, hipMemcpyHostToDevice, s));
    } else {
        double** soa[10] = {&o.mx, &o.my, &o.mz, &o.ax, &o.ay, &o.az, &o.e1x, &o.e1y, &o.e1z, &o.rho};
        for (double** f : soa) TRY(own(keep, f, 0));
        TRY(own(keep, &o.rec, 0));
        TRY(own(keep, &o.ci, 0));
        TRY(own(keep, &o.cj, 0));
        TRY(own(keep, &o.cb_off, 1));
        TRY(own(keep, &o.cb_items, 0));
        TRY(own(keep, &o.vx, 0));
        TRY(own(keep, &o.vy, 0));
        TRY(own(keep, &o.vz, 0));
        TRY(own(keep, &o.vrec, 0));
        TRY(own(keep, &o.vatoms, 0));
        HIP_TRY(hipMemsetAsync(o.cb_off, 0, 4, s));
    }
    // Status, the field's neighbour rows, thresholds.
    uint8_t* status;
    uint32_t *fcount, *foff, *fitems;
    SphereThr* thr;
    TRY(own(keep, &status, n));
    TRY(own(keep, &thr, n));
    TRY(X.temp(&fcount, n));
    TRY(own(keep, &foff, (size_t)n + 1));
    k_arr_status<<<stride_grid(n), 256, 0, s>>>(B, exposed, kind == SMG_KIND_SES ? 1 : 0, status, fcount, thr);
    HIP_TRY(hipGetLastError());
    TRY(scan_u32(X, fcount, foff, cnt, n, foff + n));
    uint32_t ft = 0;
    TRY(read_u32(d, foff + n, ft));
    TRY(own(keep, &fitems, ft));
    k_arr_rows<<<stride_grid(n), 256, 0, s>>>(B, status, foff, fitems);
    HIP_TRY(hipGetLastError());
    // The field's circle and vertex grids, and the finer gather grids.
    double rho_max = 0.0;
    Grid cshape, vshape, fs, fc, fv;
    cshape.cell = vshape.cell = cell;
    fs.cell = fc.cell = fv.cell = fcell;
    if (n) grid_shape(lo, hi, n, fcell, fs);
    if (nc) {
        double clo[3], chi[3];
        TRY(device_bounds(X, m3, nc, clo, chi, o.rho, &rho_max));
        grid_shape(clo, chi, nc, cell, cshape);
        grid_shape(clo, chi, nc, fcell, fc);
    }
    if (nv) {
        double vlo[3], vhi[3];
        TRY(device_bounds(X, v3, nv, vlo, vhi, nullptr, nullptr));
        grid_shape(vlo, vhi, nv, cell, vshape);
        grid_shape(vlo, vhi, nv, fcell, fv);
    }
    TRY(device_grid(X, keep, m3, nc, cshape, A.gc));
    TRY(device_grid(X, keep, v3, nv, vshape, A.gv));
    TRY(device_grid(X, keep, c3, n, fs, A.fs));
    TRY(device_grid(X, keep, m3, nc, fc, A.fc));
    TRY(device_grid(X, keep, v3, nv, fv, A.fv));
    A.cx = (gptr<const double>)cx;
    A.cy = (gptr<const double>)cy;
    A.cz = (gptr<const double>)cz;
    A.R = (gptr<const double>)Rx;
    A.r = (gptr<const double>)r;
    A.status = (gptr<const uint8_t>)status;
    A.orig = (gptr<const uint32_t>)orig;
    A.nb_off = (gptr<const uint32_t>)foff;
    A.nb_items = (gptr<const uint32_t>)fitems;
    A.mx = (gptr<const double>)o.mx;
    A.my = (gptr<const double>)o.my;
    A.mz = (gptr<const double>)o.mz;
    A.ax = (gptr<const double>)o.ax;
    A.ay = (gptr<const double>)o.ay;
    A.az = (gptr<const double>)o.az;
    A.e1x = (gptr<const double>)o.e1x;
    A.e1y = (gptr<const double>)o.e1y;
    A.e1z = (gptr<const double>)o.e1z;
    A.crho = (gptr<const double>)o.rho;
    A.ci = (gptr<const uint32_t>)o.ci;
    A.cj = (gptr<const uint32_t>)o.cj;
    A.cb_off = (gptr<const uint32_t>)o.cb_off;
    A.cb_items = (gptr<const uint32_t>)o.cb_items;
    A.vx = (gptr<const double>)o.vx;
    A.vy = (gptr<const double>)o.vy;
    A.vz = (gptr<const double>)o.vz;
    A.vatoms = (gptr<const uint32_t>)o.vatoms;
    A.sph = (gptr<const SphereRec>)srec;
    A.cir = (gptr<const CircleRec>)o.rec;
    A.ver = (gptr<const VertexRec>)o.vrec;
    A.thr = (gptr<const SphereThr>)thr;
    A.n_spheres = n;
    A.n_circles = nc;
    A.n_vertices = nv;
    A.kind = (uint32_t)kind;
    A.probe = probe;
    A.max_expanded = maxe;
    A.rho_max = rho_max;
    TRY(wait_stream(d, s));
    return SMG_OK;
}

}  // namespace

extern "C" int smg_resident_build(smg_device* dev, const double* xyzr, size_t n_in, double probe, int kind,
                                  smg_resident** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!dev) return fail(SMG_E_INVALID, "null device");
    if (!xyzr && n_in) return fail(SMG_E_INVALID, "no spheres");
    if (!(std::isfinite(probe) && probe >= 0.0)) return fail(SMG_E_INVALID, "probe radius must be finite and >= 0");
    if (kind < SMG_KIND_SES || kind > SMG_KIND_VDW) return fail(SMG_E_INVALID, "unknown surface kind " + std::to_string(kind));
    if (kind == SMG_KIND_SES && probe == 0.0) kind = SMG_KIND_VDW;  // field/point.rs:224-228
    if (kind == SMG_KIND_VDW) probe = 0.0;                        // SurfaceParams::effective_probe
    std::vector<uint32_t> active;
    double maxe = 0.0, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (size_t i = 0; i < n_in; i++) {
        const double* s = xyzr + 4 * i;
        if (!(std::isfinite(s[0]) && std::isfinite(s[1]) && std::isfinite(s[2])))
            return fail(SMG_E_INVALID, "sphere " + std::to_string(i) + " has a non-finite centre");
        if (!(std::isfinite(s[3]) && s[3] >= 0.0))
            return fail(SMG_E_INVALID, "sphere " + std::to_string(i) + " has an invalid radius");
        if (!(s[3] > 0.0)) continue;
        if (active.empty())
            for (int a = 0; a < 3; a++) lo[a] = hi[a] = s[a];
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], s[a]);
            hi[a] = std::fmax(hi[a], s[a]);
        }
        maxe = std::fmax(maxe, s[3] + probe);
        active.push_back((uint32_t)i);
    }
    if (active.size() >= (1u << 30)) return fail(SMG_E_LENGTH, "arrangement too large for 30-bit candidate indices");
    HIP_TRY(hipSetDevice(dev->ordinal));
    const auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<smg_resident> r(new smg_resident());
    r->dev = dev;
    r->n_spheres = active.size();
    for (int a = 0; a < 3; a++) {
        r->lo[a] = lo[a];
        r->hi[a] = hi[a];
    }
    bool undecided = false;
    const int rc = arrange_device(dev, xyzr, n_in, active, probe, kind, maxe, lo, hi, *r, undecided);
    if (rc != SMG_OK) {
        r->buf.release();
        return rc;
    }
    if (undecided) {  // the host's atan2 decides: the whole arrangement on the host
        r->buf.release();
        r.reset();
        smg_arrangement* a = nullptr;
        TRY(smg_arrangement_build_kind(xyzr, n_in, probe, kind, 0, &a));
        std::unique_ptr<smg_arrangement> ha(a);
        smg_arrangement_ref ref;
        TRY(smg_arrangement_view(a, &ref));
        smg_resident* h = nullptr;
        TRY(smg_resident_new(dev, &ref, &h));
        r.reset(h);
    } else {
        r->on_device = true;
    }
    r->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = r.release();
    return SMG_OK;
}

extern "C" int smg_resident_info_get(const smg_resident* r, smg_resident_info* info) {
    if (!r || !info) return fail(SMG_E_INVALID, "null resident or info");
    info->n_spheres = r->A.n_spheres;
    info->n_circles = r->A.n_circles;
    info->n_vertices = r->A.n_vertices;
    info->kind = (int32_t)r->A.kind;
    info->built_on_device = r->on_device ? 1 : 0;
    info->probe = r->A.probe;
    info->build_ms = r->build_ms;
    info->bytes = r->buf.bytes;
    return SMG_OK;
}

// The resident's arrangement copied back to the host (both kinds of
// resident hold the same arrays).
extern "C" int smg_resident_arrangement(const smg_resident* r, smg_arrangement** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!r) return fail(SMG_E_INVALID, "null resident");
    HIP_TRY(hipSetDevice(r->dev->ordinal));
    const DArr& A = r->A;
    std::unique_ptr<smg_arrangement> h(new smg_arrangement());
    Arrangement& a = h->a;
    const size_t n = A.n_spheres, nc = A.n_circles, nv = A.n_vertices;
    auto get = [&](auto& vec, const void* src, size_t count) -> int {
        vec.resize(count);
        if (count) HIP_TRY(hipMemcpy(vec.data(), src, count * sizeof(vec[0]), hipMemcpyDeviceToHost));
        return SMG_OK;
    };
    std::vector<double> x, y, z;
    TRY(get(x, (const void*)A.cx, n));
    TRY(get(y, (const void*)A.cy, n));
    TRY(get(z, (const void*)A.cz, n));
    a.center.resize(3 * n);
    for (size_t i = 0; i < n; i++) {
        a.center[3 * i] = x[i];
        a.center[3 * i + 1] = y[i];
        a.center[3 * i + 2] = z[i];
    }
    TRY(get(a.radius, (const void*)A.r, n));
    TRY(get(a.expanded, (const void*)A.R, n));
    TRY(get(a.status, (const void*)A.status, n));
    TRY(get(a.orig, (const void*)A.orig, n));
    TRY(get(a.nb_off, (const void*)A.nb_off, n + 1));
    TRY(get(a.nb_items, (const void*)A.nb_items, a.nb_off[n]));
    std::vector<uint32_t> ci, cj;
    TRY(get(ci, (const void*)A.ci, nc));
    TRY(get(cj, (const void*)A.cj, nc));
    a.circle_ij.resize(2 * nc);
    for (size_t c = 0; c < nc; c++) {
        a.circle_ij[2 * c] = ci[c];
        a.circle_ij[2 * c + 1] = cj[c];
    }
    const void* cf[10] = {(const void*)A.mx,  (const void*)A.my,  (const void*)A.mz,  (const void*)A.ax,
                          (const void*)A.ay,  (const void*)A.az,  (const void*)A.e1x, (const void*)A.e1y,
                          (const void*)A.e1z, (const void*)A.crho};
    a.circle.resize(10 * nc);
    for (int q = 0; q < 10; q++) {
        std::vector<double> f;
        TRY(get(f, cf[q], nc));
        for (size_t c = 0; c < nc; c++) a.circle[10 * c + q] = f[c];
    }
    TRY(get(a.cb_off, (const void*)A.cb_off, nc + 1));
    TRY(get(a.cb_items, (const void*)A.cb_items, a.cb_off[nc]));
    std::vector<double> vx, vy, vz;
    TRY(get(vx, (const void*)A.vx, nv));
    TRY(get(vy, (const void*)A.vy, nv));
    TRY(get(vz, (const void*)A.vz, nv));
    a.vertex.resize(3 * nv);
    for (size_t v = 0; v < nv; v++) {
        a.vertex[3 * v] = vx[v];
        a.vertex[3 * v + 1] = vy[v];
        a.vertex[3 * v + 2] = vz[v];
    }
    TRY(get(a.vertex_atoms, (const void*)A.vatoms, 3 * nv));
    auto grid = [&](const DGrid& g, Grid& out) -> int {
        out.origin[0] = g.ox;
        out.origin[1] = g.oy;
        out.origin[2] = g.oz;
        out.cell = g.cell;
        out.dims[0] = g.dx;
        out.dims[1] = g.dy;
        out.dims[2] = g.dz;
        const size_t cells = (size_t)g.dx * g.dy * g.dz;
        TRY(get(out.starts, (const void*)g.starts, cells + 1));
        TRY(get(out.items, (const void*)g.items, out.starts[cells]));
        return SMG_OK;
    };
    TRY(grid(A.gs, a.sphere_grid));
    TRY(grid(A.gc, a.circle_grid));
    TRY(grid(A.gv, a.vertex_grid));
    a.kind = (int)A.kind;
    a.probe = A.probe;
    a.max_expanded = A.max_expanded;
    a.rho_max = A.rho_max;
    *out = h.release();
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

// ------------------------------------------------------------------ tuning
// ------------------------------------------------------ allocation-free
// Set-up of smg_mesh_into at device open, with the device current: this
// device's address of the scratch array and the pinned table and summary
// images of a one-molecule run.
static int noalloc_setup(smg_device* d) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || !ptr) {
        // The device still opens; smg_mesh_into reports this.
        d->scratch_error = std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e);
        (void)hipGetLastError();
    } else {
        d->scratch_ptr = (char*)ptr;
    }
    HIP_TRY(hipHostMalloc((void**)&d->noalloc_stage_h, stage_layout(1).bytes, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&d->noalloc_sum_h, sum_layout(1).bytes, hipHostMallocDefault));
    return SMG_OK;
}

extern "C" int smg_mesh_into(smg_resident* r, const smg_mesh_params* p, const smg_mesh_out* out, smg_mesh_info* info) {
    if (!r || !p || !out) return fail(SMG_E_INVALID, "null resident, parameters or outputs");
    if (!out->positions || !out->normals || !out->atoms || !out->patches || !out->triangles)
        return fail(SMG_E_INVALID, "every output buffer is required");
    if (out->vertex_capacity >= (1ull << 32) || out->triangle_capacity >= (1ull << 32))
        return fail(SMG_E_LENGTH, "output capacities beyond 32-bit counts");
    smg_device* d = r->dev;
    std::lock_guard<std::mutex> lk(d->noalloc_mu);
    if (!d->scratch_ptr) return fail(SMG_E_CALL, "no scratch array on this device: " + d->scratch_error);
    HIP_TRY(hipSetDevice(d->ordinal));
    Sizing sz;
    TRY(run_sizing(d, p, 1, sz));
    Table t;
    smg_resident* const R[1] = {r};
    TRY(build_table(R, p, 1, 1, t));
    const uint32_t b = t.b, spb = (b + 1) * (b + 1) * (b + 1);
    // Surface bricks are at most the molecule's brick-box cells.
    const uint64_t nb = std::max<uint64_t>(t.dense, 1);
    if (nb * spb * 4 > SMG_NOALLOC_SCRATCH_BYTES / 2)
        return fail(SMG_E_LENGTH, "molecule too large for smg_mesh_into: its brick box (" + std::to_string(t.dense) +
                                      " bricks) needs more than the " + std::to_string(SMG_NOALLOC_SCRATCH_BYTES >> 20) +
                                      " MiB scratch holds; use smg_mesh_new");
    Extents E{t.rows, t.locals(), t.spheres(), t.words(), t.dense, (uint32_t)nb, (uint32_t)out->vertex_capacity,
              (uint32_t)out->triangle_capacity};
    Run X;
    X.d = d;
    X.s = d->stream;
    X.arena = d->scratch_ptr;
    X.arena_size = SMG_NOALLOC_SCRATCH_BYTES;
    X.scratch_wgs = 16;  // the global-scratch passes take ~0.9 MB per workgroup
    smg_mesh M;
    M.dev = d;
    M.detached = true;  // nothing here is the mesh's to free
    M.pos = out->positions;
    M.nrm = out->normals;
    M.atom = out->atoms;
    M.patch = out->patches;
    M.tris = out->triangles;
    TRY(prepare_run(d));
    int rc = pipeline(X, t, E, d->noalloc_stage_h, d->noalloc_sum_h, &M, false, sz);
    if (rc != SMG_OK) {
        (void)hipStreamSynchronize(d->stream);
        return rc;
    }
    TRY(wait_stream(d, d->stream));
    uint32_t flags = 0;
    TRY(finish(&M, t, d->noalloc_sum_h, false, nullptr, flags));
    if (flags & OVER_VERTICES) return fail(SMG_E_LENGTH, "more vertices than the output's vertex capacity");
    if (flags & OVER_TRIANGLES) return fail(SMG_E_LENGTH, "more triangles than the output's triangle capacity");
    if (flags) return fail(SMG_E_INTERNAL, "a device count exceeded its capacity");
    if (info) *info = M.info[0];
    return SMG_OK;
}

extern "C" int smg_device_set_control(smg_device* d, const smg_control* c) {
    if (!d) return fail(SMG_E_INVALID, "null device");
    d->ctl = c ? *c : smg_control{};
    return SMG_OK;
}

static int check_tuning(const smg_tuning& t) {
    if (t.sample_first_tier < 0 || t.sample_first_tier > TIERS - 2 || t.refine_first_tier < 0 ||
        t.refine_first_tier > TIERS - 2)
        return fail(SMG_E_INVALID, "first tiers must be LDS tiers 0.." + std::to_string(TIERS - 2));
    int g = t.refine_group_max;
    if (g < 1 || g > WAVE || (g & (g - 1))) return fail(SMG_E_INVALID, "refine_group_max must be a power of two <= 64");
    if (t.uniform_shared != 0 && t.uniform_shared != 1) return fail(SMG_E_INVALID, "uniform_shared is 0 or 1");
    if (t.scratch_workgroups != 0 && (t.scratch_workgroups < 16 || t.scratch_workgroups > 8192))
        return fail(SMG_E_INVALID, "scratch_workgroups is 0 (default) or 16..8192");
    return SMG_OK;
}

extern "C" int smg_device_set_tuning(smg_device* d, const smg_tuning* t) {
    if (!d) return fail(SMG_E_INVALID, "null device");
    smg_tuning v = t ? *t : smg_tuning{0, 0, 64, 1, 0};
    TRY(check_tuning(v));
    d->tune = v;
    return SMG_OK;
}

extern "C" int smg_device_get_tuning(const smg_device* d, smg_tuning* t) {
    if (!d || !t) return fail(SMG_E_INVALID, "null device");
    *t = d->tune;
    return SMG_OK;
}

// Device time (ms, best of `repeats`) of the mesh stage under tuning t.
static int time_tuning(smg_resident* const* rs, const smg_mesh_params* ps, size_t count, int repeats,
                       const smg_tuning& t, double& best) {
    smg_device* d = rs[0]->dev;
    d->tune = t;
    best = INFINITY;
    for (int r = 0; r < repeats; r++) {
        smg_mesh* m = nullptr;
        smg_mesh_times tm{};
        TRY(smg_mesh_many(rs, ps, count, &m, &tm));
        smg_mesh_free(m);
        best = std::min(best, tm.total_ms);
    }
    return SMG_OK;
}

extern "C" int smg_tune(smg_resident* const* rs, const smg_mesh_params* ps, size_t count, int repeats,
                        smg_tuning* best, double* best_ms) {
    if (!rs || !ps || count == 0 || !rs[0]) return fail(SMG_E_INVALID, "no molecules");
    if (repeats < 1 || repeats > 100) return fail(SMG_E_INVALID, "repeats out of range");
    smg_device* d = rs[0]->dev;
    smg_tuning cur = d->tune, start = cur;
    double cur_ms = 0.0;
    int rc = time_tuning(rs, ps, count, repeats, cur, cur_ms);
    // Coordinate descent over the knobs, from the device's current tuning.
    const int tiers[] = {0, 1, 2};
    const int groups[] = {64, 32, 16, 8};
    const int shared[] = {1, 0};
    const int scratch[] = {0, 256, 2048};
    for (int knob = 0; knob < 5 && rc == SMG_OK; knob++) {
        const int* vals = knob < 2 ? tiers : (knob == 2 ? groups : (knob == 3 ? shared : scratch));
        int nv = knob < 2 ? 3 : (knob == 2 ? 4 : (knob == 3 ? 2 : 3));
        for (int i = 0; i < nv && rc == SMG_OK; i++) {
            smg_tuning t = cur;
            int* field = knob == 0   ? &t.sample_first_tier
                         : knob == 1 ? &t.refine_first_tier
                         : knob == 2 ? &t.refine_group_max
                         : knob == 3 ? &t.uniform_shared
                                     : &t.scratch_workgroups;
            if (*field == vals[i]) continue;
            *field = vals[i];
            double ms = 0.0;
            rc = time_tuning(rs, ps, count, repeats, t, ms);
            if (rc == SMG_OK && ms < cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
    }
    d->tune = rc == SMG_OK ? cur : start;
    if (rc != SMG_OK) return rc;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return SMG_OK;
}

// ------------------------------------------------------------ sizing model
namespace {

int grow_dataset(smg_model* m, size_t need) {
    if (need <= m->cap) return SMG_OK;
    size_t cap = std::max<size_t>(need, m->cap * 2 + 4096);
    float *X = nullptr, *tgt = nullptr;
    uint8_t* occ = nullptr;
    HIP_TRY(hipMalloc(&X, cap * MLP_IN * sizeof(float)));
    HIP_TRY(hipMalloc(&occ, cap));
    HIP_TRY(hipMalloc(&tgt, cap * sizeof(float)));
    hipStream_t s = m->dev->stream;
    if (m->n) {
        HIP_TRY(hipMemcpyAsync(X, m->X, m->n * MLP_IN * sizeof(float), hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(occ, m->occ, m->n, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(tgt, m->tgt, m->n * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    if (m->X) (void)hipFree(m->X);
    if (m->occ) (void)hipFree(m->occ);
    if (m->tgt) (void)hipFree(m->tgt);
    m->X = X;
    m->occ = occ;
    m->tgt = tgt;
    m->cap = cap;
    return SMG_OK;
}

// Scratch of the training kernels: index list, partial gradients of up to
// `nw` workgroups, their loss terms and the summed gradient.
struct TrainBufs {
    uint32_t* idx = nullptr;
    float *G = nullptr, *S = nullptr, *grad = nullptr;
    ~TrainBufs() {
        void* ps[] = {idx, G, S, grad};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
    int alloc(size_t rows, uint32_t nw) {
        HIP_TRY(hipMalloc(&idx, std::max<size_t>(rows, 1) * 4));
        HIP_TRY(hipMalloc(&G, (size_t)nw * MLP_PARAMS * 4));
        HIP_TRY(hipMalloc(&S, (size_t)nw * MLP_STATS * 4));
        HIP_TRY(hipMalloc(&grad, MLP_PARAMS * 4));
        return SMG_OK;
    }
};

constexpr float LOAD_WEIGHT = 1.0f;  // weight of the load term in the loss

// One fwd/bwd over rows idx[0..nb) of (X, occ, tgt): summed gradient into
// B.grad (when wanted) and loss terms summed into S_host; with lr > 0, an
// Adam step.
int mlp_step(smg_model* m, const float* X, const uint8_t* occ, const float* tgt, const uint32_t* idx, uint32_t nb,
             TrainBufs& B, float lr, bool want_grad, double S_host[MLP_STATS]) {
    hipStream_t s = m->dev->stream;
    uint32_t nw = (nb + MLP_ROWS - 1) / MLP_ROWS;
    k_mlp_grad<<<nw, 256, 0, s>>>(X, occ, tgt, idx, nb, LOAD_WEIGHT, m->P, m->W1b, m->W2b, B.G, B.S);
    HIP_TRY(hipGetLastError());
    float bc1 = 1.0f, bc2 = 1.0f;
    if (lr > 0.0f) {
        m->step++;
        bc1 = 1.0f - std::pow(0.9f, (float)m->step);
        bc2 = 1.0f - std::pow(0.999f, (float)m->step);
    }
    k_mlp_adam<<<(MLP_PARAMS + 255) / 256, 256, 0, s>>>(B.G, nw, m->P, m->Mo, m->Ve, m->W1b, m->W2b,
                                                        want_grad ? B.grad : nullptr, lr, 0.9f, 0.999f, 1e-8f, bc1,
                                                        bc2);
    HIP_TRY(hipGetLastError());
    if (S_host) {
        std::vector<float> st((size_t)nw * MLP_STATS);
        HIP_TRY(hipMemcpyAsync(st.data(), B.S, st.size() * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (uint32_t w = 0; w < nw; w++)
            for (int k = 0; k < MLP_STATS; k++) S_host[k] += st[(size_t)w * MLP_STATS + k];
    }
    return SMG_OK;
}

// Deterministic initial weights: uniform in +-sqrt(6 / fan_in) (He), zero biases.
void init_params(uint64_t seed, std::vector<float>& P) {
    P.assign(MLP_PARAMS, 0.0f);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    auto next = [&]() {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        return (double)((x * 0x2545F4914F6CDD1Dull) >> 11) / (double)(1ull << 53);
    };
    auto fill = [&](int off, int fan_in, int count) {
        double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; i++) P[off + i] = (float)((2.0 * next() - 1.0) * a);
    };
    fill(P_W1, MLP_IN, MLP_IN * MLP_H);
    fill(P_W2, MLP_H, MLP_H * MLP_H);
    fill(P_W3, MLP_H, MLP_H * MLP_OUT);
}

int upload_params(smg_model* m, const float* host) {
    hipStream_t s = m->dev->stream;
    HIP_TRY(hipMemcpyAsync(m->P, host, MLP_PARAMS * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(m->Mo, 0, MLP_PARAMS * 4, s));
    HIP_TRY(hipMemsetAsync(m->Ve, 0, MLP_PARAMS * 4, s));
    k_mlp_cast<<<(MLP_H * MLP_H + 255) / 256, 256, 0, s>>>(m->P, m->W1b, m->W2b);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    m->step = 0;
    return SMG_OK;
}

}  // namespace

extern "C" int smg_model_new(smg_device* dev, uint64_t seed, smg_model** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!dev) return fail(SMG_E_INVALID, "null device");
    HIP_TRY(hipSetDevice(dev->ordinal));
    std::unique_ptr<smg_model> m(new smg_model());
    m->dev = dev;
    static std::atomic<uint64_t> next_id{1};
    m->id = next_id.fetch_add(1);
    HIP_TRY(hipMalloc(&m->P, MLP_PARAMS * 4));
    HIP_TRY(hipMalloc(&m->Mo, MLP_PARAMS * 4));
    HIP_TRY(hipMalloc(&m->Ve, MLP_PARAMS * 4));
    HIP_TRY(hipMalloc(&m->W1b, MLP_IN * MLP_H * 2));
    HIP_TRY(hipMalloc(&m->W2b, MLP_H * MLP_H * 2));
    std::vector<float> P;
    init_params(seed, P);
    TRY(upload_params(m.get(), P.data()));
    *out = m.release();
    return SMG_OK;
}

extern "C" void smg_model_free(smg_model* m) {
    if (!m) return;
    (void)hipSetDevice(m->dev->ordinal);
    (void)hipStreamSynchronize(m->dev->stream);
    if (m->dev->model == m) m->dev->model = nullptr;
    delete m;
}

extern "C" size_t smg_model_param_count(void) { return MLP_PARAMS; }

extern "C" int smg_model_get_params(const smg_model* m, float* params) {
    if (!m || !params) return fail(SMG_E_INVALID, "null model");
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    HIP_TRY(hipMemcpyAsync(params, m->P, MLP_PARAMS * 4, hipMemcpyDeviceToHost, m->dev->stream));
    HIP_TRY(hipStreamSynchronize(m->dev->stream));
    return SMG_OK;
}

extern "C" int smg_model_set_params(smg_model* m, const float* params) {
    if (!m || !params) return fail(SMG_E_INVALID, "null model");
    for (int i = 0; i < MLP_PARAMS; i++)
        if (!std::isfinite(params[i])) return fail(SMG_E_INVALID, "parameters must be finite");
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    return upload_params(m, params);
}

extern "C" int smg_model_collect(smg_model* m, smg_resident* const* rs, const smg_mesh_params* ps, size_t count) {
    if (!m || !rs || !ps || count == 0) return fail(SMG_E_INVALID, "no molecules");
    for (size_t i = 0; i < count; i++)
        if (!rs[i] || rs[i]->dev != m->dev) return fail(SMG_E_INVALID, "molecule not on the model's device");
    std::vector<smg_mesh_params> q(ps, ps + count);
    for (auto& x : q) x.sizing = 0;
    std::unique_ptr<smg_mesh> M(new smg_mesh());
    M->dev = m->dev;
    TRY(mesh_many(rs, q.data(), count, M.get(), nullptr, m));
    HIP_TRY(hipStreamSynchronize(m->dev->stream));
    const double b = (double)q[0].brick;
    for (size_t i = 0; i < M->info.size(); i++) {
        const smg_mesh_info& I = M->info[i];
        const int k = smg_model::spacing_bucket(q[i].spacing);
        m->sum_vertices[k] += (double)I.vertices;
        m->sum_triangles[k] += (double)I.triangles;
        m->sum_brick_b2[k] += (double)I.surface_bricks * b * b;
    }
    return SMG_OK;
}

extern "C" size_t smg_model_samples(const smg_model* m) { return m ? m->n : 0; }

extern "C" int smg_model_dataset(const smg_model* m, float* features, uint8_t* occupied, float* log_load, size_t cap,
                                 size_t* n) {
    if (!m || !n) return fail(SMG_E_INVALID, "null model");
    *n = m->n;
    size_t k = std::min(cap, m->n);
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    if (features && k) HIP_TRY(hipMemcpy(features, m->X, k * MLP_IN * 4, hipMemcpyDeviceToHost));
    if (occupied && k) HIP_TRY(hipMemcpy(occupied, m->occ, k, hipMemcpyDeviceToHost));
    if (log_load && k) HIP_TRY(hipMemcpy(log_load, m->tgt, k * 4, hipMemcpyDeviceToHost));
    return SMG_OK;
}

extern "C" int smg_model_forward(smg_model* m, const float* features, size_t n, float* out) {
    if (!m || (!features && n) || (!out && n)) return fail(SMG_E_INVALID, "null model or arrays");
    if (n >= (1u << 28)) return fail(SMG_E_LENGTH, "too many rows");
    if (n == 0) return SMG_OK;
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    hipStream_t s = m->dev->stream;
    Temps T(s);
    float *X, *Y;
    uint32_t* nd;
    HIP_TRY(T.get((void**)&X, n * MLP_IN * 4));
    HIP_TRY(T.get((void**)&Y, n * MLP_OUT * 4));
    HIP_TRY(T.get((void**)&nd, 4));
    uint32_t n32 = (uint32_t)n;
    HIP_TRY(hipMemcpyAsync(X, features, n * MLP_IN * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(nd, &n32, 4, hipMemcpyHostToDevice, s));
    k_mlp_forward<<<g1(std::min<size_t>((n + MLP_ROWS - 1) / MLP_ROWS, 4096)), 256, 0, s>>>(X, nd, m->P, m->W1b,
                                                                                           m->W2b, Y);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, Y, n * MLP_OUT * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SMG_OK;
}

extern "C" int smg_model_gradients(smg_model* m, const float* features, const uint8_t* occupied, const float* log_load,
                                   size_t n, float* grads, double* loss) {
    if (!m || !features || !occupied || !log_load || !grads) return fail(SMG_E_INVALID, "null model or arrays");
    if (n == 0 || n > 65536) return fail(SMG_E_LENGTH, "gradient batches hold 1..65536 rows");
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    hipStream_t s = m->dev->stream;
    Temps T(s);
    float *X, *tg;
    uint8_t* oc;
    HIP_TRY(T.get((void**)&X, n * MLP_IN * 4));
    HIP_TRY(T.get((void**)&tg, n * 4));
    HIP_TRY(T.get((void**)&oc, n));
    HIP_TRY(hipMemcpyAsync(X, features, n * MLP_IN * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(tg, log_load, n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(oc, occupied, n, hipMemcpyHostToDevice, s));
    TrainBufs B;
    TRY(B.alloc(n, (uint32_t)((n + MLP_ROWS - 1) / MLP_ROWS)));
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
    HIP_TRY(hipMemcpyAsync(B.idx, idx.data(), n * 4, hipMemcpyHostToDevice, s));
    double S[MLP_STATS] = {0, 0, 0, 0};
    TRY(mlp_step(m, X, oc, tg, B.idx, (uint32_t)n, B, 0.0f, true, S));
    HIP_TRY(hipMemcpyAsync(grads, B.grad, MLP_PARAMS * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (loss) {
        loss[0] = S[0] / (double)n;
        loss[1] = 0.5 * LOAD_WEIGHT * S[1] / (double)n;
    }
    return SMG_OK;
}

extern "C" int smg_model_train(smg_model* m, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                               smg_train_stats* st) {
    if (!m) return fail(SMG_E_INVALID, "null model");
    if (m->n == 0) return fail(SMG_E_INVALID, "the model has no training samples (smg_model_collect)");
    if (batch < MLP_ROWS || batch > (1u << 20) || !(lr > 0.0f && lr < 1.0f) || epochs > 100000)
        return fail(SMG_E_INVALID, "batch 64..2^20 rows, 0 < lr < 1");
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    hipStream_t s = m->dev->stream;
    const size_t n = m->n;
    TrainBufs B;
    TRY(B.alloc(n, (batch + MLP_ROWS - 1) / MLP_ROWS));
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        return x * 0x2545F4914F6CDD1Dull;
    };
    HIP_TRY(hipEventRecord(m->dev->ev[14], s));
    for (uint32_t e = 0; e < epochs; e++) {
        for (size_t i = n - 1; i > 0; i--) std::swap(idx[i], idx[rnd() % (i + 1)]);
        HIP_TRY(hipMemcpyAsync(B.idx, idx.data(), n * 4, hipMemcpyHostToDevice, s));
        for (size_t b0 = 0; b0 < n; b0 += batch) {
            uint32_t nb = (uint32_t)std::min<size_t>(batch, n - b0);
            TRY(mlp_step(m, m->X, m->occ, m->tgt, B.idx + b0, nb, B, lr, false, nullptr));
        }
        HIP_TRY(hipStreamSynchronize(s));  // the index list is rewritten next epoch
    }
    HIP_TRY(hipEventRecord(m->dev->ev[15], s));
    if (st) {
        // Loss over the whole set with the trained weights (no update).
        for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
        HIP_TRY(hipMemcpyAsync(B.idx, idx.data(), n * 4, hipMemcpyHostToDevice, s));
        double S[MLP_STATS] = {0, 0, 0, 0};
        for (size_t b0 = 0; b0 < n; b0 += batch) {
            uint32_t nb = (uint32_t)std::min<size_t>(batch, n - b0);
            double part[MLP_STATS] = {0, 0, 0, 0};
            TRY(mlp_step(m, m->X, m->occ, m->tgt, B.idx + b0, nb, B, 0.0f, false, part));
            // k_mlp_grad scales by the batch, the loss terms are sums.
            for (int k = 0; k < MLP_STATS; k++) S[k] += part[k];
        }
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, m->dev->ev[14], m->dev->ev[15]));
        st->samples = n;
        st->epochs = epochs;
        st->batch = batch;
        st->occupancy_loss = S[0] / (double)n;
        st->occupancy_accuracy = S[2] / (double)n;
        st->load_rmse = S[3] > 0.0 ? std::sqrt(S[1] / S[3]) : 0.0;
        st->train_ms = ms;
    }
    return SMG_OK;
}

extern "C" int smg_model_set_policy(smg_model* m, float scale, int32_t max_tier) {
    if (!m) return fail(SMG_E_INVALID, "null model");
    if (!(scale > 0.0f && scale < 100.0f) || max_tier < 0 || max_tier > TIERS - 1)
        return fail(SMG_E_INVALID, "scale in (0, 100), max_tier 0.." + std::to_string(TIERS - 1));
    m->scale = scale;
    m->max_tier = (uint32_t)max_tier;
    return SMG_OK;
}

extern "C" int smg_device_set_model(smg_device* d, smg_model* m) {
    if (!d) return fail(SMG_E_INVALID, "null device");
    if (m && m->dev != d) return fail(SMG_E_INVALID, "the model lives on another device");
    d->model = m;
    return SMG_OK;
}

// ------------------------------------------------------------------ plans
struct smg_plan {
    DevRef dev;
    smg_plan_caps caps{};
    std::vector<void*> slots;
    std::vector<size_t> slot_bytes;
    char* stage_h = nullptr;  // pinned: the table image the graph uploads
    char* sum_h = nullptr;    // pinned: the summary the graph reads back
    size_t stage_bytes = 0, sum_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    smg_plan_stats stats{};
    std::vector<smg_mesh*> views;  // meshes viewing the plan's buffers
    float* lbuf = nullptr;  // slot of the f32 copies (the table image points into it)
    // Sizing the graph was recorded with (model id, 0 for none, and its
    // placement policy, which the graph holds as kernel arguments).
    uint64_t model_id = 0;
    float model_scale = 0.0f;
    uint32_t model_max_tier = 0;
    bool timed = false;          // the graph records the phase events
    void* out[9] = {};      // slots of the mesh arrays: keys, values, pos, nrm, atom, patch, t, evals, tris
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        lbuf = nullptr;
        model_id = 0;
        model_scale = 0.0f;
        model_max_tier = 0;
        timed = false;
        for (void*& o : out) o = nullptr;
        for (void* p : slots) (void)hipFree(p);
        slots.clear();
        slot_bytes.clear();
        if (stage_h) (void)hipHostFree(stage_h);
        if (sum_h) (void)hipHostFree(sum_h);
        stage_h = sum_h = nullptr;
        stage_bytes = sum_bytes = 0;
    }
    ~smg_plan() {
        for (smg_mesh* v : views) {
            v->plan = nullptr;
            v->detached = true;
        }
        release();
    }
};

smg_mesh::~smg_mesh() {
    if (vcomp) (void)hipFree(vcomp);
    if (plan) {
        auto& v = plan->views;
        v.erase(std::remove(v.begin(), v.end(), this), v.end());
        return;
    }
    if (detached) return;
    void* ps[] = {keys, values, pos, nrm, atom, patch, t, vevals, tris};
    for (void* p : ps)
        if (p) (void)hipFree(p);
}

extern "C" void smg_mesh_free(smg_mesh* m) {
    if (!m) return;
    (void)hipSetDevice(m->dev->ordinal);
    delete m;
}

namespace {

Extents plan_extents(const smg_plan_caps& c) {
    auto u32cap = [](uint64_t v) { return (uint32_t)std::min<uint64_t>(v, NO_CAP - 1); };
    return Extents{c.molecules, c.locals, c.spheres, c.box_bricks / 32 + c.molecules, c.box_bricks,
                   u32cap(c.surface_bricks), u32cap(c.vertices), u32cap(c.triangles)};
}

int check_caps(const smg_plan_caps* c) {
    if (!c) return fail(SMG_E_INVALID, "null capacities");
    if (c->molecules == 0 || c->molecules >= (1u << 24)) return fail(SMG_E_INVALID, "plan molecules out of range");
    if (c->brick < 4 || c->brick > 32 || c->brick % 2) return fail(SMG_E_UNSUPPORTED, "even brick size in 4..=32");
    if (c->refine != SMG_REFINE_EXACT && c->refine != SMG_REFINE_FAST) return fail(SMG_E_INVALID, "refinement mode");
    if (c->locals >= (1ull << 32) || c->spheres >= (1ull << 32) || c->box_bricks >= (1ull << 32) ||
        c->surface_bricks >= (1ull << 29) || c->vertices >= (1ull << 32) || c->triangles >= (1ull << 32))
        return fail(SMG_E_LENGTH, "plan capacities beyond 32-bit counts");
    return SMG_OK;
}

// Runs the plan's pipeline once on the stream (planned sizes); with
// `capture`, records it into the plan's graph instead, with the phase
// events when `timed` (event records become graph nodes).
int plan_pipeline(smg_plan* p, Table& t, smg_mesh* M, bool capture, const Sizing& sz, bool timed) {
    smg_device* d = p->dev;
    Run X;
    X.d = d;
    X.s = d->stream;
    X.slots = &p->slots;
    X.slot_bytes = &p->slot_bytes;
    Extents E = plan_extents(p->caps);
    if (!capture) {
        TRY(pipeline(X, t, E, p->stage_h, p->sum_h, M, timed, sz));
        p->lbuf = X.lbuf;
        void* o[9] = {M->keys, M->values, M->pos, M->nrm, M->atom, M->patch, M->t, M->vevals, M->tris};
        for (int q = 0; q < 9; q++) p->out[q] = o[q];
        return SMG_OK;
    }
    HIP_TRY(hipStreamBeginCapture(d->stream, hipStreamCaptureModeRelaxed));
    int rc = pipeline(X, t, E, p->stage_h, p->sum_h, M, timed, sz);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(d->stream, &g);
    if (rc == SMG_OK && e != hipSuccess) rc = fail(SMG_E_CALL, std::string("graph capture: ") + hipGetErrorString(e));
    if (rc == SMG_OK) {
        hipGraphExec_t x = nullptr;
        hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ei == hipSuccess) {
            p->graph = g;
            p->exec = x;
            p->timed = timed;
            p->stats.captures++;
            return SMG_OK;
        }
        rc = fail(SMG_E_CALL, std::string("graph instantiation: ") + hipGetErrorString(ei));
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return rc;
}

}  // namespace

extern "C" int smg_plan_new(smg_device* dev, const smg_plan_caps* caps, smg_plan** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!dev) return fail(SMG_E_INVALID, "null device");
    TRY(check_caps(caps));
    smg_plan* p = new smg_plan();
    p->dev = dev;
    p->caps = *caps;
    *out = p;
    return SMG_OK;
}

extern "C" void smg_plan_free(smg_plan* p) {
    if (!p) return;
    (void)hipSetDevice(p->dev->ordinal);
    (void)hipStreamSynchronize(p->dev->stream);
    delete p;
}

extern "C" int smg_plan_stats_get(const smg_plan* p, smg_plan_stats* s) {
    if (!p || !s) return fail(SMG_E_INVALID, "null plan");
    *s = p->stats;
    return SMG_OK;
}

extern "C" int smg_plan_caps_for(smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                                 const smg_mesh* m, double margin, smg_plan_caps* caps) {
    if (!residents || !params || !caps || count == 0) return fail(SMG_E_INVALID, "no molecules");
    if (!(margin >= 0.0 && margin <= 100.0)) return fail(SMG_E_INVALID, "margin out of range");
    Table t;
    TRY(build_table(residents, params, count, count, t));
    smg_plan_caps c{};
    c.molecules = (uint32_t)count;
    c.brick = t.b;
    c.refine = t.fast ? SMG_REFINE_FAST : SMG_REFINE_EXACT;
    c.locals = t.locals();
    c.spheres = t.spheres();
    c.box_bricks = t.dense;
    if (m) {
        if (m->info.size() != count) return fail(SMG_E_INVALID, "mesh is not of these molecules");
        auto grow = [&](uint64_t v) { return (uint64_t)std::ceil((double)v * (1.0 + margin)) + 1; };
        c.surface_bricks = grow(m->bfirst[count]);
        c.vertices = grow(m->vfirst[count]);
        c.triangles = grow(m->tfirst[count]);
    }
    *caps = c;
    return SMG_OK;
}

namespace {
// Per molecule: the sum of sigmoid(occupancy logit) and the count of
// positive logits of its touched bricks (touched[t]: a bit of the run's
// bitmaps, molecule m holding words wofs[m] .. wofs[m + 1]).
__global__ void k_pred_occupancy(const float* Y, const uint32_t* touched, const uint32_t* n_rows,
                                 const uint64_t* wofs, uint32_t nm, double* out) {
    const uint32_t n = *n_rows;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const bool on = i < n;
        double p = 0.0, c = 0.0;
        uint32_t m = 0;
        if (on) {
            const float z = Y[2 * (size_t)i];
            p = 1.0 / (1.0 + exp(-(double)z));
            c = z > 0.0f ? 1.0 : 0.0;
            m = find_mol(wofs, nm, (uint64_t)(touched[i] / 32));
        }
        add_by_key(out, 2 * m, p, on);
        add_by_key(out, 2 * m + 1, c, on);
    }
}
}  // namespace

extern "C" int smg_plan_caps_predict(smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                                     const smg_model* model, double margin, smg_plan_caps* caps,
                                     smg_sizing_estimate* est) {
    if (!residents || !params || !caps || !model || count == 0) return fail(SMG_E_INVALID, "no molecules or model");
    if (!(margin >= 0.0 && margin <= 100.0)) return fail(SMG_E_INVALID, "margin out of range");
    bool calibrated = false;
    for (int k = 0; k < smg_model::SIZING_BUCKETS; k++) calibrated |= model->sum_brick_b2[k] > 0.0;
    if (!calibrated)
        return fail(SMG_E_INVALID, "the model has collected no meshes to calibrate its capacities (smg_model_collect)");
    TRY(smg_plan_caps_for(residents, params, count, nullptr, margin, caps));
    smg_device* d = residents[0]->dev;
    if (model->dev != d) return fail(SMG_E_INVALID, "model and molecules on different devices");
    HIP_TRY(hipSetDevice(d->ordinal));
    Table t;
    TRY(build_table(residents, params, count, count, t));
    hipStream_t s = d->stream;
    HIP_TRY(hipEventRecord(d->ev[0], s));
    Temps Tm(s);
    Run X;
    X.d = d;
    X.s = s;
    X.T = &Tm;
    const uint32_t rows = (uint32_t)t.rows;
    StageLayout SL = stage_layout(rows);
    char* stage;
    float* lbuf;
    uint32_t* cnt;
    TRY(X.temp(&stage, SL.bytes));
    TRY(X.temp(&lbuf, 4 * LOC_ARRAYS * std::max<uint64_t>(t.locals(), 1)));
    TRY(X.temp(&cnt, 2));
    std::vector<char> host(SL.bytes);
    TRY(prepare_run(d));
    write_stage(t, lbuf, host.data(), d->gen);
    HIP_TRY(hipMemcpyAsync(stage, host.data(), SL.bytes, hipMemcpyHostToDevice, s));
    const RunHdr* hdr = (const RunHdr*)stage;
    k_run_begin<<<1, 1, 0, s>>>(&hdr->gen);
    const Mol* mols = (const Mol*)(stage + SL.mols);
    const uint64_t* sph_d = (const uint64_t*)(stage + SL.sph);
    const uint64_t* wofs_d = (const uint64_t*)(stage + SL.wofs);
    uint32_t *bits, *wcount, *woff, *touched;
    TRY(X.temp(&bits, t.words()));
    TRY(X.temp(&wcount, t.words()));
    TRY(X.temp(&woff, t.words()));
    TRY(X.temp(&touched, t.dense));
    k_fill_u32<<<stride_grid(t.words()), 256, 0, s>>>(bits, 0u, 0, &hdr->nwords);
    k_mark_touched<<<stride_grid(t.spheres(), 128), 128, 0, s>>>(mols, sph_d, rows, bits);
    k_popcount<<<stride_grid(t.words()), 256, 0, s>>>(bits, wcount, &hdr->nwords);
    TRY(scan_u32(X, wcount, woff, &hdr->nwords, t.words(), &cnt[0]));
    k_bits_to_list<<<stride_grid(t.words()), 256, 0, s>>>(bits, woff, &hdr->nwords, touched);
    uint32_t ntouched = 0;
    TRY(read_u32(d, &cnt[0], ntouched));
    float *feat, *pred;
    double* sums;
    TRY(X.temp(&feat, (size_t)std::max<uint32_t>(ntouched, 1) * MLP_IN));
    TRY(X.temp(&pred, (size_t)std::max<uint32_t>(ntouched, 1) * MLP_OUT));
    TRY(X.temp(&sums, 2 * (size_t)rows));
    HIP_TRY(hipMemsetAsync(sums, 0, 16 * (size_t)rows, s));
    if (ntouched) {
        k_features<<<gridn(ntouched, WAVE), WAVE, 0, s>>>(mols, wofs_d, rows, touched, nullptr, nullptr,
                                                          Queue{nullptr, &cnt[0], nullptr}, feat);
        k_mlp_forward<<<g1(std::min<uint64_t>((ntouched + MLP_ROWS - 1) / MLP_ROWS, 4096)), 256, 0, s>>>(
            feat, &cnt[0], model->P, model->W1b, model->W2b, pred);
        k_pred_occupancy<<<stride_grid(ntouched), 256, 0, s>>>(pred, touched, &cnt[0], wofs_d, rows, sums);
        HIP_TRY(hipGetLastError());
    }
    std::vector<double> h(2 * (size_t)rows);
    HIP_TRY(hipMemcpyAsync(h.data(), sums, 16 * (size_t)rows, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(d->ev[1], s));
    HIP_TRY(hipStreamSynchronize(s));
    // Each molecule's ratios from the nearest spacing the model has seen.
    const double b = (double)t.b;
    double bricks = 0.0, occupied = 0.0, verts = 0.0, tris = 0.0;
    for (size_t i = 0; i < count; i++) {
        int k0 = smg_model::spacing_bucket(params[i].spacing), k = -1;
        for (int dk = 0; dk < smg_model::SIZING_BUCKETS && k < 0; dk++)
            for (int sgn : {-1, 1}) {
                int q = k0 + sgn * dk;
                if (k < 0 && q >= 0 && q < smg_model::SIZING_BUCKETS && model->sum_brick_b2[q] > 0.0) k = q;
            }
        const double v = h[2 * i] * b * b * model->sum_vertices[k] / model->sum_brick_b2[k];
        bricks += h[2 * i];
        occupied += h[2 * i + 1];
        verts += v;
        tris += v * model->sum_triangles[k] / model->sum_vertices[k];
    }
    auto grow = [&](double v) { return (uint64_t)std::ceil(v * (1.0 + margin)) + 1; };
    caps->surface_bricks = std::min<uint64_t>(grow(bricks), t.dense);
    caps->vertices = grow(verts);
    caps->triangles = grow(tris);
    if (est) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, d->ev[0], d->ev[1]));
        est->touched_bricks = ntouched;
        est->surface_bricks = bricks;
        est->occupied_bricks = occupied;
        est->vertices = verts;
        est->triangles = tris;
        est->device_ms = ms;
    }
    return SMG_OK;
}

extern "C" int smg_plan_run(smg_plan* p, smg_resident* const* residents, const smg_mesh_params* params, size_t count,
                            smg_mesh** out, smg_mesh_times* times) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!p || !residents || !params || count == 0) return fail(SMG_E_INVALID, "no molecules");
    for (size_t m = 0; m < count; m++)
        if (!residents[m] || residents[m]->dev != p->dev) return fail(SMG_E_INVALID, "molecule not on the plan's device");
    if (!p->views.empty()) return fail(SMG_E_INVALID, "free the mesh of the plan's previous run first");
    if (params[0].brick != p->caps.brick || params[0].refine != p->caps.refine)
        return fail(SMG_E_INVALID, "brick size and refinement mode must be the plan's");
    if (times) std::memset(times, 0, sizeof(*times));
    smg_device* d = p->dev;
    HIP_TRY(hipSetDevice(d->ordinal));
    Sizing sz;
    TRY(run_sizing(d, params, count, sz));
    {  // a graph recorded with another model or policy is recorded again
        const uint64_t id = sz.predict ? sz.predict->id : 0;
        const float scale = sz.predict ? sz.predict->scale : 0.0f;
        const uint32_t tier = sz.predict ? sz.predict->max_tier : 0;
        if (p->exec && (id != p->model_id || scale != p->model_scale || tier != p->model_max_tier)) {
            HIP_TRY(hipStreamSynchronize(d->stream));
            p->release();
        }
        p->model_id = id;
        p->model_scale = scale;
        p->model_max_tier = tier;
    }
    p->stats.runs++;
    Table t;
    uint32_t flags = 0;
    if (count > p->caps.molecules) {
        flags = OVER_TABLE;
    } else {
        TRY(build_table(residents, params, count, p->caps.molecules, t));
        if (t.locals() > p->caps.locals || t.spheres() > p->caps.spheres || t.dense > p->caps.box_bricks)
            flags = OVER_TABLE;
    }
    if (!flags) {
        if (!p->stage_h) {
            p->stage_bytes = stage_layout(t.rows).bytes;
            p->sum_bytes = sum_layout(t.rows).bytes;
            HIP_TRY(hipHostMalloc((void**)&p->stage_h, p->stage_bytes, hipHostMallocDefault));
            HIP_TRY(hipHostMalloc((void**)&p->sum_h, p->sum_bytes, hipHostMallocDefault));
        }
        std::unique_ptr<smg_mesh> V(new smg_mesh());
        V->dev = d;
        V->plan = p;  // a view: the arrays are the plan's
        p->views.push_back(V.get());
        V->spb = (t.b + 1) * (t.b + 1) * (t.b + 1);
        TRY(prepare_run(d));
        report(d, SMG_STAGE_CLASSIFY, 0, 1);
        TRY(check_cancel(d));
        HIP_TRY(hipEventRecord(d->ev[14], d->stream));
        if (p->exec) {
            // The graph uploads the table image and uses the buffers of the
            // capture: refill the image, then replay.
            write_stage(t, p->lbuf, p->stage_h, d->gen);
            HIP_TRY(hipGraphLaunch(p->exec, d->stream));
            p->stats.replays++;
            V->keys = (int32_t*)p->out[0];
            V->values = (float*)p->out[1];
            V->pos = (float*)p->out[2];
            V->nrm = (float*)p->out[3];
            V->atom = (uint32_t*)p->out[4];
            V->patch = (uint8_t*)p->out[5];
            V->t = (double*)p->out[6];
            V->vevals = (uint32_t*)p->out[7];
            V->tris = (uint32_t*)p->out[8];
        } else {
            TRY(plan_pipeline(p, t, V.get(), false, sz, times != nullptr));
        }
        const bool timed = times != nullptr && (p->exec ? p->timed : true);
        HIP_TRY(hipEventRecord(d->ev[15], d->stream));
        TRY(wait_stream(d, d->stream));
        TRY(finish(V.get(), t, p->sum_h, timed, times, flags));
        if (!flags) {
            const uint64_t nb = V->bfirst.empty() ? 0 : V->bfirst.back();
            report(d, SMG_STAGE_CLASSIFY, 1, 1);
            report(d, SMG_STAGE_MESH, 0, nb);
            report(d, SMG_STAGE_MESH, nb, nb);
            report(d, SMG_STAGE_ASSEMBLE, 0, 1);
            report(d, SMG_STAGE_ASSEMBLE, 1, 1);
        }
        if (!flags) {
            if (!p->exec) {
                // Record the graph the next runs replay (nothing executes),
                // with the phase events if the runtime records them in a
                // graph, else without.
                smg_mesh C;
                C.dev = d;
                C.plan = p;
                p->views.push_back(&C);
                if (plan_pipeline(p, t, &C, true, sz, true) != SMG_OK) TRY(plan_pipeline(p, t, &C, true, sz, false));
            }
            if (times) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, d->ev[14], d->ev[15]));
                times->total_ms = ms;
            }
            *out = V.release();
            return SMG_OK;
        }
    }
    // A count beyond the plan: the exact path makes the mesh, and the plan
    // grows to hold this run (buffers and graph rebuilt on the next run).
    p->stats.fallbacks++;
    p->stats.last_overflow = flags;
    smg_mesh* M = nullptr;
    TRY(smg_mesh_many(residents, params, count, &M, times));
    smg_plan_caps c{};
    if (smg_plan_caps_for(residents, params, count, M, 0.25, &c) == SMG_OK) {
        smg_plan_caps& P = p->caps;
        P.molecules = std::max(P.molecules, c.molecules);
        P.locals = std::max(P.locals, c.locals);
        P.spheres = std::max(P.spheres, c.spheres);
        P.box_bricks = std::max(P.box_bricks, c.box_bricks);
        P.surface_bricks = std::max(P.surface_bricks, c.surface_bricks);
        P.vertices = std::max(P.vertices, c.vertices);
        P.triangles = std::max(P.triangles, c.triangles);
        if (check_caps(&P) == SMG_OK) {
            (void)hipStreamSynchronize(d->stream);
            p->release();
        }
    }
    *out = M;
    return SMG_OK;
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
    if (m->detached) return fail(SMG_E_INVALID, "the mesh viewed a plan that has been freed");
    HIP_TRY(hipSetDevice(m->dev->ordinal));
    size_t v0 = m->vfirst[part], nv = m->vfirst[part + 1] - v0;
    size_t t0 = m->tfirst[part], nt = m->tfirst[part + 1] - t0;
    size_t b0 = m->bfirst[part], nb = m->bfirst[part + 1] - b0;
    if (positions && nv) HIP_TRY(hipMemcpy(positions, m->pos + 3 * v0, 3 * nv * 4, hipMemcpyDeviceToHost));
    if ((normals && nv && !m->nrm) || (t && nv && !m->t) || (evals && nv && !m->vevals) || (bricks && nb && !m->keys) ||
        (values && nb && !m->values))
        return fail(SMG_E_INVALID, "the mesh has no such array (an uploaded mesh holds its own arrays only)");
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
// ------------------------------------------------------------ measure stage
namespace {
inline uint32_t grid_for(uint64_t n, uint32_t bs = 256) { return g1(std::min<uint64_t>(blocks_for(n, bs), 65536)); }
}  // namespace

extern "C" int smg_mesh_measure(smg_mesh* M, double* ms) {
    if (!M) return fail(SMG_E_INVALID, "null mesh");
    if (M->detached) return fail(SMG_E_INVALID, "the mesh viewed a plan that has been freed");
    if (M->measured) return SMG_OK;  // its triangles are grouped already
    smg_device* d = M->dev;
    HIP_TRY(hipSetDevice(d->ordinal));
    const size_t P = M->info.size();
    const uint64_t V = M->vfirst[P], T = M->tfirst[P];
    if (3 * T >= 0xffffffffull) return fail(SMG_E_UNSUPPORTED, "too many triangles in one mesh to measure");
    hipStream_t s = d->stream;
    TRY(prepare_run(d));
    report(d, SMG_STAGE_MEASURE, 0, 1);
    if (ms) HIP_TRY(hipEventRecord(d->ev[0], s));
    Temps Tm(s);
    Run X;
    X.d = d;
    X.s = s;
    X.T = &Tm;
    uint32_t *vf, *tf, *cnt;
    unsigned long long* val;
    TRY(X.temp(&vf, P + 1));
    TRY(X.temp(&tf, P + 1));
    TRY(X.temp(&cnt, 8));  // V, T, components, uses, segments
    TRY(X.temp(&val, MV_N * P));
    uint32_t hc[2] = {(uint32_t)V, (uint32_t)T};
    HIP_TRY(hipMemcpyAsync(vf, M->vfirst.data(), (P + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(tf, M->tfirst.data(), (P + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(cnt, hc, 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(val, 0, MV_N * P * 8, s));
    const MeasParts MP{vf, tf, (uint32_t)P};
    uint32_t *parent, *ucount, *flag, *cid, *comp;
    TRY(X.temp(&parent, V));
    TRY(X.temp(&ucount, V));
    TRY(X.temp(&flag, V));
    TRY(X.temp(&cid, V));
    TRY(X.temp(&comp, V));
    HIP_TRY(hipMemsetAsync(ucount, 0, V * 4, s));
    k_meas_vertices<<<grid_for(V), 256, 0, s>>>(M->pos, M->nrm, (uint32_t)V, MP, parent, val);
    k_meas_triangles<<<grid_for(T), 256, 0, s>>>(M->tris, M->pos, (uint32_t)T, MP, parent, ucount, val);
    HIP_TRY(hipGetLastError());
    std::vector<unsigned long long> hv(MV_N * P);
    HIP_TRY(hipMemcpyAsync(hv.data(), val, MV_N * P * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    bool invalid = false;
    for (size_t p = 0; p < P; p++) invalid |= hv[MV_N * p + MV_INVALID] != 0;
    M->minfo.assign(P, smg_measure_info{});
    auto validation = [&](size_t p) {
        smg_validation& v = M->minfo[p].validation;
        const unsigned long long* h = &hv[MV_N * p];
        v.boundary_edges = h[MV_BOUNDARY];
        v.non_manifold_edges = h[MV_NON_MANIFOLD];
        v.misoriented_edges = h[MV_MISORIENTED];
        v.degenerate_triangles = h[MV_DEGENERATE];
        v.duplicate_triangles = h[MV_DUPLICATE];
        v.invalid_indices = h[MV_INVALID];
        v.bad_vertices = h[MV_BAD];
        if (v.invalid_indices)  // mod.rs:665-667: nothing else is measured
            v.boundary_edges = v.non_manifold_edges = v.misoriented_edges = v.degenerate_triangles =
                v.duplicate_triangles = 0;
    };
    if (invalid) {
        // A mesh with an index out of range is validated only: its
        // components are not defined and its triangles stay as they are.
        for (size_t p = 0; p < P; p++) validation(p);
        return SMG_OK;
    }
    // Components: root flags, ids in root order.
    k_meas_roots<<<grid_for(V), 256, 0, s>>>(parent, (uint32_t)V, flag);
    TRY(scan_u32(X, flag, cid, &cnt[0], V, &cnt[2]));
    uint32_t ncomp = 0;
    TRY(read_u32(d, &cnt[2], ncomp));
    uint32_t *cverts, *cedges, *ctris, *cfirst, *coff;
    TRY(X.temp(&cverts, ncomp));
    TRY(X.temp(&cedges, ncomp));
    TRY(X.temp(&ctris, ncomp));
    TRY(X.temp(&cfirst, ncomp + 1));
    TRY(X.temp(&coff, P + 1));
    HIP_TRY(hipMemsetAsync(cverts, 0, (size_t)ncomp * 4, s));
    HIP_TRY(hipMemsetAsync(cedges, 0, (size_t)ncomp * 4, s));
    HIP_TRY(hipMemsetAsync(ctris, 0, (size_t)ncomp * 4, s));
    k_meas_comp<<<grid_for(V), 256, 0, s>>>(parent, cid, (uint32_t)V, comp, cverts);
    k_meas_coff<<<grid_for(P + 1), 256, 0, s>>>(cid, &cnt[2], (uint32_t)V, MP, coff);
    // Edge uses bucketed by their smaller vertex.
    uint32_t *ustart, *unext;
    unsigned long long* uses;
    TRY(X.temp(&ustart, V));
    TRY(X.temp(&unext, V));
    TRY(X.temp(&uses, 3 * T));
    TRY(scan_u32(X, ucount, ustart, &cnt[0], V, &cnt[3]));
    HIP_TRY(hipMemcpyAsync(unext, ustart, V * 4, hipMemcpyDeviceToDevice, s));
    k_meas_uses<<<grid_for(T), 256, 0, s>>>(M->tris, (uint32_t)T, MP, unext, uses);
    k_meas_buckets<<<grid_for(V), 256, 0, s>>>(ustart, ucount, uses, (uint32_t)V, MP, comp, cedges, val);
    // Triangles grouped by component (stable), and the sums in the
    // reference's chunk order.
    uint32_t *tcomp, *idx, *skey, *sidx, *gtris, *head, *segid, *segstart;
    double *A, *Vol, *segA, *segV;
    TRY(X.temp(&tcomp, T));
    TRY(X.temp(&idx, T));
    TRY(X.temp(&skey, T));
    TRY(X.temp(&sidx, T));
    TRY(X.temp(&gtris, 3 * T));
    TRY(X.temp(&head, T));
    TRY(X.temp(&segid, T));
    TRY(X.temp(&segstart, T));
    TRY(X.temp(&A, T));
    TRY(X.temp(&Vol, T));
    TRY(X.temp(&segA, T));
    TRY(X.temp(&segV, T));
    k_meas_tcomp<<<grid_for(T), 256, 0, s>>>(M->tris, (uint32_t)T, MP, comp, tcomp, idx, ctris);
    HIP_TRY(hipGetLastError());
    TRY(scan_u32(X, ctris, cfirst, &cnt[2], ncomp, &cnt[5]));
    if (T) {
        uint32_t bits = 1;
        while (bits < 32 && (1ull << bits) < ncomp) bits++;
        size_t tmp_bytes = 0;
        HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, tcomp, skey, idx, sidx, (size_t)T, 0u, bits, s));
        void* tmp;
        TRY(X.temp((char**)&tmp, tmp_bytes));
        HIP_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, tcomp, skey, idx, sidx, (size_t)T, 0u, bits, s));
        k_meas_group<<<grid_for(T), 256, 0, s>>>(sidx, skey, M->tris, M->pos, (uint32_t)T, MP, gtris, A, Vol, head);
        TRY(scan_u32(X, head, segid, &cnt[1], T, &cnt[4]));
        k_meas_segstart<<<grid_for(T), 256, 0, s>>>(head, segid, (uint32_t)T, segstart);
        k_meas_segsum<<<grid_for(T), 256, 0, s>>>(segstart, &cnt[4], (uint32_t)T, A, Vol, segA, segV);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemsetAsync(&cnt[4], 0, 4, s));
    }
    DComp* dc;
    TRY(X.temp(&dc, ncomp));
    k_meas_comps<<<grid_for(ncomp), 256, 0, s>>>(&cnt[2], cverts, ctris, cedges, cfirst, segid, &cnt[4], (uint32_t)T,
                                                segA, segV, coff, MP, dc);
    if (!M->vcomp) HIP_TRY(hipMalloc((void**)&M->vcomp, std::max<size_t>(V, 1) * 4));
    k_meas_vcomp<<<grid_for(V), 256, 0, s>>>(comp, coff, (uint32_t)V, MP, M->vcomp);
    HIP_TRY(hipGetLastError());
    if (T) HIP_TRY(hipMemcpyAsync(M->tris, gtris, 12 * T, hipMemcpyDeviceToDevice, s));
    std::vector<DComp> hc2(ncomp);
    M->cfirst_part.assign(P + 1, 0);
    HIP_TRY(hipMemcpyAsync(hc2.data(), dc, (size_t)ncomp * sizeof(DComp), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(M->cfirst_part.data(), coff, (P + 1) * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hv.data(), val, MV_N * P * 8, hipMemcpyDeviceToHost, s));
    if (ms) HIP_TRY(hipEventRecord(d->ev[1], s));
    HIP_TRY(hipStreamSynchronize(s));
    M->comps.resize(ncomp);
    for (uint32_t c = 0; c < ncomp; c++)
        M->comps[c] = smg_component{hc2[c].vertices, hc2[c].triangles, hc2[c].area, hc2[c].volume,
                                    (int64_t)hc2[c].euler, hc2[c].first_triangle};
    for (size_t p = 0; p < P; p++) {
        validation(p);
        smg_measure_info& I = M->minfo[p];
        // surface.rs:1160-1168; Rust's f64 sums fold from -0.0
        I.area = -0.0;
        I.volume = -0.0;
        for (uint32_t c = M->cfirst_part[p]; c < M->cfirst_part[p + 1]; c++) {
            const smg_component& k = M->comps[c];
            I.area += k.area;
            I.volume += k.volume;
            I.euler += k.euler;
            I.cavities += k.volume < 0.0;
        }
        I.components = M->cfirst_part[p + 1] - M->cfirst_part[p];
        I.measured = 1;
    }
    M->measured = true;
    report(d, SMG_STAGE_MEASURE, 1, 1);
    if (ms) {
        float f = 0.0f;
        HIP_TRY(hipEventElapsedTime(&f, d->ev[0], d->ev[1]));
        *ms = f;
    }
    return SMG_OK;
}

extern "C" int smg_mesh_part_measure(const smg_mesh* m, size_t part, smg_measure_info* out) {
    if (!m || !out) return fail(SMG_E_INVALID, "null mesh or output");
    if (m->minfo.empty()) return fail(SMG_E_INVALID, "the mesh has not been measured (smg_mesh_measure)");
    if (part >= m->minfo.size()) return fail(SMG_E_INVALID, "no part " + std::to_string(part));
    *out = m->minfo[part];
    return SMG_OK;
}

extern "C" int smg_mesh_part_components(const smg_mesh* m, size_t part, smg_component* comps, size_t cap,
                                        uint32_t* vertex_component) {
    if (!m) return fail(SMG_E_INVALID, "null mesh");
    if (!m->measured) return fail(SMG_E_INVALID, "the mesh has no measured components (smg_mesh_measure)");
    if (part >= m->minfo.size()) return fail(SMG_E_INVALID, "no part " + std::to_string(part));
    const size_t c0 = m->cfirst_part[part], nc = m->cfirst_part[part + 1] - c0;
    if (comps) {
        if (cap < nc) return fail(SMG_E_LENGTH, "the part has " + std::to_string(nc) + " components");
        std::copy(m->comps.begin() + c0, m->comps.begin() + c0 + nc, comps);
    }
    const size_t v0 = m->vfirst[part], nv = m->vfirst[part + 1] - v0;
    if (vertex_component && nv) {
        HIP_TRY(hipSetDevice(m->dev->ordinal));
        HIP_TRY(hipMemcpy(vertex_component, m->vcomp + v0, nv * 4, hipMemcpyDeviceToHost));
    }
    return SMG_OK;
}

extern "C" int smg_mesh_upload(smg_device* dev, const float* positions, const float* normals, const uint32_t* atoms,
                               const uint8_t* patches, size_t n, const uint32_t* triangles, size_t t, smg_mesh** out) {
    if (!out) return fail(SMG_E_INVALID, "null output");
    *out = nullptr;
    if (!dev || (n && !positions) || (t && !triangles)) return fail(SMG_E_INVALID, "null device or arrays");
    if (n >= (1ull << 31) || t >= (1ull << 31)) return fail(SMG_E_LENGTH, "mesh too large");
    HIP_TRY(hipSetDevice(dev->ordinal));
    std::unique_ptr<smg_mesh> M(new smg_mesh());
    M->dev = dev;
    smg_mesh_info I{};
    I.vertices = n;
    I.triangles = t;
    M->info.assign(1, I);
    M->bfirst = {0, 0};
    M->vfirst = {0, (uint32_t)n};
    M->tfirst = {0, (uint32_t)t};
    HIP_TRY(hipMalloc((void**)&M->pos, std::max<size_t>(n, 1) * 12));
    HIP_TRY(hipMalloc((void**)&M->atom, std::max<size_t>(n, 1) * 4));
    HIP_TRY(hipMalloc((void**)&M->patch, std::max<size_t>(n, 1)));
    HIP_TRY(hipMalloc((void**)&M->tris, std::max<size_t>(t, 1) * 12));
    if (normals) HIP_TRY(hipMalloc((void**)&M->nrm, std::max<size_t>(n, 1) * 12));
    if (n) {
        HIP_TRY(hipMemcpy(M->pos, positions, n * 12, hipMemcpyHostToDevice));
        if (normals) HIP_TRY(hipMemcpy(M->nrm, normals, n * 12, hipMemcpyHostToDevice));
        if (atoms)
            HIP_TRY(hipMemcpy(M->atom, atoms, n * 4, hipMemcpyHostToDevice));
        else
            HIP_TRY(hipMemset(M->atom, 0, n * 4));
        if (patches)
            HIP_TRY(hipMemcpy(M->patch, patches, n, hipMemcpyHostToDevice));
        else
            HIP_TRY(hipMemset(M->patch, 0, n));
    }
    if (t) HIP_TRY(hipMemcpy(M->tris, triangles, t * 12, hipMemcpyHostToDevice));
    *out = M.release();
    return SMG_OK;
}

// ----------------------------------------------------------------- writers
extern "C" void smg_export_defaults(smg_export_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = 4;
    o->normals = 1;
}

namespace {
std::string clean_line(const char* s) {
    std::string o = s ? s : "";
    for (char& c : o)
        if (c == '\n' || c == '\r') c = ' ';
    return o;
}

// One block of text records on the device: lengths, offsets, bytes.
template <int KIND>
int text_block(Run& X, const ExportArgs& a, uint32_t count, uint32_t* cnt, char** bytes, uint32_t** off,
               uint32_t& total) {
    uint32_t* len;
    TRY(X.temp(&len, count));
    TRY(X.temp(off, (size_t)count + 1));
    HIP_TRY(hipMemcpyAsync(cnt, &count, 4, hipMemcpyHostToDevice, X.s));
    k_text_len<KIND><<<stride_grid(count), 256, 0, X.s>>>(a, count, len);
    unsigned long long* sum;
    TRY(X.temp(&sum, 1));
    HIP_TRY(hipMemsetAsync(sum, 0, 8, X.s));
    k_len_total<<<stride_grid(count), 256, 0, X.s>>>(len, count, sum);
    unsigned long long big = 0;
    HIP_TRY(hipMemcpyAsync(&big, sum, 8, hipMemcpyDeviceToHost, X.s));
    HIP_TRY(hipStreamSynchronize(X.s));
    if (big >= 0xffffffffull) return fail(SMG_E_UNSUPPORTED, "a block of the file exceeds 4 GiB");
    TRY(scan_u32(X, len, *off, cnt, count, *off + count));
    TRY(read_u32(X.d, *off + count, total));
    TRY(X.temp(bytes, total));
    k_text_write<KIND><<<stride_grid(count), 256, 0, X.s>>>(a, count, *off, *bytes);
    HIP_TRY(hipGetLastError());
    return SMG_OK;
}
}  // namespace

extern "C" int smg_export(const smg_mesh* m, size_t part, int format, const smg_export_options* opt, char* buf,
                          size_t cap, size_t* len) {
    if (!m || !len) return fail(SMG_E_INVALID, "null mesh or length");
    *len = 0;
    if (part >= m->info.size()) return fail(SMG_E_INVALID, "no part " + std::to_string(part));
    if (m->detached) return fail(SMG_E_INVALID, "the mesh viewed a plan that has been freed");
    smg_export_options o;
    smg_export_defaults(&o);
    if (opt) o = *opt;
    const uint32_t v0 = m->vfirst[part], nv = m->vfirst[part + 1] - v0;
    const uint32_t t0 = m->tfirst[part], nt = m->tfirst[part + 1] - t0;
    const bool normals = o.normals && m->nrm;
    smg_device* d = m->dev;
    HIP_TRY(hipSetDevice(d->ordinal));
    Temps Tm(d->stream);
    Run X;
    X.d = d;
    X.s = d->stream;
    X.T = &Tm;
    ExportArgs a{};
    a.pos = m->pos + 3 * (size_t)v0;
    a.nrm = m->nrm ? m->nrm + 3 * (size_t)v0 : nullptr;
    a.atom = m->atom + v0;
    a.patch = m->patch + v0;
    a.vcomp = m->measured ? m->vcomp + v0 : nullptr;
    a.tris = m->tris + 3 * (size_t)t0;
    a.nv = nv;
    a.nt = nt;
    a.prec = o.precision;
    a.normals = normals;
    if (o.vertex_colors && (format == SMG_FORMAT_OBJ || format == SMG_FORMAT_PLY)) {
        float* rgb;
        TRY(X.temp(&rgb, 3 * (size_t)nv));
        HIP_TRY(hipMemcpyAsync(rgb, o.vertex_colors, 12 * (size_t)nv, hipMemcpyHostToDevice, X.s));
        a.rgb = rgb;
    }
    uint32_t* cnt;
    unsigned long long* bad;
    TRY(X.temp(&cnt, 4));
    TRY(X.temp(&bad, 1));
    HIP_TRY(hipMemsetAsync(bad, 0, 8, X.s));
    a.bad = bad;
    std::string head;
    // (host text, device bytes, size) pieces of the file, in order
    struct Piece {
        std::string text;
        const char* dev = nullptr;
        size_t bytes = 0;
    };
    std::vector<Piece> pieces;
    const std::string version = SMG_SESMESH_VERSION;
    if (format == SMG_FORMAT_OBJ) {  // export/obj.rs:138-224
        if (o.precision < 0 || o.precision > 9)
            return fail(SMG_E_INVALID, "OBJ precision must be 0..=9, got " + std::to_string(o.precision));
        for (size_t g = 1; g < o.n_groups; g++)
            if (o.group_first[g - 1] > o.group_first[g]) return fail(SMG_E_INVALID, "OBJ groups must be sorted by first triangle");
        head = "# sesmesh " + version + "\n";
        for (size_t c = 0; c < o.n_comments; c++) head += "# " + clean_line(o.comments[c]) + "\n";
        head += "o " + clean_line(o.object_name ? o.object_name : "ses") + "\n";
        pieces.push_back({head});
        char *vb, *nb = nullptr, *fb;
        uint32_t *voff, *noff, *foff, vt = 0, ntot = 0, ft = 0;
        TRY(text_block<XK_OBJ_V>(X, a, nv, cnt, &vb, &voff, vt));
        if (normals) TRY(text_block<XK_OBJ_VN>(X, a, nv, cnt, &nb, &noff, ntot));
        TRY(text_block<XK_OBJ_F>(X, a, nt, cnt, &fb, &foff, ft));
        pieces.push_back({"", vb, vt});
        if (normals) pieces.push_back({"", nb, ntot});
        // Face ranges between group starts (obj.rs:199-222).
        std::vector<std::pair<uint32_t, const char*>> bounds = {{0u, nullptr}};
        for (size_t g = 0; g < o.n_groups; g++) {
            uint32_t first = (uint32_t)std::min<uint64_t>(o.group_first[g], nt);
            if (bounds.back().first == first) bounds.pop_back();
            bounds.push_back({first, o.group_names[g] ? o.group_names[g] : ""});
        }
        std::vector<uint32_t> at(bounds.size() + 1), byte(bounds.size() + 1);
        for (size_t k = 0; k < bounds.size(); k++) at[k] = bounds[k].first;
        at[bounds.size()] = nt;
        uint32_t *idx, *got;
        TRY(X.temp(&idx, at.size()));
        TRY(X.temp(&got, at.size()));
        HIP_TRY(hipMemcpyAsync(idx, at.data(), at.size() * 4, hipMemcpyHostToDevice, X.s));
        k_gather_u32<<<1, 256, 0, X.s>>>(foff, idx, (uint32_t)at.size(), got);
        HIP_TRY(hipMemcpyAsync(byte.data(), got, at.size() * 4, hipMemcpyDeviceToHost, X.s));
        HIP_TRY(hipStreamSynchronize(X.s));
        for (size_t k = 0; k < bounds.size(); k++) {
            if (bounds[k].second) pieces.push_back({"g " + clean_line(bounds[k].second) + "\n"});
            pieces.push_back({"", fb + byte[k], (size_t)byte[k + 1] - byte[k]});
        }
    } else if (format == SMG_FORMAT_PLY) {  // export/ply.rs:97-182
        const bool colors = a.rgb != nullptr;
        head = "ply\nformat binary_little_endian 1.0\ncomment sesmesh " + version + "\n";
        for (size_t c = 0; c < o.n_comments; c++) head += "comment " + clean_line(o.comments[c]) + "\n";
        head += "element vertex " + std::to_string(nv) + "\n";
        head += "property float x\nproperty float y\nproperty float z\n";
        if (normals) head += "property float nx\nproperty float ny\nproperty float nz\n";
        if (colors) head += "property uchar red\nproperty uchar green\nproperty uchar blue\n";
        head += "property uint atom\nproperty uchar patch\n";
        if (a.vcomp) head += "property uint component\n";
        head += "element face " + std::to_string(nt) + "\nproperty list uchar uint vertex_indices\nend_header\n";
        const uint32_t vstride = 12 + (normals ? 12 : 0) + (colors ? 3 : 0) + 4 + 1 + (a.vcomp ? 4 : 0);
        const size_t body = (size_t)nv * vstride + (size_t)nt * 13;
        char* pb;
        TRY(X.temp(&pb, body));
        k_ply_write<<<stride_grid((uint64_t)nv + nt), 256, 0, X.s>>>(a, vstride, pb);
        HIP_TRY(hipGetLastError());
        pieces.push_back({head});
        pieces.push_back({"", pb, body});
    } else if (format == SMG_FORMAT_MSMS_VERT || format == SMG_FORMAT_MSMS_FACE) {  // export/msms.rs:70-148
        if (!m->nrm) return fail(SMG_E_INVALID, "MSMS output needs normals, atoms and patches for every vertex");
        const bool vert = format == SMG_FORMAT_MSMS_VERT;
        const char* what = vert ? "vertices" : "faces";
        char line[256];
        std::snprintf(line, sizeof(line), "%7u %7llu %5.2f %5.2f\n", vert ? nv : nt, (unsigned long long)o.spheres,
                      o.density, o.probe_radius);
        head = std::string("# MSMS solvent excluded surface ") + what + " for " + clean_line(o.description) + "\n#" +
               what + " #sphere density probe_r\n" + line;
        pieces.push_back({head});
        char* b;
        uint32_t* off;
        uint32_t total = 0;
        if (vert)
            TRY(text_block<XK_MSMS_V>(X, a, nv, cnt, &b, &off, total));
        else
            TRY(text_block<XK_MSMS_F>(X, a, nt, cnt, &b, &off, total));
        pieces.push_back({"", b, total});
    } else {
        return fail(SMG_E_INVALID, "unknown format " + std::to_string(format));
    }
    unsigned long long nbad = 0;
    HIP_TRY(hipMemcpyAsync(&nbad, bad, 8, hipMemcpyDeviceToHost, X.s));
    HIP_TRY(hipStreamSynchronize(X.s));
    if (nbad) return fail(SMG_E_UNSUPPORTED, "non-finite or out-of-range values in the mesh");
    size_t total = 0;
    for (const Piece& q : pieces) total += q.dev ? q.bytes : q.text.size();
    *len = total;
    if (!buf) return SMG_OK;
    if (cap < total) return fail(SMG_E_LENGTH, "the output needs " + std::to_string(total) + " bytes");
    size_t at = 0;
    for (const Piece& q : pieces) {
        if (q.dev) {
            if (q.bytes) HIP_TRY(hipMemcpyAsync(buf + at, q.dev, q.bytes, hipMemcpyDeviceToHost, X.s));
            at += q.bytes;
        } else {
            std::memcpy(buf + at, q.text.data(), q.text.size());
            at += q.text.size();
        }
    }
    HIP_TRY(hipStreamSynchronize(X.s));
    return SMG_OK;
}

extern "C" int smg_export_file(const smg_mesh* m, size_t part, int format, const smg_export_options* o,
                               const char* path) {
    if (!path) return fail(SMG_E_INVALID, "null path");
    size_t n = 0;
    TRY(smg_export(m, part, format, o, nullptr, 0, &n));
    char* host = nullptr;
    HIP_TRY(hipHostMalloc((void**)&host, std::max<size_t>(n, 1), hipHostMallocDefault));
    int rc = smg_export(m, part, format, o, host, n, &n);
    if (rc == SMG_OK) {
        FILE* f = std::fopen(path, "wb");
        if (!f || std::fwrite(host, 1, n, f) != n || std::fclose(f) != 0) {
            if (f) std::fclose(f);
            rc = fail(SMG_E_INVALID, std::string("cannot write ") + path);
        }
    }
    (void)hipHostFree(host);
    return rc;
}

// Test helpers: the LDS plan of a brick for given static LDS sizes (the
// LdsPlan fields as 30 u32), and the static sizes this device reads.
extern "C" int smg_lds_plan(uint32_t brick, int fast, uint64_t lds_max, const uint32_t* statics, uint32_t* out) {
    static_assert(sizeof(LdsPlan) == 30 * sizeof(uint32_t), "LdsPlan is 30 u32");
    if (!statics || !out || brick < 4 || brick > 32 || brick % 2) return fail(SMG_E_INVALID, "bad arguments");
    LdsPlan L = lds_plan(brick, fast != 0, lds_max, statics);
    std::memcpy(out, &L, sizeof(L));
    return SMG_OK;
}

extern "C" int smg_device_static_lds(smg_device* d, uint32_t* out) {
    if (!d || !out) return fail(SMG_E_INVALID, "null device or output");
    HIP_TRY(hipSetDevice(d->ordinal));
    const uint32_t* st;
    TRY(device_statics(d, &st));
    std::memcpy(out, st, LK_N * sizeof(uint32_t));
    return SMG_OK;
}

extern "C" int smg_mc_table(uint16_t offset[257], uint8_t* edges, size_t cap) {
    std::call_once(g_table_once, [] { make_mc_table(g_tri_offset, g_tri_edges); });
    std::memcpy(offset, g_tri_offset, sizeof(g_tri_offset));
    std::memcpy(edges, g_tri_edges.data(), std::min(cap, g_tri_edges.size()));
    return SMG_OK;
}
