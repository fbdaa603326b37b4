// Mesh writers on the device: the records of Wavefront OBJ, MSMS .vert /
// .face (text, one record per line) and binary PLY, formatted in parallel,
// byte for byte as the reference's writers (crates/sesmesh/src/export/
// fmt.rs, obj.rs, msms.rs, ply.rs). Text records take two passes: their
// lengths (a scan gives their offsets), then their bytes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace smg {

enum : int { XK_OBJ_V, XK_OBJ_VN, XK_OBJ_F, XK_MSMS_V, XK_MSMS_F };

struct ExportArgs {
    const float* pos;    // the part's vertices (3 f32 each)
    const float* nrm;
    const uint32_t* atom;
    const uint8_t* patch;
    const uint32_t* vcomp;  // component per vertex, or null (MSMS: component 1)
    const float* rgb;       // OBJ / PLY colours (3 per vertex), or null
    const uint32_t* tris;   // the part's triangles (3 vertex indices each)
    uint32_t nv, nt;
    int prec;
    int normals;  // OBJ: "f a//a" faces; PLY: normal properties
    unsigned long long* bad;  // set when a value needs the formatter's fallback
};

__constant__ unsigned long long XPOW10[10] = {1ull,      10ull,      100ull,      1000ull,      10000ull,
                                                   100000ull, 1000000ull, 10000000ull, 100000000ull, 1000000000ull};

// Counts (out == null) or writes the bytes of a record.
struct Emit {
    char* out;
    uint32_t n = 0;
    __device__ explicit Emit(char* o) : out(o) {}
    __device__ void ch(char c) {
        if (out) out[n] = c;
        n++;
    }
    __device__ void str(const char* s) {
        while (*s) ch(*s++);
    }
    __device__ static uint32_t digits(unsigned long long v) {
        uint32_t d = 1;
        while (v >= 10) {
            v /= 10;
            d++;
        }
        return d;
    }
    // Right-aligned in `width` (fmt's `{:width}` of an integer).
    __device__ void u64(unsigned long long v, uint32_t width = 0) {
        const uint32_t d = digits(v);
        for (uint32_t k = d; k < width; k++) ch(' ');
        if (!out) {
            n += d;
            return;
        }
        for (uint32_t k = 0; k < d; k++) {
            out[n + d - 1 - k] = (char)('0' + v % 10);
            v /= 10;
        }
        n += d;
    }
    // fmt.rs:30-54: x with `prec` decimals, right-aligned in `width`. x *
    // 10^prec is exact in f64 for an f32, and rint rounds it half to even.
    __device__ void fixed(float x, int prec, uint32_t width, unsigned long long* bad) {
        const double scaled = rint((double)x * (double)XPOW10[prec]);
        if (!isfinite(scaled) || fabs(scaled) >= 1e18) {
            if (bad) atomicAdd(bad, 1ull);
            ch('?');
            return;
        }
        const bool neg = signbit(x);
        const unsigned long long m = (unsigned long long)fabs(scaled);
        const unsigned long long ip = m / XPOW10[prec], fp = m % XPOW10[prec];
        const uint32_t len = (uint32_t)neg + digits(ip) + (prec > 0 ? 1u + (uint32_t)prec : 0u);
        for (uint32_t k = len; k < width; k++) ch(' ');
        if (neg) ch('-');
        u64(ip);
        if (prec > 0) {
            ch('.');
            unsigned long long f = fp;
            if (out)
                for (int k = prec - 1; k >= 0; k--) {
                    out[n + k] = (char)('0' + f % 10);
                    f /= 10;
                }
            n += prec;
        }
    }
};

__device__ inline float clamp01(float c) { return c < 0.0f ? 0.0f : c > 1.0f ? 1.0f : c; }
__device__ inline uint32_t msms_type(uint8_t patch) { return patch == 1 ? 1u : patch == 2 ? 2u : 3u; }

// Record r of a text format: OBJ v (obj.rs:178-196), vn, f (obj.rs:211-224);
// MSMS vertex (msms.rs:95-113) and face (msms.rs:121-140) lines.
template <int KIND>
__device__ inline uint32_t text_record(const ExportArgs& a, uint32_t r, char* out) {
    Emit e(out);
    if (KIND == XK_OBJ_V || KIND == XK_OBJ_VN) {
        e.str(KIND == XK_OBJ_V ? "v " : "vn ");
        const float* v = (KIND == XK_OBJ_V ? a.pos : a.nrm) + 3 * (size_t)r;
        for (int k = 0; k < 3; k++) {
            if (k) e.ch(' ');
            e.fixed(v[k], a.prec, 0, a.bad);
        }
        if (KIND == XK_OBJ_V && a.rgb)
            for (int k = 0; k < 3; k++) {
                e.ch(' ');
                e.fixed(clamp01(a.rgb[3 * (size_t)r + k]), 4, 0, a.bad);
            }
        e.ch('\n');
    } else if (KIND == XK_OBJ_F) {
        e.ch('f');
        for (int k = 0; k < 3; k++) {
            const unsigned long long i = (unsigned long long)a.tris[3 * (size_t)r + k] + 1;
            e.ch(' ');
            e.u64(i);
            if (a.normals) {
                e.ch('/');
                e.ch('/');
                e.u64(i);
            }
        }
        e.ch('\n');
    } else if (KIND == XK_MSMS_V) {
        for (int k = 0; k < 6; k++) {
            e.fixed(k < 3 ? a.pos[3 * (size_t)r + k] : a.nrm[3 * (size_t)r + k - 3], 3, 9, a.bad);
            e.ch(' ');
        }
        e.u64(a.vcomp ? (unsigned long long)a.vcomp[r] + 1 : 1ull, 7);
        e.ch(' ');
        e.u64((unsigned long long)a.atom[r] + 1, 7);
        e.ch(' ');
        e.u64(msms_type(a.patch[r]), 2);
        e.ch('\n');
    } else {
        const uint32_t* t = a.tris + 3 * (size_t)r;
        const uint32_t t0 = msms_type(a.patch[t[0]]), t1 = msms_type(a.patch[t[1]]), t2 = msms_type(a.patch[t[2]]);
        for (int k = 0; k < 3; k++) {
            e.u64((unsigned long long)t[k] + 1, 6);
            e.ch(' ');
        }
        e.u64(t0 == t1 && t1 == t2 ? t0 : 1u, 2);
        e.ch(' ');
        e.u64(a.vcomp ? (unsigned long long)a.vcomp[t[0]] + 1 : 1ull, 6);
        e.ch('\n');
    }
    return e.n;
}

template <int KIND>
__global__ void k_text_len(ExportArgs a, uint32_t count, uint32_t* len) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x)
        len[r] = text_record<KIND>(a, r, nullptr);
}

template <int KIND>
__global__ void k_text_write(ExportArgs a, uint32_t count, const uint32_t* off, char* out) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x)
        (void)text_record<KIND>(a, r, out + off[r]);
}

// Binary PLY records (ply.rs:150-181): per vertex x, y, z, [nx, ny, nz],
// [r, g, b], atom, patch, [component], little-endian; per face 3 and the
// three indices. `vstride` is the vertex record's size.
__global__ void k_ply_write(ExportArgs a, uint32_t vstride, char* out) {
    const uint64_t total = (uint64_t)a.nv + a.nt;
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < total; r += (uint64_t)gridDim.x * blockDim.x) {
        if (r < a.nv) {
            const uint32_t i = (uint32_t)r;
            char* o = out + (size_t)i * vstride;
            auto put = [&](const void* p, int k) {
                for (int q = 0; q < k; q++) o[q] = ((const char*)p)[q];
                o += k;
            };
            put(a.pos + 3 * (size_t)i, 12);
            if (a.normals) put(a.nrm + 3 * (size_t)i, 12);
            if (a.rgb)
                for (int k = 0; k < 3; k++) {
                    const uint8_t c = (uint8_t)roundf(clamp01(a.rgb[3 * (size_t)i + k]) * 255.0f);
                    put(&c, 1);
                }
            put(a.atom + i, 4);
            put(a.patch + i, 1);
            if (a.vcomp) put(a.vcomp + i, 4);
        } else {
            const uint32_t t = (uint32_t)(r - a.nv);
            char* o = out + (size_t)a.nv * vstride + (size_t)t * 13;
            o[0] = 3;
            const char* s = (const char*)(a.tris + 3 * (size_t)t);
            for (int q = 0; q < 12; q++) o[1 + q] = s[q];
        }
    }
}

// Total bytes of `count` records (u64: the offsets are u32).
__global__ void k_len_total(const uint32_t* len, uint32_t count, unsigned long long* total) {
    unsigned long long s = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x) s += len[r];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(total, s);
}

// Offsets of the given records (group boundaries of an OBJ file).
__global__ void k_gather_u32(const uint32_t* src, const uint32_t* idx, uint32_t n, uint32_t* dst) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}

}  // namespace smg
