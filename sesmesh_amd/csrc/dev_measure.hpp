// Measure stage on the device: validation, connected components with their
// sums and Euler characteristics, and triangles grouped by component, for
// every part of a mesh at once (crates/sesmesh/src/mesh/mod.rs:91-705,
// mesh/post.rs:14-34). Parts are independent meshes (one per molecule):
// triangle indices are local to their part, components never span parts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_common.hpp"

namespace smg {

// Fixed chunks of the reference's deterministic sums (mod.rs:88).
constexpr uint32_t SUM_CHUNK_LOG2 = 14;
// Validation counters per part (u64 each), in smg_validation order.
enum : uint32_t { MV_BOUNDARY, MV_NON_MANIFOLD, MV_MISORIENTED, MV_DEGENERATE, MV_DUPLICATE, MV_INVALID, MV_BAD, MV_N = 8 };

struct MeasParts {
    const uint32_t* vfirst;  // parts + 1
    const uint32_t* tfirst;  // parts + 1
    uint32_t parts;
};

// The part holding item i (first[p] <= i < first[p + 1]).
__device__ inline uint32_t part_of(const uint32_t* first, uint32_t parts, uint32_t i) {
    uint32_t lo = 0, hi = parts;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ inline uint32_t load_parent(uint32_t* parent, uint32_t x) {
    return __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lock-free union-find as mod.rs:197-228: links go from the larger root to
// the smaller, so parents only decrease and every root is the smallest
// vertex of its component, whatever the scheduling.
__device__ inline uint32_t uf_find(uint32_t* parent, uint32_t x) {
    for (;;) {
        uint32_t p = load_parent(parent, x);
        if (p == x) return x;
        uint32_t gp = load_parent(parent, p);
        if (gp != p) atomicCAS(&parent[x], p, gp);
        x = gp;
    }
}

__device__ inline void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        uint32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
        if (ra == rb) return;
        uint32_t lo = ra < rb ? ra : rb, hi = ra < rb ? rb : ra;
        if (atomicCAS(&parent[hi], hi, lo) == hi) return;
    }
}

// arr[key] += val over the active lanes, one atomic per distinct key in
// the wavefront (neighbouring items mostly share a component).
template <class C>
__device__ inline void add_by_key(C* arr, uint32_t key, C val, bool active) {
    unsigned long long todo = __ballot(active);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t k = __shfl(key, leader);
        const bool mine = active && key == k;
        todo &= ~__ballot(mine);
        C s = mine ? val : C(0);
        for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((int)lane_id() == leader) atomicAdd(&arr[k], s);
    }
}

// mod.rs:112-124, f64 in the reference's operation order.
__device__ inline void tri_area_volume(const float* pos, uint32_t v0, uint32_t i0, uint32_t i1, uint32_t i2,
                                       const double o[3], double& area, double& volume) {
    double p[3][3];
    const uint32_t idx[3] = {i0, i1, i2};
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) p[i][k] = (double)pos[3 * (size_t)(v0 + idx[i]) + k] - o[k];
    double e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    double e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    area = 0.5 * sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double d[3] = {p[1][1] * p[2][2] - p[1][2] * p[2][1], p[1][2] * p[2][0] - p[1][0] * p[2][2],
                   p[1][0] * p[2][1] - p[1][1] * p[2][0]};
    volume = (p[0][0] * d[0] + p[0][1] * d[1] + p[0][2] * d[2]) / 6.0;
}

__device__ inline void part_origin(const float* pos, uint32_t v0, uint32_t v1, double o[3]) {
    for (int k = 0; k < 3; k++) o[k] = v1 > v0 ? (double)pos[3 * (size_t)v0 + k] : 0.0;
}

// Vertices: parents reset; non-finite positions and (with normals) normals
// that are not finite unit vectors counted (mod.rs:642-656).
__global__ void k_meas_vertices(const float* pos, const float* nrm, uint32_t V, MeasParts P, uint32_t* parent,
                                unsigned long long* val) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < V; g += gridDim.x * blockDim.x) {
        parent[g] = g;
        const float* q = pos + 3 * (size_t)g;
        uint32_t bad = !(isfinite(q[0]) && isfinite(q[1]) && isfinite(q[2]));
        if (nrm) {
            const float* m = nrm + 3 * (size_t)g;
            float len2 = m[0] * m[0];
            len2 = len2 + m[1] * m[1];
            len2 = len2 + m[2] * m[2];
            bad += !(isfinite(m[0]) && isfinite(m[1]) && isfinite(m[2]) && fabsf(len2 - 1.0f) < 1e-3f);
        }
        if (bad) atomicAdd(&val[MV_N * part_of(P.vfirst, P.parts, g) + MV_BAD], (unsigned long long)bad);
    }
}

// Triangles: invalid indices and degenerate triangles counted, components
// joined, and the uses of each edge counted at its smaller vertex.
__global__ void k_meas_triangles(const uint32_t* tris, const float* pos, uint32_t T, MeasParts P, uint32_t* parent,
                                 uint32_t* ucount, unsigned long long* val) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        const uint32_t p = part_of(P.tfirst, P.parts, t);
        const uint32_t v0 = P.vfirst[p], n = P.vfirst[p + 1] - v0;
        const uint32_t i[3] = {tris[3 * (size_t)t], tris[3 * (size_t)t + 1], tris[3 * (size_t)t + 2]};
        if (i[0] >= n || i[1] >= n || i[2] >= n) {
            atomicAdd(&val[MV_N * p + MV_INVALID], 1ull);
            continue;
        }
        bool deg = i[0] == i[1] || i[1] == i[2] || i[0] == i[2];
        if (!deg) {
            double o[3], a, v;
            part_origin(pos, v0, v0 + n, o);
            tri_area_volume(pos, v0, i[0], i[1], i[2], o, a, v);
            deg = a == 0.0;
        }
        if (deg) atomicAdd(&val[MV_N * p + MV_DEGENERATE], 1ull);
        uf_union(parent, v0 + i[0], v0 + i[1]);
        uf_union(parent, v0 + i[0], v0 + i[2]);
        for (int k = 0; k < 3; k++) {
            uint32_t a = i[k], b = i[(k + 1) % 3];
            atomicAdd(&ucount[v0 + (a < b ? a : b)], 1u);
        }
    }
}

// Roots (full path compression) and root flags.
__global__ void k_meas_roots(uint32_t* parent, uint32_t V, uint32_t* flag) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < V; g += gridDim.x * blockDim.x) {
        uint32_t r = uf_find(parent, g);
        flag[g] = r == g;
    }
}

// Global component of every vertex (ids in root order, so part-major) and
// the vertices of each component.
__global__ void k_meas_comp(uint32_t* parent, const uint32_t* cid, uint32_t V, uint32_t* comp, uint32_t* cverts) {
    const uint32_t base = blockIdx.x * blockDim.x;
    for (uint32_t g0 = base; g0 < V; g0 += gridDim.x * blockDim.x) {
        const uint32_t g = g0 + threadIdx.x;
        const bool on = g < V;
        uint32_t c = 0;
        if (on) {
            c = cid[uf_find(parent, g)];
            comp[g] = c;
        }
        add_by_key(cverts, c, 1u, on);
    }
}

// First global component of each part (parts + 1 entries, the last the total).
__global__ void k_meas_coff(const uint32_t* cid, const uint32_t* ncomp, uint32_t V, MeasParts P, uint32_t* coff) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= P.parts; p += gridDim.x * blockDim.x)
        coff[p] = P.vfirst[p] < V ? cid[P.vfirst[p]] : *ncomp;
}

// The corner k of t whose edge t[k] -> t[k + 1] joins the two smallest
// vertices with the largest as third (mod.rs:529-540).
__device__ inline int representative(const uint32_t t[3]) {
    uint32_t s0 = t[0], s1 = t[1], s2 = t[2], x;
    if (s0 > s1) x = s0, s0 = s1, s1 = x;
    if (s1 > s2) x = s1, s1 = s2, s2 = x;
    if (s0 > s1) x = s0, s0 = s1, s1 = x;
    for (int k = 0; k < 3; k++) {
        uint32_t a = t[k], b = t[(k + 1) % 3];
        if ((a < b ? a : b) == s0 && (a < b ? b : a) == s1 && t[(k + 2) % 3] == s2) return k;
    }
    return 0;
}

// Uses of each edge into the bucket of its smaller vertex: max << 33 |
// forward << 32 | representative << 31 | third (mod.rs:335-356).
__global__ void k_meas_uses(const uint32_t* tris, uint32_t T, MeasParts P, uint32_t* unext, unsigned long long* uses) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        const uint32_t p = part_of(P.tfirst, P.parts, t);
        const uint32_t v0 = P.vfirst[p];
        const uint32_t i[3] = {tris[3 * (size_t)t], tris[3 * (size_t)t + 1], tris[3 * (size_t)t + 2]};
        const int first = representative(i);
        for (int k = 0; k < 3; k++) {
            uint32_t a = i[k], b = i[(k + 1) % 3], third = i[(k + 2) % 3];
            unsigned long long u = ((unsigned long long)(a < b ? b : a) << 33) | ((unsigned long long)(a < b) << 32) |
                                   ((unsigned long long)(k == first) << 31) | (unsigned long long)third;
            uses[atomicAdd(&unext[v0 + (a < b ? a : b)], 1u)] = u;
        }
    }
}

// Per vertex: its bucket sorted, then one run per undirected edge: edges
// counted per component, boundary / misoriented / non-manifold edges and
// duplicate triangles (equal representatives) counted per part
// (mod.rs:470-498, 675-702).
__global__ void k_meas_buckets(const uint32_t* ustart, const uint32_t* ucount, unsigned long long* uses, uint32_t V,
                               MeasParts P, const uint32_t* comp, uint32_t* cedges, unsigned long long* val) {
    const uint32_t base = blockIdx.x * blockDim.x;
    for (uint32_t g0 = base; g0 < V; g0 += gridDim.x * blockDim.x) {
        const uint32_t g = g0 + threadIdx.x;
        const bool on = g < V;
        uint32_t edges = 0;
        uint32_t c = 0;
        if (on) {
            c = comp[g];
            unsigned long long* u = uses + ustart[g];
            const uint32_t n = ucount[g];
            for (uint32_t i = 1; i < n; i++) {  // insertion sort (a few uses per vertex)
                unsigned long long x = u[i];
                uint32_t j = i;
                for (; j > 0 && u[j - 1] > x; j--) u[j] = u[j - 1];
                u[j] = x;
            }
            uint32_t bnd = 0, mis = 0, nm = 0, dup = 0, lo = 0, p = 0;
            bool part_known = false;
            for (uint32_t i = 0; i < n;) {
                const unsigned long long mx = u[i] >> 33;
                uint32_t fwd = 0, bwd = 0;
                for (; i < n && (u[i] >> 33) == mx; i++) ((u[i] >> 32) & 1 ? fwd : bwd)++;
                edges++;
                if (!part_known) {
                    p = part_of(P.vfirst, P.parts, g);
                    lo = g - P.vfirst[p];
                    part_known = true;
                }
                if ((uint32_t)mx == lo || (fwd == 1 && bwd == 1)) continue;
                if (fwd + bwd == 1)
                    bnd++;
                else if (fwd + bwd == 2)
                    mis++;
                else
                    nm++;
            }
            for (uint32_t i = 0; i < n; i++) {  // representatives equal to an earlier one
                if (!((u[i] >> 31) & 1)) continue;
                const unsigned long long ki = ((u[i] >> 33) << 31) | (u[i] & 0x7fffffffull);
                for (uint32_t j = 0; j < i; j++)
                    if (((u[j] >> 31) & 1) && (((u[j] >> 33) << 31) | (u[j] & 0x7fffffffull)) == ki) {
                        dup++;
                        break;
                    }
            }
            if (bnd | mis | nm | dup) {
                unsigned long long* vp = val + MV_N * p;
                if (bnd) atomicAdd(&vp[MV_BOUNDARY], (unsigned long long)bnd);
                if (mis) atomicAdd(&vp[MV_MISORIENTED], (unsigned long long)mis);
                if (nm) atomicAdd(&vp[MV_NON_MANIFOLD], (unsigned long long)nm);
                if (dup) atomicAdd(&vp[MV_DUPLICATE], (unsigned long long)dup);
            }
        }
        add_by_key(cedges, c, edges, on && edges);
    }
}

// Component of each triangle (its first vertex's) and triangles per
// component; iota values for the stable sort by component.
__global__ void k_meas_tcomp(const uint32_t* tris, uint32_t T, MeasParts P, const uint32_t* comp, uint32_t* tcomp,
                             uint32_t* idx, uint32_t* ctris) {
    const uint32_t base = blockIdx.x * blockDim.x;
    for (uint32_t t0 = base; t0 < T; t0 += gridDim.x * blockDim.x) {
        const uint32_t t = t0 + threadIdx.x;
        const bool on = t < T;
        uint32_t c = 0;
        if (on) {
            const uint32_t p = part_of(P.tfirst, P.parts, t);
            c = comp[P.vfirst[p] + tris[3 * (size_t)t]];
            tcomp[t] = c;
            idx[t] = t;
        }
        add_by_key(ctris, c, 1u, on);
    }
}

// Grouped triangles (the stable sort's order), their areas and volumes,
// and the heads of the (component, chunk) segments: the reference sums a
// component's triangles per chunk of 2^14 of its part, in order, then the
// chunk sums in chunk order (mod.rs:231-262).
__global__ void k_meas_group(const uint32_t* sidx, const uint32_t* skey, const uint32_t* tris, const float* pos,
                             uint32_t T, MeasParts P, uint32_t* gtris, double* A, double* Vol, uint32_t* head) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t t = sidx[i];
        const uint32_t p = part_of(P.tfirst, P.parts, t);
        const uint32_t v0 = P.vfirst[p];
        const uint32_t a = tris[3 * (size_t)t], b = tris[3 * (size_t)t + 1], c = tris[3 * (size_t)t + 2];
        gtris[3 * (size_t)i] = a;
        gtris[3 * (size_t)i + 1] = b;
        gtris[3 * (size_t)i + 2] = c;
        double o[3], ar, vo;
        part_origin(pos, v0, P.vfirst[p + 1], o);
        tri_area_volume(pos, v0, a, b, c, o, ar, vo);
        A[i] = ar;
        Vol[i] = vo;
        const uint32_t chunk = (t - P.tfirst[p]) >> SUM_CHUNK_LOG2;
        // the previous item of the same component is in the same part
        head[i] = i == 0 || skey[i] != skey[i - 1] || ((sidx[i - 1] - P.tfirst[p]) >> SUM_CHUNK_LOG2) != chunk;
    }
}

__global__ void k_meas_segstart(const uint32_t* head, const uint32_t* segid, uint32_t T, uint32_t* segstart) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x)
        if (head[i]) segstart[segid[i]] = i;
}

// One segment per thread: its sums in triangle order, starting from its
// first triangle's values (the reference's first push, mod.rs:244-251).
__global__ void k_meas_segsum(const uint32_t* segstart, const uint32_t* nseg, uint32_t T, const double* A,
                              const double* Vol, double* segA, double* segV) {
    const uint32_t ns = *nseg;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        const uint32_t b = segstart[s], e = s + 1 < ns ? segstart[s + 1] : T;
        double a = A[b], v = Vol[b];
        for (uint32_t i = b + 1; i < e; i++) {
            a += A[i];
            v += Vol[i];
        }
        segA[s] = a;
        segV[s] = v;
    }
}

struct DComp {
    unsigned long long vertices, triangles;
    double area, volume;
    long long euler;
    unsigned long long first_triangle;  // within the part
};

// One component per thread: chunk sums added in chunk order to 0.0
// (mod.rs:263-271), counts, Euler characteristic V - E + F. A component's
// segments run from the segment of its first grouped triangle to the first
// segment of the next component (segid: heads before an item).
__global__ void k_meas_comps(const uint32_t* ncomp, const uint32_t* cverts, const uint32_t* ctris,
                             const uint32_t* cedges, const uint32_t* cfirst, const uint32_t* segid,
                             const uint32_t* nseg, uint32_t T, const double* segA, const double* segV,
                             const uint32_t* coff, MeasParts P, DComp* out) {
    const uint32_t nc = *ncomp;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        double a = 0.0, v = 0.0;
        if (ctris[c]) {
            const uint32_t b = cfirst[c], e = b + ctris[c];
            const uint32_t s0 = segid[b], s1 = e < T ? segid[e] : *nseg;
            for (uint32_t s = s0; s < s1; s++) {
                a += segA[s];
                v += segV[s];
            }
        }
        const uint32_t p = part_of(coff, P.parts, c);
        DComp r;
        r.vertices = cverts[c];
        r.triangles = ctris[c];
        r.area = a;
        r.volume = v;
        r.euler = (long long)cverts[c] - (long long)cedges[c] + (long long)ctris[c];
        r.first_triangle = cfirst[c] - P.tfirst[p];
        out[c] = r;
    }
}

// Component of each vertex within its part.
__global__ void k_meas_vcomp(const uint32_t* comp, const uint32_t* coff, uint32_t V, MeasParts P, uint32_t* out) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < V; g += gridDim.x * blockDim.x)
        out[g] = comp[g] - coff[part_of(P.vfirst, P.parts, g)];
}

}  // namespace smg
