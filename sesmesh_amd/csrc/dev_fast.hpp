// Fast refinement: the exact path's root search and attributes with the
// candidate evaluation in f32.
//
// The sub-block's lists are gathered exactly as in exact mode (f64
// decisions), then converted once to f32 records relative to the ball
// centre, where coordinates are a few angstroms and f32 keeps about 1e-6 A
// (far below the 1e-5 A root tolerance). Squared lengths and dot products
// are FMA chains. The Newton / Illinois / bisection iteration runs in f32
// on the edge parameter t; the vertex is placed, and its normal and atom
// found, in f64 from that t.
// Results are not bit-identical to the CPU back end; their distance from
// the exact surface is what fast mode reports. An edge the f32 Newton /
// Illinois phase does not converge on is redone by the exact f64 search,
// so it is placed exactly as in exact mode.
#pragma once
#include "dev_refine.hpp"

namespace smg {

// a . b for 3-vectors.
__device__ inline float dot3c(float ax, float ay, float az, float bx, float by, float bz) {
    return __builtin_fmaf(ax, bx, __builtin_fmaf(ay, by, az * bz));
}

// f32 lists, after the f64 ones in the same LDS / scratch slice.
struct FLayout {
    uint32_t nq, nin, cdat, ckey, bytes;
};
__host__ __device__ inline FLayout make_flayout(const RCaps& c, uint32_t base) {
    FLayout L;
    uint32_t o = base;
    auto take = [&](uint32_t bytes) {
        uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    L.nq = take(16 * c.near);   // x, y, z (local), (R - EPS)^2
    L.nin = take(4 * c.near);   // R^2
    L.cdat = take(32 * c.cand); // S (x, y, z, R); C (mx, my, mz, ax, ay, az, rho); V (x, y, z); local
    L.ckey = take(4 * c.cand);
    L.bytes = o;
    return L;
}

struct FLists {
    float4* nq;
    float* nin;
    float* cdat;
    float* ckey;
};

__device__ inline FLists fcarve(char* base, const FLayout& L) {
    FLists f;
    f.nq = (float4*)(base + L.nq);
    f.nin = (float*)(base + L.nin);
    f.cdat = (float*)(base + L.cdat);
    f.ckey = (float*)(base + L.ckey);
    return f;
}

// The f64 lists as f32 records relative to the ball centre c.
__device__ inline void to_local(const RLists& s, const FLists& f, const uint32_t* n, const double c[3]) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t n_near = n[R_NEAR], total = n[R_CAND], ns = n[R_S], nc = n[R_C];
    for (uint32_t k = tid; k < n_near; k += nt) {
        NearD q = s.nq[k];
        // blocker: |q - c| < R - EPS, with R - EPS = sqrt(blk) to f64 rounding
        f.nq[k] = make_float4((float)(q.x - c[0]), (float)(q.y - c[1]), (float)(q.z - c[2]), (float)q.blk);
        f.nin[k] = (float)s.nin[k];
    }
    for (uint32_t p = tid; p < total; p += nt) {
        const double* d = s.cdat + (size_t)CDAT * p;
        float* o = f.cdat + 8 * p;
        o[0] = (float)(d[0] - c[0]);
        o[1] = (float)(d[1] - c[1]);
        o[2] = (float)(d[2] - c[2]);
        if (p < ns) {
            o[3] = (float)d[3];
        } else if (p < ns + nc) {
            for (int q = 3; q < 7; q++) o[q] = (float)d[q];
        }
        f.ckey[p] = (float)key_value(s.ckey[p]);
    }
    __syncthreads();
}

struct EvalF {
    float value;
    uint32_t kind, id;
    float qx, qy, qz;  // local
};

__device__ inline bool blocked_fast(const RLists& s, const FLists& f, uint32_t cc, uint32_t nw, uint32_t p, float qx,
                                    float qy, float qz) {
    for (uint32_t w = 0; w < nw; w++) {
        unsigned long long m = s.cbits[w * cc + p];
        while (m) {
            float4 b = f.nq[w * 64 + (uint32_t)__builtin_ctzll(m)];
            m &= m - 1ull;
            float ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
            if (dot3c(ex, ey, ez, ex, ey, ez) < b.w) return true;
        }
    }
    return false;
}

template <int K>
__device__ inline float cand_dist_fast(const float* d, float x, float y, float z) {
    if (K == 0) {
        float vx = x - d[0], vy = y - d[1], vz = z - d[2];
        float len = sqrtf(dot3c(vx, vy, vz, vx, vy, vz));
        return len < 1e-6f ? INFINITY : fabsf(len - d[3]);
    }
    if (K == 1) {
        float yx = x - d[0], yy = y - d[1], yz = z - d[2];
        float along = dot3c(yx, yy, yz, d[3], d[4], d[5]);
        float wx = __builtin_fmaf(-d[3], along, yx), wy = __builtin_fmaf(-d[4], along, yy),
              wz = __builtin_fmaf(-d[5], along, yz);
        float wl = sqrtf(dot3c(wx, wy, wz, wx, wy, wz));
        float t = wl - d[6];
        return sqrtf(__builtin_fmaf(along, along, t * t));
    }
    float vx = x - d[0], vy = y - d[1], vz = z - d[2];
    return sqrtf(dot3c(vx, vy, vz, vx, vy, vz));
}

__device__ inline void project_fast(const DArr& A, const RLists& s, const FLists& f, uint32_t ns, uint32_t nc,
                                    uint32_t p, float x, float y, float z, float& qx, float& qy, float& qz) {
    const float* d = f.cdat + 8 * p;
    if (p < ns) {
        float vx = x - d[0], vy = y - d[1], vz = z - d[2];
        float len = sqrtf(dot3c(vx, vy, vz, vx, vy, vz));
        float sc = d[3] / len;
        qx = __builtin_fmaf(vx, sc, d[0]);
        qy = __builtin_fmaf(vy, sc, d[1]);
        qz = __builtin_fmaf(vz, sc, d[2]);
    } else if (p < ns + nc) {
        float yx = x - d[0], yy = y - d[1], yz = z - d[2];
        float along = dot3c(yx, yy, yz, d[3], d[4], d[5]);
        float wx = __builtin_fmaf(-d[3], along, yx), wy = __builtin_fmaf(-d[4], along, yy),
              wz = __builtin_fmaf(-d[5], along, yz);
        float wl = sqrtf(dot3c(wx, wy, wz, wx, wy, wz));
        float dx, dy, dz;
        if (wl > 1e-6f) {
            dx = wx / wl;
            dy = wy / wl;
            dz = wz / wl;
        } else {
            uint32_t c = s.cid[p] & ID_MASK;
            dx = (float)A.e1x[c];
            dy = (float)A.e1y[c];
            dz = (float)A.e1z[c];
        }
        qx = __builtin_fmaf(dx, d[6], d[0]);
        qy = __builtin_fmaf(dy, d[6], d[1]);
        qz = __builtin_fmaf(dz, d[6], d[2]);
    } else {
        qx = d[0];
        qy = d[1];
        qz = d[2];
    }
}

__device__ inline void group_least_f(const Group& G, float& best, uint32_t& bo) {
    for (uint32_t o = G.g >> 1; o > 0; o >>= 1) {
        float od = __shfl_xor(best, (int)o, WAVE);
        uint32_t oo = __shfl_xor(bo, (int)o, WAVE);
        if (od < best || (od == best && oo < bo)) {
            best = od;
            bo = oo;
        }
    }
}

template <int K>
__device__ inline void scan_class_fast(const DArr& A, const RLists& s, const FLists& f, uint32_t cc, uint32_t nw,
                                       uint32_t ns, uint32_t nc, uint32_t a, uint32_t e, const Group& G, float x,
                                       float y, float z, float delta, float cutoff, float& best, uint32_t& bo) {
    for (uint32_t p = a + G.r; p < e; p += G.g) {
        if (f.ckey[p] > best + delta) break;
        float dist = cand_dist_fast<K>(f.cdat + 8 * p, x, y, z);
        if (!(dist < cutoff) || !(dist < best || (dist == best && p < bo))) continue;
        if (K != 2) {
            float qx, qy, qz;
            project_fast(A, s, f, ns, nc, p, x, y, z, qx, qy, qz);
            if (blocked_fast(s, f, cc, nw, p, qx, qy, qz)) continue;
        }
        best = dist;
        bo = p;
    }
}

// eval_local at the local point (x, y, z) in f32, split over a lane group.
__device__ inline EvalF eval_group_fast(const DArr& A, const RLists& s, const FLists& f, const uint32_t* n, uint32_t cc,
                                        uint32_t nw, const Group& G, float x, float y, float z, float probe,
                                        float cutoff) {
    EvalF e;
    e.qx = x;
    e.qy = y;
    e.qz = z;
    bool ins = false;
    const uint32_t n_in = n[R_IN];
    for (uint32_t q = G.r; q < n_in; q += G.g) {
        uint32_t k = s.ins[q];
        float4 b = f.nq[k];
        float dx = x - b.x, dy = y - b.y, dz = z - b.z;
        ins |= dot3c(dx, dy, dz, dx, dy, dz) < f.nin[k];
    }
    if (!group_any(G, ins)) {
        e.value = probe;
        e.kind = F_SOLVENT;
        e.id = 0;
        return e;
    }
    const uint32_t ns = n[R_S], nc = n[R_C], total = n[R_CAND];
    // Lower bounds from the centre distances, with a margin above f32 error.
    const float delta = sqrtf(dot3c(x, y, z, x, y, z)) + 1e-4f;
    float best = cutoff;
    uint32_t bo = 0xffffffffu;
    scan_class_fast<0>(A, s, f, cc, nw, ns, nc, 0, ns, G, x, y, z, delta, cutoff, best, bo);
    group_least_f(G, best, bo);
    scan_class_fast<1>(A, s, f, cc, nw, ns, nc, ns, ns + nc, G, x, y, z, delta, cutoff, best, bo);
    group_least_f(G, best, bo);
    scan_class_fast<2>(A, s, f, cc, nw, ns, nc, ns + nc, total, G, x, y, z, delta, cutoff, best, bo);
    group_least_f(G, best, bo);
    if (bo == 0xffffffffu) {
        e.value = probe - cutoff;
        e.kind = F_BEYOND;
        e.id = 0;
        return e;
    }
    project_fast(A, s, f, ns, nc, bo, x, y, z, e.qx, e.qy, e.qz);
    e.value = probe - best;
    e.kind = bo < ns ? F_SPHERE : (bo < ns + nc ? F_CIRCLE : F_VERTEX);
    e.id = s.cid[bo] & ID_MASK;
    return e;
}

// Root search of one edge in f32 (the exact path's rules, mesh/extract.rs:
// 425-535); the edge runs from local point pa along `axis` for one step h.
struct EdgeF {
    int32_t ia[3];
    uint32_t id;
    float fa;
    uint32_t st;  // it (bits 0-6), side + 1 (7-8), bisect (9), neg0 (10), converged (11), evals (16-31)
    float t, t0, g0, t1, g1;
    EvalF last;
};

__device__ inline void edgef_start(const DRun& P, const float* val, const int64_t org[3], uint32_t id, EdgeF& E) {
    const uint32_t b = P.b;
    uint32_t axis = id % 3, cell = id / 3;
    uint32_t x = cell % b, y = (cell / b) % b, z = cell / (b * b);
    uint32_t i0 = x + P.side * (y + P.side * z);
    uint32_t stp = axis == 0 ? 1u : (axis == 1 ? P.side : P.side * P.side);
    E.ia[0] = (int32_t)(org[0] + x);
    E.ia[1] = (int32_t)(org[1] + y);
    E.ia[2] = (int32_t)(org[2] + z);
    E.id = id;
    float fa = val[i0], fb = val[i0 + stp];
    E.fa = fa;
    E.t = fminf(fmaxf(fa / (fa - fb), 0.0f), 1.0f);
    E.t0 = 0.0f;
    E.g0 = fa;
    E.t1 = 1.0f;
    E.g1 = fb;
    E.st = 1u << 7;
}

// The local point at parameter t (ball centre c).
__device__ inline void edgef_point(const EdgeF& E, double h, const double c[3], float t, float& x, float& y, float& z) {
    uint32_t axis = E.id % 3;
    float p[3];
    for (int a = 0; a < 3; a++) p[a] = (float)((double)E.ia[a] * h - c[a]);
    p[axis] = __builtin_fmaf((float)h, t, p[axis]);
    x = p[0];
    y = p[1];
    z = p[2];
}

__device__ inline bool edgef_step(EdgeF& E, const EvalF& e, float x, float y, float z, float hf) {
    E.st += 1u << 16;
    E.last = e;
    float g = e.value;
    uint32_t it = E.st & 127u;
    const uint32_t axis = E.id % 3;
    if (!(E.st & (1u << 9))) {
        if (fabsf(g) <= 1e-5f) {
            E.st |= 1u << 11;
            return true;
        }
        int side = (int)((E.st >> 7) & 3u) - 1;
        if ((g < 0.0f) == (E.g0 < 0.0f)) {
            E.t0 = E.t;
            E.g0 = g;
            if (side == -1) E.g1 *= 0.5f;
            side = -1;
        } else {
            E.t1 = E.t;
            E.g1 = g;
            if (side == 1) E.g0 *= 0.5f;
            side = 1;
        }
        float slope = 0.0f;
        if (e.kind != F_SOLVENT && e.kind != F_BEYOND) {
            float wx = e.qx - x, wy = e.qy - y, wz = e.qz - z;
            float len = sqrtf(dot3c(wx, wy, wz, wx, wy, wz));
            if (len > 1e-6f) {
                float w[3] = {wx, wy, wz};
                slope = w[axis] / len * hf;
            }
        }
        float newton = E.t - g / slope;
        if (fabsf(slope) > 1e-3f * hf && newton > E.t0 && newton < E.t1) {
            E.t = newton;
        } else {
            float sv = (E.t0 * E.g1 - E.t1 * E.g0) / (E.g1 - E.g0);
            E.t = (isfinite(sv) && sv > E.t0 && sv < E.t1) ? sv : 0.5f * (E.t0 + E.t1);
        }
        it++;
        uint32_t st = (E.st & ~(127u | (3u << 7))) | ((uint32_t)(side + 1) << 7);
        if (it == 8) {
            st |= 1u << 9;
            if (E.g0 < 0.0f) st |= 1u << 10;
            E.t = 0.5f * (E.t0 + E.t1);
        } else {
            st |= it;
        }
        E.st = st;
        return false;
    }
    if (fabsf(g) <= 1e-5f || (E.t1 - E.t0) * hf <= 2e-5f) return true;
    if ((g < 0.0f) == ((E.st & (1u << 10)) != 0))
        E.t0 = E.t;
    else
        E.t1 = E.t;
    if (++it == 64) return true;
    E.st = (E.st & ~127u) | it;
    E.t = 0.5f * (E.t0 + E.t1);
    return false;
}

// The vertex (f64 from the f32 t) and its attributes; rank-0 lane writes.
__device__ inline void edgef_finish(const DArr& A, const DRun& P, const EdgeF& E, const Group& G, const double c[3],
                                    size_t slot, const VOut& o) {
    if (G.r != 0) return;
    EdgeState X;
    X.ia[0] = E.ia[0];
    X.ia[1] = E.ia[1];
    X.ia[2] = E.ia[2];
    X.id = E.id;
    X.fa = E.fa;
    X.st = E.st;
    X.t = (double)E.t;
    X.last.value = (double)E.last.value;
    X.last.kind = E.last.kind;
    X.last.id = E.last.id;
    X.last.qx = (double)E.last.qx + c[0];
    X.last.qy = (double)E.last.qy + c[1];
    X.last.qz = (double)E.last.qz + c[2];
    edge_finish(A, P, X, G, slot, o);
}

// refine_sub in fast mode: the same work, lists and outputs.
template <bool GLOBAL>
__device__ __forceinline__ void refine_sub_fast(const Mol* __restrict__ mols, const uint32_t* bmol, const RCaps& cap,
                                                const RLayout& lay, const FLayout& flay, const int32_t* keys,
                                                const float* values, uint32_t j, const unsigned long long* mask,
                                                const uint32_t* wpre, uint32_t words, const uint32_t* vbase,
                                                const VOut& out, char* scratch, size_t slice, uint32_t* overflow,
                                                uint32_t* n_overflow, unsigned long long* evals_total,
                                                const HidList& hid) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[R_N];
    __shared__ uint32_t tmp4[WIDE / WAVE];  // block scans: one slot per wavefront
    __shared__ unsigned long long evals_sum;
    uint32_t brick = j >> 3, sb = j & 7;
    const uint32_t mol = bmol[brick];
    const DArr& A = mols[mol].A;
    const DRun& P = mols[mol].P;
    char* base = GLOBAL ? scratch + (size_t)blockIdx.x * slice : smem;
    RLists s = rcarve(base, lay);
    FLists f = fcarve(base, flay);
    const int32_t* key = keys + 3 * brick;
    const uint32_t b = P.b, sub = P.sub;
    double h = P.h, subd = (double)sub;
    if (threadIdx.x == 0) evals_sum = 0;
    double rho = 0.5 * sqrt(3.0) * subd * h * (1.0 + 1e-9) + 1e-9;
    uint32_t o3[3] = {sb & 1, (sb >> 1) & 1, sb >> 2};
    int64_t org[3];
    double c[3];
    for (int a = 0; a < 3; a++) {
        org[a] = (int64_t)key[a] * (int64_t)b;
        double basep = (double)org[a] * h;
        c[a] = basep + ((double)(o3[a] * sub) + 0.5 * subd) * h;
    }
    const uint32_t* hl = nullptr;  // the sampling's hidden candidates of this sub-block
    uint32_t hn = hid.cnt ? hid.cnt[j] : NO_LIST;
    if (hn != NO_LIST) hl = hid.ids + hid.off[j];
    const uint32_t* sl = nullptr;  // and its superset of this gather's items
    uint32_t sn = hid.sup.cnt ? hid.sup.cnt[j] : NO_LIST;
    if (sn != NO_LIST) sl = hid.sup.ids + hid.sup.off[j];
    const uint32_t* br = nullptr;  // and its blocker record
    if (hid.blk.cnt && hid.blk.cnt[j] != NO_LIST) br = hid.blk.ids + (size_t)hid.blk.slot[j] * hid.blk.words;
    bool ok = gather_refine(A, cap, s, n, tmp4, c[0], c[1], c[2], rho, P.band, true, hl, hn, sl, sn, SupList{}, 0,
                            0.0, br);
    uint32_t lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = o3[a] * sub;
        hi[a] = o3[a] ? b : sub;
    }
    uint32_t ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    uint32_t ncand = ex * ey * ez * 3;
    if (ok) {
        for (uint32_t q = threadIdx.x; q < ncand; q += blockDim.x) {
            uint32_t axis = q % 3, cell = q / 3;
            uint32_t x = lo[0] + cell % ex, y = lo[1] + (cell / ex) % ey, z = lo[2] + cell / (ex * ey);
            uint32_t id = ((z * b + y) * b + x) * 3 + axis;
            if (edge_is_crossing(mask, words, brick, id)) {
                uint32_t at = atomicAdd(&n[R_E], 1u);
                if (at < cap.edges)
                    s.ebuf[at] = id;
                else
                    n[R_OVER] = 1u;
            }
        }
        __syncthreads();
        ok = n[R_OVER] == 0u;
    }
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = j;
        return;
    }
    to_local(s, f, n, c);
    const float* val = values + (size_t)brick * P.spb;
    const uint32_t ne = n[R_E];
    uint32_t g = cap.gmax ? cap.gmax : (uint32_t)WAVE;
    while (g > 1 && blockDim.x / g < ne) g >>= 1;
    Group G;
    G.g = g;
    G.r = threadIdx.x % g;
    G.shift = (threadIdx.x % WAVE) - G.r;
    G.mask = g == WAVE ? ~0ull : ((1ull << g) - 1ull);
    const uint32_t ngroups = blockDim.x / g, nw = cap.near / 64;
    const float cutoff = (float)(A.probe + P.band), probe = (float)A.probe, hf = (float)h;
    uint32_t q = threadIdx.x / g;
    bool active = q < ne;
    EdgeF E;
    // An edge whose f32 search ends in bisection (not converged) is redone
    // by the exact f64 search from its start, over the f64 lists the
    // gather left in place, as the reference redoes unconverged device
    // edges on the CPU (mesh/extract.rs:336-358). The SAS and vdW kinds
    // (a handful of spheres per point, no blockers) take the exact search
    // for every edge.
    const bool sph = A.kind != K_SES;
    EdgeState X;
    bool exact = sph;
    auto start = [&](uint32_t id) {
        if (sph)
            edge_start(P, val, org, id, X);
        else
            edgef_start(P, val, org, id, E);
    };
    if (active) start(s.ebuf[q]);
    unsigned long long my_evals = 0;
    while (__ballot(active) != 0ull) {
        if (active) {
            bool done;
            if (exact) {
                double x, y, z;
                {
                    EdgeGeom Q = edge_geom(X, h);
                    x = Q.pa[0] + Q.d[0] * X.t;
                    y = Q.pa[1] + Q.d[1] * X.t;
                    z = Q.pa[2] + Q.d[2] * X.t;
                }
                EvalD e = eval_group(A, s, n, cap.cand, nw, G, c[0], c[1], c[2], x, y, z, A.probe + P.band, P.band);
                done = edge_step(X, e, h, sph);
                if (done) {
                    size_t slot = (size_t)vbase[brick] + edge_rank(mask, wpre, words, brick, X.id);
                    edge_finish(A, P, X, G, slot, out);
                    if (G.r == 0 && out.evals && slot < out.cap && !sph) out.evals[slot] |= 1u << 30;
                    my_evals += X.st >> 16;
                }
            } else {
                float x, y, z;
                edgef_point(E, h, c, E.t, x, y, z);
                EvalF e = eval_group_fast(A, s, f, n, cap.cand, nw, G, x, y, z, probe, cutoff);
                done = edgef_step(E, e, x, y, z, hf);
                if (done && !(E.st & (1u << 11))) {
                    edge_start(P, val, org, E.id, X);
                    exact = true;
                    done = false;
                } else if (done) {
                    size_t slot = (size_t)vbase[brick] + edge_rank(mask, wpre, words, brick, E.id);
                    edgef_finish(A, P, E, G, c, slot, out);
                    my_evals += E.st >> 16;
                }
            }
            if (done) {
                exact = sph;
                q += ngroups;
                active = q < ne;
                if (active) start(s.ebuf[q]);
            }
        }
    }
    if (evals_total) {
        if (G.r == 0 && my_evals) atomicAdd(&evals_sum, my_evals);
        __syncthreads();
        if (threadIdx.x == 0 && evals_sum) atomicAdd(&evals_total[mol], evals_sum);
    }
}

// Five wavefronts per SIMD (at most 96 VGPRs), as k_refine: the f32 search
// and the exact redo share the kernel, and unbounded they take 137 VGPRs
// (three wavefronts).
template <bool GLOBAL, bool PERSIST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_refine_fast(const Mol* __restrict__ mols, const uint32_t* bmol, RCaps cap,
                                                     RLayout lay, FLayout flay, const int32_t* keys,
                                                     const float* values, Queue q, HidList hid,
                                                     const unsigned long long* mask,
                                                     const uint32_t* wpre, uint32_t words, const uint32_t* vbase,
                                                     VOut out, char* scratch, size_t slice, uint32_t* overflow,
                                                     uint32_t* n_overflow, unsigned long long* evals_total) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t j) {
        refine_sub_fast<GLOBAL>(mols, bmol, cap, lay, flay, keys, values, j, mask, wpre, words, vbase, out, scratch,
                                slice, overflow, n_overflow, evals_total, hid);
    });
}

// k_refine_fast with WIDE threads, for the large tiers.
template <bool GLOBAL, bool PERSIST>
__global__ void __launch_bounds__(WIDE) k_refine_fast_wide(const Mol* __restrict__ mols, const uint32_t* bmol, RCaps cap,
                                                     RLayout lay, FLayout flay, const int32_t* keys,
                                                     const float* values, Queue q, HidList hid,
                                                     const unsigned long long* mask,
                                                     const uint32_t* wpre, uint32_t words, const uint32_t* vbase,
                                                     VOut out, char* scratch, size_t slice, uint32_t* overflow,
                                                     uint32_t* n_overflow, unsigned long long* evals_total) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t j) {
        refine_sub_fast<GLOBAL>(mols, bmol, cap, lay, flay, keys, values, j, mask, wpre, words, vbase, out, scratch,
                                slice, overflow, n_overflow, evals_total, hid);
    });
}

}  // namespace smg
