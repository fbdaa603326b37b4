// Candidate gathering and field sampling of one sub-block per workgroup.
//
// Restates Engine::assign / prepare_sub / gather_near / gather_ses /
// gather_spheres and the SES and sphere lane kernels (reference
// field/kernel.rs:489-751, 837-981). Membership
// decisions are made in f64 exactly as the CPU makes them; the candidate
// data stored for sampling are the CPU's frame-local f32 copies. The CPU
// sorts the class S and C candidates nearest-first; the sampled value does
// not depend on the order (it is the minimum over the valid candidates, and
// "inside" is an any()), so here they are only bucketed: class S before C,
// each in 16 distance buckets, nearest bucket first, in O(n). Near
// candidates still come early, so most later ones fail the distance test
// before their projection and blocker tests. Blockers are bitsets
// over near-list slots. Lists live in LDS (or, for sub-blocks whose lists
// overflow the LDS capacities, in a global scratch slice).
#pragma once
#include "dev_common.hpp"

namespace smg {

// Capacities of the per-sub-block lists: near spheres (a multiple of 64),
// inside group, S + C candidates (a power of two), V candidates, hash slots
// (a power of two).
struct Caps {
    uint32_t near, inside, cand, v, hash;
};

// Byte layout of the lists for given capacities (host and device agree).
struct Layout {
    uint32_t nd, id, hk, hv, cbk, cid, cdat, cbits, coff, cpre, vd, bytes;
};

__host__ __device__ inline Layout make_layout(const Caps& c) {
    Layout L;
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    L.nd = take(16 * c.near);    // x, y, z, r2tol
    L.id = take(16 * c.inside);  // x, y, z, r2
    L.hk = take(4 * c.hash);
    L.hv = take(4 * c.hash);
    L.cbk = take(4 * c.cand);
    L.cid = take(4 * c.cand);
    L.cdat = take(48 * c.cand);  // S: (x, y, z, r); C: (mx, my, mz, rho), (ux, uy, uz, -), (ex, ey, ez, -)
    L.cbits = take(c.near / 64 * 8 * c.cand);
    L.coff = take(4 * c.cand);
    L.cpre = take(4 * c.cand);
    L.vd = take(16 * c.v);
    L.bytes = o;
    return L;
}

// Engine::assign's owner of sample (lx, ly, lz) (kernel.rs:495-532): the
// first non-uniform sub-block containing it, or -1 with `positive` the sign
// of the last uniform one visited. `st` packs the 8 states, 2 bits each.
// The coordinates lie in [0, 2 sub], so the sub-block range along an axis
// is hi = min(v / sub, 1) = (v >= sub) and lo = hi except on the shared
// face v == sub, where it is 0 (v == 2 sub gives lo = hi = 1).
__device__ inline int sample_owner(uint32_t st, uint32_t sub, uint32_t lx, uint32_t ly, uint32_t lz,
                                   bool& positive) {
    uint32_t v[3] = {lx, ly, lz}, lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        hi[a] = v[a] >= sub ? 1u : 0u;
        lo[a] = v[a] == sub ? 0u : hi[a];
    }
    positive = true;
    for (uint32_t sz = lo[2]; sz <= hi[2]; sz++)
        for (uint32_t sy = lo[1]; sy <= hi[1]; sy++)
            for (uint32_t sx = lo[0]; sx <= hi[0]; sx++) {
                uint32_t s = sx + 2 * sy + 4 * sz, state = (st >> (2 * s)) & 3u;
                if (state == ST_NONE) return (int)s;
                positive = state == ST_OUT;
            }
    return -1;
}

__device__ inline uint32_t pack_states(const uint8_t* st) {
    uint32_t p = 0;
    for (int q = 0; q < 8; q++) p |= (uint32_t)st[q] << (2 * q);
    return p;
}

// Samples owned by no sub-block get +-band (kernel.rs:533-534); one
// workgroup per brick, striding over the *n_bricks bricks of the run's
// brick size b. A brick whose eight sub-blocks are all sampled owns every
// sample and is skipped.
__global__ void k_assign_uniform(const Mol* __restrict__ mols, const uint32_t* bmol, const uint8_t* st,
                                 const uint32_t* n_bricks, uint32_t b, float* values) {
    const uint32_t nb = *n_bricks;
    const uint32_t side = b + 1, plane = side * side, spb = plane * side, sub = b / 2;
    // i / d as __umulhi(i, ceil(2^32 / d)): exact for i and d below 2^16.
    const uint32_t m1 = (uint32_t)(((1ull << 32) + side - 1) / side);
    const uint32_t m2 = (uint32_t)(((1ull << 32) + plane - 1) / plane);
    for (uint32_t brick = blockIdx.x; brick < nb; brick += gridDim.x) {
        const uint32_t s = pack_states(st + 8 * brick);
        if (s == 0u) continue;  // ST_NONE in every sub-block
        const float band = mols[bmol[brick]].P.fband;
        float* out = values + (size_t)brick * spb;
        for (uint32_t i = threadIdx.x; i < spb; i += blockDim.x) {
            const uint32_t z = __umulhi(i, m2), r = i - z * plane, y = __umulhi(r, m1), x = r - y * side;
            bool positive;
            if (sample_owner(s, sub, x, y, z, positive) < 0) out[i] = positive ? band : -band;
        }
    }
}

struct Lists {
    float4* nd;
    float4* id;
    uint32_t *hk, *hv;
    uint32_t* cbk;  // candidate's key bucket (class S: 0-15, class C: 16-31)
    uint32_t* cid;
    float4* cdat;
    unsigned long long* cbits;  // near/64 rows of `cand` words
    uint32_t *coff, *cpre;
    float4* vd;
};

__device__ inline Lists carve(char* base, const Layout& L) {
    Lists s;
    s.nd = (float4*)(base + L.nd);
    s.id = (float4*)(base + L.id);
    s.hk = (uint32_t*)(base + L.hk);
    s.hv = (uint32_t*)(base + L.hv);
    s.cbk = (uint32_t*)(base + L.cbk);
    s.cid = (uint32_t*)(base + L.cid);
    s.cdat = (float4*)(base + L.cdat);
    s.cbits = (unsigned long long*)(base + L.cbits);
    s.coff = (uint32_t*)(base + L.coff);
    s.cpre = (uint32_t*)(base + L.cpre);
    s.vd = (float4*)(base + L.vd);
    return s;
}

// Counters of the list builder.
// C_DIFF: an item the refinement's gather of the sub-block (its slightly
// larger ball) would list differently.
enum { C_NEAR = 0, C_IN, C_S, C_C, C_V, C_CAND, C_OVER, C_SUP, C_SUPOVER, C_DIFF, C_N };
constexpr uint32_t HIDDEN = 0x80000000u;  // a candidate's bucket word when it is hidden

// True if (qx, qy, qz) lies inside a blocker of candidate p (kernel.rs:
// 898-906, 935-943): !(|q - c|^2 >= (r - tol)^2) for a near row member.
// The row's hint (in cpre, one of the bitset's blockers) is tried first.
__device__ inline bool blocked_f(const Lists& s, uint32_t cc, uint32_t nw, uint32_t p, float qx, float qy, float qz) {
    const uint32_t hint = s.cpre[p];
    if (hint != 0xffffffffu) {
        const float4 b = s.nd[hint];
        const float ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
        if (!(ex * ex + ey * ey + ez * ez >= b.w)) return true;
    }
    for (uint32_t w = 0; w < nw; w++) {
        unsigned long long m = s.cbits[w * cc + p];
        while (m) {
            uint32_t k[4];
            k[0] = w * 64 + (uint32_t)__builtin_ctzll(m);
            m &= m - 1ull;
            for (int j = 1; j < 4; j++) {
                k[j] = m ? w * 64 + (uint32_t)__builtin_ctzll(m) : k[0];
                m &= m - 1ull;
            }
            bool hit = false;
            for (int j = 0; j < 4; j++) {
                float4 b = s.nd[k[j]];
                float ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
                hit |= !(ex * ex + ey * ey + ez * ez >= b.w);
            }
            if (hit) return true;
        }
    }
    return false;
}

// Correctly rounded f32 square root of a finite x >= 0: the hardware root
// (within 1 ulp for x >= 2^-96) and one correction by the residuals of its
// two neighbours, the rounding step of the compiler's own sqrtf expansion,
// without that expansion's scaling of inputs below 2^-96 and its handling of
// infinities and negatives. Below 2^-96 the result is a tiny value that may
// differ from sqrtf's; the lane kernel only compares such a root against
// 1e-6 or subtracts it from a radius, where any value under 2^-40 gives the
// same outcome.
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __uint_as_float(__float_as_uint(s) - 1u), hi = __uint_as_float(__float_as_uint(s) + 1u);
    float r = s;
    if (__builtin_fmaf(-lo, s, x) <= 0.0f) r = lo;
    if (__builtin_fmaf(-hi, s, x) > 0.0f) r = hi;
    return r;
}

// One lane of the SES kernel (kernel.rs:837-952) at frame-local (x, y, z).
__device__ inline float ses_value(const Lists& s, const uint32_t* n, uint32_t cc, uint32_t nw, float x, float y,
                                  float z, float probe, float cutoff2) {
    bool inside = false;
    for (uint32_t k = 0; k < n[C_IN]; k++) {
        float4 c = s.id[k];
        float dx = x - c.x, dy = y - c.y, dz = z - c.z;
        if (dx * dx + dy * dy + dz * dz < c.w) {
            inside = true;
            break;
        }
    }
    if (!inside) return probe;
    float best2 = cutoff2;
    for (uint32_t v = 0; v < n[C_V]; v++) {
        float4 p = s.vd[v];
        float dx = x - p.x, dy = y - p.y, dz = z - p.z;
        float d2 = dx * dx + dy * dy + dz * dz;
        best2 = d2 < best2 ? d2 : best2;
    }
    // Class S and C, four candidates' distances at a time; a candidate's
    // projection and blockers are tested only if it improves on best2.
    const uint32_t ns = n[C_S], nsc = ns + n[C_C];
    for (uint32_t i0 = 0; i0 < ns; i0 += 4) {
        float lb2[4], dist[4];
        for (int j = 0; j < 4; j++) {
            uint32_t i = min(i0 + j, ns - 1);
            float4 c = s.cdat[3 * i];
            float dx = x - c.x, dy = y - c.y, dz = z - c.z;
            dist[j] = sqrt_rn(dx * dx + dy * dy + dz * dz);
            float lb = fabsf(dist[j] - c.w);
            lb2[j] = i0 + j < ns ? lb * lb : INFINITY;
        }
        for (int j = 0; j < 4; j++) {
            if (!(lb2[j] < best2 && 1e-6f < dist[j])) continue;
            uint32_t i = i0 + j;
            float4 c = s.cdat[3 * i];
            float dx = x - c.x, dy = y - c.y, dz = z - c.z;
            float sc = c.w / dist[j];
            if (!blocked_f(s, cc, nw, i, c.x + dx * sc, c.y + dy * sc, c.z + dz * sc)) best2 = lb2[j];
        }
    }
    for (uint32_t c0 = ns; c0 < nsc; c0 += 4) {
        float d2[4], wl[4];
        for (int j = 0; j < 4; j++) {
            uint32_t c = min(c0 + j, nsc - 1);
            float4 m = s.cdat[3 * c], u = s.cdat[3 * c + 1];
            float yx = x - m.x, yy = y - m.y, yz = z - m.z;
            float along = yx * u.x + yy * u.y + yz * u.z;
            float wx = yx - u.x * along, wy = yy - u.y * along, wz = yz - u.z * along;
            wl[j] = sqrt_rn(wx * wx + wy * wy + wz * wz);
            float t = wl[j] - m.w;
            d2[j] = c0 + j < nsc ? along * along + t * t : INFINITY;
        }
        for (int j = 0; j < 4; j++) {
            if (!(d2[j] < best2)) continue;
            uint32_t c = c0 + j;
            float4 m = s.cdat[3 * c], u = s.cdat[3 * c + 1];
            float rho = m.w;
            float qx, qy, qz;
            if (wl[j] < 1e-6f) {
                float4 e = s.cdat[3 * c + 2];
                qx = m.x + e.x * rho;
                qy = m.y + e.y * rho;
                qz = m.z + e.z * rho;
            } else {
                float yx = x - m.x, yy = y - m.y, yz = z - m.z;
                float along = yx * u.x + yy * u.y + yz * u.z;
                float wx = yx - u.x * along, wy = yy - u.y * along, wz = yz - u.z * along;
                float sc = rho / wl[j];
                qx = m.x + wx * sc;
                qy = m.y + wy * sc;
                qz = m.z + wz * sc;
            }
            if (!blocked_f(s, cc, nw, c, qx, qy, qz)) best2 = d2[j];
        }
    }
    return probe - sqrtf(best2);
}

// One lane of the sphere kernel of the SAS and vdW kinds (kernel.rs:953-981):
// min(band, min_k(|x - c_k| - R_k)) over the near spheres (nd: x, y, z, R).
__device__ inline float sphere_value(const Lists& s, const uint32_t* n, float x, float y, float z, float band) {
    float best = band;
    for (uint32_t k = 0; k < n[C_NEAR]; k++) {
        float4 c = s.nd[k];
        float dx = x - c.x, dy = y - c.y, dz = z - c.z;
        best = fminf(best, sqrtf(dx * dx + dy * dy + dz * dz) - c.w);
    }
    return best;
}

// The kept class S / C candidate ids of the lists (see the scatter in
// gather_sub).
__device__ inline uint32_t kept_id(const Lists& s, uint32_t ns, uint32_t p) {
    return p < ns ? __float_as_uint(s.cdat[3 * p + 1].x) : __float_as_uint(s.cdat[3 * p + 2].w);
}

// Writes the BlkRecord of sub-block j (ball centre b, the lists' ns class S
// and total class S + C kept candidates, then the class V ones) to the
// pool, or marks it absent if it does not fit. The candidates are ranked
// by the refinement's order: their f64 keys at the ball centre computed as
// gather_refine computes them (|d - R| of a sphere, the circle distance,
// the vertex distance), then the id. By the whole workgroup, after the
// bitsets; it reuses the near hash as scratch. True (in every thread) if
// it wrote a record whose lists are the refinement's own.
__device__ inline bool write_blk_record(const DArr& A, const Lists& s, const uint32_t* n, const Caps& cap,
                                        const BlkPool& blk, uint32_t j, const double b[3], uint32_t ns,
                                        uint32_t total) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t nn = n[C_NEAR], nv = n[C_V], nwr = (nn + 63) / 64, ew = blk_entry_words(nwr), cc = cap.cand;
    const uint32_t all = total + nv, m = BLK_HEAD + nn + all * ew, slot = blk.slot[j];
    const bool fits = m <= blk.words && slot < blk.slots && all <= cap.hash;  // the same in every thread
    if (tid == 0) blk.cnt[j] = fits ? m : NO_LIST;
    if (!fits) return false;
    uint32_t* o = blk.ids + (size_t)slot * blk.words;
    if (tid == 0) {
        o[0] = nn;
        o[1] = ns;
        o[2] = total - ns;
        o[3] = nwr;
        o[4] = nv;
        o[5] = n[C_DIFF] ? 0u : 1u;
    }
    for (uint32_t i = tid; i < cap.hash; i += nt) {  // near ids by slot, from the near hash
        const uint32_t key = s.hk[i];
        if (key) o[BLK_HEAD + s.hv[i]] = key - 1u;
    }
    __syncthreads();
    unsigned long long* kk = (unsigned long long*)s.hk;  // the hash is free now
    auto id_of = [&](uint32_t p) { return p < total ? kept_id(s, ns, p) : __float_as_uint(s.vd[p - total].w); };
    for (uint32_t p = tid; p < all; p += nt) {
        const uint32_t k = id_of(p);
        double key;
        if (p >= total) {
            const VertexRec V = load_rec(A.ver, k);
            key = ddist(b[0], b[1], b[2], V.x, V.y, V.z);
        } else if (p < ns) {
            const SphereRec S = load_rec(A.sph, k);
            key = fabs(ddist(b[0], b[1], b[2], S.x, S.y, S.z) - S.R);
        } else {
            const CircleRec C = load_rec(A.cir, k);
            double yx = b[0] - C.mx, yy = b[1] - C.my, yz = b[2] - C.mz;
            double ax = C.ax, ay = C.ay, az = C.az;
            double along = yx * ax + yy * ay + yz * az;
            double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
            double wl = sqrt(wx * wx + wy * wy + wz * wz);
            double rr = C.rho;
            key = sqrt(along * along + (wl - rr) * (wl - rr));
        }
        kk[p] = total_key(key);
    }
    __syncthreads();
    uint32_t* ent = o + BLK_HEAD + nn;
    // 8 lanes per candidate count the ones of its class ranked before it.
    for (uint32_t p = tid / 8; p < all; p += nt / 8) {
        const uint32_t cls = p < ns ? 0u : (p < total ? 1u : 2u);
        const uint32_t lo = cls == 0 ? 0u : (cls == 1 ? ns : total), hi = cls == 0 ? ns : (cls == 1 ? total : all);
        const uint32_t k = id_of(p);
        const unsigned long long kp = kk[p];
        uint32_t rank = 0;
        for (uint32_t u = lo + (tid & 7u); u < hi; u += 8) {
            const unsigned long long ku = kk[u];
            rank += (ku < kp || (ku == kp && id_of(u) < k)) ? 1u : 0u;
        }
        rank += __shfl_xor(rank, 1, 8);
        rank += __shfl_xor(rank, 2, 8);
        rank += __shfl_xor(rank, 4, 8);
        uint32_t* e = ent + (size_t)(lo + rank) * ew;
        const uint32_t r = tid & 7u;
        if (r == 0) {
            e[0] = k | (cls << KIND_SHIFT);
            e[1] = cls < 2 ? s.cpre[p] : 0xffffffffu;
            e[2] = (uint32_t)kp;
            e[3] = (uint32_t)(kp >> 32);
        }
        for (uint32_t w = r; w < nwr; w += 8) {
            const unsigned long long bits = cls < 2 ? s.cbits[w * cc + p] : 0ull;
            e[4 + 2 * w] = (uint32_t)bits;
            e[5 + 2 * w] = (uint32_t)(bits >> 32);
        }
    }
    return n[C_DIFF] == 0u;
}

// Gathers the lists of sub-block `sb` of brick `key` (kernel.rs:553-751).
// The SAS and vdW kinds (gather_spheres, kernel.rs:645-652) list the near
// spheres within band + rho, with their radii, and nothing else.
// Returns false (in every thread) if a capacity was exceeded. The walk
// covers the refinement's slightly larger ball (rho (1 + 1e-9) + 1e-9,
// mesh/extract.rs:172) too, and the items the refinement's gather of this
// sub-block (j = 8 brick + sb) will take go to `sup` (see SupList); the
// sampling's own tests are unchanged.
__device__ inline bool gather_sub(const DArr& A, const DLoc& Lc, const DRun& P, const Caps& cap, Lists& s,
                                  uint32_t* n, uint32_t* tmp4, const int32_t* key, uint32_t sb,
                                  const SupList& sup, uint32_t j, const uint32_t* xl, uint32_t xn,
                                  const BlkPool& blk, bool& same) {
    same = false;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    __shared__ uint32_t bfill[WAVE];
    for (uint32_t i = tid; i < cap.hash; i += nt) s.hk[i] = 0u;
    if (tid < C_N) n[tid] = 0u;
    if (tid < WAVE) bfill[tid] = 0u;
    __syncthreads();
    double h = P.h, sub = (double)P.sub, half = sub * 0.5;
    double rho = uniform_f64(0.5 * sqrt(3.0) * sub * h);
    const bool sph = A.kind != K_SES;
    double reach = (sph ? 0.0 : P.probe) + P.band + rho;
    // The refinement's ball and reach (refine_sub / gather_refine), computed
    // as it computes them.
    const double rho_r = 0.5 * sqrt(3.0) * sub * h * (1.0 + 1e-9) + 1e-9;
    const double reach_r = (sph ? 0.0 : A.probe) + P.band + rho_r;
    uint32_t* stage = (uint32_t*)s.cdat;  // free until the candidate records are stored
    const uint32_t stage_cap = sup.ids ? 12u * cap.cand : 0u;
    auto sup_add = [&](uint32_t code) {
        if (!stage_cap) return;
        const uint32_t at = atomicAdd(&n[C_SUP], 1u);
        if (at < stage_cap)
            stage[at] = code;
        else
            n[C_SUPOVER] = 1u;
    };
    double b[3];
    for (int a = 0; a < 3; a++) {
        double base = (double)((int64_t)key[a] * (int64_t)P.b) * h;
        double off = (double)((sb >> a) & 1) * sub + half;
        b[a] = uniform_f64(base + off * h);
    }
    const uint32_t hmask = cap.hash - 1;
    const double radii[3] = {reach_r + A.max_expanded, reach_r + A.rho_max, reach_r};
    const double per_bucket = KEY_BUCKETS / reach;
    auto bucket = [&](double key) { return min(KEY_BUCKETS - 1u, (uint32_t)(key * per_bucket)); };
    // The grids' items near the ball, or (xl) the brick classification's
    // superset of them; the tests are the same either way.
    auto visit = [&](uint32_t grid, uint32_t k) {
        if (grid == 0) {  // near spheres, the inside group and class S (gather_near, gather_ses)
            if (A.status[k] == S_CONTAINED) return;
            const SphereRec S = load_rec(A.sph, k);
            double R = S.R;
            double d = ddist(b[0], b[1], b[2], S.x, S.y, S.z);
            // gather_refine: SES near (d < R + reach) or class S (exposed,
            // |d - R| <= reach); SAS / vdW d - R < reach.
            if (sph ? d - R < reach_r
                    : (d < R + reach_r || (A.status[k] == S_EXPOSED && fabs(d - R) <= reach_r)))
                sup_add(k);
            if (!sph) {  // near and class S here and in the refinement's gather
                const bool ex = A.status[k] == S_EXPOSED, near_s = d < R + reach, near_r = d < R + reach_r;
                const double lb0 = fabs(d - R);
                if (near_s != near_r || (near_s && ex && lb0 <= reach) != (ex && lb0 <= reach_r)) n[C_DIFF] = 1u;
            }
            if (d >= R + reach) return;
            uint32_t at = atomicAdd(&n[C_NEAR], 1u);
            if (at >= cap.near) {
                n[C_OVER] = 1u;
                return;
            }
            const float4 p = load_f4(Lc.p4, k);
            if (sph) {
                s.nd[at] = p;
                return;
            }
            s.nd[at] = make_float4(p.x, p.y, p.z, rec_r2tol(p.w));
            hash_put(s.hk, s.hv, hmask, k, at);
            if (d < R + rho) {
                uint32_t q = atomicAdd(&n[C_IN], 1u);
                if (q >= cap.inside)
                    n[C_OVER] = 1u;
                else
                    s.id[q] = make_float4(p.x, p.y, p.z, rec_r2(p.w));
            }
            double lb = fabs(d - R);
            if (A.status[k] == S_EXPOSED && lb <= reach) {
                atomicAdd(&n[C_S], 1u);
                uint32_t q = atomicAdd(&n[C_CAND], 1u);
                if (q >= cap.cand) {
                    n[C_OVER] = 1u;
                } else {
                    s.cbk[q] = bucket(lb);
                    s.cid[q] = k;
                }
            }
        } else if (grid == 1) {  // class C (gather_ses, kernel.rs:701-736)
            const CircleRec C = load_rec(A.cir, k);
            double yx = b[0] - C.mx, yy = b[1] - C.my, yz = b[2] - C.mz;
            double ax = C.ax, ay = C.ay, az = C.az;
            double along = yx * ax + yy * ay + yz * az;
            double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
            double wl = sqrt(wx * wx + wy * wy + wz * wz);
            double rr = C.rho;
            double dist = sqrt(along * along + (wl - rr) * (wl - rr));
            if (dist <= reach_r) sup_add(k | (1u << KIND_SHIFT));
            if ((dist <= reach) != (dist <= reach_r)) n[C_DIFF] = 1u;
            if (!(dist <= reach)) return;
            atomicAdd(&n[C_C], 1u);
            uint32_t q = atomicAdd(&n[C_CAND], 1u);
            if (q >= cap.cand) {
                n[C_OVER] = 1u;
                return;
            }
            s.cbk[q] = KEY_BUCKETS + bucket(dist);
            s.cid[q] = k;
        } else {  // class V (kernel.rs:738-747)
            const VertexRec V = load_rec(A.ver, k);
            const double dv = ddist(b[0], b[1], b[2], V.x, V.y, V.z);
            if (dv <= reach_r) sup_add(k | (2u << KIND_SHIFT));
            if ((dv <= reach) != (dv <= reach_r)) n[C_DIFF] = 1u;
            if (!(dv <= reach)) return;
            uint32_t q = atomicAdd(&n[C_V], 1u);
            if (q >= cap.v) {
                n[C_OVER] = 1u;
                return;
            }
            float4 v = load_f4(Lc.v4, k);
            v.w = __uint_as_float(k);  // the lane kernel reads x, y, z
            s.vd[q] = v;
        }
    };
    if (xl) {
        for (uint32_t i = tid; i < xn; i += nt) {
            const uint32_t e = xl[i];
            visit(e >> KIND_SHIFT, e & ID_MASK);
        }
        __syncthreads();
    } else {
        gather3(A, b[0], b[1], b[2], radii, tmp4, visit);
    }
    if (n[C_OVER]) return false;
    if (stage_cap) {  // the superset to the pool (none if it overflowed)
        __shared__ uint32_t sup_at;
        const uint32_t m = n[C_SUP];
        if (threadIdx.x == 0) {
            uint32_t at = NO_LIST;
            if (!n[C_SUPOVER]) {
                at = atomicAdd(sup.used, m);
                if (at > sup.cap || m > sup.cap - at) at = NO_LIST;
            }
            sup_at = at;
            sup.off[j] = at == NO_LIST ? 0u : at;
            sup.cnt[j] = at == NO_LIST ? NO_LIST : m;
        }
        __syncthreads();
        if (sup_at != NO_LIST)
            for (uint32_t i = tid; i < m; i += nt) sup.ids[sup_at + i] = stage[i];
        __syncthreads();
    }
    const uint32_t gathered = n[C_CAND];
    // Candidates one blocker hides from the whole sub-block leave the lists.
    // The test runs on the f32 frame-local values the lane kernel uses (the
    // ball centre taken to that frame); they differ from the f64 records by
    // at most about 1.2e-4 per coordinate, far inside COVER_MARGIN, so the
    // refinement's f64 tests find these candidates blocked too.
    double bl[3];
    for (int a = 0; a < 3; a++) bl[a] = uniform_f64(b[a] - (double)P.frame[a] * P.h);
    mark_hidden(
        bl, uniform_f64(rho + COVER_PAD), gathered, n[C_NEAR], [&](uint32_t q) { return s.cbk[q] < KEY_BUCKETS ? 0u : 1u; },
        [&](uint32_t q) { return s.cid[q]; },
        [&](uint32_t kd, uint32_t k, dvec4& g0, dvec4& g1) {
            if (kd == 0u) {
                const float4 c = load_f4(Lc.p4, k);
                g0 = dvec4{c.x, c.y, c.z, c.w};
            } else {
                const float4 m = load_f4(Lc.m4, k), u = load_f4(Lc.u4, k);
                g0 = dvec4{m.x, m.y, m.z, u.x};
                g1 = dvec4{u.y, u.z, m.w, 0.0};
            }
        },
        [&](uint32_t slot) {
            const float4 k = s.nd[slot];
            return dvec4{k.x, k.y, k.z, k.w};
        },
        // The prefilter in f32: the projection rounded to f32 moves by at
        // most 1.1e-4 (coordinates below 2048), which moves the squared
        // distance of a blocker of radius below 9 by less than 2e-3, and
        // the f32 arithmetic adds a few ulps; the limit is widened by 2e-3
        // plus 1e-4 of itself, so every blocker the f64 test passes passes.
        [](double x, double y, double z) { return make_float4((float)x, (float)y, (float)z, 0.0f); },
        [&](const float4& p, uint32_t slot) {
            const float4 k = s.nd[slot];
            const float ex = k.x - p.x, ey = k.y - p.y, ez = k.z - p.z;
            return ex * ex + ey * ey + ez * ez < k.w + (2e-3f + 1e-4f * k.w);
        },
        [&](uint32_t q) { s.cbk[q] |= HIDDEN; });
    __syncthreads();
    for (uint32_t q = tid; q < gathered; q += nt)
        if (s.cbk[q] < HIDDEN) atomicAdd(&bfill[s.cbk[q]], 1u);
    __syncthreads();
    // Bucket starts: an exclusive scan of the counts by the first wavefront;
    // the kept class S candidates end at the start of the first class C
    // bucket, all kept ones at the end of the last.
    if (tid < WAVE) {
        const uint32_t c = bfill[tid];
        uint32_t inc = c;
        for (uint32_t o = 1; o < WAVE; o <<= 1) {
            uint32_t y = __shfl_up(inc, o, WAVE);
            if (tid >= o) inc += y;
        }
        bfill[tid] = inc - c;
    }
    __syncthreads();
    const uint32_t ns = bfill[KEY_BUCKETS], total = bfill[2 * KEY_BUCKETS];
    __syncthreads();
    if (tid == 0) {
        n[C_S] = ns;
        n[C_C] = total - ns;
    }
    // Each kept candidate takes the next slot p of its bucket; there its
    // record, cleared bitset and blocker row (neighbour / circle-blocker rows
    // restricted to near spheres, kernel.rs:685-735) are stored. Class S
    // fills [0, ns).
    const uint32_t cc = cap.cand, nw = cap.near / 64;
    for (uint32_t q = tid; q < gathered; q += nt) {
        if (s.cbk[q] >= HIDDEN) continue;
        const uint32_t p = atomicAdd(&bfill[s.cbk[q]], 1u);
        const uint32_t k = s.cid[q];
        // The candidate's id also goes to a slot of its record the lane
        // kernel does not read (S: the second float4, C: the third's w).
        if (p < ns) {
            s.cdat[3 * p] = load_f4(Lc.p4, k);
            s.cdat[3 * p + 1].x = __uint_as_float(k);
            s.coff[p] = A.nb_off[k];
            s.cpre[p] = A.nb_off[k + 1] - A.nb_off[k];
        } else {
            s.cdat[3 * p] = load_f4(Lc.m4, k);
            s.cdat[3 * p + 1] = load_f4(Lc.u4, k);
            float4 e = load_f4(Lc.e4, k);
            e.w = __uint_as_float(k);
            s.cdat[3 * p + 2] = e;
            s.coff[p] = A.cb_off[k];
            s.cpre[p] = A.cb_off[k + 1] - A.cb_off[k];
        }
        for (uint32_t w = 0; w < nw; w++) s.cbits[w * cc + p] = 0ull;
    }
    __syncthreads();
    // Blocker bitsets of the S and C rows (cpre holds the row length). A
    // blocker that no projection of the ball comes near stays out of the
    // bitset: it could never block a sample of the sub-block. The test runs
    // in f32 on the f32 frame-local records the lane kernel uses (the ball
    // centre rounded to f32); its limits carry COVER_MARGIN, which holds the
    // f32 rounding of the samples and of their projections, and REACH_SLACK.
    // Each row's hint (then in cpre) is the kept blocker holding the row's
    // central projection p0 deepest (least |k - p0|^2 - t^2, when
    // negative), which blocked_f() tries first.
    struct HintRow {
        ReachRow<float> w;
        float px, py, pz, best;
        uint32_t hint;
    };
    const float bf[3] = {(float)bl[0], (float)bl[1], (float)bl[2]};
    const float rbf = (float)(rho + COVER_PAD) + REACH_SLACK;
    blocker_bits(
        A.nb_items, A.cb_items, ns, total, s.coff, s.cpre, s.hk, s.hv, hmask, s.cbits, cc,
        [&](uint32_t p) {
            HintRow r;
            const float4 c = s.cdat[3 * p];
            if (p < ns) {
                r.w = cap_row(bf, rbf, c.x, c.y, c.z, c.w);
            } else {
                const float4 u = s.cdat[3 * p + 1];
                r.w = arc_row(bf, rbf, c.x, c.y, c.z, u.x, u.y, u.z, c.w);
            }
            const float vl = sqrtf(r.w.D2), sc = vl > 0.0f ? r.w.R / vl : 0.0f;
            r.px = r.w.ox + r.w.vx * sc;
            r.py = r.w.oy + r.w.vy * sc;
            r.pz = r.w.oz + r.w.vz * sc;
            r.best = 0.0f;
            r.hint = 0xffffffffu;
            return r;
        },
        [&](HintRow& r, uint32_t p, uint32_t slot) {
            const float4 k = s.nd[slot];
            const float t = sqrtf(k.w) + ((float)COVER_MARGIN + REACH_SLACK);
            const bool reach = p < ns ? cap_reaches(r.w, k.x, k.y, k.z, t) : arc_reaches(r.w, k.x, k.y, k.z, t);
            if (reach) {
                const float ex = k.x - r.px, ey = k.y - r.py, ez = k.z - r.pz;
                const float pw = ex * ex + ey * ey + ez * ez - t * t;
                if (pw < r.best) {
                    r.best = pw;
                    r.hint = slot;
                }
            }
            return reach;
        },
        [&](HintRow& r, uint32_t p) {
            auto step = [&](float ob, uint32_t oh) {
                if (ob < r.best || (ob == r.best && oh < r.hint)) {
                    r.best = ob;
                    r.hint = oh;
                }
            };
            step(__shfl_xor(r.best, 1, 8), __shfl_xor(r.hint, 1, 8));
            step(__shfl_xor(r.best, 2, 8), __shfl_xor(r.hint, 2, 8));
            step(__shfl_xor(r.best, 4, 8), __shfl_xor(r.hint, 4, 8));
            if ((threadIdx.x & 7u) == 0u) s.cpre[p] = r.hint;
        });
    __syncthreads();
    same = blk.ids && total && write_blk_record(A, s, n, cap, blk, j, b, ns, total);
    return true;
}

// Candidate load of a sub-block's sampling lists in units of the first
// tier's candidate capacity: max(S + C, 2 near, 4 SAS vertices, 4 inside).
__device__ inline uint32_t candidate_load(uint32_t cand, uint32_t near, uint32_t verts, uint32_t inside) {
    return max(max(cand, 2u * near), max(4u * verts, 4u * inside));
}

// Samples one sub-block per workgroup: j = 8 * brick + sb of a non-uniform
// sub-block. The lists live in dynamic LDS (GLOBAL false) or in scratch
// slice blockIdx.x of `slice` bytes (GLOBAL true). Sub-blocks whose lists do
// not fit are appended to `overflow`.
template <bool GLOBAL>
__device__ __forceinline__ void sample_sub(const Mol* __restrict__ mols, const uint32_t* bmol, const Caps& cap,
                                           const Layout& lay, const int32_t* keys, const uint8_t* st, uint32_t j,
                                           float* values, char* scratch, size_t slice, uint32_t* overflow,
                                           uint32_t* n_overflow, unsigned long long* n_samples, uint32_t* loads,
                                           const HidList& hid) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[C_N];
    __shared__ uint32_t tmp4[4];  // block scans: one slot per wavefront
    __shared__ uint32_t owned[512];
    __shared__ uint32_t n_owned;
    uint32_t brick = j >> 3, sb = j & 7;
    const uint32_t m = bmol[brick];
    const DArr& A = mols[m].A;
    const DLoc& Lc = mols[m].L;
    const DRun& P = mols[m].P;
    char* base = GLOBAL ? scratch + (size_t)blockIdx.x * slice : smem;
    Lists s = carve(base, lay);
    const int32_t* key = keys + 3 * brick;
    if (threadIdx.x == 0) n_owned = 0;
    const uint32_t* xl = nullptr;  // the brick's superset from its classification
    uint32_t xn = hid.bsup.cnt ? hid.bsup.cnt[brick] : NO_LIST;
    if (xn != NO_LIST) xl = hid.bsup.ids + hid.bsup.off[brick];
    bool same;  // a blocker record that gives the refinement its lists
    bool ok = gather_sub(A, Lc, P, cap, s, n, tmp4, key, sb, hid.sup, j, xl, xn, hid.blk, same);
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = j;
        return;
    }
    if (loads && threadIdx.x == 0) atomicMax(&loads[brick], candidate_load(n[C_CAND], n[C_NEAR], n[C_V], n[C_IN]));
    if (hid.ids && !same) {  // the hidden candidates, for this sub-block's refinement
        __shared__ uint32_t h_at, h_fill;
        const uint32_t gathered = n[C_CAND], nh = gathered - n[C_S] - n[C_C];
        if (threadIdx.x == 0) {
            const uint32_t at = atomicAdd(hid.used, nh);
            h_at = at <= hid.cap && nh <= hid.cap - at ? at : NO_LIST;
            h_fill = 0;
            if (h_at != NO_LIST) {
                hid.off[j] = at;
                hid.cnt[j] = nh;
            }
        }
        __syncthreads();
        if (h_at != NO_LIST)
            for (uint32_t q = threadIdx.x; q < gathered; q += blockDim.x) {
                const uint32_t bk = s.cbk[q];
                if (bk >= HIDDEN)
                    hid.ids[h_at + atomicAdd(&h_fill, 1u)] =
                        s.cid[q] | (bk - HIDDEN >= KEY_BUCKETS ? 1u << KIND_SHIFT : 0u);
            }
    }
    const uint32_t cc = cap.cand, nw = cap.near / 64;
    uint32_t stb = pack_states(st + 8 * brick);
    uint32_t sub = P.sub, box = sub + 1, side = P.side;
    uint32_t bx = (sb & 1) * sub, by = ((sb >> 1) & 1) * sub, bz = ((sb >> 2) & 1) * sub;
    int64_t org[3];
    for (int a = 0; a < 3; a++) org[a] = (int64_t)key[a] * (int64_t)P.b - P.frame[a];
    float* out = values + (size_t)brick * P.spb;
    const uint32_t nbox = box * box * box;
    // Samples this sub-block owns (Engine::assign), compacted onto the lanes
    // in chunks of at most 512.
    for (uint32_t c0 = 0; c0 < nbox; c0 += 512) {
        for (uint32_t t = c0 + threadIdx.x; t < min(nbox, c0 + 512); t += blockDim.x) {
            uint32_t lx = bx + t % box, ly = by + (t / box) % box, lz = bz + t / (box * box);
            bool positive;
            if (sample_owner(stb, sub, lx, ly, lz, positive) == (int)sb) owned[atomicAdd(&n_owned, 1u)] = t;
        }
        __syncthreads();
        uint32_t cnt = n_owned;
        for (uint32_t q = threadIdx.x; q < cnt; q += blockDim.x) {
            uint32_t t = owned[q];
            uint32_t lx = bx + t % box, ly = by + (t / box) % box, lz = bz + t / (box * box);
            // Lattice::local (lattice.rs:69-71): integer offset, then one f32 multiply.
            float x = (float)(int32_t)(org[0] + lx) * P.h32;
            float y = (float)(int32_t)(org[1] + ly) * P.h32;
            float z = (float)(int32_t)(org[2] + lz) * P.h32;
            out[lx + side * (ly + side * lz)] = A.kind != K_SES ? sphere_value(s, n, x, y, z, P.fband)
                                                                : ses_value(s, n, cc, nw, x, y, z, P.fprobe, P.fcutoff2);
        }
        if (threadIdx.x == 0 && n_samples) atomicAdd(&n_samples[m], (unsigned long long)cnt);
        __syncthreads();
        if (threadIdx.x == 0) n_owned = 0;
        __syncthreads();
    }
}

// PERSIST: residency-sized grid looping over the queue; otherwise one
// sub-block per workgroup. At most 80 VGPRs (6 wavefronts per SIMD), so six
// workgroups with the first tier's lists fit a CU: faster than five
// unbounded ones despite a few spilled registers.
template <bool GLOBAL, bool PERSIST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_sample(const Mol* __restrict__ mols, const uint32_t* bmol, Caps cap, Layout lay, const int32_t* keys,
         const uint8_t* st, Queue q, float* values, char* scratch, size_t slice, uint32_t* overflow,
         uint32_t* n_overflow, unsigned long long* n_samples, uint32_t* loads, HidList hid) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t j) {
        sample_sub<GLOBAL>(mols, bmol, cap, lay, keys, st, j, values, scratch, slice, overflow, n_overflow,
                           n_samples, loads, hid);
    });
}

}  // namespace smg
