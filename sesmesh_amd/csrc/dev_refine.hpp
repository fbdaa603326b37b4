// Crossing edges, exact f64 vertex refinement and vertex attributes.
//
// Restates the edge scan of mesh_brick (reference mesh/extract.rs:146-223),
// SesField::gather_local / eval_local (field/point.rs:419-489, 496-565,
// 686-722) and vertex / gradient / attributes (mesh/extract.rs:425-597) in
// f64 with the CPU's operations, candidate order (sorted by f64::total_cmp
// then index) and iteration rules, so every vertex, normal, atom and patch
// equals the CPU back end's bit for bit.
//
// One workgroup (4 wavefronts) per sub-block with crossing edges gathers the
// sub-block's candidate lists into LDS with their f64 data, once; then each
// wavefront refines one edge at a time, evaluating each point cooperatively:
// lanes split the candidates, and the winner is found best-first, by
// (distance, list position) over the candidates, checking the blockers of
// the nearest unchecked candidate until one is free. This selects exactly
// the candidate eval_local's sequential first-wins loop selects: the first,
// in list order, among the free candidates of least distance below the
// cutoff (a candidate taken earlier in the sequential loop is always
// replaced by it, and none after it improves on it).
#pragma once
#include "dev_common.hpp"

namespace smg {

__device__ inline bool edge_crossing(const float* val, uint32_t b, uint32_t side, uint32_t id, uint32_t& x,
                                     uint32_t& y, uint32_t& z, uint32_t& axis) {
    axis = id % 3u;
    uint32_t cell = id / 3u;
    x = cell % b;
    y = (cell / b) % b;
    z = cell / (b * b);
    uint32_t i0 = x + side * (y + side * z);
    uint32_t st = axis == 0 ? 1u : (axis == 1 ? side : side * side);
    return (val[i0] < 0.0f) != (val[i0 + st] < 0.0f);
}

// Per brick (one workgroup of 256): crossing-edge masks (edge ids
// ((z*b + y)*b + x)*3 + axis, 64 per word), per-word prefix counts, brick
// total and per-sub-block counts (mesh/extract.rs:176-216).
__global__ void __launch_bounds__(256) k_edges(DRun P, const float* values, uint32_t n_bricks, uint32_t words,
                                               unsigned long long* mask, uint32_t* wpre, uint32_t* bcount,
                                               uint32_t* scount) {
    __shared__ uint32_t wc[2048];
    __shared__ uint32_t sc[8];
    uint32_t brick = blockIdx.x;
    if (brick >= n_bricks) return;
    const float* val = values + (size_t)brick * P.spb;
    uint32_t b = P.b, nid = 3 * b * b * b, lane = lane_id(), wv = threadIdx.x / WAVE;
    if (threadIdx.x < 8) sc[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t w = wv; w < words; w += blockDim.x / WAVE) {
        uint32_t id = w * 64 + lane, x = 0, y = 0, z = 0, axis = 0;
        bool c = id < nid && edge_crossing(val, b, P.side, id, x, y, z, axis);
        unsigned long long m = __ballot(c);
        if (c) {
            uint32_t s = min(x / P.sub, 1u) + 2 * min(y / P.sub, 1u) + 4 * min(z / P.sub, 1u);
            atomicAdd(&sc[s], 1u);
        }
        if (lane == 0) {
            mask[(size_t)brick * words + w] = m;
            wc[w] = (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t w = 0; w < words; w++) {
            wpre[(size_t)brick * words + w] = acc;
            acc += wc[w];
        }
        bcount[brick] = acc;
    }
    if (threadIdx.x < 8) scount[8 * brick + threadIdx.x] = sc[threadIdx.x];
}

__device__ inline uint32_t edge_rank(const unsigned long long* mask, const uint32_t* wpre, uint32_t words,
                                     uint32_t brick, uint32_t id) {
    size_t w = (size_t)brick * words + id / 64;
    unsigned long long below = (id % 64) ? (mask[w] & ((1ull << (id % 64)) - 1ull)) : 0ull;
    return wpre[w] + (uint32_t)__popcll(below);
}

__device__ inline bool edge_is_crossing(const unsigned long long* mask, uint32_t words, uint32_t brick,
                                        uint32_t id) {
    return (mask[(size_t)brick * words + id / 64] >> (id % 64)) & 1ull;
}

// Capacities of the refinement lists (s, c, v powers of two).
struct RCaps {
    uint32_t near, inside, s, c, v, sblk, cblk, hash, edges;
};

// LDS / scratch layout of the refinement lists: f64 SoA data.
struct RLayout {
    uint32_t nx, ny, nz, nrm, nid, hk, hv, ix, iy, iz, iR, skey, sid, sx, sy, sz, sR, sb0, sb1, ckey, cid, cd, cb0,
        cb1, vkey, vid, vx, vy, vz, sblk, cblk, ebuf, dm, bytes;
};

__host__ __device__ inline uint32_t pow2_at_least(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

__host__ __device__ inline RLayout make_rlayout(const RCaps& c) {
    RLayout L;
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    L.nx = take(8 * c.near);
    L.ny = take(8 * c.near);
    L.nz = take(8 * c.near);
    L.nrm = take(8 * c.near);
    L.nid = take(4 * c.near);
    L.hk = take(4 * c.hash);
    L.hv = take(4 * c.hash);
    L.ix = take(8 * c.inside);
    L.iy = take(8 * c.inside);
    L.iz = take(8 * c.inside);
    L.iR = take(8 * c.inside);
    L.skey = take(8 * c.s);
    L.sid = take(4 * c.s);
    L.sx = take(8 * c.s);
    L.sy = take(8 * c.s);
    L.sz = take(8 * c.s);
    L.sR = take(8 * c.s);
    L.sb0 = take(4 * c.s);
    L.sb1 = take(4 * c.s);
    L.ckey = take(8 * c.c);
    L.cid = take(4 * c.c);
    L.cd = take(80 * c.c);  // 10 f64 per circle, SoA of c.c each
    L.cb0 = take(4 * c.c);
    L.cb1 = take(4 * c.c);
    L.vkey = take(8 * c.v);
    L.vid = take(4 * c.v);
    L.vx = take(8 * c.v);
    L.vy = take(8 * c.v);
    L.vz = take(8 * c.v);
    L.sblk = take(2 * c.sblk);
    L.cblk = take(2 * c.cblk);
    L.ebuf = take(4 * c.edges);
    // Candidate distances per wavefront, for lists longer than the
    // register budget (4 wavefronts per workgroup).
    uint32_t tot = c.s + c.c + c.v;
    L.dm = take(tot > 256u ? 4u * 8u * tot : 0u);
    L.bytes = o;
    return L;
}

struct RLists {
    double *nx, *ny, *nz, *nrm;
    uint32_t *nid, *hk, *hv;
    double *ix, *iy, *iz, *iR;
    unsigned long long* skey;
    uint32_t* sid;
    double *sx, *sy, *sz, *sR;
    uint32_t *sb0, *sb1;
    unsigned long long* ckey;
    uint32_t* cid;
    double* cd;
    uint32_t *cb0, *cb1;
    unsigned long long* vkey;
    uint32_t* vid;
    double *vx, *vy, *vz;
    uint16_t *sblk, *cblk;
    uint32_t* ebuf;
    double* dm;
};

__device__ inline RLists rcarve(char* base, const RLayout& L) {
    RLists s;
    s.nx = (double*)(base + L.nx);
    s.ny = (double*)(base + L.ny);
    s.nz = (double*)(base + L.nz);
    s.nrm = (double*)(base + L.nrm);
    s.nid = (uint32_t*)(base + L.nid);
    s.hk = (uint32_t*)(base + L.hk);
    s.hv = (uint32_t*)(base + L.hv);
    s.ix = (double*)(base + L.ix);
    s.iy = (double*)(base + L.iy);
    s.iz = (double*)(base + L.iz);
    s.iR = (double*)(base + L.iR);
    s.skey = (unsigned long long*)(base + L.skey);
    s.sid = (uint32_t*)(base + L.sid);
    s.sx = (double*)(base + L.sx);
    s.sy = (double*)(base + L.sy);
    s.sz = (double*)(base + L.sz);
    s.sR = (double*)(base + L.sR);
    s.sb0 = (uint32_t*)(base + L.sb0);
    s.sb1 = (uint32_t*)(base + L.sb1);
    s.ckey = (unsigned long long*)(base + L.ckey);
    s.cid = (uint32_t*)(base + L.cid);
    s.cd = (double*)(base + L.cd);
    s.cb0 = (uint32_t*)(base + L.cb0);
    s.cb1 = (uint32_t*)(base + L.cb1);
    s.vkey = (unsigned long long*)(base + L.vkey);
    s.vid = (uint32_t*)(base + L.vid);
    s.vx = (double*)(base + L.vx);
    s.vy = (double*)(base + L.vy);
    s.vz = (double*)(base + L.vz);
    s.sblk = (uint16_t*)(base + L.sblk);
    s.cblk = (uint16_t*)(base + L.cblk);
    s.ebuf = (uint32_t*)(base + L.ebuf);
    s.dm = (double*)(base + L.dm);
    return s;
}

enum { R_NEAR = 0, R_IN, R_S, R_C, R_V, R_SB, R_CB, R_E, R_OVER, R_N };

// Sorts (key, id) ascending by key then id (count <= cap, cap a power of
// two), by the whole workgroup (bitonic network).
__device__ inline void block_sort(unsigned long long* key, uint32_t* id, uint32_t count) {
    uint32_t n = pow2_at_least(count);
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t i = count + tid; i < n; i += nt) {
        key[i] = ~0ull;
        id[i] = 0xffffffffu;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= n; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < n; i += nt) {
                uint32_t l = i ^ j;
                if (l > i) {
                    unsigned long long ki = key[i], kl = key[l];
                    uint32_t ii = id[i], il = id[l];
                    bool gt = ki > kl || (ki == kl && ii > il);
                    if (gt == ((i & k) == 0)) {
                        key[i] = kl;
                        key[l] = ki;
                        id[i] = il;
                        id[l] = ii;
                    }
                }
            }
            __syncthreads();
        }
}

// gather_local (field/point.rs:496-565) of the ball (b, rho) by the
// workgroup, with the f64 data of every list in sorted order. False (in
// every thread) if a capacity was exceeded.
__device__ inline bool gather_local_block(const DArr& A, const RCaps& cap, RLists& s, uint32_t* n, double bx,
                                          double by, double bz, double rho, double band) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t i = tid; i < cap.hash; i += nt) s.hk[i] = 0u;
    if (tid < R_N) n[tid] = 0u;
    __syncthreads();
    const uint32_t hmask = cap.hash - 1;
    double reach = A.probe + band + rho;
    for_near(A.gs, bx, by, bz, reach + A.max_expanded, tid, nt, [&](uint32_t k) {
        if (A.status[k] == S_CONTAINED) return;
        double R = A.R[k];
        double d = ddist(bx, by, bz, A.cx[k], A.cy[k], A.cz[k]);
        if (d < R + reach) {
            uint32_t at = atomicAdd(&n[R_NEAR], 1u);
            if (at >= cap.near) {
                n[R_OVER] = 1u;
            } else {
                s.nx[at] = A.cx[k];
                s.ny[at] = A.cy[k];
                s.nz[at] = A.cz[k];
                s.nrm[at] = R - EPS;
                s.nid[at] = k;
                hash_put(s.hk, s.hv, hmask, k, at);
            }
        }
        if (d < R + rho) {
            uint32_t at = atomicAdd(&n[R_IN], 1u);
            if (at >= cap.inside) {
                n[R_OVER] = 1u;
            } else {
                s.ix[at] = A.cx[k];
                s.iy[at] = A.cy[k];
                s.iz[at] = A.cz[k];
                s.iR[at] = R;
            }
        }
        double lb = fabs(d - R);
        if (A.status[k] == S_EXPOSED && lb <= reach) {
            uint32_t at = atomicAdd(&n[R_S], 1u);
            if (at >= cap.s) {
                n[R_OVER] = 1u;
            } else {
                s.skey[at] = total_key(lb);
                s.sid[at] = k;
            }
        }
    });
    for_near(A.gc, bx, by, bz, reach + A.rho_max, tid, nt, [&](uint32_t c) {
        double yx = bx - A.mx[c], yy = by - A.my[c], yz = bz - A.mz[c];
        double ax = A.ax[c], ay = A.ay[c], az = A.az[c];
        double along = yx * ax + yy * ay + yz * az;
        double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
        double wl = sqrt(wx * wx + wy * wy + wz * wz);
        double rr = A.crho[c];
        double dist = sqrt(along * along + (wl - rr) * (wl - rr));
        if (!(dist <= reach)) return;
        uint32_t at = atomicAdd(&n[R_C], 1u);
        if (at >= cap.c) {
            n[R_OVER] = 1u;
            return;
        }
        s.ckey[at] = total_key(dist);
        s.cid[at] = c;
    });
    for_near(A.gv, bx, by, bz, reach, tid, nt, [&](uint32_t v) {
        double d = ddist(bx, by, bz, A.vx[v], A.vy[v], A.vz[v]);
        if (!(d <= reach)) return;
        uint32_t at = atomicAdd(&n[R_V], 1u);
        if (at >= cap.v) {
            n[R_OVER] = 1u;
            return;
        }
        s.vkey[at] = total_key(d);
        s.vid[at] = v;
    });
    __syncthreads();
    if (n[R_OVER]) return false;
    block_sort(s.skey, s.sid, n[R_S]);
    block_sort(s.ckey, s.cid, n[R_C]);
    block_sort(s.vkey, s.vid, n[R_V]);
    // Candidate data in sorted order, and blocker lists (row members that
    // are near, point.rs:152-155) as near-list slots.
    uint32_t ns = n[R_S], nc = n[R_C], nv = n[R_V];
    const uint32_t cc = cap.c;
    for (uint32_t i = tid; i < ns + nc + nv; i += nt) {
        if (i < ns) {
            uint32_t k = s.sid[i];
            s.sx[i] = A.cx[k];
            s.sy[i] = A.cy[k];
            s.sz[i] = A.cz[k];
            s.sR[i] = A.R[k];
        } else if (i < ns + nc) {
            uint32_t q = i - ns, c = s.cid[q];
            double* d = s.cd;
            d[q] = A.mx[c];
            d[cc + q] = A.my[c];
            d[2 * cc + q] = A.mz[c];
            d[3 * cc + q] = A.ax[c];
            d[4 * cc + q] = A.ay[c];
            d[5 * cc + q] = A.az[c];
            d[6 * cc + q] = A.e1x[c];
            d[7 * cc + q] = A.e1y[c];
            d[8 * cc + q] = A.e1z[c];
            d[9 * cc + q] = A.crho[c];
        } else {
            uint32_t q = i - ns - nc, v = s.vid[q];
            s.vx[q] = A.vx[v];
            s.vy[q] = A.vy[v];
            s.vz[q] = A.vz[v];
        }
        if (i < ns + nc) {
            bool is_s = i < ns;
            uint32_t id = is_s ? s.sid[i] : s.cid[i - ns];
            const uint32_t* row = is_s ? A.nb_items + A.nb_off[id] : A.cb_items + A.cb_off[id];
            uint32_t len = is_s ? A.nb_off[id + 1] - A.nb_off[id] : A.cb_off[id + 1] - A.cb_off[id];
            uint32_t cnt = 0;
            for (uint32_t t = 0; t < len; t++) cnt += hash_get(s.hk, s.hv, hmask, row[t]) != 0xffffffffu;
            uint32_t at = atomicAdd(&n[is_s ? R_SB : R_CB], cnt);
            if (at + cnt > (is_s ? cap.sblk : cap.cblk)) {
                n[R_OVER] = 1u;
                continue;
            }
            uint16_t* out = is_s ? s.sblk : s.cblk;
            uint32_t w = at;
            for (uint32_t t = 0; t < len; t++) {
                uint32_t l = hash_get(s.hk, s.hv, hmask, row[t]);
                if (l != 0xffffffffu) out[w++] = (uint16_t)l;
            }
            if (is_s) {
                s.sb0[i] = at;
                s.sb1[i] = at + cnt;
            } else {
                s.cb0[i - ns] = at;
                s.cb1[i - ns] = at + cnt;
            }
        }
    }
    __syncthreads();
    return n[R_OVER] == 0u;
}

struct EvalD {
    double value;
    uint32_t kind, id;
    double qx, qy, qz;
};

// Lexicographic minimum of (d, o) over the wave.
__device__ inline void wave_argmin(double& d, uint32_t& o) {
    for (int m = 32; m > 0; m >>= 1) {
        double od = __shfl_xor(d, m, WAVE);
        uint32_t oo = __shfl_xor(o, m, WAVE);
        if (od < d || (od == d && oo < o)) {
            d = od;
            o = oo;
        }
    }
}

constexpr int MAXC = 4;  // candidates per lane held in registers

// Distance of candidate o (list position: S, then C, then V) from (x, y, z)
// as eval_local compares it (field/point.rs:419-489), +inf where the
// candidate cannot be taken (no projection, or not below the cutoff).
__device__ inline double cand_dist(const RLists& s, uint32_t ns, uint32_t nc, uint32_t total, uint32_t ccap,
                                   uint32_t o, double x, double y, double z, double cutoff) {
    const double* cd = s.cd;
    double d = INFINITY;
    if (o < ns) {
        double vx = x - s.sx[o], vy = y - s.sy[o], vz = z - s.sz[o];
        double len = sqrt(vx * vx + vy * vy + vz * vz);
        double lb = fabs(len - s.sR[o]);
        if (!(len < 1e-12)) d = lb;
    } else if (o < ns + nc) {
        uint32_t q = o - ns;
        double yx = x - cd[q], yy = y - cd[ccap + q], yz = z - cd[2 * ccap + q];
        double ax = cd[3 * ccap + q], ay = cd[4 * ccap + q], az = cd[5 * ccap + q], rho = cd[9 * ccap + q];
        double along = yx * ax + yy * ay + yz * az;
        double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
        double wl = sqrt(wx * wx + wy * wy + wz * wz);
        d = sqrt(along * along + (wl - rho) * (wl - rho));
    } else if (o < total) {
        uint32_t q = o - ns - nc;
        d = ddist(x, y, z, s.vx[q], s.vy[q], s.vz[q]);
    }
    return d < cutoff ? d : INFINITY;
}

// eval_local (field/point.rs:686-722) at (x, y, z) by one wavefront.
// `dm` (per wavefront) holds candidate distances when the lists exceed
// WAVE * MAXC candidates; otherwise registers do.
__device__ inline EvalD eval_wave(const DArr& A, const RLists& s, const uint32_t* n, uint32_t ccap, double x,
                                  double y, double z, double band, double* dm) {
    const uint32_t lane = lane_id();
    EvalD e;
    e.qx = x;
    e.qy = y;
    e.qz = z;
    bool ins = false;
    for (uint32_t q = lane; q < n[R_IN]; q += WAVE) ins |= ddist(x, y, z, s.ix[q], s.iy[q], s.iz[q]) < s.iR[q];
    if (!wave_any(ins)) {
        e.value = A.probe;
        e.kind = F_SOLVENT;
        e.id = 0;
        return e;
    }
    const uint32_t ns = n[R_S], nc = n[R_C], nv = n[R_V], total = ns + nc + nv;
    const double cutoff = A.probe + band;
    const double* cd = s.cd;
    const bool in_regs = total <= WAVE * MAXC;
    double dist[MAXC];
    if (in_regs) {
        for (int j = 0; j < MAXC; j++) dist[j] = cand_dist(s, ns, nc, total, ccap, lane + WAVE * j, x, y, z, cutoff);
    } else {
        for (uint32_t o = lane; o < total; o += WAVE) dm[o] = cand_dist(s, ns, nc, total, ccap, o, x, y, z, cutoff);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    for (;;) {
        double bd = INFINITY;
        uint32_t bo = 0xffffffffu;
        if (in_regs) {
            for (int j = 0; j < MAXC; j++)
                if (dist[j] < bd) {
                    bd = dist[j];
                    bo = lane + WAVE * j;
                }
        } else {
            for (uint32_t o = lane; o < total; o += WAVE)
                if (dm[o] < bd) {
                    bd = dm[o];
                    bo = o;
                }
        }
        wave_argmin(bd, bo);
        if (bo == 0xffffffffu) {  // nothing free below the cutoff: Beyond
            e.value = A.probe - cutoff;
            e.kind = F_BEYOND;
            e.id = 0;
            return e;
        }
        if (bo % WAVE == lane) {
            if (in_regs)
                dist[bo / WAVE] = INFINITY;
            else
                dm[bo] = INFINITY;
        }
        double qx, qy, qz;
        uint32_t b0 = 0, b1 = 0;
        const uint16_t* blk = s.sblk;
        uint32_t kind, id;
        if (bo < ns) {
            double cx = s.sx[bo], cy = s.sy[bo], cz = s.sz[bo], Ri = s.sR[bo];
            double vx = x - cx, vy = y - cy, vz = z - cz;
            double len = sqrt(vx * vx + vy * vy + vz * vz);
            double sc = Ri / len;
            qx = cx + vx * sc;
            qy = cy + vy * sc;
            qz = cz + vz * sc;
            b0 = s.sb0[bo];
            b1 = s.sb1[bo];
            kind = F_SPHERE;
            id = s.sid[bo];
        } else if (bo < ns + nc) {
            uint32_t q = bo - ns;
            double mx = cd[q], my = cd[ccap + q], mz = cd[2 * ccap + q];
            double ax = cd[3 * ccap + q], ay = cd[4 * ccap + q], az = cd[5 * ccap + q], rho = cd[9 * ccap + q];
            double yx = x - mx, yy = y - my, yz = z - mz;
            double along = yx * ax + yy * ay + yz * az;
            double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
            double wl = sqrt(wx * wx + wy * wy + wz * wz);
            double dx, dy, dz;
            if (wl > 1e-12) {
                dx = wx / wl;
                dy = wy / wl;
                dz = wz / wl;
            } else {
                dx = cd[6 * ccap + q];
                dy = cd[7 * ccap + q];
                dz = cd[8 * ccap + q];
            }
            qx = mx + dx * rho;
            qy = my + dy * rho;
            qz = mz + dz * rho;
            b0 = s.cb0[q];
            b1 = s.cb1[q];
            blk = s.cblk;
            kind = F_CIRCLE;
            id = s.cid[q];
        } else {
            uint32_t q = bo - ns - nc;
            qx = s.vx[q];
            qy = s.vy[q];
            qz = s.vz[q];
            kind = F_VERTEX;
            id = s.vid[q];
        }
        bool blocked = false;
        for (uint32_t t = b0 + lane; t < b1; t += WAVE) {
            uint32_t k = blk[t];
            blocked |= !(ddist(qx, qy, qz, s.nx[k], s.ny[k], s.nz[k]) >= s.nrm[k]);
        }
        if (!wave_any(blocked)) {
            e.value = A.probe - bd;
            e.kind = kind;
            e.id = id;
            e.qx = qx;
            e.qy = qy;
            e.qz = qz;
            return e;
        }
    }
}

// Output arrays of the vertices.
struct VOut {
    float* pos;
    float* nrm;
    uint32_t* atom;
    uint8_t* patch;
    double* t;
    uint32_t* evals;
};

// vertex() + attributes() (mesh/extract.rs:425-535, 553-597) of the edge
// from lattice point ia along `axis`, by one wavefront; lane 0 writes.
__device__ inline uint32_t refine_vertex_wave(const DArr& A, const DRun& P, const RLists& s, const uint32_t* n,
                                              uint32_t ccap, double* dm, const int64_t ia[3], uint32_t axis,
                                              float fa32, float fb32, const VOut& o, size_t slot) {
    const double h = P.h, band = P.band;
    double pa[3] = {(double)ia[0] * h, (double)ia[1] * h, (double)ia[2] * h};
    int64_t ib[3] = {ia[0], ia[1], ia[2]};
    ib[axis] += 1;
    double pb[3] = {(double)ib[0] * h, (double)ib[1] * h, (double)ib[2] * h};
    double d[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    double fa = (double)fa32, fb = (double)fb32;
    double t = fmin(fmax(fa / (fa - fb), 0.0), 1.0);
    double t0 = 0.0, g0 = fa, t1 = 1.0, g1 = fb;
    int side = 0;
    bool converged = false;
    uint32_t evals = 0;
    EvalD last;
    double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int it = 0; it < 8; it++) {
        double x = pa[0] + d[0] * t, y = pa[1] + d[1] * t, z = pa[2] + d[2] * t;
        EvalD e = eval_wave(A, s, n, ccap, x, y, z, band, dm);
        evals++;
        last = e;
        double g = e.value;
        if (fabs(g) <= 1e-5) {
            converged = true;
            break;
        }
        if ((g < 0.0) == (g0 < 0.0)) {
            t0 = t;
            g0 = g;
            if (side == -1) g1 *= 0.5;
            side = -1;
        } else {
            t1 = t;
            g1 = g;
            if (side == 1) g0 *= 0.5;
            side = 1;
        }
        double slope = 0.0;
        if (e.kind != F_SOLVENT && e.kind != F_BEYOND) {
            double wx = e.qx - x, wy = e.qy - y, wz = e.qz - z;
            double len = sqrt(wx * wx + wy * wy + wz * wz);
            if (len > 1e-12) {
                double nx = wx / len, ny = wy / len, nz = wz / len;
                slope = nx * d[0] + ny * d[1] + nz * d[2];
            }
        }
        double newton = t - g / slope;
        if (fabs(slope) > 1e-3 * dn && newton > t0 && newton < t1) {
            t = newton;
        } else {
            double sv = (t0 * g1 - t1 * g0) / (g1 - g0);
            t = (isfinite(sv) && sv > t0 && sv < t1) ? sv : 0.5 * (t0 + t1);
        }
    }
    if (!converged) {
        bool neg0 = g0 < 0.0;
        for (int it = 0; it < 64; it++) {
            t = 0.5 * (t0 + t1);
            EvalD e = eval_wave(A, s, n, ccap, pa[0] + d[0] * t, pa[1] + d[1] * t, pa[2] + d[2] * t, band, dm);
            evals++;
            last = e;
            if (fabs(e.value) <= 1e-5 || (t1 - t0) * dn <= 2.0 * 1e-5) break;
            if ((e.value < 0.0) == neg0)
                t0 = t;
            else
                t1 = t;
        }
    }
    double tc = fmin(fmax(t, P.t_min), 1.0 - P.t_min);
    double v[3] = {pa[0] + d[0] * tc, pa[1] + d[1] * tc, pa[2] + d[2] * tc};
    double ow[3];
    for (int a = 0; a < 3; a++) ow[a] = fa < 0.0 ? d[a] : -d[a];
    double ol = sqrt(ow[0] * ow[0] + ow[1] * ow[1] + ow[2] * ow[2]);
    double outward[3] = {ow[0] / ol, ow[1] / ol, ow[2] / ol};
    double nrm[3];
    uint8_t patch = 0;
    uint32_t atom = 0;
    auto unit = [&](double wx, double wy, double wz) {
        double len = sqrt(wx * wx + wy * wy + wz * wz);
        if (len > 1e-9) {
            nrm[0] = wx / len;
            nrm[1] = wy / len;
            nrm[2] = wz / len;
        } else {
            nrm[0] = outward[0];
            nrm[1] = outward[1];
            nrm[2] = outward[2];
        }
    };
    auto nearest = [&](const uint32_t* atoms, int cnt) {
        double bd = INFINITY;
        uint32_t bk = atoms[0];
        for (int q = 0; q < cnt; q++) {
            uint32_t k = atoms[q];
            double dd = ddist(v[0], v[1], v[2], A.cx[k], A.cy[k], A.cz[k]) - A.r[k];
            if (dd < bd) {
                bd = dd;
                bk = k;
            }
        }
        return A.orig[bk];
    };
    if (last.kind == F_SPHERE) {
        unit(last.qx - v[0], last.qy - v[1], last.qz - v[2]);
        patch = 0;
        atom = A.orig[last.id];
    } else if (last.kind == F_CIRCLE) {
        unit(last.qx - v[0], last.qy - v[1], last.qz - v[2]);
        patch = 1;
        uint32_t at[2] = {A.ci[last.id], A.cj[last.id]};
        atom = nearest(at, 2);
    } else if (last.kind == F_VERTEX) {
        unit(last.qx - v[0], last.qy - v[1], last.qz - v[2]);
        patch = 2;
        uint32_t at[3] = {A.vatoms[3 * last.id], A.vatoms[3 * last.id + 1], A.vatoms[3 * last.id + 2]};
        atom = nearest(at, 3);
    } else {
        // sphere_eval(v, probe + band + 8) over the whole grid (point.rs:742-756),
        // first minimum in scan order: lane 0 alone (rare).
        nrm[0] = outward[0];
        nrm[1] = outward[1];
        nrm[2] = outward[2];
        patch = 0;
        if (lane_id() == 0) {
            double cutoff = A.probe + band + 8.0;
            double bv = cutoff;
            int64_t bi = -1;
            for_near(A.gs, v[0], v[1], v[2], cutoff + A.max_expanded, 0, 1, [&](uint32_t k) {
                if (A.status[k] == S_CONTAINED) return;
                double vv = ddist(v[0], v[1], v[2], A.cx[k], A.cy[k], A.cz[k]) - A.R[k];
                if (vv < bv) {
                    bv = vv;
                    bi = k;
                }
            });
            atom = bi < 0 ? 0u : A.orig[bi];
        }
    }
    if (lane_id() == 0) {
        for (int a = 0; a < 3; a++) {
            o.pos[3 * slot + a] = (float)v[a];
            o.nrm[3 * slot + a] = (float)nrm[a];
        }
        o.atom[slot] = atom;
        o.patch[slot] = patch;
        if (o.t) o.t[slot] = t;
        if (o.evals) o.evals[slot] = evals | (converged ? 0x80000000u : 0u);
    }
    return evals;
}

// One workgroup (4 wavefronts) per sub-block with crossing edges (`work[w]`
// = 8*brick + sb): gather_local into LDS (or scratch slice w), compact the
// crossing edges whose lower sample lies in the sub-block, then refine them
// one per wavefront. Sub-blocks whose lists overflow go to `overflow`.
__global__ void __launch_bounds__(256) k_refine(DArr A, DRun P, RCaps cap, RLayout lay, const int32_t* keys,
                                                const float* values, const uint32_t* work, uint32_t n_work,
                                                const unsigned long long* mask, const uint32_t* wpre,
                                                uint32_t words, const uint32_t* vbase, VOut out, char* scratch,
                                                size_t slice, uint32_t* overflow, uint32_t* n_overflow,
                                                unsigned long long* evals_total) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[R_N];
    uint32_t w = blockIdx.x;
    if (w >= n_work) return;
    uint32_t j = work[w], brick = j >> 3, sb = j & 7;
    char* base = scratch ? scratch + (size_t)w * slice : smem;
    RLists s = rcarve(base, lay);
    const int32_t* key = keys + 3 * brick;
    const uint32_t b = P.b, sub = P.sub;
    double h = P.h, subd = (double)sub;
    // Ball of the sub-block (mesh/extract.rs:164-205).
    double rho = 0.5 * sqrt(3.0) * subd * h * (1.0 + 1e-9) + 1e-9;
    uint32_t o3[3] = {sb & 1, (sb >> 1) & 1, sb >> 2};
    int64_t org[3];
    double c[3];
    for (int a = 0; a < 3; a++) {
        org[a] = (int64_t)key[a] * (int64_t)b;
        double basep = (double)org[a] * h;
        c[a] = basep + ((double)(o3[a] * sub) + 0.5 * subd) * h;
    }
    bool ok = gather_local_block(A, cap, s, n, c[0], c[1], c[2], rho, P.band);
    // Crossing edges whose lower sample lies in this sub-block.
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
    const float* val = values + (size_t)brick * P.spb;
    const uint32_t wv = threadIdx.x / WAVE, nwv = blockDim.x / WAVE;
    unsigned long long my_evals = 0;
    uint32_t vb = vbase[brick];
    double* dm = s.dm + (size_t)wv * (cap.s + cap.c + cap.v);
    for (uint32_t q = wv; q < n[R_E]; q += nwv) {
        uint32_t id = s.ebuf[q];
        uint32_t axis = id % 3, cell = id / 3;
        uint32_t x = cell % b, y = (cell / b) % b, z = cell / (b * b);
        uint32_t i0 = x + P.side * (y + P.side * z);
        uint32_t st = axis == 0 ? 1u : (axis == 1 ? P.side : P.side * P.side);
        int64_t ia[3] = {org[0] + x, org[1] + y, org[2] + z};
        size_t slot = (size_t)vb + edge_rank(mask, wpre, words, brick, id);
        my_evals += refine_vertex_wave(A, P, s, n, cap.c, dm, ia, axis, val[i0], val[i0 + st], out, slot);
    }
    if (evals_total && lane_id() == 0 && my_evals) atomicAdd(evals_total, my_evals);
}

}  // namespace smg
