// Crossing edges, exact f64 vertex refinement and vertex attributes.
//
// Restates the edge scan of mesh_brick (reference mesh/extract.rs:146-223),
// SesField::gather_local / eval_local (field/point.rs:419-489, 496-565,
// 686-722) and vertex / gradient / attributes (mesh/extract.rs:425-597) in
// f64 with the CPU's operation order, list order (sorted by f64::total_cmp
// then index: first candidate wins exact ties) and iteration rules, so every
// vertex, normal, atom and patch equals the CPU back end's bit for bit.
// One wavefront per sub-block with crossing edges: it gathers the sub-block's
// lists into LDS once, then refines the sub-block's edges one per lane.
#pragma once
#include "dev_common.hpp"

namespace smg {

// Edge ids of a brick: ((z*b + y)*b + x)*3 + axis, ascending; crossing
// edges as a bitmask of 64-bit words per brick with per-word prefix counts.
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

// Per brick (one workgroup of 256): crossing-edge masks, per-word prefix,
// brick total and per-sub-block counts (mesh/extract.rs:176-216).
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

// Refinement lists of one sub-block (LocalCandidates, field/point.rs:75-95).
struct RCaps {
    uint32_t near, inside, s, c, v, sblk, cblk, hash;
};
struct RLayout {
    uint32_t nid, hk, hv, in, skey, sid, soff, ckey, cid, coff, vkey, vid, sblk, cblk, ebuf, bytes;
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
    L.nid = take(4 * c.near);
    L.hk = take(4 * c.hash);
    L.hv = take(4 * c.hash);
    L.in = take(4 * c.inside);
    L.skey = take(8 * c.s);
    L.sid = take(4 * c.s);
    L.soff = take(4 * (c.s + 1));
    L.ckey = take(8 * c.c);
    L.cid = take(4 * c.c);
    L.coff = take(4 * (c.c + 1));
    L.vkey = take(8 * c.v);
    L.vid = take(4 * c.v);
    L.sblk = take(2 * c.sblk);
    L.cblk = take(2 * c.cblk);
    L.ebuf = take(4 * 128);
    L.bytes = o;
    return L;
}

struct RLists {
    uint32_t *nid, *hk, *hv, *in;
    unsigned long long* skey;
    uint32_t *sid, *soff;
    unsigned long long* ckey;
    uint32_t *cid, *coff;
    unsigned long long* vkey;
    uint32_t* vid;
    uint16_t *sblk, *cblk;
    uint32_t* ebuf;
};
__device__ inline RLists rcarve(char* base, const RLayout& L) {
    RLists s;
    s.nid = (uint32_t*)(base + L.nid);
    s.hk = (uint32_t*)(base + L.hk);
    s.hv = (uint32_t*)(base + L.hv);
    s.in = (uint32_t*)(base + L.in);
    s.skey = (unsigned long long*)(base + L.skey);
    s.sid = (uint32_t*)(base + L.sid);
    s.soff = (uint32_t*)(base + L.soff);
    s.ckey = (unsigned long long*)(base + L.ckey);
    s.cid = (uint32_t*)(base + L.cid);
    s.coff = (uint32_t*)(base + L.coff);
    s.vkey = (unsigned long long*)(base + L.vkey);
    s.vid = (uint32_t*)(base + L.vid);
    s.sblk = (uint16_t*)(base + L.sblk);
    s.cblk = (uint16_t*)(base + L.cblk);
    s.ebuf = (uint32_t*)(base + L.ebuf);
    return s;
}

enum { R_NEAR = 0, R_IN, R_S, R_C, R_V, R_OVER, R_N };

// Sorts (key, id) pairs ascending by key then id, count <= cap, by one
// wavefront (bitonic network over the next power of two).
__device__ inline void wave_sort(unsigned long long* key, uint32_t* id, uint32_t count) {
    uint32_t n = pow2_at_least(count);
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = count + lane; i < n; i += WAVE) {
        key[i] = ~0ull;
        id[i] = 0xffffffffu;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= n; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n; i += WAVE) {
                uint32_t l = i ^ j;
                if (l > i) {
                    unsigned long long ki = key[i], kl = key[l];
                    uint32_t ii = id[i], il = id[l];
                    bool gt = ki > kl || (ki == kl && ii > il);
                    bool up = (i & k) == 0;
                    if (gt == up) {
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

// gather_local (field/point.rs:496-565) by one wavefront. False if a
// capacity was exceeded.
__device__ inline bool gather_local_wave(const DArr& A, const RCaps& cap, RLists& s, uint32_t* n, double bx,
                                         double by, double bz, double rho, double band) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < cap.hash; i += WAVE) s.hk[i] = 0u;
    if (lane < R_N) n[lane] = 0u;
    __syncthreads();
    const uint32_t hmask = cap.hash - 1;
    double reach = A.probe + band + rho;
    for_near(A.gs, bx, by, bz, reach + A.max_expanded, lane, WAVE, [&](uint32_t k) {
        if (A.status[k] == S_CONTAINED) return;
        double R = A.R[k];
        double d = ddist(bx, by, bz, A.cx[k], A.cy[k], A.cz[k]);
        if (d < R + reach) {
            uint32_t at = atomicAdd(&n[R_NEAR], 1u);
            if (at >= cap.near) {
                n[R_OVER] = 1u;
            } else {
                s.nid[at] = k;
                hash_put(s.hk, s.hv, hmask, k, at);
            }
        }
        if (d < R + rho) {
            uint32_t at = atomicAdd(&n[R_IN], 1u);
            if (at >= cap.inside)
                n[R_OVER] = 1u;
            else
                s.in[at] = k;
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
    for_near(A.gc, bx, by, bz, reach + A.rho_max, lane, WAVE, [&](uint32_t c) {
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
    for_near(A.gv, bx, by, bz, reach, lane, WAVE, [&](uint32_t v) {
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
    wave_sort(s.skey, s.sid, n[R_S]);
    wave_sort(s.ckey, s.cid, n[R_C]);
    wave_sort(s.vkey, s.vid, n[R_V]);
    // Blockers in sorted candidate order: row members that are near
    // (filter_blockers, point.rs:152-155), as near-list slots.
    for (int cls = 0; cls < 2; cls++) {
        uint32_t cnt = n[cls == 0 ? R_S : R_C];
        uint32_t* ids = cls == 0 ? s.sid : s.cid;
        uint32_t* off = cls == 0 ? s.soff : s.coff;
        const uint32_t* ro = cls == 0 ? A.nb_off : A.cb_off;
        const uint32_t* ri = cls == 0 ? A.nb_items : A.cb_items;
        uint16_t* out = cls == 0 ? s.sblk : s.cblk;
        uint32_t limit = cls == 0 ? cap.sblk : cap.cblk;
        uint32_t base = 0;
        for (uint32_t c0 = 0; c0 < cnt; c0 += WAVE) {
            uint32_t i = c0 + lane, m = 0;
            if (i < cnt) {
                uint32_t id = ids[i];
                for (uint32_t t = ro[id]; t < ro[id + 1]; t++) m += hash_get(s.hk, s.hv, hmask, ri[t]) != 0xffffffffu;
            }
            // Exclusive prefix of m across the wave.
            uint32_t inc = m;
            for (int o = 1; o < WAVE; o <<= 1) {
                uint32_t y = __shfl_up(inc, o, WAVE);
                if ((int)lane >= o) inc += y;
            }
            uint32_t at = base + inc - m;
            uint32_t total = __shfl(inc, WAVE - 1, WAVE);
            if (i < cnt) {
                off[i] = at;
                if (at + m <= limit) {
                    uint32_t id = ids[i], w = at;
                    for (uint32_t t = ro[id]; t < ro[id + 1]; t++) {
                        uint32_t l = hash_get(s.hk, s.hv, hmask, ri[t]);
                        if (l != 0xffffffffu) out[w++] = (uint16_t)l;
                    }
                } else {
                    n[R_OVER] = 1u;
                }
            }
            base += total;
        }
        if (lane == 0) off[cnt] = base;
    }
    __syncthreads();
    return n[R_OVER] == 0u;
}

struct NearEvalD {
    double value;
    uint32_t kind, id;
    double qx, qy, qz;
};

// SesField::eval_local (field/point.rs:686-722) with try_sphere_with /
// try_circle_with / try_vertex (:419-489).
__device__ inline NearEvalD eval_local_d(const DArr& A, const RLists& s, const uint32_t* n, double x, double y,
                                         double z, double band) {
    NearEvalD e;
    bool inside = false;
    for (uint32_t q = 0; q < n[R_IN]; q++) {
        uint32_t k = s.in[q];
        if (ddist(x, y, z, A.cx[k], A.cy[k], A.cz[k]) < A.R[k]) {
            inside = true;
            break;
        }
    }
    e.qx = x;
    e.qy = y;
    e.qz = z;
    if (!inside) {
        e.value = A.probe;
        e.kind = F_SOLVENT;
        e.id = 0;
        return e;
    }
    double best = A.probe + band;
    e.kind = F_BEYOND;
    e.id = 0;
    for (uint32_t q = 0; q < n[R_S]; q++) {
        uint32_t i = s.sid[q];
        double cx = A.cx[i], cy = A.cy[i], cz = A.cz[i], Ri = A.R[i];
        double vx = x - cx, vy = y - cy, vz = z - cz;
        double len = sqrt(vx * vx + vy * vy + vz * vz);
        double lb = fabs(len - Ri);
        if (lb >= best || len < 1e-12) continue;
        double sc = Ri / len;
        double qx = cx + vx * sc, qy = cy + vy * sc, qz = cz + vz * sc;
        bool free = true;
        for (uint32_t t = s.soff[q], te = s.soff[q + 1]; t < te; t++) {
            uint32_t k = s.nid[s.sblk[t]];
            if (!(ddist(qx, qy, qz, A.cx[k], A.cy[k], A.cz[k]) >= A.R[k] - EPS)) {
                free = false;
                break;
            }
        }
        if (free) {
            best = lb;
            e.kind = F_SPHERE;
            e.id = i;
            e.qx = qx;
            e.qy = qy;
            e.qz = qz;
        }
    }
    for (uint32_t q = 0; q < n[R_C]; q++) {
        uint32_t c = s.cid[q];
        double mx = A.mx[c], my = A.my[c], mz = A.mz[c];
        double ax = A.ax[c], ay = A.ay[c], az = A.az[c], rho = A.crho[c];
        double yx = x - mx, yy = y - my, yz = z - mz;
        double along = yx * ax + yy * ay + yz * az;
        double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
        double wl = sqrt(wx * wx + wy * wy + wz * wz);
        double dist = sqrt(along * along + (wl - rho) * (wl - rho));
        if (dist >= best) continue;
        double dx, dy, dz;
        if (wl > 1e-12) {
            dx = wx / wl;
            dy = wy / wl;
            dz = wz / wl;
        } else {
            dx = A.e1x[c];
            dy = A.e1y[c];
            dz = A.e1z[c];
        }
        double qx = mx + dx * rho, qy = my + dy * rho, qz = mz + dz * rho;
        bool free = true;
        for (uint32_t t = s.coff[q], te = s.coff[q + 1]; t < te; t++) {
            uint32_t k = s.nid[s.cblk[t]];
            if (!(ddist(qx, qy, qz, A.cx[k], A.cy[k], A.cz[k]) >= A.R[k] - EPS)) {
                free = false;
                break;
            }
        }
        if (free) {
            best = dist;
            e.kind = F_CIRCLE;
            e.id = c;
            e.qx = qx;
            e.qy = qy;
            e.qz = qz;
        }
    }
    for (uint32_t q = 0; q < n[R_V]; q++) {
        uint32_t v = s.vid[q];
        double px = A.vx[v], py = A.vy[v], pz = A.vz[v];
        double d = ddist(x, y, z, px, py, pz);
        if (d < best) {
            best = d;
            e.kind = F_VERTEX;
            e.id = v;
            e.qx = px;
            e.qy = py;
            e.qz = pz;
        }
    }
    e.value = A.probe - best;
    return e;
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
// from lattice point ia along `axis`, sampled values fa32, fb32.
__device__ inline uint32_t refine_vertex(const DArr& A, const DRun& P, const RLists& s, const uint32_t* n,
                                     const int64_t ia[3], uint32_t axis, float fa32, float fb32, const VOut& o,
                                     size_t slot) {
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
    NearEvalD last;
    double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int it = 0; it < 8; it++) {
        double x = pa[0] + d[0] * t, y = pa[1] + d[1] * t, z = pa[2] + d[2] * t;
        NearEvalD e = eval_local_d(A, s, n, x, y, z, band);
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
            NearEvalD e = eval_local_d(A, s, n, pa[0] + d[0] * t, pa[1] + d[1] * t, pa[2] + d[2] * t, band);
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
    // Fallback normal along the edge, inside end to outside end.
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
        // sphere_eval(v, probe + band + 8) over the whole grid (point.rs:742-756).
        nrm[0] = outward[0];
        nrm[1] = outward[1];
        nrm[2] = outward[2];
        patch = 0;
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
    for (int a = 0; a < 3; a++) {
        o.pos[3 * slot + a] = (float)v[a];
        o.nrm[3 * slot + a] = (float)nrm[a];
    }
    o.atom[slot] = atom;
    o.patch[slot] = patch;
    if (o.t) o.t[slot] = t;
    if (o.evals) o.evals[slot] = evals | (converged ? 0x80000000u : 0u);
    return evals;
}

// One wavefront per sub-block with crossing edges (`work[w]` = 8*brick+sb):
// gather_local in LDS (or in scratch slice w), then one edge per lane.
// Sub-blocks whose lists overflow are appended to `overflow`.
__global__ void __launch_bounds__(64) k_refine(DArr A, DRun P, RCaps cap, RLayout lay, const int32_t* keys,
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
    const uint32_t lane = threadIdx.x, b = P.b, sub = P.sub;
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
    if (!gather_local_wave(A, cap, s, n, c[0], c[1], c[2], rho, P.band)) {
        if (lane == 0) overflow[atomicAdd(n_overflow, 1u)] = j;
        return;
    }
    const float* val = values + (size_t)brick * P.spb;
    // Edges whose lower sample lies in this sub-block: x in [lo, hi) per axis.
    uint32_t lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = o3[a] * sub;
        hi[a] = o3[a] ? b : sub;
    }
    uint32_t ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    uint32_t ncand = ex * ey * ez * 3;
    unsigned long long my_evals = 0;
    uint32_t vb = vbase[brick];
    // Crossing edges are compacted into a buffer by ballot and refined 64 at
    // a time, one per lane.
    auto refine_id = [&](uint32_t id) {
        uint32_t axis = id % 3, cell = id / 3;
        uint32_t x = cell % b, y = (cell / b) % b, z = cell / (b * b);
        uint32_t i0 = x + P.side * (y + P.side * z);
        uint32_t st = axis == 0 ? 1u : (axis == 1 ? P.side : P.side * P.side);
        int64_t ia[3] = {org[0] + x, org[1] + y, org[2] + z};
        size_t slot = (size_t)vb + edge_rank(mask, wpre, words, brick, id);
        my_evals += refine_vertex(A, P, s, n, ia, axis, val[i0], val[i0 + st], out, slot);
    };
    uint32_t fill = 0;
    for (uint32_t c0 = 0; c0 < ncand; c0 += WAVE) {
        uint32_t q = c0 + lane;
        bool cross = false;
        uint32_t id = 0;
        if (q < ncand) {
            uint32_t axis = q % 3, cell = q / 3;
            uint32_t x = lo[0] + cell % ex, y = lo[1] + (cell / ex) % ey, z = lo[2] + cell / (ex * ey);
            id = ((z * b + y) * b + x) * 3 + axis;
            cross = edge_is_crossing(mask, words, brick, id);
        }
        unsigned long long m = __ballot(cross);
        uint32_t pos = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (cross) s.ebuf[fill + pos] = id;
        fill += (uint32_t)__popcll(m);
        __syncthreads();
        if (fill >= WAVE) {
            refine_id(s.ebuf[lane]);
            uint32_t rest = fill - WAVE;
            uint32_t moved = lane < rest ? s.ebuf[WAVE + lane] : 0u;
            __syncthreads();
            if (lane < rest) s.ebuf[lane] = moved;
            __syncthreads();
            fill = rest;
        }
    }
    if (lane < fill) refine_id(s.ebuf[lane]);
    if (evals_total) {
        for (int o = 32; o > 0; o >>= 1) my_evals += __shfl_xor(my_evals, o, WAVE);
        if (lane == 0) atomicAdd(evals_total, my_evals);
    }
}

}  // namespace smg
