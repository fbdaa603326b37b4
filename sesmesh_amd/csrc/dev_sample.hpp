// Candidate gathering and field sampling of one sub-block per workgroup.
//
// Restates Engine::assign / prepare_sub / gather_near / gather_ses and the
// SES lane kernel (reference field/kernel.rs:489-751, 837-952). Membership
// decisions are made in f64 exactly as the CPU makes them; the candidate
// data stored for sampling are the CPU's frame-local f32 copies. The class S
// and C lists are sorted nearest-first, as the CPU sorts them, so most
// candidates fail the distance test before their projection and blocker
// tests; the sampled value does not depend on the order: it is the minimum
// over the valid candidates, and "inside" is an any(). Lists live in LDS
// as float4 records (or, for sub-blocks whose lists overflow the LDS
// capacities, in a global scratch slice).
#pragma once
#include "dev_common.hpp"

namespace smg {

// Capacities of the per-sub-block lists (s and c powers of two).
struct Caps {
    uint32_t near, inside, s, c, v, sblk, cblk, hash;  // hash: power of two
};

// Byte layout of the lists for given capacities (host and device agree).
struct Layout {
    uint32_t nd, nid, id, hk, hv, skey, sid, sd, sb, ckey, cid, cd, cb, vd, sblk, cblk, bytes;
};

__host__ __device__ inline Layout make_layout(const Caps& c) {
    Layout L;
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    L.nd = take(16 * c.near);  // x, y, z, r2tol
    L.nid = take(4 * c.near);
    L.id = take(16 * c.inside);  // x, y, z, r2
    L.hk = take(4 * c.hash);
    L.hv = take(4 * c.hash);
    L.skey = take(8 * c.s);
    L.sid = take(4 * c.s);
    L.sd = take(16 * c.s);  // x, y, z, r
    L.sb = take(8 * c.s);   // blocker range
    L.ckey = take(8 * c.c);
    L.cid = take(4 * c.c);
    L.cd = take(48 * c.c);  // (mx, my, mz, rho), (ux, uy, uz, -), (ex, ey, ez, -)
    L.cb = take(8 * c.c);
    L.vd = take(16 * c.v);
    L.sblk = take(2 * c.sblk);
    L.cblk = take(2 * c.cblk);
    L.bytes = o;
    return L;
}

// Open-addressing map global sphere id -> near-list slot.
__device__ inline void hash_put(uint32_t* hk, uint32_t* hv, uint32_t mask, uint32_t key, uint32_t val) {
    uint32_t h = (key * 2654435761u) & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
        uint32_t prev = atomicCAS(&hk[h], 0u, key + 1u);
        if (prev == 0u) {
            hv[h] = val;
            return;
        }
        h = (h + 1u) & mask;
    }
}
__device__ inline uint32_t hash_get(const uint32_t* hk, const uint32_t* hv, uint32_t mask, uint32_t key) {
    uint32_t h = (key * 2654435761u) & mask;
    for (uint32_t probe = 0; probe <= mask; probe++) {
        uint32_t k = hk[h];
        if (k == key + 1u) return hv[h];
        if (k == 0u) return 0xffffffffu;
        h = (h + 1u) & mask;
    }
    return 0xffffffffu;
}

// Engine::assign's owner of sample (lx, ly, lz) (kernel.rs:495-532): the
// first non-uniform sub-block containing it, or -1 with `positive` the sign
// of the last uniform one visited. `st` packs the 8 states, 2 bits each.
__device__ inline int sample_owner(uint32_t st, uint32_t sub, uint32_t lx, uint32_t ly, uint32_t lz,
                                   bool& positive) {
    uint32_t v[3] = {lx, ly, lz}, lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        hi[a] = min(v[a] / sub, 1u);
        lo[a] = (v[a] > 0 && v[a] % sub == 0) ? min(v[a] / sub - 1u, 1u) : hi[a];
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

// Samples of brick `brick` owned by no sub-block get +-band (kernel.rs:533-534).
__global__ void k_assign_uniform(DRun P, const uint8_t* st, uint32_t n_bricks, float* values) {
    uint32_t brick = blockIdx.x;
    if (brick >= n_bricks) return;
    uint32_t s = pack_states(st + 8 * brick);
    float* out = values + (size_t)brick * P.spb;
    for (uint32_t i = threadIdx.x; i < P.spb; i += blockDim.x) {
        uint32_t x = i % P.side, y = (i / P.side) % P.side, z = i / (P.side * P.side);
        bool positive;
        if (sample_owner(s, P.sub, x, y, z, positive) < 0) out[i] = positive ? P.fband : -P.fband;
    }
}

struct Lists {
    float4* nd;
    uint32_t* nid;
    float4* id;
    uint32_t *hk, *hv;
    unsigned long long* skey;
    uint32_t* sid;
    float4* sd;
    uint2* sb;
    unsigned long long* ckey;
    uint32_t* cid;
    float4* cd;
    uint2* cb;
    float4* vd;
    uint16_t *sblk, *cblk;
};

__device__ inline Lists carve(char* base, const Layout& L) {
    Lists s;
    s.nd = (float4*)(base + L.nd);
    s.nid = (uint32_t*)(base + L.nid);
    s.id = (float4*)(base + L.id);
    s.hk = (uint32_t*)(base + L.hk);
    s.hv = (uint32_t*)(base + L.hv);
    s.skey = (unsigned long long*)(base + L.skey);
    s.sid = (uint32_t*)(base + L.sid);
    s.sd = (float4*)(base + L.sd);
    s.sb = (uint2*)(base + L.sb);
    s.ckey = (unsigned long long*)(base + L.ckey);
    s.cid = (uint32_t*)(base + L.cid);
    s.cd = (float4*)(base + L.cd);
    s.cb = (uint2*)(base + L.cb);
    s.vd = (float4*)(base + L.vd);
    s.sblk = (uint16_t*)(base + L.sblk);
    s.cblk = (uint16_t*)(base + L.cblk);
    return s;
}

// Counters of the list builder.
enum { C_NEAR = 0, C_IN, C_S, C_C, C_V, C_SBLK, C_CBLK, C_OVER, C_N };

// Sorts (key, id) ascending by key then id (count <= a power-of-two
// capacity), by the whole workgroup (bitonic network).
__device__ inline void sort_pairs(unsigned long long* key, uint32_t* id, uint32_t count) {
    uint32_t n = 1;
    while (n < count) n <<= 1;
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

// One lane of the SES kernel (kernel.rs:837-952) at frame-local (x, y, z).
__device__ inline float ses_value(const Lists& s, const uint32_t* n, float x, float y, float z, float probe,
                                  float cutoff2) {
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
    for (uint32_t i = 0; i < n[C_S]; i++) {
        float4 c = s.sd[i];
        float dx = x - c.x, dy = y - c.y, dz = z - c.z;
        float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        float lb = fabsf(dist - c.w);
        float lb2 = lb * lb;
        if (!(lb2 < best2 && 1e-6f < dist)) continue;
        float sc = c.w / dist;
        float qx = c.x + dx * sc, qy = c.y + dy * sc, qz = c.z + dz * sc;
        bool valid = true;
        uint2 r = s.sb[i];
        for (uint32_t t = r.x; t < r.y; t++) {
            float4 b = s.nd[s.sblk[t]];
            float ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
            if (!(ex * ex + ey * ey + ez * ez >= b.w)) {
                valid = false;
                break;
            }
        }
        if (valid) best2 = lb2;
    }
    for (uint32_t c = 0; c < n[C_C]; c++) {
        float4 m = s.cd[3 * c], u = s.cd[3 * c + 1];
        float rho = m.w;
        float yx = x - m.x, yy = y - m.y, yz = z - m.z;
        float along = yx * u.x + yy * u.y + yz * u.z;
        float wx = yx - u.x * along, wy = yy - u.y * along, wz = yz - u.z * along;
        float wl = sqrtf(wx * wx + wy * wy + wz * wz);
        float t = wl - rho;
        float d2 = along * along + t * t;
        if (!(d2 < best2)) continue;
        float qx, qy, qz;
        if (wl < 1e-6f) {
            float4 e = s.cd[3 * c + 2];
            qx = m.x + e.x * rho;
            qy = m.y + e.y * rho;
            qz = m.z + e.z * rho;
        } else {
            float sc = rho / wl;
            qx = m.x + wx * sc;
            qy = m.y + wy * sc;
            qz = m.z + wz * sc;
        }
        bool valid = true;
        uint2 r = s.cb[c];
        for (uint32_t q = r.x; q < r.y; q++) {
            float4 b = s.nd[s.cblk[q]];
            float ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
            if (!(ex * ex + ey * ey + ez * ez >= b.w)) {
                valid = false;
                break;
            }
        }
        if (valid) best2 = d2;
    }
    return probe - sqrtf(best2);
}

// Gathers the lists of sub-block `sb` of brick `key` (kernel.rs:553-751).
// Returns false (in every thread) if a capacity was exceeded.
__device__ inline bool gather_sub(const DArr& A, const DLoc& Lc, const DRun& P, const Caps& cap, Lists& s,
                                  uint32_t* n, const int32_t* key, uint32_t sb) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t i = tid; i < cap.hash; i += nt) s.hk[i] = 0u;
    if (tid < C_N) n[tid] = 0u;
    __syncthreads();
    double h = P.h, sub = (double)P.sub, half = sub * 0.5;
    double rho = 0.5 * sqrt(3.0) * sub * h;
    double reach = P.probe + P.band + rho;
    double b[3];
    for (int a = 0; a < 3; a++) {
        double base = (double)((int64_t)key[a] * (int64_t)P.b) * h;
        double off = (double)((sb >> a) & 1) * sub + half;
        b[a] = base + off * h;
    }
    const uint32_t hmask = cap.hash - 1;
    // Near spheres, the inside group and class S (gather_near, gather_ses).
    for_near(A.gs, b[0], b[1], b[2], reach + A.max_expanded, tid, nt, [&](uint32_t k) {
        if (A.status[k] == S_CONTAINED) return;
        double R = A.R[k];
        double d = ddist(b[0], b[1], b[2], A.cx[k], A.cy[k], A.cz[k]);
        if (d >= R + reach) return;
        uint32_t at = atomicAdd(&n[C_NEAR], 1u);
        if (at >= cap.near) {
            n[C_OVER] = 1u;
            return;
        }
        s.nd[at] = make_float4(Lc.px[k], Lc.py[k], Lc.pz[k], Lc.pr2tol[k]);
        s.nid[at] = k;
        hash_put(s.hk, s.hv, hmask, k, at);
        if (d < R + rho) {
            uint32_t q = atomicAdd(&n[C_IN], 1u);
            if (q >= cap.inside)
                n[C_OVER] = 1u;
            else
                s.id[q] = make_float4(Lc.px[k], Lc.py[k], Lc.pz[k], Lc.pr2[k]);
        }
        double lb = fabs(d - R);
        if (A.status[k] == S_EXPOSED && lb <= reach) {
            uint32_t q = atomicAdd(&n[C_S], 1u);
            if (q >= cap.s) {
                n[C_OVER] = 1u;
            } else {
                s.skey[q] = total_key(lb);
                s.sid[q] = k;
            }
        }
    });
    // Class C (gather_ses, kernel.rs:701-736).
    for_near(A.gc, b[0], b[1], b[2], reach + A.rho_max, tid, nt, [&](uint32_t c) {
        double yx = b[0] - A.mx[c], yy = b[1] - A.my[c], yz = b[2] - A.mz[c];
        double ax = A.ax[c], ay = A.ay[c], az = A.az[c];
        double along = yx * ax + yy * ay + yz * az;
        double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
        double wl = sqrt(wx * wx + wy * wy + wz * wz);
        double rr = A.crho[c];
        double dist = sqrt(along * along + (wl - rr) * (wl - rr));
        if (!(dist <= reach)) return;
        uint32_t q = atomicAdd(&n[C_C], 1u);
        if (q >= cap.c) {
            n[C_OVER] = 1u;
            return;
        }
        s.ckey[q] = total_key(dist);
        s.cid[q] = c;
    });
    // Class V (kernel.rs:738-747).
    for_near(A.gv, b[0], b[1], b[2], reach, tid, nt, [&](uint32_t v) {
        if (!(ddist(b[0], b[1], b[2], A.vx[v], A.vy[v], A.vz[v]) <= reach)) return;
        uint32_t q = atomicAdd(&n[C_V], 1u);
        if (q >= cap.v) {
            n[C_OVER] = 1u;
            return;
        }
        s.vd[q] = make_float4(Lc.vx[v], Lc.vy[v], Lc.vz[v], 0.0f);
    });
    __syncthreads();
    if (n[C_OVER]) return false;
    sort_pairs(s.skey, s.sid, n[C_S]);
    sort_pairs(s.ckey, s.cid, n[C_C]);
    // Candidate records in sorted order, and blocker lists: neighbour /
    // circle-blocker rows restricted to near spheres, as near-list slots
    // (kernel.rs:685-735).
    uint32_t ns = n[C_S], nc = n[C_C];
    for (uint32_t i = tid; i < ns + nc; i += nt) {
        bool is_s = i < ns;
        uint32_t id = is_s ? s.sid[i] : s.cid[i - ns];
        if (is_s) {
            s.sd[i] = make_float4(Lc.px[id], Lc.py[id], Lc.pz[id], Lc.pr[id]);
        } else {
            uint32_t q = i - ns;
            s.cd[3 * q] = make_float4(Lc.mx[id], Lc.my[id], Lc.mz[id], Lc.rho[id]);
            s.cd[3 * q + 1] = make_float4(Lc.ux[id], Lc.uy[id], Lc.uz[id], 0.0f);
            s.cd[3 * q + 2] = make_float4(Lc.ex[id], Lc.ey[id], Lc.ez[id], 0.0f);
        }
        const uint32_t* row = is_s ? A.nb_items + A.nb_off[id] : A.cb_items + A.cb_off[id];
        uint32_t len = is_s ? A.nb_off[id + 1] - A.nb_off[id] : A.cb_off[id + 1] - A.cb_off[id];
        uint32_t cnt = 0;
        for (uint32_t t = 0; t < len; t++) cnt += hash_get(s.hk, s.hv, hmask, row[t]) != 0xffffffffu;
        uint32_t at = atomicAdd(&n[is_s ? C_SBLK : C_CBLK], cnt);
        uint32_t limit = is_s ? cap.sblk : cap.cblk;
        if (at + cnt > limit) {
            n[C_OVER] = 1u;
            continue;
        }
        uint16_t* out = is_s ? s.sblk : s.cblk;
        uint32_t w = at;
        for (uint32_t t = 0; t < len; t++) {
            uint32_t l = hash_get(s.hk, s.hv, hmask, row[t]);
            if (l != 0xffffffffu) out[w++] = (uint16_t)l;
        }
        if (is_s)
            s.sb[i] = make_uint2(at, at + cnt);
        else
            s.cb[i - ns] = make_uint2(at, at + cnt);
    }
    __syncthreads();
    return n[C_OVER] == 0u;
}

// Samples one sub-block per workgroup. `work[w]` = 8 * brick + sb of a
// non-uniform sub-block. With `scratch` == nullptr the lists live in
// dynamic LDS; otherwise in scratch slice w of `slice` bytes. Sub-blocks
// whose lists do not fit are appended to `overflow`.
__global__ void __launch_bounds__(256) k_sample(DArr A, DLoc Lc, DRun P, Caps cap, Layout lay,
                                                const int32_t* keys, const uint8_t* st, const uint32_t* work,
                                                uint32_t n_work, float* values, char* scratch, size_t slice,
                                                uint32_t* overflow, uint32_t* n_overflow,
                                                unsigned long long* n_samples) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[C_N];
    __shared__ uint32_t owned[512];
    __shared__ uint32_t n_owned;
    uint32_t w = blockIdx.x;
    if (w >= n_work) return;
    uint32_t j = work[w], brick = j >> 3, sb = j & 7;
    char* base = scratch ? scratch + (size_t)w * slice : smem;
    Lists s = carve(base, lay);
    const int32_t* key = keys + 3 * brick;
    if (threadIdx.x == 0) n_owned = 0;
    bool ok = gather_sub(A, Lc, P, cap, s, n, key, sb);
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = j;
        return;
    }
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
        uint32_t m = n_owned;
        for (uint32_t q = threadIdx.x; q < m; q += blockDim.x) {
            uint32_t t = owned[q];
            uint32_t lx = bx + t % box, ly = by + (t / box) % box, lz = bz + t / (box * box);
            // Lattice::local (lattice.rs:69-71): integer offset, then one f32 multiply.
            float x = (float)(int32_t)(org[0] + lx) * P.h32;
            float y = (float)(int32_t)(org[1] + ly) * P.h32;
            float z = (float)(int32_t)(org[2] + lz) * P.h32;
            out[lx + side * (ly + side * lz)] = ses_value(s, n, x, y, z, P.fprobe, P.fcutoff2);
        }
        if (threadIdx.x == 0 && n_samples) atomicAdd(n_samples, (unsigned long long)m);
        __syncthreads();
        if (threadIdx.x == 0) n_owned = 0;
        __syncthreads();
    }
}

}  // namespace smg
