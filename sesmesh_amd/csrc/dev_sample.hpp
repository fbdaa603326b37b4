// Candidate gathering and field sampling of one sub-block per workgroup.
//
// Restates Engine::assign / prepare_sub / gather_near / gather_ses and the
// SES lane kernel (reference field/kernel.rs:489-751, 837-952). Membership
// decisions are made in f64 exactly as the CPU makes them; the candidate
// data stored for sampling are the CPU's frame-local f32 copies. Lists live
// in LDS (or, for sub-blocks whose lists overflow the LDS capacities, in a
// global scratch slice); their order does not matter for the sampled value:
// it is the minimum over the valid candidates, and "inside" is an any().
#pragma once
#include "dev_common.hpp"

namespace smg {

// Capacities of the per-sub-block lists.
struct Caps {
    uint32_t near, inside, s, c, v, sblk, cblk, hash;  // hash: power of two
};

// Byte layout of the lists for given capacities (host and device agree).
struct Layout {
    uint32_t nx, ny, nz, nt, nid, ix, iy, iz, i2, hk, hv, sx, sy, sz, sr, sid, sb0, sb1, cf, cid, cb0, cb1, vx, vy,
        vz, sblk, cblk, bytes;
};

__host__ __device__ inline Layout make_layout(const Caps& c) {
    Layout L;
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    L.nx = take(4 * c.near);
    L.ny = take(4 * c.near);
    L.nz = take(4 * c.near);
    L.nt = take(4 * c.near);
    L.nid = take(4 * c.near);
    L.ix = take(4 * c.inside);
    L.iy = take(4 * c.inside);
    L.iz = take(4 * c.inside);
    L.i2 = take(4 * c.inside);
    L.hk = take(4 * c.hash);
    L.hv = take(4 * c.hash);
    L.sx = take(4 * c.s);
    L.sy = take(4 * c.s);
    L.sz = take(4 * c.s);
    L.sr = take(4 * c.s);
    L.sid = take(4 * c.s);
    L.sb0 = take(4 * c.s);
    L.sb1 = take(4 * c.s);
    L.cf = take(40 * c.c);  // 10 f32 per circle, SoA of c.c each
    L.cid = take(4 * c.c);
    L.cb0 = take(4 * c.c);
    L.cb1 = take(4 * c.c);
    L.vx = take(4 * c.v);
    L.vy = take(4 * c.v);
    L.vz = take(4 * c.v);
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
// of the last uniform one visited.
__device__ inline int sample_owner(const uint8_t st[8], uint32_t sub, uint32_t lx, uint32_t ly, uint32_t lz,
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
                uint32_t s = sx + 2 * sy + 4 * sz;
                if (st[s] == ST_NONE) return (int)s;
                positive = st[s] == ST_OUT;
            }
    return -1;
}

// Samples of brick `brick` owned by no sub-block get +-band (kernel.rs:533-534).
__global__ void k_assign_uniform(DRun P, const uint8_t* st, uint32_t n_bricks, float* values) {
    uint32_t brick = blockIdx.x;
    if (brick >= n_bricks) return;
    uint8_t s[8];
    for (int q = 0; q < 8; q++) s[q] = st[8 * brick + q];
    float* out = values + (size_t)brick * P.spb;
    for (uint32_t i = threadIdx.x; i < P.spb; i += blockDim.x) {
        uint32_t x = i % P.side, y = (i / P.side) % P.side, z = i / (P.side * P.side);
        bool positive;
        if (sample_owner(s, P.sub, x, y, z, positive) < 0) out[i] = positive ? P.fband : -P.fband;
    }
}

struct Lists {
    float *nx, *ny, *nz, *nt;
    uint32_t* nid;
    float *ix, *iy, *iz, *i2;
    uint32_t *hk, *hv;
    float *sx, *sy, *sz, *sr;
    uint32_t *sid, *sb0, *sb1;
    float* cf;
    uint32_t *cid, *cb0, *cb1;
    float *vx, *vy, *vz;
    uint16_t *sblk, *cblk;
};

__device__ inline Lists carve(char* base, const Layout& L) {
    Lists s;
    s.nx = (float*)(base + L.nx);
    s.ny = (float*)(base + L.ny);
    s.nz = (float*)(base + L.nz);
    s.nt = (float*)(base + L.nt);
    s.nid = (uint32_t*)(base + L.nid);
    s.ix = (float*)(base + L.ix);
    s.iy = (float*)(base + L.iy);
    s.iz = (float*)(base + L.iz);
    s.i2 = (float*)(base + L.i2);
    s.hk = (uint32_t*)(base + L.hk);
    s.hv = (uint32_t*)(base + L.hv);
    s.sx = (float*)(base + L.sx);
    s.sy = (float*)(base + L.sy);
    s.sz = (float*)(base + L.sz);
    s.sr = (float*)(base + L.sr);
    s.sid = (uint32_t*)(base + L.sid);
    s.sb0 = (uint32_t*)(base + L.sb0);
    s.sb1 = (uint32_t*)(base + L.sb1);
    s.cf = (float*)(base + L.cf);
    s.cid = (uint32_t*)(base + L.cid);
    s.cb0 = (uint32_t*)(base + L.cb0);
    s.cb1 = (uint32_t*)(base + L.cb1);
    s.vx = (float*)(base + L.vx);
    s.vy = (float*)(base + L.vy);
    s.vz = (float*)(base + L.vz);
    s.sblk = (uint16_t*)(base + L.sblk);
    s.cblk = (uint16_t*)(base + L.cblk);
    return s;
}

// Counters of the list builder.
enum { C_NEAR = 0, C_IN, C_S, C_C, C_V, C_SBLK, C_CBLK, C_OVER, C_N };

// One lane of the SES kernel (kernel.rs:837-952) at frame-local (x, y, z).
__device__ inline float ses_value(const Lists& s, const uint32_t* n, uint32_t ccap, float x, float y, float z,
                                  float probe, float cutoff2) {
    bool inside = false;
    for (uint32_t k = 0; k < n[C_IN]; k++) {
        float dx = x - s.ix[k], dy = y - s.iy[k], dz = z - s.iz[k];
        if (dx * dx + dy * dy + dz * dz < s.i2[k]) {
            inside = true;
            break;
        }
    }
    if (!inside) return probe;
    float best2 = cutoff2;
    for (uint32_t v = 0; v < n[C_V]; v++) {
        float dx = x - s.vx[v], dy = y - s.vy[v], dz = z - s.vz[v];
        float d2 = dx * dx + dy * dy + dz * dz;
        best2 = d2 < best2 ? d2 : best2;
    }
    for (uint32_t i = 0; i < n[C_S]; i++) {
        float cx = s.sx[i], cy = s.sy[i], cz = s.sz[i], r = s.sr[i];
        float dx = x - cx, dy = y - cy, dz = z - cz;
        float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        float lb = fabsf(dist - r);
        float lb2 = lb * lb;
        if (!(lb2 < best2 && 1e-6f < dist)) continue;
        float sc = r / dist;
        float qx = cx + dx * sc, qy = cy + dy * sc, qz = cz + dz * sc;
        bool valid = true;
        for (uint32_t t = s.sb0[i], e = s.sb1[i]; t < e; t++) {
            uint32_t k = s.sblk[t];
            float ex = qx - s.nx[k], ey = qy - s.ny[k], ez = qz - s.nz[k];
            if (!(ex * ex + ey * ey + ez * ez >= s.nt[k])) {
                valid = false;
                break;
            }
        }
        if (valid) best2 = lb2;
    }
    const float *mxa = s.cf, *mya = s.cf + ccap, *mza = s.cf + 2 * ccap, *uxa = s.cf + 3 * ccap,
                *uya = s.cf + 4 * ccap, *uza = s.cf + 5 * ccap, *exa = s.cf + 6 * ccap, *eya = s.cf + 7 * ccap,
                *eza = s.cf + 8 * ccap, *rha = s.cf + 9 * ccap;
    for (uint32_t c = 0; c < n[C_C]; c++) {
        float mx = mxa[c], my = mya[c], mz = mza[c];
        float ux = uxa[c], uy = uya[c], uz = uza[c], rho = rha[c];
        float yx = x - mx, yy = y - my, yz = z - mz;
        float along = yx * ux + yy * uy + yz * uz;
        float wx = yx - ux * along, wy = yy - uy * along, wz = yz - uz * along;
        float wl = sqrtf(wx * wx + wy * wy + wz * wz);
        float t = wl - rho;
        float d2 = along * along + t * t;
        if (!(d2 < best2)) continue;
        float qx, qy, qz;
        if (wl < 1e-6f) {
            qx = mx + exa[c] * rho;
            qy = my + eya[c] * rho;
            qz = mz + eza[c] * rho;
        } else {
            float sc = rho / wl;
            qx = mx + wx * sc;
            qy = my + wy * sc;
            qz = mz + wz * sc;
        }
        bool valid = true;
        for (uint32_t q = s.cb0[c], e = s.cb1[c]; q < e; q++) {
            uint32_t k = s.cblk[q];
            float ex = qx - s.nx[k], ey = qy - s.ny[k], ez = qz - s.nz[k];
            if (!(ex * ex + ey * ey + ez * ez >= s.nt[k])) {
                valid = false;
                break;
            }
        }
        if (valid) best2 = d2;
    }
    return probe - sqrtf(best2);
}

// Gathers the lists of sub-block `sb` of brick `brick` (kernel.rs:553-751).
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
        s.nx[at] = Lc.px[k];
        s.ny[at] = Lc.py[k];
        s.nz[at] = Lc.pz[k];
        s.nt[at] = Lc.pr2tol[k];
        s.nid[at] = k;
        hash_put(s.hk, s.hv, hmask, k, at);
        if (d < R + rho) {
            uint32_t q = atomicAdd(&n[C_IN], 1u);
            if (q >= cap.inside) {
                n[C_OVER] = 1u;
            } else {
                s.ix[q] = Lc.px[k];
                s.iy[q] = Lc.py[k];
                s.iz[q] = Lc.pz[k];
                s.i2[q] = Lc.pr2[k];
            }
        }
        if (A.status[k] == S_EXPOSED && fabs(d - R) <= reach) {
            uint32_t q = atomicAdd(&n[C_S], 1u);
            if (q >= cap.s) {
                n[C_OVER] = 1u;
            } else {
                s.sx[q] = Lc.px[k];
                s.sy[q] = Lc.py[k];
                s.sz[q] = Lc.pz[k];
                s.sr[q] = Lc.pr[k];
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
        float* f = s.cf;
        uint32_t cc = cap.c;
        f[q] = Lc.mx[c];
        f[cc + q] = Lc.my[c];
        f[2 * cc + q] = Lc.mz[c];
        f[3 * cc + q] = Lc.ux[c];
        f[4 * cc + q] = Lc.uy[c];
        f[5 * cc + q] = Lc.uz[c];
        f[6 * cc + q] = Lc.ex[c];
        f[7 * cc + q] = Lc.ey[c];
        f[8 * cc + q] = Lc.ez[c];
        f[9 * cc + q] = Lc.rho[c];
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
        s.vx[q] = Lc.vx[v];
        s.vy[q] = Lc.vy[v];
        s.vz[q] = Lc.vz[v];
    });
    __syncthreads();
    if (n[C_OVER]) return false;
    // Blocker lists: neighbour / circle-blocker rows restricted to near
    // spheres, as near-list slots (kernel.rs:692-698, 729-735).
    uint32_t ns = n[C_S], nc = n[C_C];
    for (uint32_t i = tid; i < ns + nc; i += nt) {
        bool is_s = i < ns;
        uint32_t id = is_s ? s.sid[i] : s.cid[i - ns];
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
        if (is_s) {
            s.sb0[i] = at;
            s.sb1[i] = at + cnt;
        } else {
            s.cb0[i - ns] = at;
            s.cb1[i - ns] = at + cnt;
        }
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
                                                uint32_t* overflow, uint32_t* n_overflow) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[C_N];
    uint32_t w = blockIdx.x;
    if (w >= n_work) return;
    uint32_t j = work[w], brick = j >> 3, sb = j & 7;
    char* base = scratch ? scratch + (size_t)w * slice : smem;
    Lists s = carve(base, lay);
    const int32_t* key = keys + 3 * brick;
    bool ok = gather_sub(A, Lc, P, cap, s, n, key, sb);
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = j;
        return;
    }
    uint8_t stb[8];
    for (int q = 0; q < 8; q++) stb[q] = st[8 * brick + q];
    uint32_t sub = P.sub, box = sub + 1, side = P.side;
    uint32_t bx = (sb & 1) * sub, by = ((sb >> 1) & 1) * sub, bz = ((sb >> 2) & 1) * sub;
    int64_t org[3];
    for (int a = 0; a < 3; a++) org[a] = (int64_t)key[a] * (int64_t)P.b - P.frame[a];
    float* out = values + (size_t)brick * P.spb;
    for (uint32_t t = threadIdx.x; t < box * box * box; t += blockDim.x) {
        uint32_t lx = bx + t % box, ly = by + (t / box) % box, lz = bz + t / (box * box);
        bool positive;
        if (sample_owner(stb, sub, lx, ly, lz, positive) != (int)sb) continue;
        // Lattice::local (lattice.rs:69-71): integer offset, then one f32 multiply.
        float x = (float)(int32_t)(org[0] + lx) * P.h32;
        float y = (float)(int32_t)(org[1] + ly) * P.h32;
        float z = (float)(int32_t)(org[2] + lz) * P.h32;
        out[lx + side * (ly + side * lz)] = ses_value(s, n, cap.c, x, y, z, P.fprobe, P.fcutoff2);
    }
}

}  // namespace smg
