// Narrow-band brick classification on the device, in f64 with the CPU's
// operations: touched bricks (reference field/lattice.rs:105-163), the
// conservative uniform-sign test (field/lattice.rs:190-227, with the exact
// ses_eval of field/point.rs:384-489) for bricks and their 2x2x2 sub-blocks
// (field/kernel.rs:467-482). One wavefront evaluates one ball; its lanes
// split the cell-grid scans and the minima are reduced across the wave (a
// minimum does not depend on the order its terms are visited in).
#pragma once
#include "dev_common.hpp"

namespace smg {

// SesField::try_sphere_with with the full neighbour row (point.rs:419-442):
// updates the lane's running `best`.
__device__ inline void try_sphere_full(const DArr& A, double x, double y, double z, uint32_t i, double& best) {
    double vx = x - A.cx[i], vy = y - A.cy[i], vz = z - A.cz[i];
    double len = sqrt(vx * vx + vy * vy + vz * vz);
    double Ri = A.R[i];
    double lb = fabs(len - Ri);
    if (lb >= best || len < 1e-12) return;
    double s = Ri / len;
    double qx = A.cx[i] + vx * s, qy = A.cy[i] + vy * s, qz = A.cz[i] + vz * s;
    for (uint32_t t = A.nb_off[i], e = A.nb_off[i + 1]; t < e; t++) {
        uint32_t k = A.nb_items[t];
        if (!(ddist(qx, qy, qz, A.cx[k], A.cy[k], A.cz[k]) >= A.R[k] - EPS)) return;
    }
    best = lb;
}

// SesField::try_circle_with with the full blocker row (point.rs:450-475).
__device__ inline void try_circle_full(const DArr& A, double x, double y, double z, uint32_t c, double& best) {
    double yx = x - A.mx[c], yy = y - A.my[c], yz = z - A.mz[c];
    double ax = A.ax[c], ay = A.ay[c], az = A.az[c];
    double along = yx * ax + yy * ay + yz * az;
    double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
    double wl = sqrt(wx * wx + wy * wy + wz * wz);
    double rho = A.crho[c];
    double dist = sqrt(along * along + (wl - rho) * (wl - rho));
    if (dist >= best) return;
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
    double qx = A.mx[c] + dx * rho, qy = A.my[c] + dy * rho, qz = A.mz[c] + dz * rho;
    for (uint32_t t = A.cb_off[c], e = A.cb_off[c + 1]; t < e; t++) {
        uint32_t k = A.cb_items[t];
        if (!(ddist(qx, qy, qz, A.cx[k], A.cy[k], A.cz[k]) >= A.R[k] - EPS)) return;
    }
    best = dist;
}

// Candidate list of one wavefront in LDS (the kernels below run one
// wavefront per workgroup).
constexpr uint32_t CL_CAP = 512;
constexpr uint32_t DIST_STEP = 4;  // candidates ses_distance_wave checks per step
struct CandList {
    double d[CL_CAP];
    uint32_t ref[CL_CAP];  // id | kind << 30 (0 sphere, 1 circle, 2 vertex)
    uint32_t n;
};

// SesField::ses_eval's distance (field/point.rs:384-411) of a point inside
// the SAS, clamped at `cutoff`, by one wavefront: every candidate's distance
// first, then blockers of the nearest unchecked candidate until one is
// free (the minimum over free candidates below the cutoff, as the
// sequential loop finds it). Falls back to each lane scanning its own
// candidates when the list outgrows the LDS.
__device__ __forceinline__ double ses_distance_wave(const DArr& A, double x, double y, double z, double cutoff) {
    __shared__ CandList L;
    const uint32_t lane = lane_id();
    if (lane == 0) L.n = 0;
    __syncthreads();
    for_near(A.gs, x, y, z, cutoff + A.max_expanded, lane, WAVE, [&](uint32_t i) {
        if (A.status[i] != S_EXPOSED) return;
        double vx = x - A.cx[i], vy = y - A.cy[i], vz = z - A.cz[i];
        double len = sqrt(vx * vx + vy * vy + vz * vz);
        double lb = fabs(len - A.R[i]);
        if (lb >= cutoff || len < 1e-12) return;
        uint32_t at = atomicAdd(&L.n, 1u);
        if (at < CL_CAP) {
            L.d[at] = lb;
            L.ref[at] = i;
        }
    });
    for_near(A.gc, x, y, z, cutoff + A.rho_max, lane, WAVE, [&](uint32_t c) {
        double yx = x - A.mx[c], yy = y - A.my[c], yz = z - A.mz[c];
        double ax = A.ax[c], ay = A.ay[c], az = A.az[c];
        double along = yx * ax + yy * ay + yz * az;
        double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
        double wl = sqrt(wx * wx + wy * wy + wz * wz);
        double rho = A.crho[c];
        double dist = sqrt(along * along + (wl - rho) * (wl - rho));
        if (dist >= cutoff) return;
        uint32_t at = atomicAdd(&L.n, 1u);
        if (at < CL_CAP) {
            L.d[at] = dist;
            L.ref[at] = c | (1u << 30);
        }
    });
    for_near(A.gv, x, y, z, cutoff, lane, WAVE, [&](uint32_t v) {
        double d = ddist(x, y, z, A.vx[v], A.vy[v], A.vz[v]);
        if (!(d < cutoff)) return;
        uint32_t at = atomicAdd(&L.n, 1u);
        if (at < CL_CAP) {
            L.d[at] = d;
            L.ref[at] = v | (2u << 30);
        }
    });
    __syncthreads();
    const uint32_t count = L.n;
    if (count > CL_CAP) {
        double best = cutoff;
        for_near(A.gs, x, y, z, cutoff + A.max_expanded, lane, WAVE, [&](uint32_t i) {
            if (A.status[i] == S_EXPOSED) try_sphere_full(A, x, y, z, i, best);
        });
        for_near(A.gc, x, y, z, cutoff + A.rho_max, lane, WAVE,
                 [&](uint32_t c) { try_circle_full(A, x, y, z, c, best); });
        for_near(A.gv, x, y, z, cutoff, lane, WAVE, [&](uint32_t v) {
            double d = ddist(x, y, z, A.vx[v], A.vy[v], A.vz[v]);
            if (d < best) best = d;
        });
        return wave_min(best);
    }
    // The DIST_STEP nearest unchecked candidates per step, in (distance,
    // slot) order, each with WAVE / DIST_STEP lanes scanning its blocker row
    // (rows are mostly short, so one candidate at a time left most lanes
    // idle); the first free one in that order is the answer.
    constexpr uint32_t NC = DIST_STEP, LPC = WAVE / DIST_STEP;
    for (;;) {
        double bd[NC];
        uint32_t bo[NC];
        for (uint32_t c = 0; c < NC; c++) {
            double d = INFINITY;
            uint32_t oi = 0xffffffffu;
            for (uint32_t o = lane; o < count; o += WAVE) {
                bool taken = false;
                for (uint32_t e = 0; e < c; e++) taken |= o == bo[e];
                if (!taken && L.d[o] < d) {
                    d = L.d[o];
                    oi = o;
                }
            }
            for (int m = 32; m > 0; m >>= 1) {
                double od = __shfl_xor(d, m, WAVE);
                uint32_t oo = __shfl_xor(oi, m, WAVE);
                if (od < d || (od == d && oo < oi)) {
                    d = od;
                    oi = oo;
                }
            }
            bd[c] = d;
            bo[c] = oi;
        }
        if (bo[0] == 0xffffffffu) return cutoff;
        __syncthreads();
        for (uint32_t c = 0; c < NC; c++)
            if (bo[c] != 0xffffffffu && lane == bo[c] % WAVE) L.d[bo[c]] = INFINITY;
        __syncthreads();
        const uint32_t o = bo[lane / LPC];
        bool blocked = false;
        if (o != 0xffffffffu) {
            const uint32_t ref = L.ref[o], id = ref & 0x3fffffffu, kind = ref >> 30;
            if (kind != 2) {  // vertices are never blocked
                double qx, qy, qz;
                gptr<const uint32_t> row, end;
                if (kind == 0) {
                    double vx = x - A.cx[id], vy = y - A.cy[id], vz = z - A.cz[id];
                    double len = sqrt(vx * vx + vy * vy + vz * vz);
                    double s = A.R[id] / len;
                    qx = A.cx[id] + vx * s;
                    qy = A.cy[id] + vy * s;
                    qz = A.cz[id] + vz * s;
                    row = A.nb_items + A.nb_off[id];
                    end = A.nb_items + A.nb_off[id + 1];
                } else {
                    double yx = x - A.mx[id], yy = y - A.my[id], yz = z - A.mz[id];
                    double ax = A.ax[id], ay = A.ay[id], az = A.az[id];
                    double along = yx * ax + yy * ay + yz * az;
                    double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
                    double wl = sqrt(wx * wx + wy * wy + wz * wz);
                    double dx, dy, dz;
                    if (wl > 1e-12) {
                        dx = wx / wl;
                        dy = wy / wl;
                        dz = wz / wl;
                    } else {
                        dx = A.e1x[id];
                        dy = A.e1y[id];
                        dz = A.e1z[id];
                    }
                    double rho = A.crho[id];
                    qx = A.mx[id] + dx * rho;
                    qy = A.my[id] + dy * rho;
                    qz = A.mz[id] + dz * rho;
                    row = A.cb_items + A.cb_off[id];
                    end = A.cb_items + A.cb_off[id + 1];
                }
                for (gptr<const uint32_t> p = row + lane % LPC; p < end; p += LPC) {
                    uint32_t k = *p;
                    blocked |= !(ddist(qx, qy, qz, A.cx[k], A.cy[k], A.cz[k]) >= A.R[k] - EPS);
                }
            }
        }
        const unsigned long long bm = __ballot(blocked);
        for (uint32_t c = 0; c < NC; c++) {
            if (bo[c] == 0xffffffffu) return cutoff;
            if (((bm >> (c * LPC)) & (LPC == 64 ? ~0ull : ((1ull << LPC) - 1ull))) == 0ull) return bd[c];
        }
    }
}

// lattice::uniform_sign, one wavefront (the workgroup). Returns
// ST_NONE/OUT/IN in every lane. The SAS and vdW kinds decide by the sign
// of the least SAS distance alone (expanded radii are the vdW radii for
// the vdW kind).
__device__ __forceinline__ uint8_t uniform_sign_wave(const DArr& A, double x, double y, double z, double rho) {
    const uint32_t lane = lane_id();
    double margin = rho + CLASSIFY_EPS;
    double s_sas = INFINITY, s_vdw = INFINITY;
    for_near(A.gs, x, y, z, margin + A.max_expanded, lane, WAVE, [&](uint32_t k) {
        if (A.status[k] != S_CONTAINED) {
            double d = ddist(x, y, z, A.cx[k], A.cy[k], A.cz[k]);
            s_sas = fmin(s_sas, d - A.R[k]);
            s_vdw = fmin(s_vdw, d - A.r[k]);
        }
    });
    s_sas = wave_min(s_sas);
    s_vdw = wave_min(s_vdw);
    if (A.kind != K_SES) return s_sas > margin ? ST_OUT : (s_sas < -margin ? ST_IN : ST_NONE);
    if (s_sas >= margin) return ST_OUT;
    if (s_vdw <= -margin) return ST_IN;
    double cutoff = A.probe + margin + CLASSIFY_EPS;
    // ses_eval(center, cutoff): inside the SAS at all?
    bool inside = false;
    for_near(A.gs, x, y, z, A.max_expanded, lane, WAVE, [&](uint32_t k) {
        if (!inside && A.status[k] != S_CONTAINED && ddist(x, y, z, A.cx[k], A.cy[k], A.cz[k]) < A.R[k])
            inside = true;
    });
    double dist = 0.0;
    if (wave_any(inside)) dist = ses_distance_wave(A, x, y, z, cutoff);
    double v = A.probe - dist;
    if (v > margin) return ST_OUT;
    if (v < -margin) return ST_IN;
    return ST_NONE;
}

// Lattice::brick_of (lattice.rs:88-90).
__device__ inline int64_t brick_of(double v, double bh) { return (int64_t)floor(v / bh); }

// Marks the touched bricks of every non-contained sphere (lattice.rs:
// 119-139) in its molecule's dense bitmap over a box holding all of them:
// set bits in ascending linear order are the sorted packed keys. One
// thread per sphere of the run (sph_off: first sphere of each molecule).
__global__ void k_mark_touched(const Mol* __restrict__ mols, const uint64_t* sph_off, uint32_t nm,
                               uint32_t* bits) {
    const uint64_t n = sph_off[nm];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t m = find_mol(sph_off, nm, i);
        const Mol& M = mols[m];
        const DArr& A = M.A;
        const DRun& P = M.P;
        uint32_t k = (uint32_t)(i - sph_off[m]);
        if (A.status[k] == S_CONTAINED) continue;
        double bh = (double)P.b * P.h;
        double r = A.R[k] + CLASSIFY_EPS;
        double c[3] = {A.cx[k], A.cy[k], A.cz[k]};
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = brick_of(c[a] - r, bh) - P.blo[a];
            hi[a] = brick_of(c[a] + r, bh) - P.blo[a];
        }
        uint32_t* mb = bits + M.word_off;
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    uint64_t lin = ((uint64_t)z * P.bdim[1] + (uint64_t)y) * P.bdim[0] + (uint64_t)x;
                    atomicOr(&mb[lin >> 5], 1u << (lin & 31));
                }
    }
}

__device__ inline void key_of_lin(const DRun& P, uint64_t lin, int32_t key[3]) {
    uint64_t x = lin % P.bdim[0];
    uint64_t t = lin / P.bdim[0];
    key[0] = (int32_t)((int64_t)x + P.blo[0]);
    key[1] = (int32_t)((int64_t)(t % P.bdim[1]) + P.blo[1]);
    key[2] = (int32_t)((int64_t)(t / P.bdim[1]) + P.blo[2]);
}

// Lattice::brick_center (lattice.rs:73-76).
__device__ inline void brick_center(const DRun& P, const int32_t key[3], double c[3]) {
    for (int a = 0; a < 3; a++) c[a] = ((double)key[a] + 0.5) * (double)P.b * P.h;
}

// Classification of touched bricks (lattice.rs:168-183): one wavefront per
// brick, item blockIdx.x of the queue q of touched-list indices; keep[t] =
// 1 where the surface may pass. touched[t] is a bit index of the run's
// concatenated bitmaps (wofs: first word of each molecule's). Most touched
// bricks are settled by the sphere pass alone, so a wavefront per brick
// beats a shared gather here.
__global__ void __launch_bounds__(64) k_classify_bricks(const Mol* __restrict__ mols, const uint64_t* wofs,
                                                        uint32_t nm, const uint32_t* touched, Queue q,
                                                        uint32_t* keep) {
    for_each_item<false>(q, nullptr, [&](uint32_t w) {
        uint32_t bit = touched[w];
        uint32_t m = find_mol(wofs, nm, (uint64_t)(bit >> 5));
        const Mol& M = mols[m];
        int32_t key[3];
        key_of_lin(M.P, bit - 32 * wofs[m], key);
        double c[3];
        brick_center(M.P, key, c);
        double rho = 0.5 * sqrt(3.0) * (double)M.P.b * M.P.h;
        uint8_t s = uniform_sign_wave(M.A, c[0], c[1], c[2], rho);
        if (lane_id() == 0) keep[w] = s == ST_NONE ? 1u : 0u;
    });
}

// Brick classification and sub-block states of the touched bricks whose
// shared-gather lists overflow (k_classify_touched): one wavefront per
// brick of the queue q (persistent): keep[w], and for a surface brick its
// eight sub-block states in stt[8 w + sb].
__global__ void __launch_bounds__(64) k_classify_touched_wave(const Mol* __restrict__ mols, const uint64_t* wofs,
                                                              uint32_t nm, const uint32_t* touched, Queue q,
                                                              uint32_t* keep, uint8_t* stt) {
    __shared__ uint32_t ticket;
    for_each_item<true>(q, &ticket, [&](uint32_t w) {
        const uint32_t bit = touched[w];
        const uint32_t m = find_mol(wofs, nm, (uint64_t)(bit >> 5));
        const Mol& M = mols[m];
        const DRun& P = M.P;
        int32_t key[3];
        key_of_lin(P, bit - 32 * wofs[m], key);
        double c[3];
        brick_center(P, key, c);
        const uint8_t bs = uniform_sign_wave(M.A, c[0], c[1], c[2], 0.5 * sqrt(3.0) * (double)P.b * P.h);
        if (lane_id() == 0) keep[w] = bs == ST_NONE ? 1u : 0u;
        if (bs != ST_NONE) return;
        const double sub = (double)P.sub, h = P.h, rho = 0.5 * sqrt(3.0) * sub * h, half = sub * 0.5;
        for (uint32_t i = 0; i < 8; i++) {
            for (int a = 0; a < 3; a++) {
                const double base = (double)((int64_t)key[a] * (int64_t)P.b) * h;
                c[a] = base + ((double)((i >> a) & 1) * sub + half) * h;
            }
            const uint8_t st = uniform_sign_wave(M.A, c[0], c[1], c[2], rho);
            if (lane_id() == 0) stt[8 * (size_t)w + i] = st;
        }
    });
}

// Sub-block states of a surface brick, gathered from the touched-brick
// order: st[8 * off[w] + i] = stt[8 * w + i] for every kept w; likewise its
// superset list (bt, touched order, into bs, surface order) when kept.
__global__ void k_gather_states(const uint32_t* keep, const uint32_t* off, const uint32_t* n, uint32_t cap,
                                const uint8_t* stt, uint8_t* st, SupList bt, SupList bs) {
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < *n; w += gridDim.x * blockDim.x) {
        if (!keep[w] || off[w] >= cap) continue;
        for (int i = 0; i < 8; i++) st[8 * (size_t)off[w] + i] = stt[8 * (size_t)w + i];
        if (bs.cnt) {
            bs.off[off[w]] = bt.off[w];
            bs.cnt[off[w]] = bt.cnt[w];
        }
    }
}

// Sub-block states of surface bricks (kernel.rs:467-482): one wavefront per
// brick of the queue q (persistent), its eight sub-blocks in turn; st[8 *
// brick + sb]. Used for bricks whose shared-gather lists overflow.
__global__ void __launch_bounds__(64) k_sub_states(const Mol* __restrict__ mols, const uint32_t* bmol,
                                                   const int32_t* keys, Queue q, uint8_t* st) {
    __shared__ uint32_t ticket;
    for_each_item<true>(q, &ticket, [&](uint32_t brick) {
        const Mol& M = mols[bmol[brick]];
        const DRun& P = M.P;
        double sub = (double)P.sub, h = P.h;
        double rho = 0.5 * sqrt(3.0) * sub * h;
        double half = sub * 0.5;
        for (uint32_t i = 0; i < 8; i++) {
            double c[3];
            for (int a = 0; a < 3; a++) {
                double base = (double)((int64_t)keys[3 * brick + a] * (int64_t)P.b) * h;
                double off = (double)((i >> a) & 1) * sub + half;
                c[a] = base + off * h;
            }
            uint8_t s = uniform_sign_wave(M.A, c[0], c[1], c[2], rho);
            if (lane_id() == 0) st[8 * brick + i] = s;
        }
    });
}

}  // namespace smg
