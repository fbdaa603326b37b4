// Sub-block classification of surface bricks, one shared gather per brick.
//
// lattice::uniform_sign (reference field/lattice.rs:190-227) of a ball
// (centre c, radius rho): with margin = rho + CLASSIFY_EPS, OUT if the
// least SAS distance min(d - R) >= margin, IN if the least van der Waals
// distance min(d - r) <= -margin, else by the exact clamped SES value
// ses_eval(c, probe + margin + CLASSIFY_EPS) (point.rs:384-489). The eight
// sub-block balls of a surface brick (kernel.rs:467-482) share one gather
// about the brick centre, grown by the largest centre offset: every sphere
// and candidate any of the eight tests can see is in it, and the extra
// ones cannot change a test (their distances exceed the margin and the
// cutoff). Eight lane groups then run the eight tests at once, the SES
// value with eval_group.
#pragma once
#include "dev_classify.hpp"
#include "dev_refine.hpp"

namespace smg {

// One workgroup (256 threads) per surface brick of the queue q: st[8 *
// brick + i] = state of sub-block i. Bricks whose lists overflow the
// capacities go to `overflow`. Lists live in dynamic LDS (GLOBAL false) or
// in scratch slice blockIdx.x of `slice` bytes (GLOBAL true).
template <bool GLOBAL>
__device__ __forceinline__ void uniform_brick(const Mol* __restrict__ mols, const uint32_t* bmol,
                                              const int32_t* keys, const RCaps& cap, const RLayout& lay,
                                              uint32_t unit, uint8_t* st, char* scratch, size_t slice,
                                              uint32_t* overflow, uint32_t* n_overflow) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[R_N];
    __shared__ uint32_t tmp4[4];
    const uint32_t m = bmol[unit];
    int32_t key[3];
    for (int a = 0; a < 3; a++) key[a] = keys[3 * unit + a];
    const DArr& A = mols[m].A;
    const DRun& P = mols[m].P;
    const double h = P.h;
    // Balls: centres, radius, margin and cutoff; the gather centre C and the
    // largest offset of a ball centre from it.
    double C[3];
    brick_center(P, key, C);
    const double sub = (double)P.sub;
    const double rho = 0.5 * sqrt(3.0) * sub * h;
    const double off = 0.5 * sqrt(3.0) * sub * h * (1.0 + 1e-9) + 1e-9;
    const double margin = rho + CLASSIFY_EPS;
    const double cutoff = A.probe + margin + CLASSIFY_EPS;
    char* base = GLOBAL ? scratch + (size_t)blockIdx.x * slice : smem;
    RLists s = rcarve(base, lay);
    bool ok = gather_refine(A, cap, s, n, tmp4, C[0], C[1], C[2], off, margin + CLASSIFY_EPS, false);
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = unit;
        return;
    }
    // Lane groups: 8 of 32 lanes, one per sub-block.
    const uint32_t g = 32u;
    const uint32_t ball = threadIdx.x / g;
    Group G;
    G.g = g;
    G.r = threadIdx.x % g;
    G.shift = (threadIdx.x % WAVE) - G.r;
    G.mask = (1ull << g) - 1ull;
    double c[3];
    {  // sub-block centres as Engine::sub_states computes them
        double half = sub * 0.5;
        for (int a = 0; a < 3; a++) {
            double basep = (double)((int64_t)key[a] * (int64_t)P.b) * h;
            double o = (double)((ball >> a) & 1) * sub + half;
            c[a] = basep + o * h;
        }
    }
    // Least SAS and van der Waals distances over the near spheres.
    double s_sas = INFINITY, s_vdw = INFINITY;
    const uint32_t n_near = n[R_NEAR];
    for (uint32_t k = G.r; k < n_near; k += g) {
        NearD q = s.nq[k];
        double d = ddist(c[0], c[1], c[2], q.x, q.y, q.z);
        s_sas = fmin(s_sas, d - s.nR[k]);
        s_vdw = fmin(s_vdw, d - s.nr[k]);
    }
    for (uint32_t o = g >> 1; o > 0; o >>= 1) {
        s_sas = fmin(s_sas, __shfl_xor(s_sas, (int)o, WAVE));
        s_vdw = fmin(s_vdw, __shfl_xor(s_vdw, (int)o, WAVE));
    }
    uint8_t state;
    if (A.kind != K_SES) {
        state = s_sas > margin ? ST_OUT : (s_sas < -margin ? ST_IN : ST_NONE);
    } else if (s_sas >= margin) {
        state = ST_OUT;
    } else if (s_vdw <= -margin) {
        state = ST_IN;
    } else {
        EvalD e = eval_group(A, s, n, cap.cand, cap.near / 64, G, C[0], C[1], C[2], c[0], c[1], c[2], cutoff,
                             margin + CLASSIFY_EPS);
        double v = e.value;
        state = v > margin ? ST_OUT : (v < -margin ? ST_IN : ST_NONE);
    }
    if (G.r == 0) st[8 * unit + ball] = state;
}

// Brick classification (lattice.rs:168-183) and, for a surface brick, its
// sub-block states (kernel.rs:467-482) from one shared gather: one
// workgroup per touched brick w of the queue (`touched` bit indices, wofs
// the first bitmap word of each molecule). The brick's own ball (centre C,
// the brick radius) lies inside the gathered one: its reach, probe + brick
// radius + 2 CLASSIFY_EPS, is below the gather's probe + (sub-block radius +
// CLASSIFY_EPS) + CLASSIFY_EPS + the sub-block offset by the offset's 1e-9
// growth, so the least SAS and vdW distances over the near spheres and the
// exact SES value at C from the lists decide as uniform_sign's scans of the
// whole grid do. keep[w] = 1 for a surface brick, whose states go to
// stt[8 w + i]. Bricks whose lists overflow go to `overflow`.
__device__ __forceinline__ void classify_touched(const Mol* __restrict__ mols, const uint64_t* wofs, uint32_t nm,
                                                 const uint32_t* touched, const RCaps& cap, const RLayout& lay,
                                                 uint32_t w, uint32_t* keep, uint8_t* stt, uint32_t* overflow,
                                                 uint32_t* n_overflow, const SupList& bsup) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[R_N];
    __shared__ uint32_t tmp4[4];
    __shared__ double red[2][4];
    __shared__ uint32_t brick_state;
    const uint32_t bit = touched[w];
    const uint32_t m = find_mol(wofs, nm, (uint64_t)(bit >> 5));
    const Mol& M = mols[m];
    const DArr& A = M.A;
    const DRun& P = M.P;
    int32_t key[3];
    key_of_lin(P, bit - 32 * wofs[m], key);
    const double h = P.h;
    double C[3];
    brick_center(P, key, C);
    const double sub = (double)P.sub;
    const double rho = 0.5 * sqrt(3.0) * sub * h;
    const double off = 0.5 * sqrt(3.0) * sub * h * (1.0 + 1e-9) + 1e-9;
    const double margin = rho + CLASSIFY_EPS;
    const double cutoff = A.probe + margin + CLASSIFY_EPS;
    RLists s = rcarve(smem, lay);
    // The walk also lists what the brick's eight sampling walks (and the
    // refinement supersets they collect) can take: items within the
    // refinement's reach of a sub-block centre, which lies within the
    // grown offset `off` of C (the 1e-6 A margin covers the rounding of the
    // centres and distances).
    const double reach_r = (A.kind != K_SES ? 0.0 : A.probe) + P.band + off;
    bool ok = gather_refine(A, cap, s, n, tmp4, C[0], C[1], C[2], off, margin + CLASSIFY_EPS, false, nullptr, NO_LIST,
                            nullptr, NO_LIST, bsup, w, reach_r + off + 1e-6);
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = w;
        return;
    }
    // The brick's own test (uniform_sign at C with the brick radius).
    const double rho_b = 0.5 * sqrt(3.0) * (double)P.b * h;
    const double margin_b = rho_b + CLASSIFY_EPS;
    {
        double ss = INFINITY, sv = INFINITY;
        for (uint32_t k = threadIdx.x; k < n[R_NEAR]; k += blockDim.x) {
            NearD q = s.nq[k];
            double d = ddist(C[0], C[1], C[2], q.x, q.y, q.z);
            ss = fmin(ss, d - s.nR[k]);
            sv = fmin(sv, d - s.nr[k]);
        }
        ss = wave_min(ss);
        sv = wave_min(sv);
        if (lane_id() == 0) {
            red[0][threadIdx.x / WAVE] = ss;
            red[1][threadIdx.x / WAVE] = sv;
        }
        __syncthreads();
        if (threadIdx.x < WAVE) {
            double s_sas = fmin(fmin(red[0][0], red[0][1]), fmin(red[0][2], red[0][3]));
            double s_vdw = fmin(fmin(red[1][0], red[1][1]), fmin(red[1][2], red[1][3]));
            uint32_t state;
            if (A.kind != K_SES) {
                state = s_sas > margin_b ? ST_OUT : (s_sas < -margin_b ? ST_IN : ST_NONE);
            } else if (s_sas >= margin_b) {
                state = ST_OUT;
            } else if (s_vdw <= -margin_b) {
                state = ST_IN;
            } else {
                Group G;
                G.g = WAVE;
                G.r = threadIdx.x;
                G.shift = 0;
                G.mask = ~0ull;
                const double cut_b = A.probe + margin_b + CLASSIFY_EPS;
                EvalD e = eval_group(A, s, n, cap.cand, cap.near / 64, G, C[0], C[1], C[2], C[0], C[1], C[2], cut_b,
                                     margin_b + CLASSIFY_EPS);
                const double v = e.value;
                state = v > margin_b ? ST_OUT : (v < -margin_b ? ST_IN : ST_NONE);
            }
            if (threadIdx.x == 0) {
                brick_state = state;
                keep[w] = state == ST_NONE ? 1u : 0u;
            }
        }
        __syncthreads();
    }
    if (brick_state != ST_NONE) return;
    // The eight sub-block tests, as uniform_brick.
    const uint32_t g = 32u;
    const uint32_t ball = threadIdx.x / g;
    Group G;
    G.g = g;
    G.r = threadIdx.x % g;
    G.shift = (threadIdx.x % WAVE) - G.r;
    G.mask = (1ull << g) - 1ull;
    double c[3];
    {
        double half = sub * 0.5;
        for (int a = 0; a < 3; a++) {
            double basep = (double)((int64_t)key[a] * (int64_t)P.b) * h;
            double o = (double)((ball >> a) & 1) * sub + half;
            c[a] = basep + o * h;
        }
    }
    double s_sas = INFINITY, s_vdw = INFINITY;
    const uint32_t n_near = n[R_NEAR];
    for (uint32_t k = G.r; k < n_near; k += g) {
        NearD q = s.nq[k];
        double d = ddist(c[0], c[1], c[2], q.x, q.y, q.z);
        s_sas = fmin(s_sas, d - s.nR[k]);
        s_vdw = fmin(s_vdw, d - s.nr[k]);
    }
    for (uint32_t o = g >> 1; o > 0; o >>= 1) {
        s_sas = fmin(s_sas, __shfl_xor(s_sas, (int)o, WAVE));
        s_vdw = fmin(s_vdw, __shfl_xor(s_vdw, (int)o, WAVE));
    }
    uint8_t state;
    if (A.kind != K_SES) {
        state = s_sas > margin ? ST_OUT : (s_sas < -margin ? ST_IN : ST_NONE);
    } else if (s_sas >= margin) {
        state = ST_OUT;
    } else if (s_vdw <= -margin) {
        state = ST_IN;
    } else {
        EvalD e = eval_group(A, s, n, cap.cand, cap.near / 64, G, C[0], C[1], C[2], c[0], c[1], c[2], cutoff,
                             margin + CLASSIFY_EPS);
        double v = e.value;
        state = v > margin ? ST_OUT : (v < -margin ? ST_IN : ST_NONE);
    }
    if (G.r == 0) stt[8 * (size_t)w + ball] = state;
}

template <bool PERSIST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_classify_touched(const Mol* __restrict__ mols, const uint64_t* wofs,
                                                          uint32_t nm, const uint32_t* touched, RCaps cap, RLayout lay,
                                                          Queue q, uint32_t* keep, uint8_t* stt, uint32_t* overflow,
                                                          uint32_t* n_overflow, SupList bsup) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t w) {
        classify_touched(mols, wofs, nm, touched, cap, lay, w, keep, stt, overflow, n_overflow, bsup);
    });
}

// One workgroup per surface brick of the queue q (PERSIST: a residency-
// sized grid looping over it; otherwise item blockIdx.x).
template <bool PERSIST>
__global__ void __launch_bounds__(256) k_uniform(const Mol* __restrict__ mols, const uint32_t* bmol,
                                                 const int32_t* keys, RCaps cap, RLayout lay, Queue q, uint8_t* st,
                                                 uint32_t* overflow, uint32_t* n_overflow) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t unit) {
        uniform_brick<false>(mols, bmol, keys, cap, lay, unit, st, nullptr, 0, overflow, n_overflow);
    });
}

}  // namespace smg
