// Brick and sub-block classification with one shared gather per brick.
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
// value with eval_group. A touched brick (lattice.rs:168-183) is one ball.
#pragma once
#include "dev_classify.hpp"
#include "dev_refine.hpp"

namespace smg {

enum : uint32_t { UNIFORM_TOUCHED = 0, UNIFORM_SUBBLOCKS = 1 };

// One workgroup (256 threads) per unit: a touched brick (MODE
// UNIFORM_TOUCHED: work item t of the touched list, keep[t] = 1 where the
// surface may pass) or a surface brick (UNIFORM_SUBBLOCKS: st[8 * brick +
// i] = state of sub-block i). `work` lists the units (nullptr: unit = w);
// units whose lists overflow the capacities go to `overflow`.
template <uint32_t MODE, bool GLOBAL>
__global__ void __launch_bounds__(256) k_uniform(const Mol* __restrict__ mols, const uint64_t* wofs, uint32_t nm,
                                                 const uint32_t* touched, const uint32_t* bmol, const int32_t* keys,
                                                 RCaps cap, RLayout lay, const uint32_t* work, uint32_t n_work,
                                                 uint32_t* out, uint8_t* st, char* scratch, size_t slice,
                                                 uint32_t* overflow, uint32_t* n_overflow) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[R_N];
    __shared__ uint32_t tmp4[4];
    uint32_t w = blockIdx.x;
    if (w >= n_work) return;
    const uint32_t unit = work ? work[w] : w;
    uint32_t m;
    int32_t key[3];
    if (MODE == UNIFORM_TOUCHED) {
        uint32_t bit = touched[unit];
        m = find_mol(wofs, nm, (uint64_t)(bit >> 5));
        key_of_lin(mols[m].P, bit - 32 * wofs[m], key);
    } else {
        m = bmol[unit];
        for (int a = 0; a < 3; a++) key[a] = keys[3 * unit + a];
    }
    const DArr& A = mols[m].A;
    const DRun& P = mols[m].P;
    const double h = P.h;
    // Balls: centres, radius, margin and cutoff; the gather centre C and the
    // largest offset of a ball centre from it.
    double C[3], rho, off;
    if (MODE == UNIFORM_TOUCHED) {
        brick_center(P, key, C);
        rho = 0.5 * sqrt(3.0) * (double)P.b * h;
        off = 0.0;
    } else {
        brick_center(P, key, C);
        double sub = (double)P.sub;
        rho = 0.5 * sqrt(3.0) * sub * h;
        off = 0.5 * sqrt(3.0) * sub * h * (1.0 + 1e-9) + 1e-9;
    }
    const double margin = rho + CLASSIFY_EPS;
    const double cutoff = A.probe + margin + CLASSIFY_EPS;
    char* base = GLOBAL ? scratch + (size_t)w * slice : smem;
    RLists s = rcarve(base, lay);
    bool ok = gather_refine(A, cap, s, n, tmp4, C[0], C[1], C[2], off, margin + CLASSIFY_EPS);
    if (!ok) {
        if (threadIdx.x == 0) overflow[atomicAdd(n_overflow, 1u)] = unit;
        return;
    }
    // Lane groups: 8 of 32 lanes (one per sub-block), or one wavefront.
    const uint32_t nballs = MODE == UNIFORM_TOUCHED ? 1u : 8u;
    const uint32_t g = MODE == UNIFORM_TOUCHED ? (uint32_t)WAVE : 32u;
    const uint32_t ball = threadIdx.x / g;
    if (ball >= nballs) return;
    Group G;
    G.g = g;
    G.r = threadIdx.x % g;
    G.shift = (threadIdx.x % WAVE) - G.r;
    G.mask = g == WAVE ? ~0ull : ((1ull << g) - 1ull);
    double c[3];
    if (MODE == UNIFORM_TOUCHED) {
        c[0] = C[0];
        c[1] = C[1];
        c[2] = C[2];
    } else {  // sub-block centres as Engine::sub_states computes them
        double sub = (double)P.sub, half = sub * 0.5;
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
    if (s_sas >= margin) {
        state = ST_OUT;
    } else if (s_vdw <= -margin) {
        state = ST_IN;
    } else {
        EvalD e = eval_group(A, s, n, cap.cand, cap.near / 64, G, C[0], C[1], C[2], c[0], c[1], c[2], cutoff);
        double v = e.value;
        state = v > margin ? ST_OUT : (v < -margin ? ST_IN : ST_NONE);
    }
    if (G.r == 0) {
        if (MODE == UNIFORM_TOUCHED)
            out[unit] = state == ST_NONE ? 1u : 0u;
        else
            st[8 * unit + ball] = state;
    }
}

}  // namespace smg
