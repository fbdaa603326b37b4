// Learned per-brick sizing: a 64 -> 128 -> 128 -> 2 MLP (ReLU) whose two
// outputs are an occupancy logit (does the surface pass through the touched
// brick) and the log of the brick's candidate load (how large its sampling
// and refinement lists are). Layers 1 and 2 run on the BF16 matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation); layer 3 and every bias,
// activation and loss term are FP32. The predictions only choose the LDS
// capacity a sub-block is first tried at; the exact gathers check every
// list, so a wrong prediction costs a retry and never changes a mesh.
//
// No counterpart in the reference: it sizes its lists by a count pass.
#pragma once
#include "dev_common.hpp"

namespace smg {

constexpr int MLP_IN = 64, MLP_H = 128, MLP_OUT = 2;
// Parameter vector (FP32 master copy), weights stored [input][output].
constexpr int P_W1 = 0, P_B1 = P_W1 + MLP_IN * MLP_H, P_W2 = P_B1 + MLP_H, P_B2 = P_W2 + MLP_H * MLP_H,
              P_W3 = P_B2 + MLP_H, P_B3 = P_W3 + MLP_H * MLP_OUT, MLP_PARAMS = P_B3 + MLP_OUT;  // 25090
constexpr int MLP_ROWS = 64;  // rows (bricks) per workgroup: 4 wavefronts of 16
constexpr int MLP_STATS = 4;  // per-workgroup loss terms: BCE sum, squared-error sum, correct, surface rows

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__host__ __device__ inline uint16_t to_bf16(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

// One 16x16x16 BF16 product-accumulate. Lane l supplies A[l % 16][4 (l / 16)
// + j] and B[4 (l / 16) + j][l % 16] (j = 0..3) and holds D[4 (l / 16) + j][l % 16].
__device__ inline v4f mfma16(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Per-brick features (64): spheres within reach L = brick half-diagonal +
// 2 probe + largest expanded radius of the brick centre c, counted in 4
// radial shells (shell min(3, floor(4 d / L))) x 16 directions (octant of
// x - c, times whether |dz| is the largest component), each count as its
// square root. Distances are f64, so the counts are exact.
__device__ inline uint32_t feature_bin(double dx, double dy, double dz, double d, double L) {
    double f = floor(4.0 * d / L);
    uint32_t shell = f >= 3.0 ? 3u : (uint32_t)f;
    uint32_t oct = (dx >= 0.0 ? 1u : 0u) + (dy >= 0.0 ? 2u : 0u) + (dz >= 0.0 ? 4u : 0u);
    uint32_t polar = (fabs(dz) > fabs(dx) && fabs(dz) > fabs(dy)) ? 8u : 0u;
    return shell * 16u + oct + polar;
}

// One wavefront per brick, item blockIdx.x of q: out[t * 64 + f]. With
// `touched`, item t is a touched-list index (a bit of the run's bitmaps);
// without, a surface brick (keys, bmol).
__global__ void __launch_bounds__(64) k_features(const Mol* __restrict__ mols, const uint64_t* wofs, uint32_t nm,
                                                 const uint32_t* touched, const int32_t* keys, const uint32_t* bmol,
                                                 Queue q, float* out) {
    __shared__ uint32_t cnt[MLP_IN];
    for_each_item<false>(q, nullptr, [&](uint32_t t) {
        uint32_t m;
        int32_t key[3];
        if (touched) {
            uint32_t bit = touched[t];
            m = find_mol(wofs, nm, (uint64_t)(bit >> 5));
            key_of_lin(mols[m].P, bit - 32 * wofs[m], key);
        } else {
            m = bmol[t];
            for (int a = 0; a < 3; a++) key[a] = keys[3 * t + a];
        }
        const Mol& M = mols[m];
        const DArr& A = M.A;
        double c[3];
        brick_center(M.P, key, c);
        const double L = 0.5 * sqrt(3.0) * (double)M.P.b * M.P.h + 2.0 * A.probe + A.max_expanded;
        const uint32_t lane = lane_id();
        cnt[lane] = 0u;
        __syncthreads();
        for_near(A.gs, c[0], c[1], c[2], L, lane, WAVE, [&](uint32_t i) {
            double dx = A.cx[i] - c[0], dy = A.cy[i] - c[1], dz = A.cz[i] - c[2];
            double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (d < L) atomicAdd(&cnt[feature_bin(dx, dy, dz, d, L)], 1u);
        });
        __syncthreads();
        out[(size_t)t * MLP_IN + lane] = sqrtf((float)cnt[lane]);
        __syncthreads();
    });
}

// Where a pass's sub-blocks go: the lists of the passes that run, the pass
// that runs each capacity tier, and each tier's capacity in load units.
constexpr int SPLIT_TIERS = 5;
struct TierSplit {
    uint32_t* list[SPLIT_TIERS];
    uint32_t pass[SPLIT_TIERS];
    float cap[SPLIT_TIERS];
    float scale;        // predicted load is scaled by this before the tier choice
    uint32_t max_tier;  // highest tier a prediction may choose
};

// Sub-blocks of `work` (*n items, j = 8 brick + sb) into the list of the
// first pass whose capacity holds their brick's predicted load (pred[2
// brick + 1] = log(1 + load)) times T.scale, up to tier T.max_tier; larger
// tiers are reached by overflow only. The exact gathers still check every
// list: a sub-block sent too low overflows to the next pass. Appends are
// aggregated per wavefront (one atomic per pass and wavefront).
__global__ void k_split_tiers(const uint32_t* work, const uint32_t* n, const float* pred, TierSplit T,
                              uint32_t* counts, uint32_t* predicted) {
    const uint32_t cnt = *n, lane = lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~(uint32_t)(WAVE - 1)); i0 < cnt;
         i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + lane;
        uint32_t j = 0, p = SPLIT_TIERS;
        if (i < cnt) {
            j = work[i];
            float load = expm1f(fmaxf(pred[2 * (size_t)(j >> 3) + 1], 0.0f)) * T.scale;
            uint32_t t = 0;
            while (t < T.max_tier && load > T.cap[t]) t++;
            p = T.pass[t];
        }
        for (uint32_t q = 0; q < (uint32_t)SPLIT_TIERS; q++) {
            unsigned long long mask = __ballot(p == q);
            if (!mask) continue;
            uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1u, base = 0;
            if (lane == leader) {
                uint32_t k = (uint32_t)__popcll(mask);
                base = atomicAdd(&counts[q], k);
                atomicAdd(&predicted[q], k);
            }
            base = __shfl(base, (int)leader, WAVE);
            if (p == q) T.list[q][base + (uint32_t)__popcll(mask & below)] = j;
        }
    }
}

// Training labels of a run's touched bricks: occupancy (the exact
// classifier's keep) and log(1 + candidate load) of surface bricks (the
// largest over their sampled sub-blocks; 0 elsewhere).
__global__ void k_labels(const uint32_t* keep, const uint32_t* koff, const uint32_t* ntouched, const uint32_t* loads,
                         const uint32_t* nb, uint8_t* occ, float* target) {
    const uint32_t n = *ntouched, nbv = *nb;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        bool k = keep[t] != 0u && koff[t] < nbv;
        occ[t] = k ? 1 : 0;
        target[t] = k ? log1pf((float)loads[koff[t]]) : 0.0f;
    }
}

// Forward pass of one wavefront's 16 rows (rows r0 .. r0 + 15 of the
// workgroup's tile). xs: BF16 inputs [64][MLP_IN]; writes h1 (BF16, after
// ReLU) and h2 (FP32, after ReLU) of those rows.
__device__ inline void mlp_forward_rows(const float* __restrict__ P, const uint16_t* __restrict__ W1b,
                                        const uint16_t* __restrict__ W2b, const uint16_t* xs, uint16_t* h1,
                                        float* h2, uint32_t r0) {
    const uint32_t l = lane_id(), col = l % 16, kq = 4 * (l / 16);
    for (uint32_t nt = 0; nt < MLP_H / 16; nt++) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t kt = 0; kt < MLP_IN / 16; kt++) {
            v4s a, b;
            for (uint32_t j = 0; j < 4; j++) {
                a[j] = (short)xs[(r0 + col) * MLP_IN + kt * 16 + kq + j];
                b[j] = (short)W1b[(kt * 16 + kq + j) * MLP_H + nt * 16 + col];
            }
            acc = mfma16(a, b, acc);
        }
        for (uint32_t j = 0; j < 4; j++) {
            uint32_t n = nt * 16 + col;
            float v = fmaxf(acc[j] + P[P_B1 + n], 0.0f);
            h1[(r0 + kq + j) * MLP_H + n] = to_bf16(v);
        }
    }
    __syncthreads();  // layer 2 reads rows other lanes wrote
    for (uint32_t nt = 0; nt < MLP_H / 16; nt++) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t kt = 0; kt < MLP_H / 16; kt++) {
            v4s a, b;
            for (uint32_t j = 0; j < 4; j++) {
                a[j] = (short)h1[(r0 + col) * MLP_H + kt * 16 + kq + j];
                b[j] = (short)W2b[(kt * 16 + kq + j) * MLP_H + nt * 16 + col];
            }
            acc = mfma16(a, b, acc);
        }
        for (uint32_t j = 0; j < 4; j++) {
            uint32_t n = nt * 16 + col;
            h2[(r0 + kq + j) * MLP_H + n] = fmaxf(acc[j] + P[P_B2 + n], 0.0f);
        }
    }
}

// Layer 3 of row r: y[o] = sum_n h2[r][n] W3[n][o] + b3[o], n ascending.
__device__ inline float mlp_out(const float* __restrict__ P, const float* h2, uint32_t r, uint32_t o) {
    float y = 0.0f;
    for (uint32_t n = 0; n < MLP_H; n++) y += h2[r * MLP_H + n] * P[P_W3 + n * MLP_OUT + o];
    return y + P[P_B3 + o];
}

// Loads rows of X (FP32 [.][MLP_IN]) into the BF16 tile; rows past the
// count are zero.
__device__ inline void mlp_load_rows(const float* X, const uint32_t* idx, uint32_t base, uint32_t n, uint16_t* xs) {
    for (uint32_t e = threadIdx.x; e < MLP_ROWS * MLP_IN; e += blockDim.x) {
        uint32_t r = e / MLP_IN, f = e % MLP_IN, g = base + r;
        float v = 0.0f;
        if (g < n) v = X[(size_t)(idx ? idx[g] : g) * MLP_IN + f];
        xs[e] = to_bf16(v);
    }
}

// Inference: Y[i][o] for the rows i < *n of X. One workgroup (256) per 64 rows.
__global__ void __launch_bounds__(256) k_mlp_forward(const float* X, const uint32_t* n_rows, const float* P,
                                                     const uint16_t* W1b, const uint16_t* W2b, float* Y) {
    __shared__ uint16_t xs[MLP_ROWS * MLP_IN];
    __shared__ uint16_t h1[MLP_ROWS * MLP_H];
    __shared__ float h2[MLP_ROWS * MLP_H];
    const uint32_t n = *n_rows;
    for (uint32_t base = blockIdx.x * MLP_ROWS; base < n; base += gridDim.x * MLP_ROWS) {
        mlp_load_rows(X, nullptr, base, n, xs);
        __syncthreads();
        mlp_forward_rows(P, W1b, W2b, xs, h1, h2, 16 * (threadIdx.x / WAVE));
        __syncthreads();
        if (threadIdx.x < MLP_ROWS * MLP_OUT) {
            uint32_t r = threadIdx.x / MLP_OUT, o = threadIdx.x % MLP_OUT;
            if (base + r < n) Y[(size_t)(base + r) * MLP_OUT + o] = mlp_out(P, h2, r, o);
        }
        __syncthreads();
    }
}

// One training step's gradients: workgroup w takes batch rows [64 w, 64 w +
// 64) of idx (rows past nb are padding) and writes its partial parameter
// gradients to G[w][.] and its loss terms to S[w][.]. Loss per row, mean
// over the nb rows: BCE(y0, occ) + 0.5 * lambda * occ * (y1 - target)^2.
__global__ void __launch_bounds__(256) k_mlp_grad(const float* X, const uint8_t* occ, const float* target,
                                                  const uint32_t* idx, uint32_t nb, float lambda, const float* P,
                                                  const uint16_t* W1b, const uint16_t* W2b, float* G, float* S) {
    __shared__ uint16_t xs[MLP_ROWS * MLP_IN];
    __shared__ uint16_t h1[MLP_ROWS * MLP_H];
    __shared__ float h2[MLP_ROWS * MLP_H];  // later dH1 (FP32)
    __shared__ uint16_t dh2[MLP_ROWS * MLP_H];
    __shared__ uint16_t dh1[MLP_ROWS * MLP_H];
    __shared__ float dy[MLP_ROWS * MLP_OUT];
    __shared__ float st[MLP_STATS];
    const uint32_t base = blockIdx.x * MLP_ROWS, tid = threadIdx.x, w = tid / WAVE;
    const uint32_t l = lane_id(), col = l % 16, kq = 4 * (l / 16);
    float* g = G + (size_t)blockIdx.x * MLP_PARAMS;
    mlp_load_rows(X, idx, base, nb, xs);
    if (tid < MLP_STATS) st[tid] = 0.0f;
    __syncthreads();
    mlp_forward_rows(P, W1b, W2b, xs, h1, h2, 16 * w);
    __syncthreads();
    // Output gradients and loss terms.
    if (tid < MLP_ROWS) {
        uint32_t r = tid, gi = base + r;
        float d0 = 0.0f, d1 = 0.0f;
        if (gi < nb) {
            uint32_t row = idx[gi];
            float z0 = mlp_out(P, h2, r, 0), z1 = mlp_out(P, h2, r, 1);
            float t0 = occ[row] ? 1.0f : 0.0f, t1 = target[row];
            float p = 1.0f / (1.0f + expf(-z0));
            d0 = (p - t0) / (float)nb;
            float e = z1 - t1;
            d1 = lambda * t0 * e / (float)nb;
            float bce = fmaxf(z0, 0.0f) - z0 * t0 + log1pf(expf(-fabsf(z0)));
            atomicAdd(&st[0], bce);
            atomicAdd(&st[1], t0 * e * e);
            atomicAdd(&st[2], ((z0 > 0.0f) == (t0 > 0.5f)) ? 1.0f : 0.0f);
            atomicAdd(&st[3], t0);
        }
        dy[r * MLP_OUT] = d0;
        dy[r * MLP_OUT + 1] = d1;
    }
    __syncthreads();
    // Layer 3: dW3, db3 (thread = (n, o)).
    {
        uint32_t nn = tid / MLP_OUT, o = tid % MLP_OUT;
        float s = 0.0f;
        for (uint32_t r = 0; r < MLP_ROWS; r++) s += h2[r * MLP_H + nn] * dy[r * MLP_OUT + o];
        g[P_W3 + nn * MLP_OUT + o] = s;
        if (tid < MLP_OUT) {
            float b = 0.0f;
            for (uint32_t r = 0; r < MLP_ROWS; r++) b += dy[r * MLP_OUT + tid];
            g[P_B3 + tid] = b;
        }
    }
    // dH2 = dY W3^T masked by ReLU (BF16 for the products), db2 in FP32.
    for (uint32_t e = tid; e < MLP_ROWS * MLP_H; e += blockDim.x) {
        uint32_t r = e / MLP_H, nn = e % MLP_H;
        float v = dy[r * MLP_OUT] * P[P_W3 + nn * MLP_OUT] + dy[r * MLP_OUT + 1] * P[P_W3 + nn * MLP_OUT + 1];
        dh2[e] = to_bf16(h2[e] > 0.0f ? v : 0.0f);
    }
    if (tid < MLP_H) {
        float b = 0.0f;
        for (uint32_t r = 0; r < MLP_ROWS; r++) {
            float v = dy[r * MLP_OUT] * P[P_W3 + tid * MLP_OUT] + dy[r * MLP_OUT + 1] * P[P_W3 + tid * MLP_OUT + 1];
            b += h2[r * MLP_H + tid] > 0.0f ? v : 0.0f;
        }
        g[P_B2 + tid] = b;
    }
    __syncthreads();
    // dW2[i][j] = sum_r h1[r][i] dh2[r][j]: 64 tiles of 16x16, 16 per wavefront.
    for (uint32_t tile = w; tile < (MLP_H / 16) * (MLP_H / 16); tile += 4) {
        uint32_t it = tile / (MLP_H / 16), jt = tile % (MLP_H / 16);
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t kt = 0; kt < MLP_ROWS / 16; kt++) {
            v4s a, b;
            for (uint32_t j = 0; j < 4; j++) {
                a[j] = (short)h1[(kt * 16 + kq + j) * MLP_H + it * 16 + col];
                b[j] = (short)dh2[(kt * 16 + kq + j) * MLP_H + jt * 16 + col];
            }
            acc = mfma16(a, b, acc);
        }
        for (uint32_t j = 0; j < 4; j++) g[P_W2 + (it * 16 + kq + j) * MLP_H + jt * 16 + col] = acc[j];
    }
    // dH1 = dH2 W2^T masked by ReLU, per wavefront on its 16 rows; FP32 into
    // the h2 buffer (no longer needed), BF16 into dh1.
    float* dh1f = h2;
    __syncthreads();
    {
        uint32_t r0 = 16 * w;
        for (uint32_t nt = 0; nt < MLP_H / 16; nt++) {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t kt = 0; kt < MLP_H / 16; kt++) {
                v4s a, b;
                for (uint32_t j = 0; j < 4; j++) {
                    a[j] = (short)dh2[(r0 + col) * MLP_H + kt * 16 + kq + j];
                    b[j] = (short)W2b[(nt * 16 + col) * MLP_H + kt * 16 + kq + j];
                }
                acc = mfma16(a, b, acc);
            }
            for (uint32_t j = 0; j < 4; j++) {
                uint32_t r = r0 + kq + j, k = nt * 16 + col;
                float v = from_bf16(h1[r * MLP_H + k]) > 0.0f ? acc[j] : 0.0f;
                dh1f[r * MLP_H + k] = v;
                dh1[r * MLP_H + k] = to_bf16(v);
            }
        }
    }
    __syncthreads();
    if (tid < MLP_H) {
        float b = 0.0f;
        for (uint32_t r = 0; r < MLP_ROWS; r++) b += dh1f[r * MLP_H + tid];
        g[P_B1 + tid] = b;
    }
    // dW1[i][j] = sum_r x[r][i] dh1[r][j]: 32 tiles, 8 per wavefront.
    for (uint32_t tile = w; tile < (MLP_IN / 16) * (MLP_H / 16); tile += 4) {
        uint32_t it = tile / (MLP_H / 16), jt = tile % (MLP_H / 16);
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t kt = 0; kt < MLP_ROWS / 16; kt++) {
            v4s a, b;
            for (uint32_t j = 0; j < 4; j++) {
                a[j] = (short)xs[(kt * 16 + kq + j) * MLP_IN + it * 16 + col];
                b[j] = (short)dh1[(kt * 16 + kq + j) * MLP_H + jt * 16 + col];
            }
            acc = mfma16(a, b, acc);
        }
        for (uint32_t j = 0; j < 4; j++) g[P_W1 + (it * 16 + kq + j) * MLP_H + jt * 16 + col] = acc[j];
    }
    if (tid < MLP_STATS) S[(size_t)blockIdx.x * MLP_STATS + tid] = st[tid];
}

// Sums the partial gradients of `nw` workgroups (workgroup order) into grad
// and, with lr > 0, takes one Adam step (FP32 master weights, BF16 copies
// of W1 and W2 refreshed).
__global__ void k_mlp_adam(const float* G, uint32_t nw, float* P, float* Mo, float* Ve, uint16_t* W1b, uint16_t* W2b,
                           float* grad, float lr, float beta1, float beta2, float eps, float bc1, float bc2) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < (uint32_t)MLP_PARAMS; p += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (uint32_t w = 0; w < nw; w++) s += G[(size_t)w * MLP_PARAMS + p];
        if (grad) grad[p] = s;
        if (!(lr > 0.0f)) continue;
        float m = beta1 * Mo[p] + (1.0f - beta1) * s;
        float v = beta2 * Ve[p] + (1.0f - beta2) * s * s;
        Mo[p] = m;
        Ve[p] = v;
        float x = P[p] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
        P[p] = x;
        if (p < (uint32_t)P_B1) W1b[p - P_W1] = to_bf16(x);
        if (p >= (uint32_t)P_W2 && p < (uint32_t)P_B2) W2b[p - P_W2] = to_bf16(x);
    }
}

// BF16 copies of W1 and W2 from the master parameters.
__global__ void k_mlp_cast(const float* P, uint16_t* W1b, uint16_t* W2b) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)(MLP_H * MLP_H); i += gridDim.x * blockDim.x) {
        if (i < (uint32_t)(MLP_IN * MLP_H)) W1b[i] = to_bf16(P[P_W1 + i]);
        W2b[i] = to_bf16(P[P_W2 + i]);
    }
}

}  // namespace smg
