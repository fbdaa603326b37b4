// Cell pass and assembly on the device: Marching Cubes cases from sample
// signs, triangles as edge references resolved to vertex indices local to
// the brick's molecule (reference mesh/extract.rs:226-265 and :606-691).
// Vertex indices are the brick's vertex base, less the molecule's first
// vertex, plus the edge's rank among the target brick's crossing edges, so
// triangles come out in the CPU's brick / cell / table order.
#pragma once
#include "dev_common.hpp"

namespace smg {

__constant__ uint16_t c_tri_offset[257];
__constant__ uint8_t c_tri_edges[820 * 3];
__constant__ uint8_t c_edge_lower[12];

// The Marching Cubes tables in LDS: indexed by each lane's own case, the
// constant-memory copies are vector loads; one copy per workgroup.
struct McTables {
    uint16_t off[257];
    uint8_t edges[820 * 3];
    uint8_t lower[12];
};
__device__ inline void load_tables(McTables& t) {
    for (uint32_t i = threadIdx.x; i < 257; i += blockDim.x) t.off[i] = c_tri_offset[i];
    for (uint32_t i = threadIdx.x; i < 820 * 3; i += blockDim.x) t.edges[i] = c_tri_edges[i];
    if (threadIdx.x < 12) t.lower[threadIdx.x] = c_edge_lower[threadIdx.x];
    __syncthreads();
}

// A brick's cell index c (< b^3) split into (x, y, z), x fastest, without
// integer division: c / d as __umulhi(c, ceil(2^32 / d)), exact for c and d
// below 2^16 (b <= 32 gives c < 32768 and d = b^2 <= 1024).
struct CellSplit {
    uint32_t b, bb, m1, m2;
};
__host__ __device__ inline CellSplit cell_split(uint32_t b) {
    const uint64_t one = 1ull << 32;
    return CellSplit{b, b * b, (uint32_t)((one + b - 1) / b), (uint32_t)((one + b * b - 1) / (b * b))};
}
__device__ inline void split_cell(const CellSplit& S, uint32_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    z = __umulhi(c, S.m2);
    const uint32_t r = c - z * S.bb;
    y = __umulhi(r, S.m1);
    x = r - y * S.b;
}

__device__ inline uint32_t cell_case(const float* val, uint32_t side, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t i0 = x + side * (y + side * z), cs = 0;
    for (uint32_t c = 0; c < 8; c++) {
        uint32_t st = (c & 1) + ((c >> 1) & 1) * side + ((c >> 2) & 1) * side * side;
        if (val[i0 + st] < 0.0f) cs |= 1u << c;
    }
    return cs;
}

// Sign bits of a brick's spb samples: bit i of sb is val[i] < 0 (the
// comparisons of cell_case and the edge test), one ballot per 64 samples;
// word (spb + 63) / 64, past the last sample, is written zero.
__device__ inline void sign_bits(const float* val, uint32_t spb, unsigned long long* sb) {
    const uint32_t nsw = (spb + 63) / 64, lane = lane_id();
    for (uint32_t k = threadIdx.x / WAVE; k <= nsw; k += blockDim.x / WAVE) {
        const uint32_t i = 64 * k + lane;
        const unsigned long long m = __ballot(i < spb && val[i] < 0.0f);
        if (lane == 0) sb[k] = m;
    }
}

// n <= 33 sign bits from bit i (a row of samples along x); the zero word
// past the last sample covers the read of word j + 1.
__device__ inline unsigned long long sign_run(const unsigned long long* sb, uint32_t i, uint32_t n) {
    const uint32_t j = i >> 6, off = i & 63u;
    const unsigned long long v = (sb[j] >> off) | ((sb[j + 1] << 1) << (63u - off));
    return v & ((1ull << n) - 1ull);
}

// The sample rows a row of cells (y, z) reads: (y, z), (y + 1, z), (y, z + 1)
// and (y + 1, z + 1), bit x the sample at x.
struct CellRows {
    unsigned long long m0, m1, m2, m3;
};
__device__ inline CellRows cell_rows(const unsigned long long* sb, uint32_t side, uint32_t y, uint32_t z) {
    const uint32_t i0 = side * (y + side * z), pl = side * side;
    return CellRows{sign_run(sb, i0, side), sign_run(sb, i0 + side, side), sign_run(sb, i0 + pl, side),
                    sign_run(sb, i0 + side + pl, side)};
}
// cell_case of cell x of the row: corner c = (dx, dy, dz) is bit dx + x of
// the row (y + dy, z + dz).
__device__ inline uint32_t row_case(const CellRows& R, uint32_t x) {
    return (uint32_t)(((R.m0 >> x) & 3ull) | (((R.m1 >> x) & 3ull) << 2) | (((R.m2 >> x) & 3ull) << 4) |
                      (((R.m3 >> x) & 3ull) << 6));
}

// Bit i of x (i < 21) to bit 3 i.
__device__ inline unsigned long long spread3(unsigned long long x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Per brick (one workgroup of 256, striding over the *n_bricks bricks):
// crossing-edge masks (edge ids ((z*b + y)*b + x)*3 + axis, 64 per word),
// per-word prefix counts, the brick's vertex total and per-sub-block counts
// (mesh/extract.rs:176-216), and its triangle count from the cells'
// Marching Cubes cases (extract.rs:245-260), which sizes the cell pass.
// The samples are read once, into sign bits (sign_bits); one thread per
// row of cells (y, z) then takes the row's x, y and z crossings as xors of
// sample rows (bit x: (x, y, z) against (x + 1, y, z), (x, y + 1, z),
// (x, y, z + 1)), interleaves them into the row's 3 b edge bits (at edge
// id 3 b (z b + y), 16 cells at a time) and ORs them into the words, and
// counts its cells' triangles from row_case. Dynamic LDS: the sign bits,
// (b + 1)^3 / 64 + 1 words, then the `words` mask words.
__global__ void __launch_bounds__(256) k_edges(uint32_t b, const float* values, const uint32_t* n_bricks,
                                               uint32_t words, unsigned long long* mask, uint32_t* wpre,
                                               uint32_t* bcount, uint32_t* scount, uint32_t* tcount) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t sc[8];
    __shared__ uint32_t tmp[4];
    __shared__ uint16_t toff[257];
    for (uint32_t i = threadIdx.x; i < 257; i += blockDim.x) toff[i] = c_tri_offset[i];
    const uint32_t nb = *n_bricks;
    const uint32_t side = b + 1, sub = b / 2, spb = side * side * side, nsw = (spb + 63) / 64;
    unsigned long long* sb = (unsigned long long*)smem;
    unsigned long long* wm = sb + nsw + 1;
    const unsigned long long lowb = (1ull << b) - 1ull, lows = (1ull << sub) - 1ull;
    for (uint32_t brick = blockIdx.x; brick < nb; brick += gridDim.x) {
        sign_bits(values + (size_t)brick * spb, spb, sb);
        for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) wm[w] = 0ull;
        if (threadIdx.x < 8) sc[threadIdx.x] = 0;
        __syncthreads();
        uint32_t mine = 0;
        for (uint32_t r = threadIdx.x; r < b * b; r += blockDim.x) {
            const uint32_t y = r % b, z = r / b;
            const CellRows R = cell_rows(sb, side, y, z);
            const unsigned long long ex = (R.m0 ^ (R.m0 >> 1)) & lowb, ey = (R.m0 ^ R.m1) & lowb,
                                     ez = (R.m0 ^ R.m2) & lowb;
            // sub-blocks of the row's lower (x < sub) and upper cells
            const uint32_t s0 = (y >= sub ? 2u : 0u) + (z >= sub ? 4u : 0u);
            const uint32_t nlo = __popcll(ex & lows) + __popcll(ey & lows) + __popcll(ez & lows);
            const uint32_t nhi = __popcll(ex >> sub) + __popcll(ey >> sub) + __popcll(ez >> sub);
            if (nlo) atomicAdd(&sc[s0], nlo);
            if (nhi) atomicAdd(&sc[s0 + 1], nhi);
            for (uint32_t k = 0; 16 * k < b; k++) {
                const uint32_t sh = 16 * k;
                const unsigned long long v = spread3((ex >> sh) & 0xffffull) | (spread3((ey >> sh) & 0xffffull) << 1) |
                                             (spread3((ez >> sh) & 0xffffull) << 2);
                if (!v) continue;
                // at most 48 bits from edge bit P: word j, and j + 1 past bit 64
                const uint32_t P = 3 * b * r + 48 * k, j = P >> 6, off = P & 63u;
                atomicOr(&wm[j], v << off);
                if (off > 16u) {
                    const unsigned long long hi = v >> (64u - off);
                    if (hi) atomicOr(&wm[j + 1], hi);
                }
            }
            for (uint32_t x = 0; x < b; x++) {
                const uint32_t cs = row_case(R, x);
                mine += toff[cs + 1] - toff[cs];
            }
        }
        __syncthreads();
        uint32_t total;
        block_scan_256(mine, tmp, total);
        uint32_t carry = 0;
        for (uint32_t base = 0; base < words; base += blockDim.x) {
            const uint32_t w = base + threadIdx.x;
            const unsigned long long m = w < words ? wm[w] : 0ull;
            uint32_t all;
            const uint32_t before = block_scan_256((uint32_t)__popcll(m), tmp, all);
            if (w < words) {
                mask[(size_t)brick * words + w] = m;
                wpre[(size_t)brick * words + w] = carry + before;
            }
            carry += all;
        }
        if (threadIdx.x == 0) {
            bcount[brick] = carry;
            tcount[brick] = total;
        }
        if (threadIdx.x < 8) scount[8 * brick + threadIdx.x] = sc[threadIdx.x];
        __syncthreads();
    }
}

__device__ inline int64_t dense_index(const DRun& P, int64_t kx, int64_t ky, int64_t kz) {
    int64_t x = kx - P.blo[0], y = ky - P.blo[1], z = kz - P.blo[2];
    if (x < 0 || y < 0 || z < 0 || x >= P.bdim[0] || y >= P.bdim[1] || z >= P.bdim[2]) return -1;
    return (z * (int64_t)P.bdim[1] + y) * (int64_t)P.bdim[0] + x;
}

// Triangles of a brick that k_cell_emit lays out in LDS before writing
// them; a brick with more takes the per-thread path.
constexpr uint32_t TRI_SLOTS = 1024;

// The bricks whose edges a brick's cells reference (mesh/extract.rs:629-667),
// one thread per (brick, code): tgt[8 brick + code] for code 1..7 is the
// brick at key + (code bits, one per axis) in the molecule's dense index
// grid, -1 if that is no surface brick; tgt[8 brick] is the molecule's
// first brick fbrick[m] (the brick itself is its own code-0 target), whose
// vertex base the molecule's local vertex indices start from. Resolved
// here, in parallel over every brick, the cell pass reads them in one load
// instead of four dependent ones per brick.
__global__ void k_cell_targets(const Mol* __restrict__ mols, const uint32_t* bmol, const uint32_t* fbrick,
                               const int32_t* keys, const uint32_t* dense, const uint32_t* n_bricks, int32_t* tgt) {
    const uint64_t n = 8ull * *n_bricks;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t brick = (uint32_t)(i / 8), code = (uint32_t)(i % 8), m = bmol[brick];
        int64_t tb = fbrick[m];
        if (code) {
            const int64_t di = dense_index(mols[m].P, (int64_t)keys[3 * brick] + (code & 1),
                                           (int64_t)keys[3 * brick + 1] + ((code >> 1) & 1),
                                           (int64_t)keys[3 * brick + 2] + ((code >> 2) & 1));
            const uint32_t* mdense = dense + mols[m].dense_off;
            tb = (di < 0 || mdense[di] == 0xffffffffu) ? -1 : (int64_t)mdense[di];
        }
        tgt[i] = (int32_t)tb;
    }
}

// Emits the triangles of every brick at tbase[brick] (mesh/extract.rs:245-260
// and :629-667), with vertex indices local to the brick's molecule (its
// first vertex is that of its first brick), the referenced bricks from
// k_cell_targets. A reference that does not resolve sets *err; triangles
// at or beyond `tcap` are not written (the run's triangle count exceeds the
// buffer and is reported by its caller).
// With `staged`, the brick's sample signs (sign_bits, (b+1)^3 / 64 + 1 u64),
// then the edge masks and per-word prefix counts of the brick and of its
// seven upper neighbours, are first laid out in dynamic LDS (8 `words` u64
// and 8 `words` u32), so every cell case and triangle corner resolves from
// LDS (a thread's cells are consecutive, so it reads each row of sample
// signs once); each thread then lists
// its cells' triangles (cell, table row) at its offset in TRI_SLOTS more
// u32, and the threads take one listed triangle each, so a wavefront no
// longer waits for the lane whose cells hold the most triangles.
__global__ void __launch_bounds__(256) k_cell_emit(uint32_t brick_size, const int32_t* tgt, const float* values,
                                                   const uint32_t* n_bricks, const uint32_t* tbase,
                                                   const uint32_t* vbase, const unsigned long long* mask,
                                                   const uint32_t* wpre, uint32_t words, uint32_t* tris, size_t tcap,
                                                   uint32_t* err, uint32_t staged) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t tmp[4];
    __shared__ int32_t target[8];  // brick at key + (code bits), -1 if none
    __shared__ uint32_t tvbase[8];  // its vertex base less the molecule's first vertex
    __shared__ McTables tab;
    load_tables(tab);
    const uint32_t nb = *n_bricks;
    for (uint32_t brick = blockIdx.x; brick < nb; brick += gridDim.x) {
        // The brick size, opaque to the compiler within each brick: known
        // across the loop, the thread's cell coordinates and LDS addresses
        // would be hoisted out of it and held in registers (about 190 VGPRs).
        uint32_t b = brick_size;
        asm volatile("" : "+s"(b));
        const uint32_t side = b + 1, spb = side * side * side;
        const float* val = values + (size_t)brick * spb;
        const unsigned long long* cmask = mask;  // edge masks and prefix counts, by brick
        const uint32_t* cpre = wpre;
        if (threadIdx.x < 8) {  // the eight bricks whose edges the cells reference, once per brick
            const uint32_t code = threadIdx.x;
            const int32_t first = tgt[8 * (size_t)brick], tb = code ? tgt[8 * (size_t)brick + code] : (int32_t)brick;
            target[code] = tb;
            tvbase[code] = tb < 0 ? 0u : vbase[tb] - vbase[first];
        }
        unsigned long long* sb = (unsigned long long*)smem;
        if (staged) sign_bits(val, spb, sb);
        __syncthreads();
        if (staged) {
            unsigned long long* sm = sb + (spb + 63) / 64 + 1;
            uint32_t* sp = (uint32_t*)(sm + 8 * words);
            for (uint32_t i = threadIdx.x; i < 8 * words; i += blockDim.x) {
                const int32_t tb = target[i / words];
                if (tb >= 0) {
                    sm[i] = mask[(size_t)tb * words + i % words];
                    sp[i] = wpre[(size_t)tb * words + i % words];
                }
            }
            cmask = sm;
            cpre = sp;
        }
        __syncthreads();
        const uint32_t ncell = b * b * b, per = (ncell + blockDim.x - 1) / blockDim.x;
        uint32_t c_lo = threadIdx.x * per, c_hi = min(ncell, (threadIdx.x + 1) * per);
        const CellSplit S = cell_split(b);
        // The case of cell c = (x, y, z): from the sign rows of its row of
        // cells (y, z), or from the samples in global memory. The first walk
        // over the thread's cells reads each row once and keeps the cases of
        // its first 8 cells, a byte each, for the second.
        uint32_t mine = 0, pk0 = 0u, pk1 = 0u;
        {
            CellRows rows{};
            uint32_t row_at = 0xffffffffu;
            for (uint32_t c = c_lo; c < c_hi; c++) {
                uint32_t x, y, z;
                split_cell(S, c, x, y, z);
                uint32_t cs;
                if (staged) {
                    if (z * b + y != row_at) {
                        rows = cell_rows(sb, side, y, z);
                        row_at = z * b + y;
                    }
                    cs = row_case(rows, x);
                } else {
                    cs = cell_case(val, side, x, y, z);
                }
                const uint32_t q = c - c_lo;
                if (q < 4u)
                    pk0 |= cs << (8u * q);
                else if (q < 8u)
                    pk1 |= cs << (8u * (q - 4u));
                mine += tab.off[cs + 1] - tab.off[cs];
            }
        }
        auto case_again = [&](uint32_t c) {
            const uint32_t q = c - c_lo;
            if (q < 8u) return ((q < 4u ? pk0 : pk1) >> (8u * (q & 3u))) & 255u;
            uint32_t x, y, z;
            split_cell(S, c, x, y, z);
            return staged ? row_case(cell_rows(sb, side, y, z), x) : cell_case(val, side, x, y, z);
        };
        uint32_t total;
        const uint32_t first = block_scan_256(mine, tmp, total);
        const size_t base = (size_t)tbase[brick];
        // The vertex of edge e of table row's cell (x, y, z).
        auto corner = [&](uint32_t x, uint32_t y, uint32_t z, uint32_t e) {
            uint32_t axis = e / 4, lo = tab.lower[e];
            uint32_t l[3] = {x + (lo & 1), y + ((lo >> 1) & 1), z + ((lo >> 2) & 1)};
            uint32_t code = 0, lc[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++) {
                if (l[a] == b)
                    code |= 1u << a;
                else
                    lc[a] = l[a];
            }
            uint32_t id = ((lc[2] * b + lc[1]) * b + lc[0]) * 3 + axis;
            const int32_t tb = target[code];
            const uint32_t at_b = staged ? code : (uint32_t)tb;  // row of cmask / cpre
            uint32_t vi = 0;
            if (tb < 0 || !edge_is_crossing(cmask, words, at_b, id))
                atomicOr(err, 1u);
            else
                vi = tvbase[code] + edge_rank(cmask, cpre, words, at_b, id);
            return vi;
        };
        if (staged && total <= TRI_SLOTS) {
            uint32_t* slots = (uint32_t*)(cpre + 8 * words);
            uint32_t k = first;
            for (uint32_t c = c_lo; c < c_hi; c++) {
                const uint32_t cs = case_again(c);
                for (uint32_t t = tab.off[cs]; t < tab.off[cs + 1]; t++) slots[k++] = (c << 10) | t;
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
                const uint32_t c = slots[i] >> 10, t = slots[i] & 1023u;
                uint32_t x, y, z;
                split_cell(S, c, x, y, z);
                const size_t at = base + i;
                for (int q = 0; q < 3; q++) {
                    const uint32_t vi = corner(x, y, z, tab.edges[3 * t + q]);
                    if (at < tcap) tris[3 * at + q] = vi;
                }
            }
        } else {
            size_t at = base + first;
            for (uint32_t c = c_lo; c < c_hi; c++) {
                uint32_t x, y, z;
                split_cell(S, c, x, y, z);
                const uint32_t cs = case_again(c);
                for (uint32_t t = tab.off[cs]; t < tab.off[cs + 1]; t++) {
                    for (int q = 0; q < 3; q++) {
                        const uint32_t vi = corner(x, y, z, tab.edges[3 * t + q]);
                        if (at < tcap) tris[3 * at + q] = vi;
                    }
                    at++;
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace smg
