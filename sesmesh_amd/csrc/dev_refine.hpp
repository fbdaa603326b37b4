// Crossing edges, exact f64 vertex refinement and vertex attributes.
//
// Restates the edge scan of mesh_brick (reference mesh/extract.rs:146-223),
// SesField::gather_local / eval_local (field/point.rs:419-489, 496-581,
// 686-738) and vertex / gradient / attributes (mesh/extract.rs:425-597) in
// f64 with the CPU's operations and iteration rules, so every vertex,
// normal, atom and patch equals the CPU back end's bit for bit. The SAS and
// vdW kinds list every sphere within reach as class S candidates, nearest
// first by their signed distance at the ball centre, without blockers.
//
// One workgroup (256 threads) per sub-block with crossing edges. It gathers
// the sub-block's lists into LDS once: the near spheres, the inside group,
// and the S, C and V candidates sorted as the CPU orders them (class, then
// f64::total_cmp of the distance at the ball centre, then index), each
// with its f64 data and its blockers as a bitset over near-list slots.
// The crossing edges are then refined by lane groups: g lanes per edge
// (g = 64 / 2^k, as many groups as edges allow). An evaluation point is
// split over the group's lanes: each lane scans its share of the
// candidates in list order, keeping the least (distance, list position)
// among the free ones below the cutoff, and the group takes the least of
// its lanes. That is the candidate eval_local's sequential loop keeps: the
// first, in list order, of the free candidates at the least distance. A
// lane stops a class's scan once the candidate's centre distance minus the
// point's offset from the centre (a lower bound of its distance at the
// point, the distances being 1-Lipschitz) exceeds its best by a margin far
// above f64 rounding: no later candidate can be taken.
#pragma once
#include "dev_common.hpp"

namespace smg {

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

// Capacities of the refinement lists: near spheres (a multiple of 64: the
// blocker bitsets hold near/64 words), inside group, candidates (a power
// of two), hash slots (a power of two) and crossing edges; `radii` 1 keeps
// each near sphere's expanded and van der Waals radii too. `gmax` bounds
// the lanes per edge (0: a wavefront); any power of two gives the same
// vertices.
struct RCaps {
    uint32_t near, inside, cand, hash, edges, radii, gmax;
};

// LDS / scratch layout of the refinement lists (byte offsets).
struct RLayout {
    uint32_t nq, nin, nR, nr, nt, hk, hv, ins, ckey, cid, cdat, cbits, coff, cpre, ebuf, bytes;
};

// Near sphere record: centre and the blocking threshold of R - EPS.
struct alignas(16) NearD {
    double x, y, z, blk;
};

constexpr uint32_t CDAT = 8;  // doubles per candidate record

__host__ __device__ inline RLayout make_rlayout(const RCaps& c) {
    RLayout L;
    uint32_t o = 0;
    auto take = [&](uint32_t bytes) {
        uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    L.nq = take(32 * c.near);
    L.nin = take(8 * c.near);  // inside threshold of R
    L.nR = take(c.radii ? 8 * c.near : 0);
    L.nr = take(c.radii ? 8 * c.near : 0);
    L.nt = take(c.radii ? 0 : 4 * c.near);  // reach limit of the near spheres as blockers
    L.hk = take(4 * c.hash);
    L.hv = take(4 * c.hash);
    L.ins = take(2 * c.inside);
    L.ckey = take(8 * c.cand);
    L.cid = take(4 * c.cand);
    L.cdat = take(8 * CDAT * c.cand);
    L.cbits = take(c.near / 64 * 8 * c.cand);
    L.coff = take(4 * c.cand);
    L.cpre = take(4 * c.cand);
    L.ebuf = take(4 * c.edges);
    L.bytes = o;
    return L;
}

struct RLists {
    NearD* nq;
    double* nin;
    double *nR, *nr;  // radii of near spheres (when gathered)
    float* nt;        // sqrt(blk) + COVER_MARGIN + REACH_SLACK of near spheres (refinement lists)
    uint32_t *hk, *hv;
    uint16_t* ins;
    unsigned long long* ckey;
    uint32_t* cid;
    // Candidate records: S (x, y, z, R), C (mx, my, mz, ax, ay, az, rho),
    // V (x, y, z).
    double* cdat;
    unsigned long long* cbits;  // near/64 rows of `cand` words
    uint32_t *coff, *cpre;  // blocker row offset and length; once the bitsets are built, cpre holds the hint
    uint32_t* ebuf;
};

constexpr uint32_t NO_HINT = 0xffffffffu;

__device__ inline RLists rcarve(char* base, const RLayout& L) {
    RLists s;
    s.nq = (NearD*)(base + L.nq);
    s.nin = (double*)(base + L.nin);
    s.nR = (double*)(base + L.nR);
    s.nr = (double*)(base + L.nr);
    s.nt = (float*)(base + L.nt);
    s.hk = (uint32_t*)(base + L.hk);
    s.hv = (uint32_t*)(base + L.hv);
    s.ins = (uint16_t*)(base + L.ins);
    s.ckey = (unsigned long long*)(base + L.ckey);
    s.cid = (uint32_t*)(base + L.cid);
    s.cdat = (double*)(base + L.cdat);
    s.cbits = (unsigned long long*)(base + L.cbits);
    s.coff = (uint32_t*)(base + L.coff);
    s.cpre = (uint32_t*)(base + L.cpre);
    s.ebuf = (uint32_t*)(base + L.ebuf);
    return s;
}

enum { R_NEAR = 0, R_IN, R_S, R_C, R_V, R_CAND, R_E, R_OVER, R_X, R_XOVER, R_N };

// gather_local (field/point.rs:496-565) of the ball (b, rho) by the
// workgroup: near spheres with their thresholds, the inside group as
// near-list slots, the S / C / V candidates in the CPU's list order with
// their f64 data and blocker bitsets (row members that are near,
// point.rs:152-155). False (in every thread) if a capacity was exceeded.
// With `hide`, candidates that one blocker hides from the whole ball are
// left out: for callers that evaluate many points of the ball. The hidden
// ones are taken from `hl` (hn entries, see HidList) when the sampling of
// the same sub-block listed them, else found with mark_hidden. With xs,
// the walk also lists (SupList xs, entry xj) every item within reach_x of
// the ball centre (spheres: d - R, or for the SES kind d < R + reach_x),
// whatever the tests decide: a superset for walks of nearby balls.
__device__ inline bool gather_refine(const DArr& A, const RCaps& cap, RLists& s, uint32_t* n, uint32_t* tmp4,
                                     double bx, double by, double bz, double rho, double band, bool hide,
                                     const uint32_t* hl = nullptr, uint32_t hn = NO_LIST,
                                     const uint32_t* sl = nullptr, uint32_t sn = NO_LIST, SupList xs = SupList{},
                                     uint32_t xj = 0, double reach_x = 0.0, const uint32_t* br = nullptr) {
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    __shared__ uint32_t bstart[WAVE], bfill[WAVE];  // per key bucket: 16 for each of S, C, V
    for (uint32_t i = tid; i < cap.hash; i += nt) s.hk[i] = 0u;
    if (tid < R_N) n[tid] = 0u;
    if (tid < WAVE) bstart[tid] = 0u;
    __syncthreads();
    const uint32_t hmask = cap.hash - 1;
    // SAS / vdW (point.rs:566-578): reach band + rho, keys d - R (signed,
    // at least -max_expanded), bucketed from -max_expanded.
    const bool sph = A.kind != K_SES;
    hide = hide && !sph;
    // A blocker record whose lists are this gather's own (`same`, see
    // BlkRecord) gives them directly: the near spheres in its slot order
    // with their records, the inside group tested here, the kept candidates
    // in list order with their keys, records, bitsets and hints.
    if (br && hide && br[5] && br[0] <= cap.near && br[1] + br[2] + br[4] <= cap.cand &&
        br[3] <= cap.near / 64) {
        const uint32_t nn = br[0], ns = br[1], nc = br[2], nwr = br[3], nv = br[4], total = ns + nc + nv;
        const uint32_t ew = blk_entry_words(nwr), cc = cap.cand, nw = cap.near / 64;
        const uint32_t* ent = br + BLK_HEAD + nn;
        for (uint32_t i = tid; i < nn; i += nt) {
            const uint32_t k = br[BLK_HEAD + i];
            const SphereRec S = load_rec(A.sph, k);
            const SphereThr T = load_rec(A.thr, k);
            s.nq[i] = NearD{S.x, S.y, S.z, T.blk};
            s.nin[i] = T.nin;
            if (ddist(bx, by, bz, S.x, S.y, S.z) < S.R + rho) {
                const uint32_t q = atomicAdd(&n[R_IN], 1u);
                if (q >= cap.inside)
                    n[R_OVER] = 1u;
                else
                    s.ins[q] = (uint16_t)i;
            }
        }
        for (uint32_t p = tid; p < total; p += nt) {
            const uint32_t* e = ent + (size_t)p * ew;
            const uint32_t k = e[0] & ID_MASK;
            s.ckey[p] = (unsigned long long)e[2] | ((unsigned long long)e[3] << 32);
            s.cid[p] = e[0];
            double* d = s.cdat + (size_t)CDAT * p;
            if (p < ns) {
                const SphereRec S = load_rec(A.sph, k);
                d[0] = S.x;
                d[1] = S.y;
                d[2] = S.z;
                d[3] = S.R;
                s.cpre[p] = e[1];
            } else if (p < ns + nc) {
                const CircleRec C = load_rec(A.cir, k);
                d[0] = C.mx;
                d[1] = C.my;
                d[2] = C.mz;
                d[3] = C.ax;
                d[4] = C.ay;
                d[5] = C.az;
                d[6] = C.rho;
                s.cpre[p] = e[1];
            } else {
                const VertexRec V = load_rec(A.ver, k);
                d[0] = V.x;
                d[1] = V.y;
                d[2] = V.z;
            }
            for (uint32_t w = 0; w < nw; w++)
                s.cbits[w * cc + p] =
                    w < nwr ? (unsigned long long)e[4 + 2 * w] | ((unsigned long long)e[5 + 2 * w] << 32) : 0ull;
        }
        if (tid == 0) {
            n[R_NEAR] = nn;
            n[R_S] = ns;
            n[R_C] = nc;
            n[R_V] = nv;
            n[R_CAND] = total;
        }
        __syncthreads();
        return n[R_OVER] == 0u;
    }
    const double reach = (sph ? 0.0 : A.probe) + band + rho;
    const double rw = xs.ids ? fmax(reach, reach_x) : reach;  // the walk's reach
    const double radii[3] = {rw + A.max_expanded, rw + A.rho_max, rw};
    const double key0 = sph ? A.max_expanded : 0.0;
    uint32_t* xstage = (uint32_t*)s.cdat;  // free until the walk has ended
    const uint32_t xcap = xs.ids ? 2u * CDAT * cap.cand : 0u;
    auto xadd = [&](uint32_t code) {
        const uint32_t at = atomicAdd(&n[R_X], 1u);
        if (at < xcap)
            xstage[at] = code;
        else
            n[R_XOVER] = 1u;
    };
    const double per_bucket = KEY_BUCKETS / (reach + key0);
    auto bucket = [&](uint32_t cls, double key) {
        return cls * KEY_BUCKETS + min(KEY_BUCKETS - 1u, (uint32_t)(fmax(key + key0, 0.0) * per_bucket));
    };
    // The grids' items near the ball, or (sl) the sampling's list of the
    // items these tests take (SupList); the tests are the same either way.
    auto visit = [&](uint32_t grid, uint32_t k) {
        if (grid == 0 && sph) {
            if (A.status[k] == S_CONTAINED) return;
            const SphereRec S = load_rec(A.sph, k);
            const double v = ddist(bx, by, bz, S.x, S.y, S.z) - S.R;
            if (xcap && v < reach_x) xadd(k);
            if (!(v < reach)) return;
            uint32_t at = atomicAdd(&n[R_NEAR], 1u);
            uint32_t q = atomicAdd(&n[R_CAND], 1u);
            if (at >= cap.near || q >= cap.cand) {
                n[R_OVER] = 1u;
                return;
            }
            s.nq[at] = NearD{S.x, S.y, S.z, 0.0};
            if (cap.radii) {
                s.nR[at] = S.R;
                s.nr[at] = A.r[k];
            }
            s.ckey[q] = total_key(v);
            s.cid[q] = k;
        } else if (grid == 0) {
            if (A.status[k] == S_CONTAINED) return;
            const SphereRec S = load_rec(A.sph, k);
            double R = S.R, cx = S.x, cy = S.y, cz = S.z;
            double d = ddist(bx, by, bz, cx, cy, cz);
            if (xcap && d < R + reach_x) xadd(k);
            if (d < R + reach) {
                uint32_t at = atomicAdd(&n[R_NEAR], 1u);
                if (at >= cap.near) {
                    n[R_OVER] = 1u;
                    return;
                }
                NearD q;
                q.x = cx;
                q.y = cy;
                q.z = cz;
                const SphereThr t = load_rec(A.thr, k);
                q.blk = t.blk;
                s.nq[at] = q;
                s.nin[at] = t.nin;
                if (hide) s.nt[at] = (float)(sqrt(t.blk) + COVER_MARGIN) + REACH_SLACK;
                if (cap.radii) {
                    s.nR[at] = R;
                    s.nr[at] = A.r[k];
                }
                hash_put(s.hk, s.hv, hmask, k, at);
                if (d < R + rho) {
                    uint32_t q = atomicAdd(&n[R_IN], 1u);
                    if (q >= cap.inside)
                        n[R_OVER] = 1u;
                    else
                        s.ins[q] = (uint16_t)at;
                }
            }
            double lb = fabs(d - R);
            if (A.status[k] == S_EXPOSED && lb <= reach) {
                atomicAdd(&n[R_S], 1u);
                uint32_t q = atomicAdd(&n[R_CAND], 1u);
                if (q >= cap.cand) {
                    n[R_OVER] = 1u;
                } else {
                    s.ckey[q] = total_key(lb);
                    s.cid[q] = k;
                }
            }
        } else if (grid == 1) {
            const CircleRec C = load_rec(A.cir, k);
            double yx = bx - C.mx, yy = by - C.my, yz = bz - C.mz;
            double ax = C.ax, ay = C.ay, az = C.az;
            double along = yx * ax + yy * ay + yz * az;
            double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
            double wl = sqrt(wx * wx + wy * wy + wz * wz);
            double rr = C.rho;
            double dist = sqrt(along * along + (wl - rr) * (wl - rr));
            if (xcap && dist <= reach_x) xadd(k | (1u << KIND_SHIFT));
            if (!(dist <= reach)) return;
            atomicAdd(&n[R_C], 1u);
            uint32_t q = atomicAdd(&n[R_CAND], 1u);
            if (q >= cap.cand) {
                n[R_OVER] = 1u;
                return;
            }
            s.ckey[q] = total_key(dist);
            s.cid[q] = k | (1u << KIND_SHIFT);
        } else {
            const VertexRec V = load_rec(A.ver, k);
            double d = ddist(bx, by, bz, V.x, V.y, V.z);
            if (xcap && d <= reach_x) xadd(k | (2u << KIND_SHIFT));
            if (!(d <= reach)) return;
            atomicAdd(&n[R_V], 1u);
            uint32_t q = atomicAdd(&n[R_CAND], 1u);
            if (q >= cap.cand) {
                n[R_OVER] = 1u;
                return;
            }
            s.ckey[q] = total_key(d);
            s.cid[q] = k | (2u << KIND_SHIFT);
        }
    };
    if (sl) {
        for (uint32_t i = tid; i < sn; i += nt) {
            const uint32_t e = sl[i];
            visit(e >> KIND_SHIFT, e & ID_MASK);
        }
        __syncthreads();
    } else {
        gather3(A, bx, by, bz, radii, tmp4, visit);
    }
    if (n[R_OVER]) return false;
    if (xcap) {  // the superset to its pool (none if it overflowed)
        __shared__ uint32_t x_at;
        const uint32_t m = n[R_X];
        if (tid == 0) {
            uint32_t at = NO_LIST;
            if (!n[R_XOVER]) {
                at = atomicAdd(xs.used, m);
                if (at > xs.cap || m > xs.cap - at) at = NO_LIST;
            }
            x_at = at;
            xs.off[xj] = at == NO_LIST ? 0u : at;
            xs.cnt[xj] = at == NO_LIST ? NO_LIST : m;
        }
        __syncthreads();
        if (x_at != NO_LIST)
            for (uint32_t i = tid; i < m; i += nt) xs.ids[x_at + i] = xstage[i];
        __syncthreads();
    }
    const uint32_t gathered = n[R_CAND];
    // Candidates one blocker hides from the whole ball leave the lists (their
    // key becomes all ones); the list order of the others is unchanged.
    constexpr unsigned long long HIDDEN_KEY = ~0ull;
    if (hide && hn != NO_LIST) {
        // Through the near hash, under keys with the top bit set (no sphere
        // id has it), while the table stays at most half full; a listed
        // candidate that is not entered is kept, which is always safe.
        const uint32_t used = n[R_NEAR], room = cap.hash / 2 > used ? cap.hash / 2 - used : 0u;
        for (uint32_t i = tid; i < min(hn, room); i += nt) {
            const uint32_t key = hl[i] | 0x80000000u;
            if (key != 0xffffffffu) hash_put(s.hk, s.hv, hmask, key, 0u);
        }
        __syncthreads();
        for (uint32_t q = tid; q < gathered; q += nt) {
            const uint32_t id = s.cid[q], key = id | 0x80000000u;
            if ((id >> KIND_SHIFT) <= 1u && key != 0xffffffffu && hash_get(s.hk, s.hv, hmask, key) != NO_LIST)
                s.ckey[q] = HIDDEN_KEY;
        }
    } else if (hide) {
        const double bc[3] = {bx, by, bz};
        mark_hidden(
            bc, rho + COVER_PAD, gathered, n[R_NEAR], [&](uint32_t q) { return s.cid[q] >> KIND_SHIFT; },
            [&](uint32_t q) { return s.cid[q] & ID_MASK; },
            [&](uint32_t kd, uint32_t k, dvec4& g0, dvec4& g1) {
                if (kd == 0u) {
                    g0 = ((gptr<const dvec4>)A.sph)[k];
                } else {
                    gptr<const dvec4> cr = (gptr<const dvec4>)A.cir + 2 * (size_t)k;
                    g0 = cr[0];
                    g1 = cr[1];
                }
            },
            [&](uint32_t slot) {
                const NearD k = s.nq[slot];
                return dvec4{k.x, k.y, k.z, k.blk};
            },
            [](double x, double y, double z) { return dvec4{x, y, z, 0.0}; },
            [&](const dvec4& p, uint32_t slot) {
                const NearD k = s.nq[slot];
                const double ex = k.x - p.x, ey = k.y - p.y, ez = k.z - p.z;
                return ex * ex + ey * ey + ez * ez < k.blk;
            },
            [&](uint32_t q) { s.ckey[q] = HIDDEN_KEY; });
    }
    __syncthreads();
    auto bucket_of = [&](unsigned long long key, uint32_t id) {
        return bucket(id >> KIND_SHIFT, sph ? key_value_any(key) : key_value(key));
    };
    for (uint32_t q = tid; q < gathered; q += nt)
        if (s.ckey[q] != HIDDEN_KEY) atomicAdd(&bstart[bucket_of(s.ckey[q], s.cid[q])], 1u);
    __syncthreads();
    // The CPU's list order, (class, key, index) by cand_less, sorted exactly
    // through the buckets: a candidate's bucket is monotone in (class, key),
    // so the candidates are first grouped by bucket (in any order within
    // one, into the free record area), then each is ranked against its own
    // bucket only.
    if (tid < WAVE) {
        const uint32_t c = bstart[tid];
        uint32_t inc = c;
        for (uint32_t o = 1; o < WAVE; o <<= 1) {
            uint32_t y = __shfl_up(inc, o, WAVE);
            if (tid >= o) inc += y;
        }
        bstart[tid] = inc - c;
        bfill[tid] = inc - c;
    }
    __syncthreads();
    const uint32_t ns = bstart[KEY_BUCKETS], nc = bstart[2 * KEY_BUCKETS] - ns, total = bstart[3 * KEY_BUCKETS];
    if (tid == 0) {
        n[R_S] = ns;
        n[R_C] = nc;
        n[R_CAND] = total;
    }
    unsigned long long* tkey = (unsigned long long*)s.cdat;
    uint32_t* tcid = (uint32_t*)(tkey + cap.cand);
    for (uint32_t q = tid; q < gathered; q += nt) {
        const unsigned long long key = s.ckey[q];
        if (key == HIDDEN_KEY) continue;
        const uint32_t id = s.cid[q];
        const uint32_t t = atomicAdd(&bfill[bucket_of(key, id)], 1u);
        tkey[t] = key;
        tcid[t] = id;
    }
    __syncthreads();
    for (uint32_t t = tid; t < total; t += nt) {
        const unsigned long long key = tkey[t];
        const uint32_t id = tcid[t], b = bucket_of(key, id);
        uint32_t rank = bstart[b];
        const uint32_t end = bstart[b + 1];
        for (uint32_t u = bstart[b]; u < end; u++) rank += cand_less(tkey[u], tcid[u], key, id) ? 1u : 0u;
        s.ckey[rank] = key;
        s.cid[rank] = id;
    }
    __syncthreads();
    // Candidate data, cleared bitsets and blocker row lengths (in cpre).
    const uint32_t cc = cap.cand, nw = cap.near / 64;
    for (uint32_t p = tid; p < total; p += nt) {
        uint32_t k = s.cid[p] & ID_MASK;
        double* d = s.cdat + (size_t)CDAT * p;
        if (p < ns) {
            const SphereRec S = load_rec(A.sph, k);
            d[0] = S.x;
            d[1] = S.y;
            d[2] = S.z;
            d[3] = S.R;
            s.coff[p] = A.nb_off[k];
            s.cpre[p] = sph ? 0u : A.nb_off[k + 1] - A.nb_off[k];
        } else if (p < ns + nc) {
            const CircleRec C = load_rec(A.cir, k);
            d[0] = C.mx;
            d[1] = C.my;
            d[2] = C.mz;
            d[3] = C.ax;
            d[4] = C.ay;
            d[5] = C.az;
            d[6] = C.rho;
            s.coff[p] = A.cb_off[k];
            s.cpre[p] = A.cb_off[k + 1] - A.cb_off[k];
        } else {
            const VertexRec V = load_rec(A.ver, k);
            d[0] = V.x;
            d[1] = V.y;
            d[2] = V.z;
        }
        for (uint32_t w = 0; w < nw; w++) s.cbits[w * cc + p] = 0ull;
    }
    __syncthreads();
    // Blocker bitsets of the S and C rows (cpre holds the row length).
    // With `hide`, a blocker that no projection of the ball comes near stays
    // out of the bitset: it could never block there. The test runs in f32 on
    // the records relative to the ball centre (lengths of a few A), its
    // limits widened by REACH_SLACK. Each row's hint (then in cpre) is the
    // kept blocker holding the row's central projection deepest (least
    // power distance |k - p0|^2 - t^2, when negative): projections of nearby
    // points are likely inside it too, so blocked() tries it first.
    //
    // The sampling of the same sub-block may have left its blocker record
    // (br, see BlkRecord): when its near spheres and kept candidates are
    // this list's, in the same order, its bitsets (a superset of what this
    // reach test keeps) and hints are taken instead, mapped to this list's
    // near-list slots through the hash.
    __shared__ uint32_t rec_head, rec_bad;
    if (tid == 0) {
        rec_head = br && hide && br[0] == n[R_NEAR] && br[1] == ns && br[2] == nc && br[3] == (n[R_NEAR] + 63) / 64;
        rec_bad = 0u;
    }
    __syncthreads();
    if (rec_head) {
        const uint32_t nn = br[0], nwr = br[3], ew = blk_entry_words(nwr);
        const uint32_t* ent = br + BLK_HEAD + nn;
        uint16_t* rslot = (uint16_t*)s.nt;  // the reach limits are not needed on this path
        for (uint32_t i = tid; i < nn; i += nt) {
            const uint32_t slot = hash_get(s.hk, s.hv, hmask, br[BLK_HEAD + i]);
            if (slot == 0xffffffffu)
                rec_bad = 1u;
            else
                rslot[i] = (uint16_t)slot;
        }
        for (uint32_t p = tid; p < ns + nc; p += nt)
            if (ent[(size_t)p * ew] != s.cid[p]) rec_bad = 1u;
        __syncthreads();
        if (!rec_bad) {
            for (uint32_t q = tid; q < (ns + nc) * nwr; q += nt) {
                const uint32_t p = q / nwr, w = q % nwr;
                const uint32_t* e = ent + (size_t)p * ew;
                unsigned long long bits = (unsigned long long)e[4 + 2 * w] | ((unsigned long long)e[5 + 2 * w] << 32);
                while (bits) {
                    const uint32_t sl = rslot[w * 64 + (uint32_t)__builtin_ctzll(bits)];
                    bits &= bits - 1ull;
                    atomicOr(&s.cbits[(sl >> 6) * cc + p], 1ull << (sl & 63));
                }
                if (w == 0) s.cpre[p] = e[1] == NO_HINT ? NO_HINT : (uint32_t)rslot[e[1]];
            }
            __syncthreads();
            return true;
        }
    }
    struct HintRow {
        ReachRow<float> w;
        float px, py, pz, best;
        uint32_t hint;
    };
    const float b0[3] = {0.0f, 0.0f, 0.0f};
    const float rbf = (float)(rho + COVER_PAD) + REACH_SLACK;
    blocker_bits(
        A.nb_items, A.cb_items, ns, ns + nc, s.coff, s.cpre, s.hk, s.hv, hmask, s.cbits, cc,
        [&](uint32_t p) {
            HintRow r{};
            r.best = 0.0f;
            r.hint = NO_HINT;
            if (!hide) return r;
            const double* d = s.cdat + (size_t)CDAT * p;
            const float ox = (float)(d[0] - bx), oy = (float)(d[1] - by), oz = (float)(d[2] - bz);
            r.w = p < ns ? cap_row(b0, rbf, ox, oy, oz, (float)d[3])
                         : arc_row(b0, rbf, ox, oy, oz, (float)d[3], (float)d[4], (float)d[5], (float)d[6]);
            // p0 = c + R v / |v| (v = b - c, or the in-plane w for a circle)
            const float vl = sqrtf(r.w.D2), sc = vl > 0.0f ? r.w.R / vl : 0.0f;
            r.px = r.w.ox + r.w.vx * sc;
            r.py = r.w.oy + r.w.vy * sc;
            r.pz = r.w.oz + r.w.vz * sc;
            return r;
        },
        [&](HintRow& r, uint32_t p, uint32_t slot) {
            if (!hide) return true;
            const NearD k = s.nq[slot];
            const float kx = (float)(k.x - bx), ky = (float)(k.y - by), kz = (float)(k.z - bz);
            const float t = s.nt[slot];
            const bool reach = p < ns ? cap_reaches(r.w, kx, ky, kz, t) : arc_reaches(r.w, kx, ky, kz, t);
            if (reach) {
                const float ex = kx - r.px, ey = ky - r.py, ez = kz - r.pz;
                const float pw = ex * ex + ey * ey + ez * ez - t * t;
                if (pw < r.best) {
                    r.best = pw;
                    r.hint = slot;
                }
            }
            return reach;
        },
        [&](HintRow& r, uint32_t p) {
            // the least (best, hint) of the row's 8 lanes (total order: power, then slot)
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
    return true;
}

struct EvalD {
    double value;
    uint32_t kind, id;
    double qx, qy, qz;
};

// True if (qx, qy, qz) lies inside a blocker of candidate p:
// !(ddist(q, c) >= R - EPS) for a near row member.
__device__ inline bool blocked(const RLists& s, uint32_t cc, uint32_t nw, uint32_t p, double qx, double qy,
                               double qz) {
    const uint32_t hint = s.cpre[p];  // one of the bitset's blockers, tried first
    if (hint != NO_HINT) {
        const NearD b = s.nq[hint];
        const double ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
        if (ex * ex + ey * ey + ez * ez < b.blk) return true;
    }
    for (uint32_t w = 0; w < nw; w++) {
        unsigned long long m = s.cbits[w * cc + p];
        while (m) {
            NearD b = s.nq[w * 64 + (uint32_t)__builtin_ctzll(m)];
            m &= m - 1ull;
            double ex = qx - b.x, ey = qy - b.y, ez = qz - b.z;
            if (ex * ex + ey * ey + ez * ez < b.blk) return true;
        }
    }
    return false;
}

// Takes (d, p) if it precedes (best, bo): less distance, or the same and
// earlier in list order.
__device__ inline bool precedes(double d, uint32_t p, double best, uint32_t bo) {
    return d < best || (d == best && p < bo);
}

// Distance from (x, y, z) of a class-K candidate (record d) as eval_local
// compares it (field/point.rs:419-489); +inf for a sphere through the point,
// and for a candidate whose squared distances already show it farther than
// `best` by more than 1e-9 A (the square roots are then skipped). That
// margin is far above the rounding of the squares and roots (about 1e-15
// A here), so the distance computed in full would exceed `best` too: the
// candidate could not be taken either way.
//
// Where the distance is finite, `pv` keeps what project() needs of it:
// class S the offset x - c and its length, class C the in-plane part w of
// x - m and |w|, computed with project()'s operations.
template <int K>
__device__ inline double cand_dist(const double* d, double x, double y, double z, double best, dvec4& pv) {
    if (K == 0) {
        double vx = x - d[0], vy = y - d[1], vz = z - d[2];
        const double s = vx * vx + vy * vy + vz * vz, R = d[3];
        const double hi = R + best + 1e-9, lo = R - best - 1e-9;
        if (s >= hi * hi || (lo > 0.0 && s <= lo * lo)) return INFINITY;
        double len = sqrt(s);
        pv = dvec4{vx, vy, vz, len};
        return len < 1e-12 ? INFINITY : fabs(len - R);
    }
    const double b2 = (best + 1e-9) * (best + 1e-9);
    if (K == 1) {
        double yx = x - d[0], yy = y - d[1], yz = z - d[2];
        double ax = d[3], ay = d[4], az = d[5], rho = d[6];
        double along = yx * ax + yy * ay + yz * az;
        if (along * along >= b2) return INFINITY;  // the sum below is at least along^2
        double wx = yx - ax * along, wy = yy - ay * along, wz = yz - az * along;
        double wl = sqrt(wx * wx + wy * wy + wz * wz);
        pv = dvec4{wx, wy, wz, wl};
        const double D = along * along + (wl - rho) * (wl - rho);
        return D >= b2 ? INFINITY : sqrt(D);
    }
    double dx = x - d[0], dy = y - d[1], dz = z - d[2];
    const double s = dx * dx + dy * dy + dz * dz;
    return s >= b2 ? INFINITY : sqrt(s);
}

// Projection of x onto candidate p (its probe position, point.rs:419-489).
__device__ inline void project(const DArr& A, const RLists& s, uint32_t ns, uint32_t nc, uint32_t p, double x,
                               double y, double z, double& qx, double& qy, double& qz) {
    const double* d = s.cdat + (size_t)CDAT * p;
    if (p < ns) {
        double cx = d[0], cy = d[1], cz = d[2], R = d[3];
        double vx = x - cx, vy = y - cy, vz = z - cz;
        double len = sqrt(vx * vx + vy * vy + vz * vz);
        double sc = R / len;
        qx = cx + vx * sc;
        qy = cy + vy * sc;
        qz = cz + vz * sc;
    } else if (p < ns + nc) {
        double mx = d[0], my = d[1], mz = d[2], ax = d[3], ay = d[4], az = d[5], rho = d[6];
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
            uint32_t c = s.cid[p] & ID_MASK;
            dx = A.e1x[c];
            dy = A.e1y[c];
            dz = A.e1z[c];
        }
        qx = mx + dx * rho;
        qy = my + dy * rho;
        qz = mz + dz * rho;
    } else {
        qx = d[0];
        qy = d[1];
        qz = d[2];
    }
}

// project() of candidate p of class K from cand_dist's `pv` at the same
// point: the same operations on the same values, so the same bits.
template <int K>
__device__ inline void project_kept(const DArr& A, const RLists& s, uint32_t p, const dvec4& pv, double& qx,
                                    double& qy, double& qz) {
    const double* d = s.cdat + (size_t)CDAT * p;
    if (K == 0) {
        double cx = d[0], cy = d[1], cz = d[2], R = d[3];
        double sc = R / pv.w;
        qx = cx + pv.x * sc;
        qy = cy + pv.y * sc;
        qz = cz + pv.z * sc;
    } else {
        double mx = d[0], my = d[1], mz = d[2], rho = d[6];
        double dx, dy, dz;
        if (pv.w > 1e-12) {
            dx = pv.x / pv.w;
            dy = pv.y / pv.w;
            dz = pv.z / pv.w;
        } else {
            uint32_t c = s.cid[p] & ID_MASK;
            dx = A.e1x[c];
            dy = A.e1y[c];
            dz = A.e1z[c];
        }
        qx = mx + dx * rho;
        qy = my + dy * rho;
        qz = mz + dz * rho;
    }
}

// Lane group of a workgroup: g lanes (a power of two, at most 64) starting
// at lane `shift` of the wavefront; r is the lane's rank in the group.
struct Group {
    uint32_t g, r, shift;
    unsigned long long mask;
};

__device__ inline bool group_any(const Group& G, bool p) { return ((__ballot(p) >> G.shift) & G.mask) != 0ull; }

// One step of group_least with a DPP lane pattern (a VALU move, no LDS
// round trip): each lane takes the least of its pair and the lane CTRL
// names, which lies in the same group.
template <int CTRL>
__device__ inline void least_dpp(double& best, uint32_t& bo) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(best);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
    const double od = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    const uint32_t oo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)bo, CTRL, 0xF, 0xF, false);
    if (precedes(od, oo, best, bo)) {
        best = od;
        bo = oo;
    }
}

// The least (best, bo) over the group's lanes, in every lane of the group
// (every lane of the group must be active). precedes() is a total order on
// (distance, list position), so the least does not depend on how lanes are
// paired: within 16 lanes by DPP quad permutes (lane ^ 1, lane ^ 2) and the
// half-row and row mirrors (lane i with 7 - i, then 15 - i, in its 8 and 16
// lanes), beyond by shuffles.
__device__ inline void group_least(const Group& G, double& best, uint32_t& bo) {
    if (G.g >= 2) least_dpp<0xB1>(best, bo);   // quad_perm [1, 0, 3, 2]
    if (G.g >= 4) least_dpp<0x4E>(best, bo);   // quad_perm [2, 3, 0, 1]
    if (G.g >= 8) least_dpp<0x141>(best, bo);  // row_half_mirror
    if (G.g >= 16) least_dpp<0x140>(best, bo); // row_mirror
    for (uint32_t o = 16; o < G.g; o <<= 1) {
        double od = __shfl_xor(best, (int)o, WAVE);
        uint32_t oo = __shfl_xor(bo, (int)o, WAVE);
        if (precedes(od, oo, best, bo)) {
            best = od;
            bo = oo;
        }
    }
}

// One class's candidates [a, e) of this lane (a + r, a + r + g, ...),
// W distances at a time, then taken in list order: a candidate is checked
// against its blockers only if it precedes the best so far. The scan stops
// at the first step whose lower bound key - delta exceeds best. The calls
// use W = 1: on gfx950 four wavefronts per SIMD (at most 128 VGPRs) hide
// the LDS latency better than wider batches do.
template <int K, int W>
__device__ inline void scan_class(const DArr& A, const RLists& s, uint32_t cc, uint32_t nw, uint32_t ns, uint32_t nc,
                                  uint32_t a, uint32_t e, const Group& G, double x, double y, double z,
                                  double delta, double cutoff, double& best, uint32_t& bo) {
    for (uint32_t p0 = a + G.r; p0 < e; p0 += W * G.g) {
        double lim = best + delta;
        double dist[W];
        dvec4 pv[W];
        bool stop = false;
        for (int j = 0; j < W; j++) {
            uint32_t p = p0 + j * G.g;
            bool use = p < e && !stop && !(key_value(s.ckey[p]) > lim);
            stop |= !use;
            dist[j] = use ? cand_dist<K>(s.cdat + (size_t)CDAT * p, x, y, z, best, pv[j]) : INFINITY;
        }
        for (int j = 0; j < W; j++) {
            uint32_t p = p0 + j * G.g;
            if (!(dist[j] < cutoff) || !precedes(dist[j], p, best, bo)) continue;
            if (K != 2) {
                double qx, qy, qz;
                project_kept<K>(A, s, p, pv[j], qx, qy, qz);
                if (blocked(s, cc, nw, p, qx, qy, qz)) continue;
            }
            best = dist[j];
            bo = p;
        }
        if (stop) break;
    }
}

// The scans' bound on the point's offset o from the gathered ball's centre:
// a scan stops at the first candidate whose key (its value at the centre)
// exceeds best + bound, which is sound for every bound of at least
// |o| + 1e-7 (the values are 1-Lipschitz; 1e-7 A covers the rounding). It is
// taken in f32 with the hardware square root (within 1 ulp) and scaled by
// 1 + 1.0e-5, far above the f32 roundings of the square, the root and the
// product (together below 2.2e-7), then 1.01e-7 is added in f64: the result
// is never below the f64 sqrt(|o|^2) + 1e-7, so a scan stops no earlier
// than with that bound and takes the same candidate. Where |o|^2 is below
// the f32 normal range, |o| < 1.1e-19 and the 1.01e-7 alone covers it.
__device__ inline double offset_bound(double ox, double oy, double oz) {
    const float s = (float)(ox * ox + oy * oy + oz * oz);
    return (double)(__builtin_amdgcn_sqrtf(s) * 1.00001f) + 1.01e-7;
}

// eval_local of the SAS and vdW kinds (field/point.rs:723-736): the first
// sphere, in list order, of the least |x - c| - R below band. Each lane
// scans its share in list order and stops where the key (the value at the
// ball centre) minus the point's offset exceeds its best: the values are
// 1-Lipschitz, so no later sphere is below that best. The record keeps the
// sphere's centre in q, from which the gradient and the normal are taken.
__device__ inline EvalD eval_spheres(const RLists& s, const uint32_t* n, const Group& G, double cx, double cy,
                                     double cz, double x, double y, double z, double band) {
    const uint32_t total = n[R_CAND];
    const double delta = offset_bound(x - cx, y - cy, z - cz);
    double best = band;
    uint32_t bo = 0xffffffffu;
    for (uint32_t p = G.r; p < total; p += G.g) {
        if (key_value_any(s.ckey[p]) > best + delta) break;
        const double* d = s.cdat + (size_t)CDAT * p;
        const double v = ddist(x, y, z, d[0], d[1], d[2]) - d[3];
        if (v < best) {
            best = v;
            bo = p;
        }
    }
    group_least(G, best, bo);
    EvalD e;
    e.value = best;
    if (bo == 0xffffffffu) {
        e.kind = F_BEYOND;
        e.id = 0;
        e.qx = x;
        e.qy = y;
        e.qz = z;
    } else {
        const double* d = s.cdat + (size_t)CDAT * bo;
        e.kind = F_SPHERE;
        e.id = s.cid[bo] & ID_MASK;
        e.qx = d[0];
        e.qy = d[1];
        e.qz = d[2];
    }
    return e;
}

// eval_local (field/point.rs:686-722) at (x, y, z), split over a lane group.
// (cx, cy, cz) is the centre of the gathered ball. The group shares its
// best after each class, which only tightens later tests. The SAS and vdW
// kinds take eval_spheres with `band`.
__device__ inline EvalD eval_group(const DArr& A, const RLists& s, const uint32_t* n, uint32_t cc, uint32_t nw,
                                   const Group& G, double cx, double cy, double cz, double x, double y, double z,
                                   double cutoff, double band) {
    if (A.kind != K_SES) return eval_spheres(s, n, G, cx, cy, cz, x, y, z, band);
    EvalD e;
    e.qx = x;
    e.qy = y;
    e.qz = z;
    bool ins = false;
    const uint32_t n_in = n[R_IN];
    for (uint32_t q = G.r; q < n_in; q += G.g) {
        uint32_t k = s.ins[q];
        NearD b = s.nq[k];
        double dx = x - b.x, dy = y - b.y, dz = z - b.z;
        ins |= dx * dx + dy * dy + dz * dz < s.nin[k];
    }
    if (!group_any(G, ins)) {
        e.value = A.probe;
        e.kind = F_SOLVENT;
        e.id = 0;
        return e;
    }
    const uint32_t ns = n[R_S], nc = n[R_C], total = n[R_CAND];
    const double delta = offset_bound(x - cx, y - cy, z - cz);
    double best = cutoff;
    uint32_t bo = 0xffffffffu;
    scan_class<0, 1>(A, s, cc, nw, ns, nc, 0, ns, G, x, y, z, delta, cutoff, best, bo);
    group_least(G, best, bo);
    scan_class<1, 1>(A, s, cc, nw, ns, nc, ns, ns + nc, G, x, y, z, delta, cutoff, best, bo);
    group_least(G, best, bo);
    scan_class<2, 1>(A, s, cc, nw, ns, nc, ns + nc, total, G, x, y, z, delta, cutoff, best, bo);
    group_least(G, best, bo);
    if (bo == 0xffffffffu) {  // nothing free below the cutoff: Beyond
        e.value = A.probe - cutoff;
        e.kind = F_BEYOND;
        e.id = 0;
        return e;
    }
    project(A, s, ns, nc, bo, x, y, z, e.qx, e.qy, e.qz);
    e.value = A.probe - best;
    e.kind = bo < ns ? F_SPHERE : (bo < ns + nc ? F_CIRCLE : F_VERTEX);
    e.id = s.cid[bo] & ID_MASK;
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
    size_t cap;  // vertices the arrays hold; slots beyond are not written
};

// State of one edge's root search (mesh/extract.rs:425-535): up to 8
// Newton / Illinois steps, then up to 64 bisections. The end points and
// direction are recomputed from the lattice point when needed, with the
// same operations, so they keep their bits.
struct EdgeState {
    int32_t ia[3];  // lower lattice point
    uint32_t id;    // edge id in its brick; axis = id % 3
    float fa;       // sample value at the lower point
    uint32_t st;    // it (bits 0-6), side + 1 (7-8), bisect (9), neg0 (10), converged (11), evals (16-31)
    double t, t0, g0, t1, g1;
    EvalD last;
};

struct EdgeGeom {
    double pa[3], d[3], dn;
};

__device__ inline EdgeGeom edge_geom(const EdgeState& E, double h) {
    EdgeGeom G;
    uint32_t axis = E.id % 3;
    for (int a = 0; a < 3; a++) {
        int64_t ib = (int64_t)E.ia[a] + (a == (int)axis ? 1 : 0);
        G.pa[a] = (double)E.ia[a] * h;
        G.d[a] = (double)ib * h - G.pa[a];
    }
    G.dn = sqrt(G.d[0] * G.d[0] + G.d[1] * G.d[1] + G.d[2] * G.d[2]);
    return G;
}

__device__ inline void edge_start(const DRun& P, const float* val, const int64_t org[3], uint32_t id, EdgeState& E) {
    const uint32_t b = P.b;
    uint32_t axis = id % 3, cell = id / 3;
    uint32_t x = cell % b, y = (cell / b) % b, z = cell / (b * b);
    uint32_t i0 = x + P.side * (y + P.side * z);
    uint32_t stp = axis == 0 ? 1u : (axis == 1 ? P.side : P.side * P.side);
    E.ia[0] = (int32_t)(org[0] + x);
    E.ia[1] = (int32_t)(org[1] + y);
    E.ia[2] = (int32_t)(org[2] + z);
    E.id = id;
    float fa32 = val[i0], fb32 = val[i0 + stp];
    E.fa = fa32;
    double fa = (double)fa32, fb = (double)fb32;
    E.t = fmin(fmax(fa / (fa - fb), 0.0), 1.0);
    E.t0 = 0.0;
    E.g0 = fa;
    E.t1 = 1.0;
    E.g1 = fb;
    E.st = 1u << 7;  // side 0
}

// Applies evaluation e at E.t; true when the search is over. The gradient
// (mesh/extract.rs:539-549) points from x to the closest point for the SES
// kind, and from the sphere's centre (held in e.q) to x for the others.
__device__ inline bool edge_step(EdgeState& E, const EvalD& e, double h, bool sph) {
    E.st += 1u << 16;  // evals
    E.last = e;
    double g = e.value;
    uint32_t it = E.st & 127u;
    if (!(E.st & (1u << 9))) {
        if (fabs(g) <= 1e-5) {
            E.st |= 1u << 11;  // converged
            return true;
        }
        int side = (int)((E.st >> 7) & 3u) - 1;
        if ((g < 0.0) == (E.g0 < 0.0)) {
            E.t0 = E.t;
            E.g0 = g;
            if (side == -1) E.g1 *= 0.5;
            side = -1;
        } else {
            E.t1 = E.t;
            E.g1 = g;
            if (side == 1) E.g0 *= 0.5;
            side = 1;
        }
        EdgeGeom Q = edge_geom(E, h);
        double slope = 0.0;
        if (e.kind != F_SOLVENT && e.kind != F_BEYOND) {
            double x = Q.pa[0] + Q.d[0] * E.t, y = Q.pa[1] + Q.d[1] * E.t, z = Q.pa[2] + Q.d[2] * E.t;
            double wx = e.qx - x, wy = e.qy - y, wz = e.qz - z;
            if (sph) {
                wx = x - e.qx;
                wy = y - e.qy;
                wz = z - e.qz;
            }
            double len = sqrt(wx * wx + wy * wy + wz * wz);
            if (len > 1e-12) {
                // n . d with d zero off the edge's axis: the other two
                // products are +-0 and leave the sum's bits, except the sign
                // of a zero slope, which never takes the Newton step below.
                const uint32_t axis = E.id % 3;
                const double wa = axis == 0 ? wx : (axis == 1 ? wy : wz);
                const double da = axis == 0 ? Q.d[0] : (axis == 1 ? Q.d[1] : Q.d[2]);
                slope = (wa / len) * da;
            }
        }
        double newton = E.t - g / slope;
        if (fabs(slope) > 1e-3 * Q.dn && newton > E.t0 && newton < E.t1) {
            E.t = newton;
        } else {
            double sv = (E.t0 * E.g1 - E.t1 * E.g0) / (E.g1 - E.g0);
            E.t = (isfinite(sv) && sv > E.t0 && sv < E.t1) ? sv : 0.5 * (E.t0 + E.t1);
        }
        it++;
        uint32_t st = (E.st & ~(127u | (3u << 7))) | ((uint32_t)(side + 1) << 7);
        if (it == 8) {
            st |= 1u << 9;                      // bisect
            if (E.g0 < 0.0) st |= 1u << 10;     // neg0
            E.t = 0.5 * (E.t0 + E.t1);
        } else {
            st |= it;
        }
        E.st = st;
        return false;
    }
    double dn = edge_geom(E, h).dn;
    if (fabs(g) <= 1e-5 || (E.t1 - E.t0) * dn <= 2.0 * 1e-5) return true;
    if ((g < 0.0) == ((E.st & (1u << 10)) != 0))
        E.t0 = E.t;
    else
        E.t1 = E.t;
    if (++it == 64) return true;
    E.st = (E.st & ~127u) | it;
    E.t = 0.5 * (E.t0 + E.t1);
    return false;
}

// The vertex and its attributes (mesh/extract.rs:535, 553-597) into `slot`;
// the group's rank-0 lane writes them.
__device__ inline void edge_finish(const DArr& A, const DRun& P, const EdgeState& E, const Group& G, size_t slot,
                                   const VOut& o) {
    if (G.r != 0 || slot >= o.cap) return;
    const double t = E.t;
    const EvalD& last = E.last;
    EdgeGeom Q = edge_geom(E, P.h);
    double tc = fmin(fmax(t, P.t_min), 1.0 - P.t_min);
    double v[3] = {Q.pa[0] + Q.d[0] * tc, Q.pa[1] + Q.d[1] * tc, Q.pa[2] + Q.d[2] * tc};
    double nrm[3];
    uint8_t patch = 0;
    uint32_t atom = 0;
    // The edge's outward direction, the normal where the closest point
    // gives none (computed only there).
    auto outward = [&]() {
        double ow[3];
        for (int a = 0; a < 3; a++) ow[a] = (double)E.fa < 0.0 ? Q.d[a] : -Q.d[a];
        double ol = sqrt(ow[0] * ow[0] + ow[1] * ow[1] + ow[2] * ow[2]);
        for (int a = 0; a < 3; a++) nrm[a] = ow[a] / ol;
    };
    auto unit = [&](double wx, double wy, double wz) {
        double len = sqrt(wx * wx + wy * wy + wz * wz);
        if (len > 1e-9) {
            nrm[0] = wx / len;
            nrm[1] = wy / len;
            nrm[2] = wz / len;
        } else {
            outward();
        }
    };
    auto nearest = [&](const uint32_t* atoms, int cnt) {
        double bd = INFINITY;
        uint32_t bk = atoms[0];
        for (int q = 0; q < cnt; q++) {
            uint32_t k = atoms[q];
            const SphereRec S = load_rec(A.sph, k);  // the centre's x, y, z in one load
            double dd = ddist(v[0], v[1], v[2], S.x, S.y, S.z) - A.r[k];
            if (dd < bd) {
                bd = dd;
                bk = k;
            }
        }
        return A.orig[bk];
    };
    if (last.kind == F_SPHERE && A.kind != K_SES) {  // q: the sphere's centre
        unit(v[0] - last.qx, v[1] - last.qy, v[2] - last.qz);
        patch = 0;
        atom = A.orig[last.id];
    } else if (last.kind == F_SPHERE) {
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
        // first minimum in scan order (rare).
        outward();
        patch = 0;
        double cutoff = A.probe + P.band + 8.0;
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
    if (o.evals) o.evals[slot] = (E.st >> 16) | ((E.st & (1u << 11)) ? 0x80000000u : 0u);
}

// One workgroup (256 threads) per sub-block with crossing edges (j =
// 8*brick + sb): gather_local into LDS (GLOBAL false) or into scratch slice
// blockIdx.x (GLOBAL true), compact the crossing edges whose lower sample
// lies in the sub-block, then refine them by lane groups. Sub-blocks whose
// lists overflow go to `overflow`.
template <bool GLOBAL>
__device__ __forceinline__ void refine_sub(const Mol* __restrict__ mols, const uint32_t* bmol, const RCaps& cap,
                                           const RLayout& lay, const int32_t* keys, const float* values, uint32_t j,
                                           const unsigned long long* mask, const uint32_t* wpre, uint32_t words,
                                           const uint32_t* vbase, const VOut& out, char* scratch, size_t slice,
                                           uint32_t* overflow, uint32_t* n_overflow,
                                           unsigned long long* evals_total, const HidList& hid) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ uint32_t n[R_N];
    __shared__ uint32_t tmp4[WIDE / WAVE];  // block scans: one slot per wavefront
    __shared__ unsigned long long evals_sum;
    uint32_t brick = j >> 3, sb = j & 7;
    const uint32_t mol = bmol[brick];
    const DArr& A = mols[mol].A;
    const DRun& P = mols[mol].P;
    char* base = GLOBAL ? scratch + (size_t)blockIdx.x * slice : smem;
    RLists s = rcarve(base, lay);
    const int32_t* key = keys + 3 * brick;
    const uint32_t b = P.b, sub = P.sub;
    double h = P.h, subd = (double)sub;
    if (threadIdx.x == 0) evals_sum = 0;
    // Ball of the sub-block (mesh/extract.rs:164-205).
    double rho = uniform_f64(0.5 * sqrt(3.0) * subd * h * (1.0 + 1e-9) + 1e-9);
    uint32_t o3[3] = {sb & 1, (sb >> 1) & 1, sb >> 2};
    int64_t org[3];
    double c[3];
    for (int a = 0; a < 3; a++) {
        org[a] = (int64_t)key[a] * (int64_t)b;
        double basep = (double)org[a] * h;
        c[a] = uniform_f64(basep + ((double)(o3[a] * sub) + 0.5 * subd) * h);
    }
    const uint32_t* hl = nullptr;  // the sampling's hidden candidates of this sub-block
    uint32_t hn = hid.cnt ? hid.cnt[j] : NO_LIST;
    if (hn != NO_LIST) hl = hid.ids + hid.off[j];
    const uint32_t* sl = nullptr;  // and its superset of this gather's items
    uint32_t sn = hid.sup.cnt ? hid.sup.cnt[j] : NO_LIST;
    if (sn != NO_LIST) sl = hid.sup.ids + hid.sup.off[j];
    const uint32_t* br = nullptr;  // and its blocker record
    if (hid.blk.cnt && hid.blk.cnt[j] != NO_LIST) br = hid.blk.ids + (size_t)hid.blk.slot[j] * hid.blk.words;
    bool ok = gather_refine(A, cap, s, n, tmp4, c[0], c[1], c[2], rho, P.band, true, hl, hn, sl, sn, SupList{}, 0,
                            0.0, br);
    // Crossing edges whose lower sample lies in this sub-block, ascending.
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
    const uint32_t ne = n[R_E];
    // Lane groups: the largest g <= gmax with at least `ne` groups.
    uint32_t g = cap.gmax ? cap.gmax : (uint32_t)WAVE;
    while (g > 1 && blockDim.x / g < ne) g >>= 1;
    Group G;
    G.g = g;
    G.r = threadIdx.x % g;
    G.shift = (threadIdx.x % WAVE) - G.r;
    G.mask = g == WAVE ? ~0ull : ((1ull << g) - 1ull);
    const uint32_t ngroups = blockDim.x / g, nw = cap.near / 64;
    const double cutoff = A.probe + P.band;
    const bool sph = A.kind != K_SES;
    uint32_t q = threadIdx.x / g;
    bool active = q < ne;
    EdgeState E;
    if (active) edge_start(P, val, org, s.ebuf[q], E);
    unsigned long long my_evals = 0;
    while (__ballot(active) != 0ull) {
        if (active) {
            double x, y, z;
            {
                EdgeGeom Q = edge_geom(E, h);
                x = Q.pa[0] + Q.d[0] * E.t;
                y = Q.pa[1] + Q.d[1] * E.t;
                z = Q.pa[2] + Q.d[2] * E.t;
            }
            // The lane's rank, opaque to the compiler here: the list
            // addresses derived from it are recomputed in each evaluation
            // (two ALU operations) instead of being hoisted out of the loop
            // and spilled for want of registers.
            asm volatile("" : "+v"(G.r));
            EvalD e = eval_group(A, s, n, cap.cand, nw, G, c[0], c[1], c[2], x, y, z, cutoff, P.band);
            if (edge_step(E, e, h, sph)) {
                size_t slot = (size_t)vbase[brick] + edge_rank(mask, wpre, words, brick, E.id);
                edge_finish(A, P, E, G, slot, out);
                my_evals += E.st >> 16;
                q += ngroups;
                active = q < ne;
                if (active) edge_start(P, val, org, s.ebuf[q], E);
            }
        }
    }
    if (evals_total) {
        if (G.r == 0 && my_evals) atomicAdd(&evals_sum, my_evals);
        __syncthreads();
        if (threadIdx.x == 0 && evals_sum) atomicAdd(&evals_total[mol], evals_sum);
    }
}

// PERSIST: residency-sized grid looping over the queue; otherwise one
// sub-block per workgroup. Five wavefronts per SIMD (at most 96 VGPRs, a
// few spilled) beat four without spills: the candidate and blocker loops
// wait on LDS and want wavefronts to hide it.
template <bool GLOBAL, bool PERSIST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_refine(const Mol* __restrict__ mols, const uint32_t* bmol, RCaps cap,
                                                RLayout lay, const int32_t* keys, const float* values, Queue q,
                                                HidList hid, const unsigned long long* mask, const uint32_t* wpre,
                                                uint32_t words, const uint32_t* vbase, VOut out, char* scratch,
                                                size_t slice, uint32_t* overflow, uint32_t* n_overflow,
                                                unsigned long long* evals_total) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t j) {
        refine_sub<GLOBAL>(mols, bmol, cap, lay, keys, values, j, mask, wpre, words, vbase, out, scratch, slice,
                           overflow, n_overflow, evals_total, hid);
    });
}

// k_refine with WIDE threads, for the large tiers.
template <bool GLOBAL, bool PERSIST>
__global__ void __launch_bounds__(WIDE) k_refine_wide(const Mol* __restrict__ mols, const uint32_t* bmol, RCaps cap,
                                                RLayout lay, const int32_t* keys, const float* values, Queue q,
                                                HidList hid, const unsigned long long* mask, const uint32_t* wpre,
                                                uint32_t words, const uint32_t* vbase, VOut out, char* scratch,
                                                size_t slice, uint32_t* overflow, uint32_t* n_overflow,
                                                unsigned long long* evals_total) {
    __shared__ uint32_t ticket;
    for_each_item<PERSIST>(q, &ticket, [&](uint32_t j) {
        refine_sub<GLOBAL>(mols, bmol, cap, lay, keys, values, j, mask, wpre, words, vbase, out, scratch, slice,
                           overflow, n_overflow, evals_total, hid);
    });
}

}  // namespace smg
