// The SAS arrangement built on the device, f64, with the reference's
// decisions and orders (crates/sesmesh/src/geom/arrangement.rs:437-763,
// geom/cells.rs:27-77, field/point.rs:213-274):
//
// * spheres in Morton order of their cells (ties: input index) by a stable
//   radix sort of (key, index);
// * cell grids filed by a stable radix sort of (cell, index), so a cell
//   lists its points in index order;
// * containment and neighbour rows, one thread per sphere, rows sorted;
// * circles (i, j > i) with their blockers and triple points, one thread per
//   pair: a count pass, scans, then a fill pass that writes every record at
//   the position the CPU's sphere-ordered concatenation gives it.
//
// The exposed-arc test compares angles from atan2. The device's atan2 is
// not the host C library's to the last bit, so a gap whose angle lies within
// GAP_MARGIN (far above either's error) of the test's threshold, or of a
// wrap of the angle, is not decided here: the pair is flagged and the caller
// builds that arrangement on the host instead.
#pragma once
#include "dev_common.hpp"

namespace smg {

namespace arrb {
constexpr double ANG_EPS = 1e-12;     // geom/arrangement.rs:29
constexpr double PERIOD = 4.0;        // geom/arrangement.rs:275 (diamond angle period)
constexpr double PSEUDO_EPS = 1e-12;  // geom/arrangement.rs:380-392
constexpr double TAU = 6.283185307179586;
constexpr double GAP_MARGIN = 1e-13;
constexpr uint32_t MAX_IV = 256;  // buried intervals one circle may hold here (more: the host builds it)
enum : uint32_t { F_UNCERTAIN = 1u, F_OVERFLOW = 2u };
}  // namespace arrb

// The arrangement under construction (device pointers).
struct ABuild {
    const double *cx, *cy, *cz, *R;
    const uint32_t* orig;
    const uint8_t* contained;
    const uint32_t *nb_off, *nb_items;  // full neighbour rows
    uint32_t n;
    double max_expanded;
};

__device__ inline uint64_t spread21_dev(uint32_t v) {
    uint64_t x = (uint64_t)v & 0x1fffffull;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Morton key of each active sphere's cell (arrangement.rs:478-489): per
// axis `(p - lo) / cell` truncated, saturated at 2^21 - 1 (NaN and
// negatives give 0).
__global__ void k_arr_morton(const double* xyzr, const uint32_t* active, uint32_t n, double lx, double ly, double lz,
                             double cell, uint64_t* keys, uint32_t* vals) {
    const double lo[3] = {lx, ly, lz};
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = active[k];
        uint64_t m = 0;
        for (int a = 0; a < 3; a++) {
            const double v = (xyzr[4 * (size_t)i + a] - lo[a]) / cell;
            const uint32_t t = !(v > 0.0) ? 0u : (v >= 2097151.0 ? 0x1fffffu : (uint32_t)v);
            m |= spread21_dev(t) << a;
        }
        keys[k] = m;
        vals[k] = i;
    }
}

// The spheres in their sorted order: centres (SoA and interleaved), radii,
// expanded radii and the packed records.
__global__ void k_arr_gather(const double* xyzr, const uint32_t* orig, uint32_t n, double probe, double* cx,
                             double* cy, double* cz, double* c3, double* r, double* R, SphereRec* rec) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const double* s = xyzr + 4 * (size_t)orig[k];
        const double x = s[0], y = s[1], z = s[2], rr = s[3], ex = rr + probe;
        cx[k] = x;
        cy[k] = y;
        cz[k] = z;
        c3[3 * (size_t)k] = x;
        c3[3 * (size_t)k + 1] = y;
        c3[3 * (size_t)k + 2] = z;
        r[k] = rr;
        R[k] = ex;
        rec[k] = SphereRec{x, y, z, ex};
    }
}

// Cell of each point (cells.rs:62-77) and the count per cell.
__global__ void k_arr_cells(const double* p3, uint32_t n, DGrid g, uint32_t* key, uint32_t* idx, uint32_t* count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t x = gcoord(p3[3 * (size_t)i], g.ox, g.cell, g.dx);
        const uint32_t y = gcoord(p3[3 * (size_t)i + 1], g.oy, g.cell, g.dy);
        const uint32_t z = gcoord(p3[3 * (size_t)i + 2], g.oz, g.cell, g.dz);
        const uint32_t c = (z * g.dy + y) * g.dx + x;
        key[i] = c;
        idx[i] = i;
        atomicAdd(&count[c], 1u);
    }
}

// Containment (arrangement.rs:499-516): i is contained if another sphere's
// expanded ball holds its expanded ball; of exact duplicates the lower input
// index survives. One thread per sphere.
__global__ void k_arr_contain(ABuild B, DGrid gs, uint8_t* contained) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.n; i += gridDim.x * blockDim.x) {
        const double xi = B.cx[i], yi = B.cy[i], zi = B.cz[i], Ri = B.R[i];
        bool found = false;
        for_near(gs, xi, yi, zi, Ri + B.max_expanded, 0, 1, [&](uint32_t j) {
            if (j == i || found) return;
            const double d = ddist(xi, yi, zi, B.cx[j], B.cy[j], B.cz[j]);
            const double Rj = B.R[j];
            if (d + Ri <= Rj + EPS && (Rj - Ri > EPS || d > EPS || B.orig[j] < B.orig[i])) found = true;
        });
        contained[i] = found ? 1 : 0;
    }
}

// Neighbour rows (arrangement.rs:519-550): the spheres, not contained,
// whose expanded balls overlap by more than EPS. FILL false counts them;
// FILL true writes and sorts each row.
template <bool FILL>
__global__ void k_arr_neighbours(ABuild B, DGrid gs, uint32_t* count, const uint32_t* off, uint32_t* items) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.n; i += gridDim.x * blockDim.x) {
        uint32_t c = 0;
        if (!B.contained[i]) {
            const double xi = B.cx[i], yi = B.cy[i], zi = B.cz[i], Ri = B.R[i];
            uint32_t* row = FILL ? items + off[i] : nullptr;
            for_near(gs, xi, yi, zi, Ri + B.max_expanded, 0, 1, [&](uint32_t j) {
                if (j == i || B.contained[j]) return;
                if (ddist(xi, yi, zi, B.cx[j], B.cy[j], B.cz[j]) < Ri + B.R[j] - EPS) {
                    if (FILL) row[c] = j;
                    c++;
                }
            });
            if (FILL)
                for (uint32_t a = 1; a < c; a++) {
                    const uint32_t v = row[a];
                    uint32_t b = a;
                    while (b > 0 && row[b - 1] > v) {
                        row[b] = row[b - 1];
                        b--;
                    }
                    row[b] = v;
                }
        }
        if (!FILL) count[i] = c;
    }
}

// Circle pairs: per sphere, the row members j > i (the row is sorted).
template <bool FILL>
__global__ void k_arr_pairs(ABuild B, uint32_t* count, const uint32_t* off, uint32_t* pi, uint32_t* pj) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.n; i += gridDim.x * blockDim.x) {
        const uint32_t a = B.nb_off[i], e = B.nb_off[i + 1];
        uint32_t q = a;
        while (q < e && B.nb_items[q] <= i) q++;
        if (!FILL) {
            count[i] = e - q;
            continue;
        }
        uint32_t at = off[i];
        for (; q < e; q++, at++) {
            pi[at] = i;
            pj[at] = B.nb_items[q];
        }
    }
}

struct D3v {
    double x, y, z;
};
__device__ inline D3v dsub(D3v a, D3v b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3v dadd(D3v a, D3v b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3v dmul(D3v a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline D3v ddiv(D3v a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double ddot(D3v a, D3v b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3v dcross(D3v a, D3v b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline double dnorm(D3v a) { return sqrt(ddot(a, a)); }

// A circle of intersection: centre, axis, in-plane basis, radius.
struct CircleG {
    D3v cc, axis, e1, e2;
    double radius;
};

// Circle::cut (arrangement.rs:251-272): 0 outside, 1 the whole circle
// buried, 2 an arc (enter, leave as (cos, sin) in the e1, e2 frame).
__device__ inline int cut_dev(const CircleG& C, D3v center, double expanded, double enter[2], double leave[2]) {
    const D3v g = dsub(C.cc, center);
    const D3v e2 = dcross(C.axis, C.e1);
    const double a = 2.0 * C.radius * ddot(C.e1, g);
    const double b = 2.0 * C.radius * ddot(e2, g);
    const double rhs = expanded * expanded - ddot(g, g) - C.radius * C.radius;
    const double amp = sqrt(a * a + b * b);
    if (amp <= rhs) return 1;
    if (amp <= -rhs || amp == 0.0) return 0;
    const double cp = a / amp, sp = b / amp, ca = rhs / amp;
    const double sa = sqrt(fmax(1.0 - ca * ca, 0.0));
    enter[0] = cp * ca - sp * sa;
    enter[1] = sp * ca + cp * sa;
    leave[0] = cp * ca + sp * sa;
    leave[1] = sp * ca - cp * sa;
    return 2;
}

// The pseudo-angle of a unit vector in [0, 4) (arrangement.rs:282-295).
__device__ inline double diamond_dev(const double v[2]) {
    const double x = v[0], y = v[1];
    if (y >= 0.0) return x >= 0.0 ? y / (x + y) : 1.0 - x / (y - x);
    return x < 0.0 ? 2.0 - y / (-x - y) : 3.0 + x / (x - y);
}

__device__ inline double rem_tau_dev(double a) {
    const double r = fmod(a, arrb::TAU);
    return r < 0.0 ? r + arrb::TAU : r;
}

__device__ inline D3v centre_of(const ABuild& B, uint32_t k) { return {B.cx[k], B.cy[k], B.cz[k]}; }

// Whether an arc of the circle stays exposed (arrangement.rs:340-410), from
// its buried intervals sorted by (start, end descending, sphere): 1 yes,
// 0 no, 2 undecided (an angle gap within GAP_MARGIN of the threshold or of
// a wrap; the host decides those). One lane. The intervals (st, en, and
// first == last their sphere) are merged into runs in place, run r written
// after interval r is read; a run keeps its first and last sphere, whose
// cuts give its enter and leave points.
__device__ inline int exposed_arc_runs(const ABuild& B, const CircleG& C, double* st, double* en, uint32_t* fi,
                                       uint32_t* la, uint32_t m) {
    if (m == 0) return 1;
    uint32_t nr = 0;
    for (uint32_t q = 0; q < m; q++) {
        const double bs = st[q], be = en[q];
        const uint32_t bf = fi[q], bl = la[q];
        if (nr && bs <= en[nr - 1]) {
            if (be > en[nr - 1]) {
                en[nr - 1] = be;
                la[nr - 1] = bl;
            }
        } else {
            st[nr] = bs;
            en[nr] = be;
            fi[nr] = bf;
            la[nr] = bl;
            nr++;
        }
    }
    uint32_t head = 0;
    while (nr - head > 1) {
        const double fs = st[head], fe = en[head];
        const uint32_t fl = la[head], l = nr - 1;
        if (en[l] - arrb::PERIOD < fs) break;
        if (fe + arrb::PERIOD > en[l]) {
            en[l] = fe + arrb::PERIOD;
            la[l] = fl;
        }
        head++;
    }
    const uint32_t m2 = nr - head;
    bool undecided = false;
    for (uint32_t k = 0; k < m2; k++) {
        const uint32_t c = head + k, x = head + (k + 1) % m2;
        const double ns = k + 1 < m2 ? st[x] : st[x] + arrb::PERIOD;
        if (!(ns - en[c] > arrb::PSEUDO_EPS)) continue;
        double ent[2], lv[2], tmp[2];
        (void)cut_dev(C, centre_of(B, la[c]), B.R[la[c]], tmp, lv);
        (void)cut_dev(C, centre_of(B, fi[x]), B.R[fi[x]], ent, tmp);
        const double a0 = rem_tau_dev(atan2(lv[1], lv[0]));
        const double a1 = atan2(ent[1], ent[0]);
        const double g = rem_tau_dev(a1 - a0);
        if (g > arrb::ANG_EPS + arrb::GAP_MARGIN && g < arrb::TAU - arrb::GAP_MARGIN) return 1;
        undecided = true;
    }
    return undecided ? 2 : 0;
}

// Outputs of the circle fill pass.
struct CircleOut {
    double *mx, *my, *mz, *ax, *ay, *az, *e1x, *e1y, *e1z, *rho;
    double* m3;  // centres, interleaved
    CircleRec* rec;
    uint32_t *ci, *cj;
    uint32_t *cb_off, *cb_items;
    double *vx, *vy, *vz, *v3;
    VertexRec* vrec;
    uint32_t* vatoms;
    uint32_t* exposed;
};

// Orders LDS traffic between the lanes of one wavefront.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline bool row_has(const uint32_t* row, uint32_t n, uint32_t k) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (row[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && row[lo] == k;
}

constexpr uint32_t ARR_WAVES = 4;  // wavefronts per workgroup of k_arr_circles

// The circle of pair p = (i, j > i) (arrangement.rs:672-751), one wavefront
// per pair: kept if no common neighbour buries it and an arc stays exposed;
// its blockers are the common neighbours cutting it (ascending), its triple
// points those of blockers k > j (leave, then enter) that no other
// neighbour of i holds. The lanes take the common neighbours 64 at a time;
// the cutting ones are compacted into LDS in ascending order, ranked there
// by (start, end descending, sphere), and one lane merges them into runs.
// COUNT writes keep / blocker / vertex counts per pair; FILL writes the
// records at the scanned offsets.
template <bool FILL>
__global__ void __launch_bounds__(256) k_arr_circles(ABuild B, const uint32_t* pi, const uint32_t* pj, uint32_t npairs,
                                                     uint32_t* keep, uint32_t* nblk, uint32_t* nvert,
                                                     const uint32_t* coff, const uint32_t* boff, const uint32_t* voff,
                                                     CircleOut o, uint32_t* flags) {
    constexpr uint32_t M = arrb::MAX_IV;
    __shared__ double u_st[ARR_WAVES][M], u_en[ARR_WAVES][M], s_st[ARR_WAVES][M], s_en[ARR_WAVES][M];
    __shared__ uint32_t u_k[ARR_WAVES][M], s_fi[ARR_WAVES][M], s_la[ARR_WAVES][M];
    __shared__ int s_res[ARR_WAVES];
    const uint32_t w = threadIdx.x / WAVE, lane = lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const uint64_t waves = (uint64_t)gridDim.x * ARR_WAVES;
    for (uint64_t p = (uint64_t)blockIdx.x * ARR_WAVES + w; p < npairs; p += waves) {
        if (FILL && !keep[p]) continue;
        auto drop = [&]() {
            if (!FILL && lane == 0) keep[p] = nblk[p] = nvert[p] = 0;
        };
        const uint32_t i = pi[p], j = pj[p];
        const D3v ci = centre_of(B, i), cj = centre_of(B, j);
        const double Ri = B.R[i], Rj = B.R[j];
        const double d = dnorm(dsub(ci, cj));
        const D3v axis = ddiv(dsub(cj, ci), d);
        const double t = (d * d + Ri * Ri - Rj * Rj) / (2.0 * d);
        const double rho2 = Ri * Ri - t * t;
        if (rho2 <= 0.0) {
            drop();
            continue;
        }
        CircleG C;
        C.axis = axis;
        {  // geom/vec.rs:112-124: helper along the smallest axis component
            const double x = fabs(axis.x), y = fabs(axis.y), z = fabs(axis.z);
            const D3v h = (x <= y && x <= z) ? D3v{1, 0, 0} : (y <= z ? D3v{0, 1, 0} : D3v{0, 0, 1});
            const D3v e1 = dcross(h, axis);
            C.e1 = ddiv(e1, dnorm(e1));
        }
        C.e2 = dcross(axis, C.e1);
        C.cc = dadd(ci, dmul(axis, t));
        C.radius = sqrt(rho2);
        const uint32_t* rowi = B.nb_items + B.nb_off[i];
        const uint32_t ni = B.nb_off[i + 1] - B.nb_off[i];
        const uint32_t* rowj = B.nb_items + B.nb_off[j];
        const uint32_t nj = B.nb_off[j + 1] - B.nb_off[j];
        // Common neighbours, ascending: a burying one drops the circle; the
        // cutting ones go to u_* in order (after the last pair's reads).
        wave_sync();
        uint32_t m = 0;
        bool buried = false, over = false;
        for (uint32_t base = 0; base < ni; base += WAVE) {
            const uint32_t q = base + lane;
            const uint32_t k = q < ni ? rowi[q] : 0u;
            int c = 0;
            double en[2], lv[2];
            if (q < ni && row_has(rowj, nj, k)) c = cut_dev(C, centre_of(B, k), B.R[k], en, lv);
            if (__ballot(c == 1)) {
                buried = true;
                break;
            }
            const unsigned long long bm = __ballot(c == 2);
            const uint32_t cnt = (uint32_t)__popcll(bm);
            if (m + cnt > M) {
                over = true;
                break;
            }
            if (c == 2) {
                const uint32_t slot = m + (uint32_t)__popcll(bm & lt);
                const double s0 = diamond_dev(en);
                double e0 = diamond_dev(lv);
                if (e0 <= s0) e0 += arrb::PERIOD;
                u_st[w][slot] = s0;
                u_en[w][slot] = e0;
                u_k[w][slot] = k;
            }
            m += cnt;
        }
        if (over) {
            if (lane == 0) atomicOr(flags, arrb::F_OVERFLOW);
            drop();
            continue;
        }
        if (buried) {
            drop();
            continue;
        }
        wave_sync();
        if (!FILL) {
            // Rank each interval by (start, end descending, sphere) in the
            // total order; the spheres differ, so the ranks are a permutation.
            for (uint32_t e = lane; e < m; e += WAVE) {
                const uint64_t ks = total_key(u_st[w][e]), ke = total_key(u_en[w][e]);
                const uint32_t kk = u_k[w][e];
                uint32_t r = 0;
                for (uint32_t f = 0; f < m; f++) {
                    const uint64_t fs = total_key(u_st[w][f]), fe = total_key(u_en[w][f]);
                    r += (fs < ks || (fs == ks && (ke < fe || (fe == ke && u_k[w][f] < kk)))) ? 1u : 0u;
                }
                s_st[w][r] = u_st[w][e];
                s_en[w][r] = u_en[w][e];
                s_fi[w][r] = kk;
                s_la[w][r] = kk;
            }
            wave_sync();
            if (lane == 0) s_res[w] = exposed_arc_runs(B, C, s_st[w], s_en[w], s_fi[w], s_la[w], m);
            wave_sync();
            const int ex = s_res[w];
            if (ex == 2 && lane == 0) atomicOr(flags, arrb::F_UNCERTAIN);
            if (ex != 1) {
                drop();
                continue;
            }
        }
        // Triple points of the cutting neighbours k > j: leave, then enter,
        // if no other neighbour of i holds them.
        uint32_t nv = 0;
        const uint32_t vw = FILL ? voff[p] : 0u;
        for (uint32_t base = 0; base < m; base += WAVE) {
            const uint32_t e = base + lane;
            const uint32_t k = e < m ? u_k[w][e] : 0u;
            bool f[2] = {false, false};
            D3v qv[2];
            if (e < m && k > j) {
                double en[2], lv[2];
                (void)cut_dev(C, centre_of(B, k), B.R[k], en, lv);
                const double* pts[2] = {lv, en};
                for (int s2 = 0; s2 < 2; s2++) {
                    qv[s2] = dadd(C.cc, dmul(dadd(dmul(C.e1, pts[s2][0]), dmul(C.e2, pts[s2][1])), C.radius));
                    bool free = true;
                    for (uint32_t u = 0; u < ni && free; u++) {
                        const uint32_t l = rowi[u];
                        if (l == j || l == k) continue;
                        if (!(dnorm(dsub(qv[s2], centre_of(B, l))) >= B.R[l] - EPS)) free = false;
                    }
                    f[s2] = free;
                }
            }
            const unsigned long long b0 = __ballot(f[0]), b1 = __ballot(f[1]);
            if (FILL) {
                uint32_t at = vw + nv + (uint32_t)(__popcll(b0 & lt) + __popcll(b1 & lt));
                for (int s2 = 0; s2 < 2; s2++) {
                    if (!f[s2]) continue;
                    o.vx[at] = qv[s2].x;
                    o.vy[at] = qv[s2].y;
                    o.vz[at] = qv[s2].z;
                    o.v3[3 * (size_t)at] = qv[s2].x;
                    o.v3[3 * (size_t)at + 1] = qv[s2].y;
                    o.v3[3 * (size_t)at + 2] = qv[s2].z;
                    o.vrec[at] = VertexRec{qv[s2].x, qv[s2].y, qv[s2].z, 0.0};
                    o.vatoms[3 * (size_t)at] = i;
                    o.vatoms[3 * (size_t)at + 1] = j;
                    o.vatoms[3 * (size_t)at + 2] = k;
                    at++;
                }
            }
            nv += (uint32_t)(__popcll(b0) + __popcll(b1));
        }
        if (!FILL) {
            if (lane == 0) {
                keep[p] = 1;
                nblk[p] = m;
                nvert[p] = nv;
            }
            continue;
        }
        for (uint32_t e = lane; e < m; e += WAVE) o.cb_items[boff[p] + e] = u_k[w][e];
        if (lane == 0) {
            const uint32_t c = coff[p];
            o.mx[c] = C.cc.x;
            o.my[c] = C.cc.y;
            o.mz[c] = C.cc.z;
            o.ax[c] = axis.x;
            o.ay[c] = axis.y;
            o.az[c] = axis.z;
            o.e1x[c] = C.e1.x;
            o.e1y[c] = C.e1.y;
            o.e1z[c] = C.e1.z;
            o.rho[c] = C.radius;
            o.m3[3 * (size_t)c] = C.cc.x;
            o.m3[3 * (size_t)c + 1] = C.cc.y;
            o.m3[3 * (size_t)c + 2] = C.cc.z;
            o.rec[c] = CircleRec{C.cc.x, C.cc.y, C.cc.z, axis.x, axis.y, axis.z, C.radius, 0.0};
            o.ci[c] = i;
            o.cj[c] = j;
            o.cb_off[c] = boff[p];
            o.exposed[i] = 1u;
            o.exposed[j] = 1u;
        }
    }
}

// Status of each sphere (arrangement.rs:566-568, 620-630): contained;
// exposed when not contained and without neighbours, on a kept circle, or
// for an arrangement without circles; else buried. Also the length of the
// row the field keeps and the sphere's refinement thresholds.
__global__ void k_arr_status(ABuild B, const uint32_t* exposed, int with_circles, uint8_t* status, uint32_t* fcount,
                             SphereThr* thr) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.n; i += gridDim.x * blockDim.x) {
        const uint32_t len = B.nb_off[i + 1] - B.nb_off[i];
        uint8_t st;
        if (B.contained[i])
            st = S_CONTAINED;
        else if (!with_circles || len == 0 || exposed[i])
            st = S_EXPOSED;
        else
            st = S_BURIED;
        status[i] = st;
        fcount[i] = st == S_EXPOSED ? len : 0u;  // the field keeps exposed rows (point.rs:243-244)
        thr[i] = SphereThr{sqrt_threshold(B.R[i] - EPS), sqrt_threshold(B.R[i])};
    }
}

__global__ void k_arr_rows(ABuild B, const uint8_t* status, const uint32_t* foff, uint32_t* fitems) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B.n; i += gridDim.x * blockDim.x) {
        if (status[i] != S_EXPOSED) continue;
        const uint32_t a = B.nb_off[i], e = B.nb_off[i + 1];
        for (uint32_t q = a; q < e; q++) fitems[foff[i] + q - a] = B.nb_items[q];
    }
}

// Bounds of n interleaved points, one partial (lo xyz, hi xyz) per
// workgroup, and the largest of `vals` (circle radii, all positive).
__global__ void __launch_bounds__(256) k_arr_bounds(const double* p3, uint32_t n, double* part, const double* vals,
                                                    unsigned long long* vmax) {
    __shared__ double red[6][256];
    double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; a++) {
            const double v = p3[3 * (size_t)i + a];
            b[a] = fmin(b[a], v);
            b[3 + a] = fmax(b[3 + a], v);
        }
        if (vals) atomicMax(vmax, (unsigned long long)__double_as_longlong(vals[i]));
    }
    for (int a = 0; a < 6; a++) red[a][threadIdx.x] = b[a];
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 6; a++)
                red[a][threadIdx.x] = a < 3 ? fmin(red[a][threadIdx.x], red[a][threadIdx.x + s])
                                            : fmax(red[a][threadIdx.x], red[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace smg
