// Host-side SAS arrangement for the mesh stage (f64, multithreaded).
//
// The reference builds this on the CPU before its mesh stage
// (crates/sesmesh/src/geom/arrangement.rs:437-645, with circles and triple
// points, without arcs, and crates/sesmesh/src/field/point.rs:213-274 for the
// neighbour-row release and the circle/vertex cell grids). It is restated
// here with the same f64 operations and orders, so the device's decisions
// made from it are the CPU's. Work is split over threads by contiguous
// sphere ranges whose outputs are concatenated in sphere order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "host.hpp"

namespace smg {

namespace {

constexpr double EPS = 1e-9;        // geom/arrangement.rs:26
constexpr double ANG_EPS = 1e-12;   // geom/arrangement.rs:29
constexpr double PERIOD = 4.0;      // geom/arrangement.rs:275
constexpr double PSEUDO_EPS = 1e-12;
constexpr double TAU = 6.283185307179586;

struct D3 {
    double x, y, z;
};
inline D3 sub3(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add3(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 mul3(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 div3(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross3(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm3(D3 a) { return std::sqrt(dot3(a, a)); }
inline double dist3(D3 a, D3 b) { return norm3(sub3(a, b)); }

inline bool key_lt(double a, double b) {  // f64::total_cmp "<"
    int64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    ia ^= (int64_t)(((uint64_t)(ia >> 63)) >> 1);
    ib ^= (int64_t)(((uint64_t)(ib >> 63)) >> 1);
    return ia < ib;
}

template <class F>
void parallel_for(size_t n, int threads, F&& f) {
    if (threads <= 1 || n < 256) {
        f(0, n, 0);
        return;
    }
    size_t per = (n + threads - 1) / threads;
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; w++) {
        size_t a = w * per, b = std::min(n, a + per);
        if (a >= b) break;
        pool.emplace_back([&f, a, b, w] { f(a, b, w); });
    }
    for (auto& t : pool) t.join();
}

uint64_t spread21(uint32_t v) {
    uint64_t x = (uint64_t)v & 0x1fffffULL;
    x = (x | (x << 32)) & 0x001f00000000ffffULL;
    x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

struct CutResult {
    int kind;  // 0 empty, 1 full, 2 interval
    double enter[2], leave[2];
};

// geom/arrangement.rs:251-272
CutResult cut(D3 cc, D3 axis, D3 e1, double radius, D3 center, double expanded) {
    D3 g = sub3(cc, center);
    D3 e2 = cross3(axis, e1);
    double a = 2.0 * radius * dot3(e1, g);
    double b = 2.0 * radius * dot3(e2, g);
    double rhs = expanded * expanded - dot3(g, g) - radius * radius;
    double amp = std::sqrt(a * a + b * b);
    CutResult r{};
    if (amp <= rhs) {
        r.kind = 1;
    } else if (amp <= -rhs || amp == 0.0) {
        r.kind = 0;
    } else {
        double cp = a / amp, sp = b / amp, ca = rhs / amp;
        double sa = std::sqrt(std::fmax(1.0 - ca * ca, 0.0));
        r.kind = 2;
        r.enter[0] = cp * ca - sp * sa;
        r.enter[1] = sp * ca + cp * sa;
        r.leave[0] = cp * ca + sp * sa;
        r.leave[1] = sp * ca - cp * sa;
    }
    return r;
}

// geom/arrangement.rs:282-295
double diamond(const double v[2]) {
    double x = v[0], y = v[1];
    if (y >= 0.0) return x >= 0.0 ? y / (x + y) : 1.0 - x / (y - x);
    return x < 0.0 ? 2.0 - y / (-x - y) : 3.0 + x / (x - y);
}

struct Interval {
    double start, end;
    uint32_t sphere;
    double enter[2], leave[2];
};
struct Run {
    double start, end;
    uint32_t first, last;
    double enter[2], leave[2];
};

double rem_tau(double a) {
    double r = std::fmod(a, TAU);
    return r < 0.0 ? r + TAU : r;
}

// geom/arrangement.rs:340-410: whether any arc of the circle stays exposed.
bool has_exposed_arc(std::vector<Interval>& iv, std::vector<Run>& runs) {
    if (iv.empty()) return true;
    std::sort(iv.begin(), iv.end(), [](const Interval& x, const Interval& y) {
        if (key_lt(x.start, y.start)) return true;
        if (key_lt(y.start, x.start)) return false;
        if (key_lt(y.end, x.end)) return true;
        if (key_lt(x.end, y.end)) return false;
        return x.sphere < y.sphere;
    });
    runs.clear();
    for (const Interval& b : iv) {
        if (!runs.empty() && b.start <= runs.back().end) {
            Run& l = runs.back();
            if (b.end > l.end) {
                l.end = b.end;
                l.last = b.sphere;
                l.leave[0] = b.leave[0];
                l.leave[1] = b.leave[1];
            }
        } else {
            runs.push_back({b.start, b.end, b.sphere, b.sphere, {b.enter[0], b.enter[1]}, {b.leave[0], b.leave[1]}});
        }
    }
    size_t head = 0;  // runs[head..] are live (the reference removes from the front)
    while (runs.size() - head > 1) {
        Run first = runs[head];
        Run& l = runs.back();
        if (l.end - PERIOD < first.start) break;
        if (first.end + PERIOD > l.end) {
            l.end = first.end + PERIOD;
            l.last = first.last;
            l.leave[0] = first.leave[0];
            l.leave[1] = first.leave[1];
        }
        head++;
    }
    size_t m = runs.size() - head;
    for (size_t k = 0; k < m; k++) {
        const Run& cur = runs[head + k];
        const Run& nx = runs[head + (k + 1) % m];
        double ns = k + 1 < m ? nx.start : nx.start + PERIOD;
        if (ns - cur.end > PSEUDO_EPS) {
            double a0 = rem_tau(std::atan2(cur.leave[1], cur.leave[0]));
            double a1 = std::atan2(nx.enter[1], nx.enter[0]);
            if (rem_tau(a1 - a0) > ANG_EPS) return true;
        }
    }
    return false;
}

}  // namespace

// Dimensions of the cell grid of n points with bounds lo..hi
// (geom/cells.rs:27-60): the cell edge doubles until the cells number at
// most max(8 n, 2^16).
void grid_shape(const double lo[3], const double hi[3], size_t n, double cell, Grid& g) {
    g.cell = cell;
    size_t max_cells = std::max<size_t>(8 * n, size_t(1) << 16);
    for (;;) {
        unsigned __int128 total = 1;
        bool ok = true;
        size_t d[3];
        for (int a = 0; a < 3; a++) {
            double f = std::floor((hi[a] - lo[a]) / g.cell);
            if (!(f < 1.8e19)) {
                ok = false;
                d[a] = 0;
            } else {
                d[a] = (size_t)f + 1;
                total *= d[a];
            }
        }
        if (ok && total <= max_cells) {
            for (int a = 0; a < 3; a++) g.dims[a] = (uint32_t)d[a];
            break;
        }
        g.cell *= 2.0;
    }
    for (int a = 0; a < 3; a++) g.origin[a] = lo[a];
}

// Cell grid of geom/cells.rs:27-77.
void build_grid(const double* pts, size_t n, double cell, Grid& g) {
    g.cell = cell;
    if (n == 0) {
        g.origin[0] = g.origin[1] = g.origin[2] = 0.0;
        g.dims[0] = g.dims[1] = g.dims[2] = 1;
        g.starts.assign(2, 0);
        g.items.clear();
        return;
    }
    double lo[3] = {pts[0], pts[1], pts[2]}, hi[3] = {pts[0], pts[1], pts[2]};
    for (size_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], pts[3 * i + a]);
            hi[a] = std::fmax(hi[a], pts[3 * i + a]);
        }
    grid_shape(lo, hi, n, cell, g);
    size_t total = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    std::vector<size_t> key(n);
    for (size_t i = 0; i < n; i++) {
        size_t c[3];
        for (int a = 0; a < 3; a++) {
            double f = std::floor((pts[3 * i + a] - g.origin[a]) / g.cell);
            c[a] = (f <= 0.0 || std::isnan(f)) ? 0 : std::min<size_t>((size_t)f, g.dims[a] - 1);
        }
        key[i] = (c[2] * g.dims[1] + c[1]) * g.dims[0] + c[0];
    }
    g.starts.assign(total + 1, 0);
    for (size_t k : key) g.starts[k + 1]++;
    for (size_t i = 0; i < total; i++) g.starts[i + 1] += g.starts[i];
    std::vector<uint32_t> cursor(g.starts.begin(), g.starts.end() - 1);
    g.items.assign(n, 0);
    for (size_t i = 0; i < n; i++) g.items[cursor[key[i]]++] = (uint32_t)i;
}

template <class F>
static void grid_near(const Grid& g, const double c[3], double r, F&& f) {
    size_t lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        double fl = std::floor((c[a] - r - g.origin[a]) / g.cell);
        double fh = std::floor((c[a] + r - g.origin[a]) / g.cell);
        lo[a] = (fl <= 0.0 || std::isnan(fl)) ? 0 : std::min<size_t>((size_t)fl, g.dims[a] - 1);
        hi[a] = (fh <= 0.0 || std::isnan(fh)) ? 0 : std::min<size_t>((size_t)fh, g.dims[a] - 1);
    }
    for (size_t z = lo[2]; z <= hi[2]; z++)
        for (size_t y = lo[1]; y <= hi[1]; y++) {
            size_t row = (z * g.dims[1] + y) * g.dims[0];
            for (uint32_t i = g.starts[row + lo[0]]; i < g.starts[row + hi[0] + 1]; i++) f(g.items[i]);
        }
}

std::string build_arrangement(const double* xyzr, size_t n_in, double probe, int kind, int threads, Arrangement& A) {
    if (!(std::isfinite(probe) && probe >= 0.0)) return "probe radius must be finite and >= 0";
    if (kind < 0 || kind > 2) return "unknown surface kind " + std::to_string(kind);
    if (kind == 0 && probe == 0.0) kind = 2;  // field/point.rs:224-228
    if (kind == 2) probe = 0.0;               // SurfaceParams::effective_probe (surface.rs:411-417)
    const bool with_circles = kind == 0;
    for (size_t i = 0; i < n_in; i++) {
        if (!(std::isfinite(xyzr[4 * i]) && std::isfinite(xyzr[4 * i + 1]) && std::isfinite(xyzr[4 * i + 2])))
            return "sphere " + std::to_string(i) + " has a non-finite centre";
        double r = xyzr[4 * i + 3];
        if (!(std::isfinite(r) && r >= 0.0)) return "sphere " + std::to_string(i) + " has an invalid radius";
    }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    A = Arrangement();
    A.kind = kind;
    A.probe = probe;
    std::vector<uint32_t> active;
    for (size_t i = 0; i < n_in; i++)
        if (xyzr[4 * i + 3] > 0.0) active.push_back((uint32_t)i);
    double max_expanded = 0.0;
    for (uint32_t i : active) max_expanded = std::fmax(max_expanded, xyzr[4 * i + 3] + probe);
    double cell = std::fmax(2.0 * max_expanded, 1e-6);
    if (!active.empty()) {
        double lo[3] = {xyzr[4 * active[0]], xyzr[4 * active[0] + 1], xyzr[4 * active[0] + 2]};
        for (uint32_t i : active)
            for (int a = 0; a < 3; a++) lo[a] = std::fmin(lo[a], xyzr[4 * i + a]);
        std::vector<std::pair<uint64_t, uint32_t>> keyed(active.size());
        for (size_t k = 0; k < active.size(); k++) {
            uint32_t i = active[k];
            uint64_t m = 0;
            for (int a = 0; a < 3; a++) {
                double v = (xyzr[4 * i + a] - lo[a]) / cell;
                uint64_t t = !(v > 0.0) ? 0 : (v >= 1.8446744073709552e19 ? UINT64_MAX : (uint64_t)v);
                m |= spread21((uint32_t)std::min<uint64_t>(t, 0x1fffff)) << a;
            }
            keyed[k] = {m, i};
        }
        std::sort(keyed.begin(), keyed.end());
        for (size_t k = 0; k < active.size(); k++) active[k] = keyed[k].second;
    }
    const size_t n = active.size();
    A.center.resize(3 * n);
    A.radius.resize(n);
    A.expanded.resize(n);
    for (size_t k = 0; k < n; k++) {
        for (int a = 0; a < 3; a++) A.center[3 * k + a] = xyzr[4 * active[k] + a];
        A.radius[k] = xyzr[4 * active[k] + 3];
        A.expanded[k] = A.radius[k] + probe;
    }
    A.orig = active;
    A.max_expanded = max_expanded;
    build_grid(A.center.data(), n, cell, A.sphere_grid);
    auto C = [&](size_t k) { return D3{A.center[3 * k], A.center[3 * k + 1], A.center[3 * k + 2]}; };
    const auto& R = A.expanded;

    // Containment (geom/arrangement.rs:499-516).
    std::vector<uint8_t> contained(n, 0);
    parallel_for(n, threads, [&](size_t a, size_t b, int) {
        for (size_t i = a; i < b; i++) {
            bool found = false;
            D3 ci = C(i);
            grid_near(A.sphere_grid, &A.center[3 * i], R[i] + max_expanded, [&](uint32_t j) {
                if (j == i || found) return;
                double d = dist3(ci, C(j));
                if (d + R[i] <= R[j] + EPS && (R[j] - R[i] > EPS || d > EPS || active[j] < active[i])) found = true;
            });
            contained[i] = found;
        }
    });

    // Neighbours (geom/arrangement.rs:519-550), per-thread parts in order.
    int T = threads;
    std::vector<std::vector<uint32_t>> nb_items(T), nb_len(T);
    std::vector<size_t> part_first(T + 1, n);
    {
        size_t per = (n + T - 1) / std::max(T, 1);
        for (int w = 0; w <= T; w++) part_first[w] = std::min(n, (size_t)w * per);
    }
    auto by_part = [&](auto&& f) {
        std::vector<std::thread> pool;
        for (int w = 0; w < T; w++)
            if (part_first[w] < part_first[w + 1]) pool.emplace_back([&, w] { f(part_first[w], part_first[w + 1], w); });
        for (auto& t : pool) t.join();
    };
    by_part([&](size_t a, size_t b, int w) {
        auto& items = nb_items[w];
        auto& lens = nb_len[w];
        for (size_t i = a; i < b; i++) {
            size_t start = items.size();
            if (!contained[i]) {
                D3 ci = C(i);
                grid_near(A.sphere_grid, &A.center[3 * i], R[i] + max_expanded, [&](uint32_t j) {
                    if (j != i && !contained[j] && dist3(ci, C(j)) < R[i] + R[j] - EPS) items.push_back(j);
                });
                std::sort(items.begin() + start, items.end());
            }
            lens.push_back((uint32_t)(items.size() - start));
        }
    });
    std::vector<uint32_t> nb_off(n + 1, 0), nb_all;
    {
        size_t k = 0;
        for (int w = 0; w < T; w++) {
            for (uint32_t l : nb_len[w]) {
                nb_off[k + 1] = nb_off[k] + l;
                k++;
            }
            nb_all.insert(nb_all.end(), nb_items[w].begin(), nb_items[w].end());
            std::vector<uint32_t>().swap(nb_items[w]);
        }
    }

    // Circles, blockers and triple points per sphere, SES kind only (geom/arrangement.rs:552-611, 672-751).
    struct Part {
        std::vector<double> circle;  // 10 per circle
        std::vector<uint32_t> ij, cb_len, cb_items;
        std::vector<double> vpos;
        std::vector<uint32_t> vatoms;
        std::vector<uint32_t> exposed;  // spheres made exposed by a circle
    };
    std::vector<Part> parts(T);
    by_part([&](size_t a, size_t b, int w) {
        Part& P = parts[w];
        std::vector<Interval> iv;
        std::vector<Run> runs;
        std::vector<uint32_t> blk;
        for (size_t i = a; i < b && with_circles; i++) {
            if (contained[i]) continue;
            const uint32_t* rowi = nb_all.data() + nb_off[i];
            size_t ni = nb_off[i + 1] - nb_off[i];
            for (size_t q = 0; q < ni; q++) {
                uint32_t j = rowi[q];
                if (j <= i) continue;
                D3 ci = C(i), cj = C(j);
                double d = dist3(ci, cj);
                D3 axis = div3(sub3(cj, ci), d);
                double t = (d * d + R[i] * R[i] - R[j] * R[j]) / (2.0 * d);
                double rho2 = R[i] * R[i] - t * t;
                if (rho2 <= 0.0) continue;
                // geom/vec.rs:112-124
                double ax = std::fabs(axis.x), ay = std::fabs(axis.y), az = std::fabs(axis.z);
                D3 helper = (ax <= ay && ax <= az) ? D3{1, 0, 0} : (ay <= az ? D3{0, 1, 0} : D3{0, 0, 1});
                D3 e1 = cross3(helper, axis);
                e1 = div3(e1, norm3(e1));
                D3 e2 = cross3(axis, e1);
                D3 cc = add3(ci, mul3(axis, t));
                double radius = std::sqrt(rho2);
                iv.clear();
                blk.clear();
                bool buried = false;
                const uint32_t* rowj = nb_all.data() + nb_off[j];
                size_t nj = nb_off[j + 1] - nb_off[j], jj = 0;
                for (size_t p = 0; p < ni && !buried; p++) {
                    uint32_t k = rowi[p];
                    while (jj < nj && rowj[jj] < k) jj++;
                    if (!(jj < nj && rowj[jj] == k)) continue;
                    CutResult c = cut(cc, axis, e1, radius, C(k), R[k]);
                    if (c.kind == 1) {
                        buried = true;
                    } else if (c.kind == 2) {
                        Interval it;
                        it.start = diamond(c.enter);
                        it.end = diamond(c.leave);
                        if (it.end <= it.start) it.end += PERIOD;
                        it.sphere = k;
                        it.enter[0] = c.enter[0];
                        it.enter[1] = c.enter[1];
                        it.leave[0] = c.leave[0];
                        it.leave[1] = c.leave[1];
                        iv.push_back(it);
                        blk.push_back(k);
                    }
                }
                if (buried || !has_exposed_arc(iv, runs)) continue;
                for (uint32_t k : blk) {
                    if (k <= j) continue;
                    CutResult c = cut(cc, axis, e1, radius, C(k), R[k]);
                    if (c.kind != 2) continue;
                    const double* pts[2] = {c.leave, c.enter};
                    for (int s = 0; s < 2; s++) {
                        D3 qv = add3(cc, mul3(add3(mul3(e1, pts[s][0]), mul3(e2, pts[s][1])), radius));
                        bool free = true;
                        for (size_t p = 0; p < ni && free; p++) {
                            uint32_t l = rowi[p];
                            if (l == j || l == k) continue;
                            if (!(dist3(qv, C(l)) >= R[l] - EPS)) free = false;
                        }
                        if (free) {
                            P.vpos.insert(P.vpos.end(), {qv.x, qv.y, qv.z});
                            P.vatoms.insert(P.vatoms.end(), {(uint32_t)i, j, k});
                        }
                    }
                }
                P.exposed.push_back((uint32_t)i);
                P.exposed.push_back(j);
                P.ij.insert(P.ij.end(), {(uint32_t)i, j});
                P.circle.insert(P.circle.end(), {cc.x, cc.y, cc.z, axis.x, axis.y, axis.z, e1.x, e1.y, e1.z, radius});
                P.cb_len.push_back((uint32_t)blk.size());
                P.cb_items.insert(P.cb_items.end(), blk.begin(), blk.end());
            }
        }
    });
    std::vector<uint8_t> exposed(n);
    for (size_t i = 0; i < n; i++) exposed[i] = !contained[i] && (!with_circles || nb_off[i + 1] == nb_off[i]);
    A.cb_off.assign(1, 0);
    for (auto& P : parts) {
        for (uint32_t s : P.exposed) exposed[s] = 1;
        A.circle.insert(A.circle.end(), P.circle.begin(), P.circle.end());
        A.circle_ij.insert(A.circle_ij.end(), P.ij.begin(), P.ij.end());
        for (uint32_t l : P.cb_len) A.cb_off.push_back(A.cb_off.back() + l);
        A.cb_items.insert(A.cb_items.end(), P.cb_items.begin(), P.cb_items.end());
        A.vertex.insert(A.vertex.end(), P.vpos.begin(), P.vpos.end());
        A.vertex_atoms.insert(A.vertex_atoms.end(), P.vatoms.begin(), P.vatoms.end());
        Part().exposed.swap(P.exposed);
    }
    A.status.resize(n);
    for (size_t i = 0; i < n; i++) A.status[i] = contained[i] ? 2 : (exposed[i] ? 0 : 1);

    // The field keeps neighbour rows of exposed spheres only (point.rs:243-244).
    A.nb_off.assign(n + 1, 0);
    A.nb_items.clear();
    for (size_t i = 0; i < n; i++) {
        if (A.status[i] == 0) A.nb_items.insert(A.nb_items.end(), nb_all.begin() + nb_off[i], nb_all.begin() + nb_off[i + 1]);
        A.nb_off[i + 1] = (uint32_t)A.nb_items.size();
    }
    size_t nc = A.circle.size() / 10, nv = A.vertex.size() / 3;
    std::vector<double> pts(3 * nc);
    A.rho_max = 0.0;
    for (size_t c = 0; c < nc; c++) {
        for (int a = 0; a < 3; a++) pts[3 * c + a] = A.circle[10 * c + a];
        A.rho_max = std::fmax(A.rho_max, A.circle[10 * c + 9]);
    }
    build_grid(pts.data(), nc, cell, A.circle_grid);
    build_grid(A.vertex.data(), nv, cell, A.vertex_grid);
    return "";
}

}  // namespace smg
