// Oracle: vectors, cell grids and the SAS arrangement (test infrastructure).
// Restates geom/vec.rs, geom/cells.rs and geom/arrangement.rs of the
// reference (crates/sesmesh/src/).
#include <algorithm>
#include <cstdio>

#include "oracle.hpp"

namespace orc {

// geom/vec.rs:112-124
void orthonormal_basis(const V3& u, V3& e1, V3& e2) {
    double ax = std::fabs(u.x), ay = std::fabs(u.y), az = std::fabs(u.z);
    V3 helper;
    if (ax <= ay && ax <= az)
        helper = V3(1, 0, 0);
    else if (ay <= az)
        helper = V3(0, 1, 0);
    else
        helper = V3(0, 0, 1);
    e1 = helper.cross(u).normalized();
    e2 = u.cross(e1);
}

// geom/cells.rs:27-77
void CellGrid::build(const std::vector<V3>& points, double cell_in) {
    cell = cell_in;
    if (points.empty()) {
        origin = V3();
        dims[0] = dims[1] = dims[2] = 1;
        starts.assign(2, 0);
        items.clear();
        return;
    }
    V3 lo = points[0], hi = points[0];
    for (const V3& p : points) {
        lo = lo.vmin(p);
        hi = hi.vmax(p);
    }
    V3 extent = hi - lo;
    size_t max_cells = std::max<size_t>(8 * points.size(), size_t(1) << 16);
    for (;;) {
        size_t d[3];
        bool ok = true;
        unsigned __int128 total = 1;
        for (int a = 0; a < 3; a++) {
            double f = std::floor(extent[a] / cell);
            // `as usize` saturates; such sizes never fit anyway.
            if (!(f < 1.8e19)) {
                ok = false;
                d[a] = 0;
            } else {
                d[a] = (size_t)f + 1;
                total *= d[a];
            }
        }
        if (ok && total <= max_cells) {
            for (int a = 0; a < 3; a++) dims[a] = d[a];
            break;
        }
        cell *= 2.0;
    }
    origin = lo;
    size_t total = dims[0] * dims[1] * dims[2];
    std::vector<size_t> keys(points.size());
    for (size_t i = 0; i < points.size(); i++) {
        size_t c[3];
        coord(points[i], c);
        keys[i] = flat(c);
    }
    starts.assign(total + 1, 0);
    for (size_t k : keys) starts[k + 1]++;
    for (size_t i = 0; i < total; i++) starts[i + 1] += starts[i];
    std::vector<u32> cursor(starts.begin(), starts.end() - 1);
    items.assign(points.size(), 0);
    for (size_t i = 0; i < keys.size(); i++) items[cursor[keys[i]]++] = (u32)i;
}

// geom/cells.rs:91-105
void CellGrid::coord(const V3& p, size_t c[3]) const {
    double rel[3] = {p.x - origin.x, p.y - origin.y, p.z - origin.z};
    for (int a = 0; a < 3; a++) {
        double f = std::floor(rel[a] / cell);
        if (f <= 0.0 || std::isnan(f)) {
            c[a] = 0;
        } else {
            size_t v = f >= 1.8e19 ? SIZE_MAX : (size_t)f;
            c[a] = std::min(v, dims[a] - 1);
        }
    }
}

// geom/cells.rs:132-146
static u64 spread(u32 v) {
    u64 x = (u64)v & 0x1fffffULL;
    x = (x | (x << 32)) & 0x001f00000000ffffULL;
    x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}
u64 morton(const u32 c[3]) { return spread(c[0]) | (spread(c[1]) << 1) | (spread(c[2]) << 2); }

// ------------------------------------------------------------ circle stage
namespace {

constexpr double EPS = 1e-9;       // geom/arrangement.rs:26
constexpr double ANG_EPS = 1e-12;  // geom/arrangement.rs:29
constexpr double PSEUDO_PERIOD = 4.0;
constexpr double PSEUDO_EPS = 1e-12;
constexpr double TAU = 6.283185307179586;

// geom/arrangement.rs:238-272. kind 0 empty, 1 full, 2 interval.
struct Cut {
    int kind;
    double enter[2], leave[2];
};

Cut cut(const Circle& circle, const V3& center, double expanded) {
    V3 g = circle.center - center;
    V3 e2 = circle.axis.cross(circle.e1);
    double a = 2.0 * circle.radius * circle.e1.dot(g);
    double b = 2.0 * circle.radius * e2.dot(g);
    double rhs = expanded * expanded - g.norm2() - circle.radius * circle.radius;
    double amp = std::sqrt(a * a + b * b);
    Cut c{};
    if (amp <= rhs) {
        c.kind = 1;
    } else if (amp <= -rhs || amp == 0.0) {
        c.kind = 0;
    } else {
        double cp = a / amp, sp = b / amp;
        double ca = rhs / amp;
        double sa = std::sqrt(std::fmax(1.0 - ca * ca, 0.0));
        c.kind = 2;
        c.enter[0] = cp * ca - sp * sa;
        c.enter[1] = sp * ca + cp * sa;
        c.leave[0] = cp * ca + sp * sa;
        c.leave[1] = sp * ca - cp * sa;
    }
    return c;
}

// geom/arrangement.rs:279-295
double pseudo_angle(const double v[2]) {
    double x = v[0], y = v[1];
    if (y >= 0.0) {
        if (x >= 0.0) return y / (x + y);
        return 1.0 - x / (y - x);
    }
    if (x < 0.0) return 2.0 - y / (-x - y);
    return 3.0 + x / (x - y);
}

struct Buried {
    double start, end;
    u32 sphere;
    double enter[2], leave[2];
};

Buried make_buried(const double enter[2], const double leave[2], u32 sphere) {
    Buried b;
    b.start = pseudo_angle(enter);
    b.end = pseudo_angle(leave);
    if (b.end <= b.start) b.end += PSEUDO_PERIOD;
    b.sphere = sphere;
    b.enter[0] = enter[0];
    b.enter[1] = enter[1];
    b.leave[0] = leave[0];
    b.leave[1] = leave[1];
    return b;
}

struct Merged {
    double start, end;
    u32 first, last;
    double enter[2], leave[2];
};

double rem_euclid(double a, double b) {
    double r = std::fmod(a, b);
    return r < 0.0 ? r + std::fabs(b) : r;
}

// geom/arrangement.rs:337-410: true if some arc is exposed (or no interval).
bool exposed_arcs(std::vector<Buried>& iv, std::vector<Merged>& merged) {
    if (iv.empty()) return true;
    std::sort(iv.begin(), iv.end(), [](const Buried& x, const Buried& y) {
        if (total_lt(x.start, y.start)) return true;
        if (total_lt(y.start, x.start)) return false;
        if (total_lt(y.end, x.end)) return true;
        if (total_lt(x.end, y.end)) return false;
        return x.sphere < y.sphere;
    });
    merged.clear();
    for (const Buried& b : iv) {
        if (!merged.empty() && b.start <= merged.back().end) {
            Merged& last = merged.back();
            if (b.end > last.end) {
                last.end = b.end;
                last.last = b.sphere;
                last.leave[0] = b.leave[0];
                last.leave[1] = b.leave[1];
            }
        } else {
            merged.push_back(Merged{b.start, b.end, b.sphere, b.sphere, {b.enter[0], b.enter[1]},
                                    {b.leave[0], b.leave[1]}});
        }
    }
    while (merged.size() > 1) {
        Merged first = merged[0];
        Merged& last = merged.back();
        if (last.end - PSEUDO_PERIOD < first.start) break;
        if (first.end + PSEUDO_PERIOD > last.end) {
            last.end = first.end + PSEUDO_PERIOD;
            last.last = first.last;
            last.leave[0] = first.leave[0];
            last.leave[1] = first.leave[1];
        }
        merged.erase(merged.begin());
    }
    size_t m = merged.size();
    bool any = false;
    for (size_t idx = 0; idx < m; idx++) {
        const Merged& cur = merged[idx];
        const Merged& next = merged[(idx + 1) % m];
        double next_start = idx + 1 < m ? next.start : next.start + PSEUDO_PERIOD;
        if (next_start - cur.end > PSEUDO_EPS) {
            double a0 = rem_euclid(std::atan2(cur.leave[1], cur.leave[0]), TAU);
            double a1 = std::atan2(next.enter[1], next.enter[0]);
            double len = rem_euclid(a1 - a0, TAU);
            if (len > ANG_EPS) any = true;
        }
    }
    return any;
}

}  // namespace

// geom/arrangement.rs:437-645 and :672-763
std::string build_arrangement(const std::vector<Sphere>& spheres, double probe, Arrangement& A, bool with_circles) {
    if (!(std::isfinite(probe) && probe >= 0.0)) return "probe radius must be finite and >= 0";
    for (size_t i = 0; i < spheres.size(); i++) {
        const Sphere& s = spheres[i];
        if (!(std::isfinite(s.c[0]) && std::isfinite(s.c[1]) && std::isfinite(s.c[2])))
            return "sphere " + std::to_string(i) + " has a non-finite centre";
        if (!(std::isfinite(s.r) && s.r >= 0.0))
            return "sphere " + std::to_string(i) + " has an invalid radius";
    }
    A = Arrangement();
    A.probe = probe;
    std::vector<u32> active;
    for (size_t i = 0; i < spheres.size(); i++)
        if (spheres[i].r > 0.0) active.push_back((u32)i);
    double max_expanded = 0.0;
    for (u32 i : active) max_expanded = std::fmax(max_expanded, spheres[i].r + probe);
    double cell = std::fmax(2.0 * max_expanded, 1e-6);

    auto C = [&](u32 i) { return V3(spheres[i].c[0], spheres[i].c[1], spheres[i].c[2]); };
    if (!active.empty()) {
        V3 lo = C(active[0]);
        for (u32 i : active) lo = lo.vmin(C(i));
        std::vector<std::pair<u64, u32>> keyed;
        keyed.reserve(active.size());
        for (u32 i : active) {
            V3 p = (C(i) - lo) / cell;
            u32 c[3];
            for (int a = 0; a < 3; a++) {
                double v = p[a];
                // `v as u64` truncates toward zero and saturates (NaN -> 0).
                u64 t = (!(v > 0.0)) ? 0 : (v >= 1.8446744073709552e19 ? UINT64_MAX : (u64)v);
                c[a] = (u32)std::min<u64>(t, 0x1fffff);
            }
            keyed.push_back({morton(c), i});
        }
        std::sort(keyed.begin(), keyed.end());
        for (size_t k = 0; k < active.size(); k++) active[k] = keyed[k].second;
    }
    size_t n = active.size();
    A.centers.resize(n);
    A.radii.resize(n);
    A.expanded.resize(n);
    for (size_t k = 0; k < n; k++) {
        A.centers[k] = C(active[k]);
        A.radii[k] = spheres[active[k]].r;
        A.expanded[k] = A.radii[k] + probe;
    }
    A.cells.build(A.centers, cell);
    const auto& centers = A.centers;
    const auto& expanded = A.expanded;

    // Containment (geom/arrangement.rs:499-516)
    std::vector<uint8_t> contained(n, 0);
    for (size_t i = 0; i < n; i++) {
        bool found = false;
        A.cells.for_each_near(centers[i], expanded[i] + max_expanded, [&](u32 ju) {
            size_t j = ju;
            if (j != i && !found) {
                double d = centers[i].dist(centers[j]);
                if (d + expanded[i] <= expanded[j] + EPS &&
                    (expanded[j] - expanded[i] > EPS || d > EPS || active[j] < active[i]))
                    found = true;
            }
        });
        contained[i] = found;
        A.n_contained += found;
    }

    // Neighbours (geom/arrangement.rs:519-550)
    for (size_t i = 0; i < n; i++) {
        size_t start = A.neighbors.items.size();
        if (!contained[i]) {
            A.cells.for_each_near(centers[i], expanded[i] + max_expanded, [&](u32 j) {
                if (j != i && !contained[j] && centers[i].dist(centers[j]) < expanded[i] + expanded[j] - EPS)
                    A.neighbors.items.push_back(j);
            });
            std::sort(A.neighbors.items.begin() + start, A.neighbors.items.end());
        }
        A.neighbors.close_row();
    }

    // Circles, blockers and triple points (geom/arrangement.rs:552-611, 672-763)
    std::vector<uint8_t> exposed(n);
    for (size_t i = 0; i < n; i++) {
        size_t nr;
        A.neighbors.row(i, nr);
        exposed[i] = !contained[i] && (!with_circles || nr == 0);  // arrangement.rs:566-568
    }
    std::vector<Buried> intervals;
    std::vector<Merged> merged;
    std::vector<u32> blockers;
    for (size_t i = 0; i < n && with_circles; i++) {
        if (contained[i]) continue;
        size_t ni;
        const u32* rowi = A.neighbors.row(i, ni);
        for (size_t q = 0; q < ni; q++) {
            u32 j = rowi[q];
            if (j <= i) continue;
            double d = centers[i].dist(centers[j]);
            V3 axis = (centers[j] - centers[i]) / d;
            double t = (d * d + expanded[i] * expanded[i] - expanded[j] * expanded[j]) / (2.0 * d);
            double rho2 = expanded[i] * expanded[i] - t * t;
            if (rho2 <= 0.0) continue;
            Circle circle;
            circle.i = (u32)i;
            circle.j = j;
            circle.center = centers[i] + axis * t;
            circle.axis = axis;
            V3 e1, e2;
            orthonormal_basis(axis, e1, e2);
            circle.e1 = e1;
            circle.radius = std::sqrt(rho2);
            intervals.clear();
            blockers.clear();
            bool buried = false;
            size_t nj;
            const u32* rowj = A.neighbors.row(j, nj);
            // sorted_intersection (geom/arrangement.rs:754-763)
            size_t jj = 0;
            for (size_t p = 0; p < ni && !buried; p++) {
                u32 k = rowi[p];
                while (jj < nj && rowj[jj] < k) jj++;
                if (!(jj < nj && rowj[jj] == k)) continue;
                Cut c = cut(circle, centers[k], expanded[k]);
                if (c.kind == 1) {
                    buried = true;
                } else if (c.kind == 2) {
                    intervals.push_back(make_buried(c.enter, c.leave, k));
                    blockers.push_back(k);
                }
            }
            if (buried) continue;
            if (!exposed_arcs(intervals, merged)) continue;
            for (u32 k : blockers) {
                if (k <= j) continue;
                Cut c = cut(circle, centers[k], expanded[k]);
                if (c.kind != 2) continue;
                const double* pts[2] = {c.leave, c.enter};
                for (int w = 0; w < 2; w++) {
                    double cc = pts[w][0], s = pts[w][1];
                    V3 qv = circle.center + (circle.e1 * cc + e2 * s) * circle.radius;
                    bool free = true;
                    for (size_t p = 0; p < ni; p++) {
                        u32 l = rowi[p];
                        if (l == j || l == k) continue;
                        if (!(qv.dist(centers[l]) >= expanded[l] - EPS)) {
                            free = false;
                            break;
                        }
                    }
                    if (free) {
                        SasVertex v;
                        v.position = qv;
                        v.atoms[0] = (u32)i;
                        v.atoms[1] = j;
                        v.atoms[2] = k;
                        A.vertices.push_back(v);
                    }
                }
            }
            exposed[i] = 1;
            exposed[j] = 1;
            A.circles.push_back(circle);
            for (u32 k : blockers) A.circle_blockers.items.push_back(k);
            A.circle_blockers.close_row();
        }
    }
    A.status.resize(n);
    for (size_t i = 0; i < n; i++) {
        A.status[i] = contained[i] ? CONTAINED : (exposed[i] ? EXPOSED : BURIED);
        A.n_exposed += A.status[i] == EXPOSED;
    }
    A.orig = active;
    A.max_expanded = max_expanded;
    return "";
}

}  // namespace orc
