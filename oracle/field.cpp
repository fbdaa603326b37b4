// Oracle: exact f64 point field (test infrastructure).
// Restates field/point.rs of the reference (crates/sesmesh/src/): the SES
// kind, and the SAS and vdW kinds (unions of spheres).
#include <algorithm>

#include "oracle.hpp"

namespace orc {

thread_local OpCounts g_ops;
thread_local bool g_in_eval_local = false;

namespace {
inline void count64(uint64_t n) {
    if (g_ops.on && g_in_eval_local) g_ops.f64_ops += n;
}
constexpr double EPS = 1e-9;  // geom/arrangement.rs:26

// field/point.rs:158-161
void sort_nearest(std::vector<std::pair<double, u32>>& v) {
    std::sort(v.begin(), v.end(), [](const std::pair<double, u32>& a, const std::pair<double, u32>& b) {
        if (total_lt(a.first, b.first)) return true;
        if (total_lt(b.first, a.first)) return false;
        return a.second < b.second;
    });
}

// field/point.rs:152-155
void filter_blockers(const std::vector<u32>& near, const u32* row, size_t n, std::vector<u32>& out) {
    for (size_t i = 0; i < n; i++)
        if (std::binary_search(near.begin(), near.end(), row[i])) out.push_back(row[i]);
}
}  // namespace

void LocalCandidates::clear() {
    inside.clear();
    near.clear();
    spheres.clear();
    s_off.assign(1, 0);
    s_blk.clear();
    circles.clear();
    c_off.assign(1, 0);
    c_blk.clear();
    vertices.clear();
    order.clear();
}

// field/point.rs:213-274
std::string SesField::build(const std::vector<Sphere>& spheres, double p, int k) {
    if (!(std::isfinite(p) && p >= 0.0)) return "probe radius must be finite and >= 0";
    if (k != K_SES && k != K_SAS && k != K_VDW) return "unknown surface kind";
    if (k == K_SES && p == 0.0) k = K_VDW;
    std::string err = build_arrangement(spheres, k == K_VDW ? 0.0 : p, arr, k == K_SES);
    if (!err.empty()) return err;
    kind = k;
    probe = p;
    // Drop neighbour rows of non-exposed spheres (field/point.rs:243-244).
    {
        Csr& nb = arr.neighbors;
        size_t start = 0, write = 0;
        for (size_t i = 0; i + 1 < nb.off.size(); i++) {
            size_t end = nb.off[i + 1];
            if (arr.status[i] == EXPOSED) {
                std::copy(nb.items.begin() + start, nb.items.begin() + end, nb.items.begin() + write);
                write += end - start;
            }
            nb.off[i + 1] = (u32)write;
            start = end;
        }
        nb.items.resize(write);
    }
    double cell = std::fmax(2.0 * arr.max_expanded, 1e-6);
    std::vector<V3> pts;
    for (const Circle& c : arr.circles) pts.push_back(c.center);
    circle_cells.build(pts, cell);
    pts.clear();
    for (const SasVertex& v : arr.vertices) pts.push_back(v.position);
    vertex_cells.build(pts, cell);
    rho_max = 0.0;
    for (const Circle& c : arr.circles) rho_max = std::fmax(rho_max, c.radius);
    if (arr.centers.empty()) {
        extent = 0.0;
    } else {
        V3 lo = arr.centers[0], hi = arr.centers[0];
        for (const V3& c : arr.centers) {
            lo = lo.vmin(c);
            hi = hi.vmax(c);
        }
        extent = (hi - lo).norm() + 4.0 * arr.max_expanded + 4.0 * p + 1.0;
    }
    return "";
}

// field/point.rs:371-382
bool SesField::inside_sas(const V3& x) const {
    const Arrangement& a = arr;
    bool inside = false;
    a.cells.for_each_near(x, a.max_expanded, [&](u32 k) {
        if (!inside && a.status[k] != CONTAINED && x.dist(a.centers[k]) < a.expanded[k]) inside = true;
    });
    return inside;
}

// field/point.rs:384-411
PointEval SesField::ses_eval(const V3& x, double cutoff) const {
    const Arrangement& a = arr;
    if (!inside_sas(x)) return PointEval{0.0, x, Feature{F_SOLVENT, 0}};
    PointEval best{cutoff, x, Feature{F_BEYOND, 0}};
    a.cells.for_each_near(x, cutoff + a.max_expanded, [&](u32 i) {
        if (a.status[i] == EXPOSED) {
            size_t nb;
            const u32* row = a.neighbors.row(i, nb);
            try_sphere_with(x, i, row, nb, best);
        }
    });
    circle_cells.for_each_near(x, cutoff + rho_max, [&](u32 ci) {
        size_t nb;
        const u32* row = a.circle_blockers.row(ci, nb);
        try_circle_with(x, ci, row, nb, best);
    });
    vertex_cells.for_each_near(x, cutoff, [&](u32 vi) { try_vertex(x, vi, best); });
    return best;
}

// field/point.rs:419-442
void SesField::try_sphere_with(const V3& x, u32 i, const u32* blk, size_t nb, PointEval& best) const {
    const Arrangement& a = arr;
    V3 v = x - a.centers[i];
    double len = v.norm();
    double lb = std::fabs(len - a.expanded[i]);
    count64(12);
    if (lb >= best.distance || len < 1e-12) return;
    V3 q = a.centers[i] + v * (a.expanded[i] / len);
    count64(7);
    for (size_t t = 0; t < nb; t++) {
        u32 k = blk[t];
        count64(11);
        if (!(q.dist(a.centers[k]) >= a.expanded[k] - EPS)) return;
    }
    best = PointEval{lb, q, Feature{F_SPHERE, i}};
}

// field/point.rs:450-475
void SesField::try_circle_with(const V3& x, u32 ci, const u32* blk, size_t nb, PointEval& best) const {
    const Arrangement& a = arr;
    const Circle& c = a.circles[ci];
    V3 y = x - c.center;
    double along = y.dot(c.axis);
    V3 w = y - c.axis * along;
    double wl = w.norm();
    double dist = std::sqrt(along * along + (wl - c.radius) * (wl - c.radius));
    count64(27);
    if (dist >= best.distance) return;
    V3 dir = wl > 1e-12 ? w / wl : c.e1;
    V3 q = c.center + dir * c.radius;
    count64(wl > 1e-12 ? 10 : 4);
    for (size_t t = 0; t < nb; t++) {
        u32 k = blk[t];
        count64(11);
        if (!(q.dist(a.centers[k]) >= a.expanded[k] - EPS)) return;
    }
    best = PointEval{dist, q, Feature{F_CIRCLE, ci}};
}

// field/point.rs:477-489
void SesField::try_vertex(const V3& x, u32 vi, PointEval& best) const {
    const V3& p = arr.vertices[vi].position;
    double d = x.dist(p);
    count64(10);
    if (d < best.distance) best = PointEval{d, p, Feature{F_VERTEX, vi}};
}

// field/point.rs:496-581
void SesField::gather_local(const V3& b, double rho, double band, LocalCandidates& out) const {
    const Arrangement& a = arr;
    out.clear();
    if (kind != K_SES) {  // point.rs:566-578: every sphere within reach, nearest first
        double reach = band + rho;
        a.cells.for_each_near(b, reach + a.max_expanded, [&](u32 k) {
            if (a.status[k] != CONTAINED && b.dist(a.centers[k]) - a.expanded[k] < reach)
                out.order.push_back({b.dist(a.centers[k]) - a.expanded[k], k});
        });
        sort_nearest(out.order);
        for (auto& o : out.order) out.inside.push_back(o.second);
        out.order.clear();
        return;
    }
    double reach = probe + band + rho;
    a.cells.for_each_near(b, reach + a.max_expanded, [&](u32 k) {
        if (a.status[k] == CONTAINED) return;
        double d = b.dist(a.centers[k]);
        if (d < a.expanded[k] + reach) out.near.push_back(k);
        if (d < a.expanded[k] + rho) out.inside.push_back(k);
        double lb = std::fabs(d - a.expanded[k]);
        if (a.status[k] == EXPOSED && lb <= reach) out.order.push_back({lb, k});
    });
    std::sort(out.near.begin(), out.near.end());
    sort_nearest(out.order);
    for (auto& o : out.order) {
        out.spheres.push_back(o.second);
        size_t nb;
        const u32* row = a.neighbors.row(o.second, nb);
        filter_blockers(out.near, row, nb, out.s_blk);
        out.s_off.push_back((u32)out.s_blk.size());
    }
    out.order.clear();
    circle_cells.for_each_near(b, reach + rho_max, [&](u32 ci) {
        const Circle& c = a.circles[ci];
        V3 y = b - c.center;
        double along = y.dot(c.axis);
        double wl = (y - c.axis * along).norm();
        double dist = std::sqrt(along * along + (wl - c.radius) * (wl - c.radius));
        if (dist <= reach) out.order.push_back({dist, ci});
    });
    sort_nearest(out.order);
    for (auto& o : out.order) {
        out.circles.push_back(o.second);
        size_t nb;
        const u32* row = a.circle_blockers.row(o.second, nb);
        filter_blockers(out.near, row, nb, out.c_blk);
        out.c_off.push_back((u32)out.c_blk.size());
    }
    out.order.clear();
    vertex_cells.for_each_near(b, reach, [&](u32 vi) {
        double d = b.dist(a.vertices[vi].position);
        if (d <= reach) out.order.push_back({d, vi});
    });
    sort_nearest(out.order);
    for (auto& o : out.order) out.vertices.push_back(o.second);
    out.order.clear();
}

// field/point.rs:686-738
NearEval SesField::eval_local(const V3& x, double band, const LocalCandidates& l) const {
    const Arrangement& a = arr;
    struct Scope {
        Scope() { g_in_eval_local = true; }
        ~Scope() { g_in_eval_local = false; }
    } scope;
    if (g_ops.on) g_ops.evals++;
    if (kind != K_SES) {  // point.rs:723-736: closest point is x itself
        double best = band;
        int64_t bi = -1;
        for (u32 k : l.inside) {
            double v = x.dist(a.centers[k]) - a.expanded[k];
            count64(11);
            if (v < best) {
                best = v;
                bi = k;
            }
        }
        return NearEval{best, bi < 0 ? Feature{F_BEYOND, 0} : Feature{F_SPHERE, (u32)bi}, x};
    }
    bool inside = false;
    for (u32 k : l.inside) {
        count64(10);
        if (x.dist(a.centers[k]) < a.expanded[k]) {
            inside = true;
            break;
        }
    }
    if (!inside) return NearEval{probe, Feature{F_SOLVENT, 0}, x};
    PointEval best{probe + band, x, Feature{F_BEYOND, 0}};
    for (size_t k = 0; k < l.spheres.size(); k++)
        try_sphere_with(x, l.spheres[k], l.s_blk.data() + l.s_off[k], l.s_off[k + 1] - l.s_off[k], best);
    for (size_t k = 0; k < l.circles.size(); k++)
        try_circle_with(x, l.circles[k], l.c_blk.data() + l.c_off[k], l.c_off[k + 1] - l.c_off[k], best);
    for (u32 vi : l.vertices) try_vertex(x, vi, best);
    return NearEval{probe - best.distance, best.feature, best.closest};
}

// field/point.rs:742-756
std::pair<double, int64_t> SesField::sphere_eval(const V3& x, double cutoff) const {
    const Arrangement& a = arr;
    std::pair<double, int64_t> best{cutoff, -1};
    a.cells.for_each_near(x, cutoff + a.max_expanded, [&](u32 k) {
        if (a.status[k] == CONTAINED) return;
        double v = x.dist(a.centers[k]) - a.expanded[k];
        if (v < best.first) best = {v, (int64_t)k};
    });
    return best;
}

// field/point.rs:292-369
double SesField::sample(const V3& x, V3& closest, Feature& f) const {
    if (arr.centers.empty()) {
        closest = x;
        f = Feature{F_SOLVENT, 0};
        return kind == K_SES ? probe : INFINITY;
    }
    if (kind != K_SES) {
        double cutoff = 4.0;
        std::pair<double, int64_t> e;
        for (;;) {
            e = sphere_eval(x, cutoff);
            if (e.second >= 0 || cutoff > extent) break;
            cutoff *= 2.0;
        }
        if (e.second < 0) {
            closest = x;
            f = Feature{F_SOLVENT, 0};
            return e.first;
        }
        const V3& c = arr.centers[e.second];
        V3 dir = x - c;
        double len = dir.norm();
        V3 unit = len > 1e-12 ? dir / len : V3(1.0, 0.0, 0.0);
        closest = c + unit * arr.expanded[e.second];
        f = Feature{F_SPHERE, (u32)e.second};
        return e.first;
    }
    double cutoff = probe + 4.0;
    PointEval e;
    for (;;) {
        e = ses_eval(x, cutoff);
        if (e.feature.kind != F_BEYOND || cutoff > extent) break;
        cutoff *= 2.0;
    }
    closest = e.closest;
    f = e.feature;
    return probe - e.distance;
}

// field/point.rs:760-765
double SesField::clamped_value(const V3& x, double band) const {
    if (kind == K_SES) return probe - ses_eval(x, probe + band).distance;
    return sphere_eval(x, band).first;
}

}  // namespace orc
