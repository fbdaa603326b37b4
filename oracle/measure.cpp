// Measure stage of the oracle (test infrastructure only): validation,
// connected components with their sums and Euler characteristics, and the
// grouping of triangles by component, as the reference runs them after
// assembly (crates/sesmesh/src/surface.rs:1081-1095).
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace orc {

namespace {

// mesh/mod.rs:112-124
void area_volume_of(const std::vector<std::array<float, 3>>& pos, const std::array<u32, 3>& t, const double o[3],
                    double& area, double& volume) {
    double p[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) p[i][k] = (double)pos[t[i]][k] - o[k];
    V3 a(p[0][0], p[0][1], p[0][2]), b(p[1][0], p[1][1], p[1][2]), c(p[2][0], p[2][1], p[2][2]);
    V3 e1 = b - a, e2 = c - a;
    area = 0.5 * e1.cross(e2).norm();
    volume = a.dot(b.cross(c)) / 6.0;
}

// An undirected edge's uses: (min, max, forward uses, backward uses), in
// increasing (min, max) order (mod.rs:311-333: the content of
// Edges::par_runs does not depend on how the uses are bucketed).
struct Run {
    u32 a, b;
    u64 fwd, bwd;
};

std::vector<Run> edge_runs(const std::vector<std::array<u32, 3>>& tris) {
    std::vector<u64> keys;
    keys.reserve(3 * tris.size());
    for (const auto& t : tris)
        for (int k = 0; k < 3; k++) {
            u32 a = t[k], b = t[(k + 1) % 3];
            u32 lo = std::min(a, b), hi = std::max(a, b);
            // (min, max, forward) in one sortable key: min 31 bits, max 31 bits, forward 1 bit
            keys.push_back(((u64)lo << 32) | ((u64)hi << 1) | (u64)(a < b));
        }
    std::sort(keys.begin(), keys.end());
    std::vector<Run> runs;
    for (u64 k : keys) {
        u32 a = (u32)(k >> 32), b = (u32)((k >> 1) & 0x7fffffffu);
        bool fwd = k & 1;
        if (runs.empty() || runs.back().a != a || runs.back().b != b) runs.push_back({a, b, 0, 0});
        (fwd ? runs.back().fwd : runs.back().bwd)++;
    }
    return runs;
}

}  // namespace

// mesh/mod.rs:632-705 (validate), 181-300 (components_with), post.rs:14-34
// (group_by_component), surface.rs:1160-1168 (the stats sums).
Measured measure_full(const std::vector<std::array<float, 3>>& pos, const std::vector<std::array<float, 3>>& nrm,
                      const std::vector<std::array<u32, 3>>& tris) {
    Measured M;
    const size_t n = pos.size();
    Validation& v = M.validation;
    for (const auto& p : pos)
        if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]))) v.bad_vertices++;
    if (!nrm.empty()) {
        for (const auto& q : nrm) {
            float len2 = q[0] * q[0];
            len2 = len2 + q[1] * q[1];
            len2 = len2 + q[2] * q[2];
            bool ok = std::isfinite(q[0]) && std::isfinite(q[1]) && std::isfinite(q[2]) &&
                      std::fabs(len2 - 1.0f) < 1e-3f;
            if (!ok) v.bad_vertices++;
        }
        v.bad_vertices += n > nrm.size() ? n - nrm.size() : nrm.size() - n;
    }
    for (const auto& t : tris)
        if (t[0] >= n || t[1] >= n || t[2] >= n) v.invalid_indices++;
    if (v.invalid_indices) return M;  // the rest is not measured (mod.rs:665-667)
    M.indices_valid = true;
    double origin[3] = {0, 0, 0};
    if (n)
        for (int k = 0; k < 3; k++) origin[k] = (double)pos[0][k];
    for (const auto& t : tris) {
        double a, vol;
        area_volume_of(pos, t, origin, a, vol);
        if (t[0] == t[1] || t[1] == t[2] || t[0] == t[2] || a == 0.0) v.degenerate_triangles++;
    }
    {  // mod.rs:545-559 duplicates_by_sort
        std::vector<std::array<u32, 3>> s(tris);
        for (auto& t : s) std::sort(t.begin(), t.end());
        std::sort(s.begin(), s.end());
        for (size_t i = 1; i < s.size(); i++) v.duplicate_triangles += s[i] == s[i - 1];
    }
    const std::vector<Run> runs = edge_runs(tris);
    for (const Run& r : runs) {
        if (r.a == r.b || (r.fwd == 1 && r.bwd == 1)) continue;
        u64 uses = r.fwd + r.bwd;
        if (uses == 1)
            v.boundary_edges++;
        else if (uses == 2)
            v.misoriented_edges++;
        else
            v.non_manifold_edges++;
    }
    // Components: the root of each vertex is the smallest vertex of its
    // component; ids in order of the roots.
    std::vector<u32> parent(n);
    for (size_t i = 0; i < n; i++) parent[i] = (u32)i;
    auto find = [&](u32 x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    auto unite = [&](u32 a, u32 b) {
        u32 ra = find(a), rb = find(b);
        if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
    };
    for (const auto& t : tris) {
        unite(t[0], t[1]);
        unite(t[0], t[2]);
    }
    std::vector<u32>& id = M.vertex_component;
    id.assign(n, 0);
    std::vector<Component>& comps = M.components;
    for (size_t i = 0; i < n; i++) {
        u32 r = find((u32)i);
        if (r == i) {
            id[i] = (u32)comps.size();
            comps.push_back(Component{});
        } else {
            id[i] = id[r];
        }
        comps[id[i]].vertices++;
    }
    // Per fixed chunk of 2^14 triangles, per component, a sum in triangle
    // order; chunk sums merged in chunk order (mod.rs:231-262).
    const size_t CH = 1 << 14;
    for (size_t c0 = 0; c0 < tris.size(); c0 += CH) {
        struct Acc {
            u32 comp;
            u64 count;
            double a, v;
        };
        std::vector<Acc> acc;
        for (size_t i = c0; i < std::min(c0 + CH, tris.size()); i++) {
            u32 comp = id[tris[i][0]];
            double a, vol;
            area_volume_of(pos, tris[i], origin, a, vol);
            auto it = std::find_if(acc.begin(), acc.end(), [&](const Acc& e) { return e.comp == comp; });
            if (it == acc.end()) {
                acc.push_back({comp, 1, a, vol});
            } else {
                it->count++;
                it->a += a;
                it->v += vol;
            }
        }
        for (const Acc& e : acc) {
            comps[e.comp].triangles += e.count;
            comps[e.comp].area += e.a;
            comps[e.comp].volume += e.v;
        }
    }
    std::vector<int64_t> edges(comps.size(), 0);
    for (const Run& r : runs) edges[id[r.a]]++;  // every undirected edge once, self-edges included
    for (size_t c = 0; c < comps.size(); c++)
        comps[c].euler = (int64_t)comps[c].vertices - edges[c] + (int64_t)comps[c].triangles;
    // post.rs:14-34: triangles grouped by component, in their order.
    std::vector<u64> next(comps.size());
    u64 start = 0;
    for (size_t c = 0; c < comps.size(); c++) {
        comps[c].first_triangle = start;
        next[c] = start;
        start += comps[c].triangles;
    }
    M.grouped.assign(tris.size(), {0, 0, 0});
    for (const auto& t : tris) M.grouped[next[id[t[0]]]++] = t;
    // surface.rs:1160-1168: float sums fold from -0.0 (Rust's Sum for f64).
    M.area = -0.0;
    M.volume = -0.0;
    for (const Component& c : comps) {
        M.area += c.area;
        M.volume += c.volume;
        M.euler += c.euler;
        M.cavities += c.volume < 0.0;
    }
    return M;
}

}  // namespace orc
