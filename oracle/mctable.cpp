// Oracle: Marching Cubes case table generator (test infrastructure).
//
// Restates the generator rules of mesh/tables.rs:7-22 and :47-274 of the
// reference (crates/sesmesh/src/): per case, segments on each cube face
// under the uniform face rule, chained into loops, each loop triangulated by
// the enumeration-order-first triangulation with no face diagonal, fewest
// inverted vertices, least area (rounded at 1e-9), largest minimum angle.
// The reference ships the generated table as mesh/mc_table.rs; the tests pin
// this generator against a SHA-256 of that table.
#include <algorithm>
#include <mutex>

#include "oracle.hpp"

namespace orc {

namespace {
using P = std::array<double, 3>;

P psub(P a, P b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
double pdot(P a, P b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
P pcross(P a, P b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
double pnorm(P a) { return std::sqrt(pdot(a, a)); }

P corner_pos(int c) { return {(double)(c & 1), (double)((c >> 1) & 1), (double)((c >> 2) & 1)}; }
void edge_corners(int e, int& lo, int& hi) {
    int axis = e / 4;
    lo = EDGE_LOWER[axis][e % 4];
    hi = lo | (1 << axis);
}
P edge_mid(int e) {
    int a, b;
    edge_corners(e, a, b);
    P pa = corner_pos(a), pb = corner_pos(b);
    return {0.5 * (pa[0] + pb[0]), 0.5 * (pa[1] + pb[1]), 0.5 * (pa[2] + pb[2])};
}
int edge_between(int c0, int c1) {
    int axis = __builtin_ctz((unsigned)(c0 ^ c1));
    int lo = std::min(c0, c1);
    int k = 0;
    while (EDGE_LOWER[axis][k] != lo) k++;
    return axis * 4 + k;
}
void face_corners(int f, int cs[4]) {
    int a = f / 2, s = f % 2;
    int others[2], n = 0;
    for (int q = 0; q < 3; q++)
        if (q != a) others[n++] = q;
    int b = others[0], c = others[1];
    const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int i = 0; i < 4; i++) cs[i] = (s << a) | (uv[i][0] << b) | (uv[i][1] << c);
}
void face_edges(int f, int fe[4]) {
    int cs[4];
    face_corners(f, cs);
    for (int i = 0; i < 4; i++) fe[i] = edge_between(cs[i], cs[(i + 1) % 4]);
}
P face_normal(int f) {
    P n = {0, 0, 0};
    n[f / 2] = f % 2 == 1 ? 1.0 : -1.0;
    return n;
}
bool inside(int cs, int c) { return ((cs >> c) & 1) == 1; }

std::vector<std::pair<int, int>> face_segments(int cs_, int f) {
    int cs[4];
    face_corners(f, cs);
    auto edge = [&](int i) { return edge_between(cs[i % 4], cs[(i + 1) % 4]); };
    std::vector<int> crossing;
    for (int i = 0; i < 4; i++)
        if (inside(cs_, cs[i]) != inside(cs_, cs[(i + 1) % 4])) crossing.push_back(edge(i));
    std::vector<std::pair<int, int>> raw;
    if (crossing.size() == 2) {
        raw.push_back({crossing[0], crossing[1]});
    } else if (crossing.size() == 4) {
        for (int i = 0; i < 4; i++)
            if (inside(cs_, cs[i])) raw.push_back({edge(i + 3), edge(i)});
    }
    P n = face_normal(f);
    std::vector<std::pair<int, int>> out;
    for (auto [p, q] : raw) {
        P t = pcross(n, psub(edge_mid(q), edge_mid(p)));
        int lo, hi;
        edge_corners(p, lo, hi);
        P o = psub(corner_pos(hi), corner_pos(lo));
        double s = inside(cs_, lo) ? 1.0 : -1.0;
        if (pdot(t, o) * s > 0.0)
            out.push_back({p, q});
        else
            out.push_back({q, p});
    }
    return out;
}

using Tri = std::array<int, 3>;

std::vector<std::vector<Tri>> triangulations(int i, int j) {
    if (j < i + 2) return {std::vector<Tri>()};
    std::vector<std::vector<Tri>> all;
    for (int k = i + 1; k < j; k++) {
        auto left = triangulations(i, k);
        auto right = triangulations(k, j);
        for (auto& l : left)
            for (auto& r : right) {
                std::vector<Tri> t = {{i, k, j}};
                t.insert(t.end(), l.begin(), l.end());
                t.insert(t.end(), r.begin(), r.end());
                all.push_back(t);
            }
    }
    return all;
}

double min_angle(P a, P b, P c) {
    auto ang = [](P p, P q, P r) {
        P u = psub(q, p), v = psub(r, p);
        double x = pdot(u, v) / (pnorm(u) * pnorm(v));
        x = std::fmin(std::fmax(x, -1.0), 1.0);
        return std::acos(x);
    };
    return std::fmin(std::fmin(ang(a, b, c), ang(b, c, a)), ang(c, a, b));
}

bool on_one_face(int a, int b) {
    for (int f = 0; f < 6; f++) {
        int fe[4];
        face_edges(f, fe);
        bool ha = false, hb = false;
        for (int q = 0; q < 4; q++) {
            ha |= fe[q] == a;
            hb |= fe[q] == b;
        }
        if (ha && hb) return true;
    }
    return false;
}

std::vector<Tri> best_triangulation(int cs, const std::vector<int>& ring) {
    int n = (int)ring.size();
    bool have = false;
    double bs0 = 0, bs1 = 0, bs2 = 0;
    std::vector<Tri> best;
    for (auto& tri : triangulations(0, n - 1)) {
        bool diag = false;
        for (auto& t : tri)
            for (int k = 0; k < 3 && !diag; k++) {
                int i = std::min(t[k], t[(k + 1) % 3]), j = std::max(t[k], t[(k + 1) % 3]);
                bool interior = j - i != 1 && !(i == 0 && j == n - 1);
                if (interior && on_one_face(ring[i], ring[j])) diag = true;
            }
        if (diag) continue;
        std::vector<Tri> mapped;
        for (auto& t : tri) mapped.push_back({ring[t[0]], ring[t[1]], ring[t[2]]});
        double acc[12][3] = {};
        double area = 0.0, angle = INFINITY;
        for (auto& t : mapped) {
            P pa = edge_mid(t[0]), pb = edge_mid(t[1]), pc = edge_mid(t[2]);
            P nrm = pcross(psub(pb, pa), psub(pc, pa));
            area += 0.5 * pnorm(nrm);
            angle = std::fmin(angle, min_angle(pa, pb, pc));
            for (int e : t)
                for (int k = 0; k < 3; k++) acc[e][k] += nrm[k];
        }
        int bad = 0;
        for (int e : ring) {
            int lo, hi;
            edge_corners(e, lo, hi);
            P o = psub(corner_pos(hi), corner_pos(lo));
            double s = inside(cs, lo) ? 1.0 : -1.0;
            if (pdot({acc[e][0], acc[e][1], acc[e][2]}, o) * s < 0.0) bad++;
        }
        double s0 = bad, s1 = std::round(area * 1e9), s2 = -angle;
        bool better = !have || s0 < bs0 || (s0 == bs0 && (s1 < bs1 || (s1 == bs1 && s2 < bs2 - 1e-12)));
        if (better) {
            have = true;
            bs0 = s0;
            bs1 = s1;
            bs2 = s2;
            best = mapped;
        }
    }
    return best;
}
}  // namespace

void generate_mc_table(McTable& T) {
    T.tris.clear();
    T.offset[0] = 0;
    for (int cs = 0; cs < 256; cs++) {
        int next[12];
        for (int& v : next) v = -1;
        for (int f = 0; f < 6; f++)
            for (auto [p, q] : face_segments(cs, f)) next[p] = q;
        bool seen[12] = {};
        for (int start = 0; start < 12; start++) {
            if (next[start] < 0 || seen[start]) continue;
            std::vector<int> ring;
            int e = start;
            while (!seen[e]) {
                seen[e] = true;
                ring.push_back(e);
                e = next[e];
            }
            for (auto& t : best_triangulation(cs, ring))
                T.tris.push_back({(uint8_t)t[0], (uint8_t)t[1], (uint8_t)t[2]});
        }
        T.offset[cs + 1] = (uint16_t)T.tris.size();
    }
}

const McTable& mc_table() {
    static McTable table;
    static std::once_flag once;
    std::call_once(once, [] { generate_mc_table(table); });
    return table;
}

}  // namespace orc
