// Marching Cubes case table of the mesh stage, generated at load time.
//
// The reference ships its table as generated data (crates/sesmesh/src/mesh/
// mc_table.rs) made by the generator in mesh/tables.rs:47-274; the cell pass
// must use the same triangulation for the same mesh. This generator applies
// the same rules: per case, face segments under the uniform face rule
// (inside corners on a diagonal face are cut off separately), oriented so
// the outside is on their right seen along the face normal, chained into
// loops; each loop takes the first triangulation in enumeration order with
// no diagonal on a cube face, then fewest inverted vertices, least area
// (rounded to 1e-9) and largest minimum angle.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace smg {

namespace {
const uint8_t LOWER[3][4] = {{0, 2, 4, 6}, {0, 1, 4, 5}, {0, 1, 2, 3}};
using P3 = std::array<double, 3>;

P3 corner(int c) { return {(double)(c & 1), (double)((c >> 1) & 1), (double)((c >> 2) & 1)}; }
int lower_of(int e) { return LOWER[e / 4][e % 4]; }
int upper_of(int e) { return lower_of(e) | (1 << (e / 4)); }
P3 mid(int e) {
    P3 a = corner(lower_of(e)), b = corner(upper_of(e));
    return {0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]), 0.5 * (a[2] + b[2])};
}
P3 minus(P3 a, P3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
double dotp(P3 a, P3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
P3 crossp(P3 a, P3 b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
double len(P3 a) { return std::sqrt(dotp(a, a)); }
bool in(int cs, int c) { return (cs >> c) & 1; }

int edge_of(int c0, int c1) {
    int axis = __builtin_ctz((unsigned)(c0 ^ c1)), lo = std::min(c0, c1), k = 0;
    while (LOWER[axis][k] != lo) k++;
    return 4 * axis + k;
}
// Corners of face f (axis f/2, side f%2) in cyclic order.
std::array<int, 4> face(int f) {
    int a = f / 2, s = f % 2, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
    return {(s << a), (s << a) | (1 << u), (s << a) | (1 << u) | (1 << v), (s << a) | (1 << v)};
}
bool same_face(int e0, int e1) {
    for (int f = 0; f < 6; f++) {
        auto c = face(f);
        bool h0 = false, h1 = false;
        for (int i = 0; i < 4; i++) {
            int e = edge_of(c[i], c[(i + 1) % 4]);
            h0 |= e == e0;
            h1 |= e == e1;
        }
        if (h0 && h1) return true;
    }
    return false;
}

using Tri = std::array<int, 3>;
void all_triangulations(int i, int j, std::vector<std::vector<Tri>>& out) {
    out.clear();
    if (j < i + 2) {
        out.push_back({});
        return;
    }
    for (int k = i + 1; k < j; k++) {
        std::vector<std::vector<Tri>> l, r;
        all_triangulations(i, k, l);
        all_triangulations(k, j, r);
        for (auto& a : l)
            for (auto& b : r) {
                std::vector<Tri> t{{i, k, j}};
                t.insert(t.end(), a.begin(), a.end());
                t.insert(t.end(), b.begin(), b.end());
                out.push_back(std::move(t));
            }
    }
}

double smallest_angle(P3 a, P3 b, P3 c) {
    auto ang = [](P3 p, P3 q, P3 r) {
        P3 u = minus(q, p), v = minus(r, p);
        return std::acos(std::fmin(std::fmax(dotp(u, v) / (len(u) * len(v)), -1.0), 1.0));
    };
    return std::fmin(std::fmin(ang(a, b, c), ang(b, c, a)), ang(c, a, b));
}

void triangulate(int cs, const std::vector<int>& ring, std::vector<Tri>& out) {
    int n = (int)ring.size();
    std::vector<std::vector<Tri>> cands;
    all_triangulations(0, n - 1, cands);
    bool have = false;
    double best[3] = {0, 0, 0};
    std::vector<Tri> pick;
    for (auto& tri : cands) {
        bool diag = false;
        for (auto& t : tri)
            for (int k = 0; k < 3; k++) {
                int i = std::min(t[k], t[(k + 1) % 3]), j = std::max(t[k], t[(k + 1) % 3]);
                if (j - i != 1 && !(i == 0 && j == n - 1) && same_face(ring[i], ring[j])) diag = true;
            }
        if (diag) continue;
        double acc[12][3] = {}, area = 0.0, angle = INFINITY;
        std::vector<Tri> m;
        for (auto& t : tri) {
            Tri e{ring[t[0]], ring[t[1]], ring[t[2]]};
            P3 a = mid(e[0]), b = mid(e[1]), c = mid(e[2]);
            P3 nn = crossp(minus(b, a), minus(c, a));
            area += 0.5 * len(nn);
            angle = std::fmin(angle, smallest_angle(a, b, c));
            for (int q : e)
                for (int k = 0; k < 3; k++) acc[q][k] += nn[k];
            m.push_back(e);
        }
        int bad = 0;
        for (int e : ring) {
            P3 out_dir = minus(corner(upper_of(e)), corner(lower_of(e)));
            double sgn = in(cs, lower_of(e)) ? 1.0 : -1.0;
            if (dotp({acc[e][0], acc[e][1], acc[e][2]}, out_dir) * sgn < 0.0) bad++;
        }
        double sc[3] = {(double)bad, std::round(area * 1e9), -angle};
        bool better = !have || sc[0] < best[0] ||
                      (sc[0] == best[0] && (sc[1] < best[1] || (sc[1] == best[1] && sc[2] < best[2] - 1e-12)));
        if (better) {
            have = true;
            for (int k = 0; k < 3; k++) best[k] = sc[k];
            pick = m;
        }
    }
    out.insert(out.end(), pick.begin(), pick.end());
}
}  // namespace

void make_mc_table(uint16_t offset[257], std::vector<uint8_t>& edges) {
    edges.clear();
    offset[0] = 0;
    size_t ntri = 0;
    for (int cs = 0; cs < 256; cs++) {
        int next[12];
        std::fill(next, next + 12, -1);
        for (int f = 0; f < 6; f++) {
            auto c = face(f);
            auto fe = [&](int i) { return edge_of(c[i % 4], c[(i + 1) % 4]); };
            std::vector<std::pair<int, int>> seg;
            std::vector<int> cross;
            for (int i = 0; i < 4; i++)
                if (in(cs, c[i]) != in(cs, c[(i + 1) % 4])) cross.push_back(fe(i));
            if (cross.size() == 2)
                seg.push_back({cross[0], cross[1]});
            else if (cross.size() == 4)
                for (int i = 0; i < 4; i++)
                    if (in(cs, c[i])) seg.push_back({fe(i + 3), fe(i)});
            P3 nrm = {0, 0, 0};
            nrm[f / 2] = f % 2 == 1 ? 1.0 : -1.0;
            for (auto [p, q] : seg) {
                P3 t = crossp(nrm, minus(mid(q), mid(p)));
                P3 o = minus(corner(upper_of(p)), corner(lower_of(p)));
                double sgn = in(cs, lower_of(p)) ? 1.0 : -1.0;
                if (dotp(t, o) * sgn > 0.0)
                    next[p] = q;
                else
                    next[q] = p;
            }
        }
        bool seen[12] = {};
        std::vector<Tri> tris;
        for (int s = 0; s < 12; s++) {
            if (next[s] < 0 || seen[s]) continue;
            std::vector<int> ring;
            for (int e = s; !seen[e]; e = next[e]) {
                seen[e] = true;
                ring.push_back(e);
            }
            triangulate(cs, ring, tris);
        }
        for (auto& t : tris)
            for (int k = 0; k < 3; k++) edges.push_back((uint8_t)t[k]);
        ntri += tris.size();
        offset[cs + 1] = (uint16_t)ntri;
    }
}

}  // namespace smg
