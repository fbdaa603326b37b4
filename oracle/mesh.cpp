// Oracle: per-brick Marching Cubes with exact f64 edge refinement, the cell
// pass, assembly, component sums and the pipeline driver (test
// infrastructure). Restates mesh/extract.rs, mesh/mod.rs:88-300 and
// surface.rs:953-1079 of the reference (crates/sesmesh/src/).
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <thread>
#include <unordered_map>

#include "oracle.hpp"

namespace orc {

// par.rs:66-68
size_t chunk_len(size_t n) { return std::min<size_t>(std::max<size_t>(n / 256, 1), 4096); }

// mesh/extract.rs:47-51
double edge_clamp(const Engine& e) {
    double cm = std::fmin(std::fmax(e.coord_max, 1.0), (double)FLT_MAX / 4.0);
    float m = (float)cm;
    uint32_t bits;
    std::memcpy(&bits, &m, 4);
    bits += 1;
    float next;
    std::memcpy(&next, &bits, 4);
    double ulp = (double)(next - m);
    return std::fmin(std::fmax(4.0 * ulp / e.lattice.h, T_MIN), 0.25);
}

namespace {

// mesh/extract.rs:539-549
bool gradient(const SesField& f, const V3& x, const NearEval& e, V3& out) {
    if (e.feature.kind == F_SOLVENT || e.feature.kind == F_BEYOND) return false;
    V3 w;
    if (f.kind == K_SES)
        w = e.closest - x;
    else if (e.feature.kind == F_SPHERE)
        w = x - f.arr.centers[e.feature.id];
    else
        return false;
    double len = w.norm();
    if (!(len > 1e-12)) return false;
    out = w / len;
    return true;
}

// mesh/extract.rs:553-597
void attributes(const Engine& eng, const V3& v, const NearEval& e, const V3& outward, V3& normal,
                uint8_t& patch, u32& atom) {
    const SesField& field = *eng.field;
    const Arrangement& a = field.arr;
    auto unit = [&](const V3& w) {
        double len = w.norm();
        return len > 1e-9 ? w / len : outward;
    };
    auto nearest = [&](const u32* atoms, int n) {
        double bd = INFINITY;
        u32 bk = atoms[0];
        for (int q = 0; q < n; q++) {
            u32 k = atoms[q];
            double d = v.dist(a.centers[k]) - a.radii[k];
            if (d < bd) {
                bd = d;
                bk = k;
            }
        }
        return a.orig[bk];
    };
    if (field.kind != K_SES && e.feature.kind == F_SPHERE) {
        normal = unit(v - a.centers[e.feature.id]);
        patch = CONVEX;
        atom = a.orig[e.feature.id];
        return;
    }
    switch (field.kind == K_SES ? e.feature.kind : (u32)F_BEYOND) {
        case F_SPHERE:
            normal = unit(e.closest - v);
            patch = CONVEX;
            atom = a.orig[e.feature.id];
            return;
        case F_CIRCLE: {
            const Circle& c = a.circles[e.feature.id];
            u32 at[2] = {c.i, c.j};
            normal = unit(e.closest - v);
            patch = SADDLE;
            atom = nearest(at, 2);
            return;
        }
        case F_VERTEX:
            normal = unit(e.closest - v);
            patch = CONCAVE;
            atom = nearest(a.vertices[e.feature.id].atoms, 3);
            return;
        default: {
            auto se = field.sphere_eval(v, field.probe + eng.band + 8.0);
            normal = outward;
            patch = CONVEX;
            atom = se.second < 0 ? 0 : a.orig[se.second];
            return;
        }
    }
}

// mesh/extract.rs:425-535 (exact refinement)
void vertex(const Engine& eng, const V3& pa, const V3& pb, double fa, double fb, double t_min,
            const LocalCandidates& local, BrickMesh& out) {
    const SesField& field = *eng.field;
    double band = eng.band;
    V3 d = pb - pa;
    double t = std::fmin(std::fmax(fa / (fa - fb), 0.0), 1.0);
    bool have_last = false;
    NearEval last{};
    u32 evals = 0;
    double t0 = 0.0, g0 = fa, t1 = 1.0, g1 = fb;
    int side = 0;
    bool converged = false;
    for (int it = 0; it < MAX_ITER; it++) {
        V3 x = pa + d * t;
        NearEval e = field.eval_local(x, band, local);
        evals++;
        last = e;
        have_last = true;
        double g = e.value;
        if (std::fabs(g) <= TOL) {
            converged = true;
            break;
        }
        if ((g < 0.0) == (g0 < 0.0)) {
            t0 = t;
            g0 = g;
            if (side == -1) g1 *= 0.5;
            side = -1;
        } else {
            t1 = t;
            g1 = g;
            if (side == 1) g0 *= 0.5;
            side = 1;
        }
        V3 n;
        double slope = gradient(field, pa + d * t, e, n) ? n.dot(d) : 0.0;
        double newton = t - g / slope;
        if (std::fabs(slope) > 1e-3 * d.norm() && newton > t0 && newton < t1) {
            t = newton;
        } else {
            double s = (t0 * g1 - t1 * g0) / (g1 - g0);
            t = (std::isfinite(s) && s > t0 && s < t1) ? s : 0.5 * (t0 + t1);
        }
    }
    if (!converged) {
        double len = d.norm();
        bool neg0 = g0 < 0.0;
        for (int it = 0; it < MAX_BISECT; it++) {
            t = 0.5 * (t0 + t1);
            NearEval e = field.eval_local(pa + d * t, band, local);
            evals++;
            last = e;
            have_last = true;
            if (std::fabs(e.value) <= TOL || (t1 - t0) * len <= 2.0 * TOL) break;
            if ((e.value < 0.0) == neg0)
                t0 = t;
            else
                t1 = t;
        }
    }
    V3 v = pa + d * std::fmin(std::fmax(t, t_min), 1.0 - t_min);
    if (!have_last) {
        evals++;
        last = field.eval_local(v, band, local);
    }
    V3 outward = (fa < 0.0 ? d : -d).normalized();
    V3 normal;
    uint8_t patch;
    u32 atom;
    attributes(eng, v, last, outward, normal, patch, atom);
    out.positions.push_back({(float)v.x, (float)v.y, (float)v.z});
    out.normals.push_back({(float)normal.x, (float)normal.y, (float)normal.z});
    out.atoms.push_back(atom);
    out.patches.push_back(patch);
    out.evals += evals;
    out.t.push_back(t);
    out.vevals.push_back(evals | (converged ? 0x80000000u : 0u));
}

}  // namespace

const uint8_t EDGE_LOWER[3][4] = {{0, 2, 4, 6}, {0, 1, 4, 5}, {0, 1, 2, 3}};

// mesh/extract.rs:146-223
void mesh_brick(const Engine& eng, const int32_t key[3], const float* values, BrickMesh& out) {
    const Lattice& lat = eng.lattice;
    size_t b = (size_t)lat.b, s = b + 1;
    size_t stride[3] = {1, s, s * s};
    int64_t origin[3];
    lat.brick_origin(key, origin);
    size_t sub = (b % 2 == 0 && b >= 4) ? b / 2 : b;
    size_t nsub = b / sub;
    double h = lat.h;
    V3 base = lat.point(origin);
    double rho = 0.5 * std::sqrt(3.0) * (double)sub * h * (1.0 + 1e-9) + 1e-9;
    LocalCandidates locals[8];
    bool have[8] = {false, false, false, false, false, false, false, false};
    double t_min = edge_clamp(eng);
    for (size_t z = 0; z < b; z++)
        for (size_t y = 0; y < b; y++)
            for (size_t x = 0; x < b; x++) {
                size_t i = (z * s + y) * s + x;
                float va = values[i];
                for (int axis = 0; axis < 3; axis++) {
                    float vb = values[i + stride[axis]];
                    if ((va < 0.0f) == (vb < 0.0f)) continue;
                    int64_t ia[3] = {origin[0] + (int64_t)x, origin[1] + (int64_t)y, origin[2] + (int64_t)z};
                    int64_t ib[3] = {ia[0], ia[1], ia[2]};
                    ib[axis] += 1;
                    V3 pa = lat.point(ia), pb = lat.point(ib);
                    size_t sb3[3] = {std::min(x / sub, nsub - 1), std::min(y / sub, nsub - 1),
                                     std::min(z / sub, nsub - 1)};
                    size_t sb = sb3[0] + 2 * sb3[1] + 4 * sb3[2];
                    if (!have[sb]) {
                        V3 center = base + V3((double)(sb3[0] * sub) + 0.5 * (double)sub,
                                              (double)(sb3[1] * sub) + 0.5 * (double)sub,
                                              (double)(sb3[2] * sub) + 0.5 * (double)sub) *
                                               h;
                        eng.field->gather_local(center, rho, eng.band, locals[sb]);
                        have[sb] = true;
                    }
                    vertex(eng, pa, pb, (double)va, (double)vb, t_min, locals[sb], out);
                    out.edges.push_back((u32)(((z * b + y) * b + x) * 3 + axis));
                }
            }
    cells(b, values, out);
}

// mesh/extract.rs:226-265
void cells(size_t b, const float* values, BrickMesh& out) {
    const McTable& T = mc_table();
    size_t s = b + 1;
    size_t corner_step[8];
    for (int c = 0; c < 8; c++) corner_step[c] = (c & 1) + ((c >> 1) & 1) * s + ((c >> 2) & 1) * s * s;
    for (size_t z = 0; z < b; z++)
        for (size_t y = 0; y < b; y++)
            for (size_t x = 0; x < b; x++) {
                size_t i0 = (z * s + y) * s + x;
                int cs = 0;
                for (int c = 0; c < 8; c++)
                    if (values[i0 + corner_step[c]] < 0.0f) cs |= 1 << c;
                if (cs == 0 || cs == 255) continue;
                for (int t = T.offset[cs]; t < T.offset[cs + 1]; t++) {
                    std::array<u32, 3> tri;
                    for (int q = 0; q < 3; q++) {
                        int e = T.tris[t][q];
                        int axis = e / 4, lo = EDGE_LOWER[axis][e % 4];
                        size_t l[3] = {x + (size_t)(lo & 1), y + (size_t)((lo >> 1) & 1), z + (size_t)((lo >> 2) & 1)};
                        u32 code = 0;
                        size_t lc[3] = {0, 0, 0};
                        for (int a = 0; a < 3; a++) {
                            if (l[a] == b)
                                code |= 1u << a;
                            else
                                lc[a] = l[a];
                        }
                        tri[q] = (code << CODE_SHIFT) | (u32)(((lc[2] * b + lc[1]) * b + lc[0]) * 3 + (size_t)axis);
                    }
                    out.tris.push_back(tri);
                }
            }
}

// mesh/extract.rs:606-691. `parts[c]` holds chunk c (bricks c*chunk..).
std::string assemble(const std::vector<std::array<int32_t, 3>>& bricks, std::vector<BrickMesh>& parts,
                     size_t chunk, Mesh& mesh) {
    std::vector<size_t> base;
    base.reserve(bricks.size());
    size_t total = 0;
    for (auto& p : parts) {
        for (size_t l = 0; l < p.vertex_end.size(); l++) base.push_back(total + (l == 0 ? 0 : p.vertex_end[l - 1]));
        total += p.edges.size();
    }
    if (total > UINT32_MAX) return "mesh exceeds u32 vertex indices";
    std::unordered_map<u64, u32> index;
    index.reserve(bricks.size() * 2);
    for (size_t i = 0; i < bricks.size(); i++) index[pack(bricks[i].data())] = (u32)i;
    auto edges_of = [&](size_t bi, size_t& n) -> const u32* {
        const BrickMesh& p = parts[bi / chunk];
        size_t l = bi % chunk;
        size_t s = l == 0 ? 0 : p.vertex_end[l - 1];
        n = p.vertex_end[l] - s;
        return p.edges.data() + s;
    };
    mesh = Mesh();
    for (size_t c = 0; c < parts.size(); c++) {
        const BrickMesh& part = parts[c];
        for (size_t l = 0; l < part.tri_end.size(); l++) {
            size_t bi = c * chunk + l;
            const auto& key = bricks[bi];
            int64_t cache[8];
            for (int q = 0; q < 8; q++) cache[q] = -1;
            cache[0] = (int64_t)bi;
            size_t t0 = l == 0 ? 0 : part.tri_end[l - 1];
            for (size_t t = t0; t < part.tri_end[l]; t++) {
                std::array<u32, 3> o;
                for (int q = 0; q < 3; q++) {
                    u32 r = part.tris[t][q];
                    u32 code = r >> CODE_SHIFT;
                    if (cache[code] < 0) {
                        int32_t nk[3];
                        for (int a = 0; a < 3; a++) nk[a] = key[a] + (int32_t)((code >> a) & 1);
                        auto it = index.find(pack(nk));
                        if (it == index.end()) return "triangle references an edge of a missing brick";
                        cache[code] = it->second;
                    }
                    size_t target = (size_t)cache[code];
                    u32 id = r & ID_MASK;
                    size_t n;
                    const u32* e = edges_of(target, n);
                    const u32* f = std::lower_bound(e, e + n, id);
                    if (f == e + n || *f != id) return "triangle references an edge without vertex";
                    o[q] = (u32)(base[target] + (size_t)(f - e));
                }
                mesh.triangles.push_back(o);
            }
        }
    }
    for (auto& p : parts) {
        mesh.positions.insert(mesh.positions.end(), p.positions.begin(), p.positions.end());
        mesh.normals.insert(mesh.normals.end(), p.normals.begin(), p.normals.end());
        mesh.vertex_atom.insert(mesh.vertex_atom.end(), p.atoms.begin(), p.atoms.end());
        mesh.vertex_patch.insert(mesh.vertex_patch.end(), p.patches.begin(), p.patches.end());
        mesh.t.insert(mesh.t.end(), p.t.begin(), p.t.end());
        mesh.vevals.insert(mesh.vevals.end(), p.vevals.begin(), p.vevals.end());
        mesh.evals += p.evals;
    }
    return "";
}

// mesh/mod.rs:181-300 and surface.rs:1160-1168: the stats of the measure
// stage (components, cavities, area, volume).
ComponentStats measure(const Mesh& m) {
    Measured M = measure_full(m.positions, m.normals, m.triangles);
    ComponentStats s;
    s.components = M.components.size();
    s.cavities = M.cavities;
    s.area = M.area;
    s.volume = M.volume;
    return s;
}

// Brick classification (field/kernel.rs:301-305, lattice.rs:105-183) with
// `threads` threads over the touched bricks; the order of the result is the
// sorted key order whatever the thread count.
std::string classify_bricks(Surface& S, int threads) {
    std::vector<u64> keys;
    std::string err = touched_bricks(S.field, S.engine.lattice, keys);
    if (!err.empty()) return err;
    S.touched = keys.size();
    if (threads < 1) threads = 1;
    std::vector<uint8_t> keep(keys.size());
    double rho = S.engine.lattice.brick_radius();
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; w++)
        pool.emplace_back([&] {
            for (;;) {
                size_t c = next.fetch_add(1024);
                if (c >= keys.size()) break;
                for (size_t i = c; i < std::min(keys.size(), c + 1024); i++) {
                    int32_t key[3];
                    unpack(keys[i], key);
                    keep[i] = uniform_sign(S.field, S.engine.lattice.brick_center(key), rho) == SS_NONE;
                }
            }
        });
    for (auto& t : pool) t.join();
    S.bricks.clear();
    for (size_t i = 0; i < keys.size(); i++)
        if (keep[i]) {
            int32_t key[3];
            unpack(keys[i], key);
            S.bricks.push_back({key[0], key[1], key[2]});
        }
    return "";
}

// surface.rs:1041-1079 (CPU back end, no memory guard, exact refinement)
std::string compute_surface(const std::vector<Sphere>& spheres, double probe, double spacing, size_t brick,
                            int threads, Surface& S, int kind) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    // SurfaceParams::effective_probe (surface.rs:411-417): 0 for vdW.
    std::string err = S.field.build(spheres, kind == K_VDW ? 0.0 : probe, kind);
    if (!err.empty()) return err;
    err = S.engine.init(&S.field, spacing, brick);
    if (!err.empty()) return err;
    auto t1 = clk::now();
    err = classify_bricks(S, threads);
    if (!err.empty()) return err;
    if (threads < 1) threads = 1;
    auto t2 = clk::now();
    size_t chunk = chunk_len(S.bricks.size());
    size_t nchunks = (S.bricks.size() + chunk - 1) / chunk;
    std::vector<BrickMesh> parts(nchunks);
    {
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (int w = 0; w < threads; w++)
            pool.emplace_back([&] {
                BrickScratch s = S.engine.scratch();
                for (;;) {
                    size_t c = next.fetch_add(1);
                    if (c >= nchunks) break;
                    BrickMesh& part = parts[c];
                    size_t end = std::min(S.bricks.size(), (c + 1) * chunk);
                    for (size_t i = c * chunk; i < end; i++) {
                        S.engine.evaluate(S.bricks[i].data(), s);
                        mesh_brick(S.engine, S.bricks[i].data(), s.values.data(), part);
                        part.vertex_end.push_back((u32)part.edges.size());
                        part.tri_end.push_back((u32)part.tris.size());
                    }
                }
            });
        for (auto& t : pool) t.join();
    }
    auto t3 = clk::now();
    err = assemble(S.bricks, parts, chunk, S.mesh);
    auto t4 = clk::now();
    auto sec = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    S.t_geometry = sec(t0, t1);
    S.t_classify = sec(t1, t2);
    S.t_mesh = sec(t2, t3);
    S.t_assemble = sec(t3, t4);
    return err;
}

}  // namespace orc
