// Oracle C API for the Python tests and the CPU baseline of bench.py (test
// infrastructure only; the product never links this library).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>

#include "oracle.hpp"

using namespace orc;

namespace {
void set_err(char* err, size_t len, const std::string& s) {
    if (err && len) {
        std::snprintf(err, len, "%s", s.c_str());
    }
}
}  // namespace

extern "C" {

// Builds the field and engine of surface kind `kind` (0 SES, 1 SAS, 2 vdW),
// then (stage >= 2) classifies bricks and (stage >= 3) meshes them with
// `threads` threads. Returns NULL on error.
void* orc_surface_new(const double* xyzr, size_t n, double probe, double spacing, size_t brick, int threads,
                      int stage, int kind, char* err, size_t errlen) {
    std::vector<Sphere> spheres(n);
    for (size_t i = 0; i < n; i++) {
        spheres[i].c[0] = xyzr[4 * i];
        spheres[i].c[1] = xyzr[4 * i + 1];
        spheres[i].c[2] = xyzr[4 * i + 2];
        spheres[i].r = xyzr[4 * i + 3];
    }
    Surface* S = new Surface();
    std::string e;
    if (stage >= 3) {
        e = compute_surface(spheres, probe, spacing, brick, threads, *S, kind);
    } else {
        e = S->field.build(spheres, kind == K_VDW ? 0.0 : probe, kind);
        if (e.empty()) e = S->engine.init(&S->field, spacing, brick);
        if (e.empty() && stage >= 2) e = classify_bricks(*S, threads);
    }
    if (!e.empty()) {
        set_err(err, errlen, e);
        delete S;
        return nullptr;
    }
    return S;
}

void orc_surface_free(void* s) { delete static_cast<Surface*>(s); }

// [touched, bricks, vertices, triangles, evals, spheres, circles, sas_vertices,
//  contained, exposed, f32_ok, frame x, y, z]
void orc_surface_counts(void* p, int64_t out[14]) {
    Surface* S = static_cast<Surface*>(p);
    const Arrangement& a = S->field.arr;
    out[0] = (int64_t)S->touched;
    out[1] = (int64_t)S->bricks.size();
    out[2] = (int64_t)S->mesh.positions.size();
    out[3] = (int64_t)S->mesh.triangles.size();
    out[4] = (int64_t)S->mesh.evals;
    out[5] = (int64_t)a.centers.size();
    out[6] = (int64_t)a.circles.size();
    out[7] = (int64_t)a.vertices.size();
    out[8] = (int64_t)a.n_contained;
    out[9] = (int64_t)a.n_exposed;
    out[10] = S->engine.f32_ok ? 1 : 0;
    out[11] = S->engine.lattice.frame[0];
    out[12] = S->engine.lattice.frame[1];
    out[13] = S->engine.lattice.frame[2];
}

// [band, coord_max, edge_clamp, rho_max, max_expanded, extent]
void orc_surface_params(void* p, double out[6]) {
    Surface* S = static_cast<Surface*>(p);
    out[0] = S->engine.band;
    out[1] = S->engine.coord_max;
    out[2] = edge_clamp(S->engine);
    out[3] = S->field.rho_max;
    out[4] = S->field.arr.max_expanded;
    out[5] = S->field.extent;
}

void orc_surface_times(void* p, double out[4]) {
    Surface* S = static_cast<Surface*>(p);
    out[0] = S->t_geometry;
    out[1] = S->t_classify;
    out[2] = S->t_mesh;
    out[3] = S->t_assemble;
}

void orc_surface_bricks(void* p, int32_t* keys) {
    Surface* S = static_cast<Surface*>(p);
    for (size_t i = 0; i < S->bricks.size(); i++)
        for (int a = 0; a < 3; a++) keys[3 * i + a] = S->bricks[i][a];
}

void orc_surface_mesh(void* p, float* pos, float* nrm, uint32_t* atoms, uint8_t* patches, uint32_t* tris,
                      double* t, uint32_t* vevals) {
    Surface* S = static_cast<Surface*>(p);
    const Mesh& m = S->mesh;
    for (size_t i = 0; i < m.positions.size(); i++)
        for (int k = 0; k < 3; k++) {
            if (pos) pos[3 * i + k] = m.positions[i][k];
            if (nrm) nrm[3 * i + k] = m.normals[i][k];
        }
    for (size_t i = 0; i < m.positions.size(); i++) {
        if (atoms) atoms[i] = m.vertex_atom[i];
        if (patches) patches[i] = m.vertex_patch[i];
        if (t) t[i] = m.t[i];
        if (vevals) vevals[i] = m.vevals[i];
    }
    if (tris)
        for (size_t i = 0; i < m.triangles.size(); i++)
            for (int k = 0; k < 3; k++) tris[3 * i + k] = m.triangles[i][k];
}

// [components, cavities, area, volume]
void orc_surface_measure(void* p, double out[4]) {
    ComponentStats s = measure(static_cast<Surface*>(p)->mesh);
    out[0] = (double)s.components;
    out[1] = (double)s.cavities;
    out[2] = s.area;
    out[3] = s.volume;
}

// The measure stage of a mesh given as arrays (normals may be NULL), or of a
// Surface's mesh (orc_surface_measure_new). Returns a handle to read with
// orc_measure_* and free with orc_measure_free.
void* orc_measure_new(const float* pos, size_t n, const float* nrm, size_t n_nrm, const uint32_t* tris, size_t t) {
    std::vector<std::array<float, 3>> P(n), N(nrm ? n_nrm : 0);
    std::vector<std::array<u32, 3>> T(t);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) P[i][k] = pos[3 * i + k];
    for (size_t i = 0; i < N.size(); i++)
        for (int k = 0; k < 3; k++) N[i][k] = nrm[3 * i + k];
    for (size_t i = 0; i < t; i++)
        for (int k = 0; k < 3; k++) T[i][k] = tris[3 * i + k];
    return new Measured(measure_full(P, N, T));
}

void* orc_surface_measure_new(void* p) {
    const Mesh& m = static_cast<Surface*>(p)->mesh;
    return new Measured(measure_full(m.positions, m.normals, m.triangles));
}

void orc_measure_free(void* h) { delete static_cast<Measured*>(h); }

// validation[7]: boundary, non-manifold, misoriented, degenerate, duplicate,
// invalid, bad vertices; stats[4]: area, volume, components, cavities;
// counts[2]: euler, indices valid.
void orc_measure_get(void* h, uint64_t validation[7], double stats[4], int64_t counts[2]) {
    const Measured& M = *static_cast<Measured*>(h);
    const Validation& v = M.validation;
    uint64_t vals[7] = {v.boundary_edges, v.non_manifold_edges, v.misoriented_edges, v.degenerate_triangles,
                        v.duplicate_triangles, v.invalid_indices, v.bad_vertices};
    std::copy(vals, vals + 7, validation);
    stats[0] = M.area;
    stats[1] = M.volume;
    stats[2] = (double)M.components.size();
    stats[3] = (double)M.cavities;
    counts[0] = M.euler;
    counts[1] = M.indices_valid;
}

// Per component: vt[2] vertices, triangles; av[2] area, volume; euler; first triangle.
void orc_measure_components(void* h, uint64_t* vt, double* av, int64_t* euler, uint64_t* first) {
    const Measured& M = *static_cast<Measured*>(h);
    for (size_t c = 0; c < M.components.size(); c++) {
        const Component& k = M.components[c];
        vt[2 * c] = k.vertices;
        vt[2 * c + 1] = k.triangles;
        av[2 * c] = k.area;
        av[2 * c + 1] = k.volume;
        euler[c] = k.euler;
        first[c] = k.first_triangle;
    }
}

void orc_measure_arrays(void* h, uint32_t* vertex_component, uint32_t* grouped) {
    const Measured& M = *static_cast<Measured*>(h);
    if (vertex_component) std::copy(M.vertex_component.begin(), M.vertex_component.end(), vertex_component);
    if (grouped)
        for (size_t i = 0; i < M.grouped.size(); i++)
            for (int k = 0; k < 3; k++) grouped[3 * i + k] = M.grouped[i][k];
}

// Writers. The mesh is given as arrays (normals, atoms, patches, vertex
// components and colours may be NULL); strings are NUL-separated lists of
// `count` entries. The result is a handle to read with orc_bytes.
namespace {
ExportMesh export_mesh(const float* pos, const float* nrm, const uint32_t* atoms, const uint8_t* patches,
                       const uint32_t* vcomp, size_t n, const uint32_t* tris, size_t t) {
    ExportMesh m;
    m.positions.resize(n);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) m.positions[i][k] = pos[3 * i + k];
    if (nrm) {
        m.normals.resize(n);
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) m.normals[i][k] = nrm[3 * i + k];
    }
    m.vertex_atom.assign(n, 0);
    m.vertex_patch.assign(n, 0);
    if (atoms) m.vertex_atom.assign(atoms, atoms + n);
    if (patches) m.vertex_patch.assign(patches, patches + n);
    if (vcomp) m.vertex_component.assign(vcomp, vcomp + n);
    m.triangles.resize(t);
    for (size_t i = 0; i < t; i++)
        for (int k = 0; k < 3; k++) m.triangles[i][k] = tris[3 * i + k];
    return m;
}
std::vector<std::string> split(const char* s, size_t count) {
    std::vector<std::string> out;
    for (size_t i = 0; i < count; i++) {
        out.emplace_back(s);
        s += out.back().size() + 1;
    }
    return out;
}
std::vector<std::array<float, 3>> colors(const float* c, size_t n) {
    std::vector<std::array<float, 3>> out;
    if (c) {
        out.resize(n);
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) out[i][k] = c[3 * i + k];
    }
    return out;
}
}  // namespace

void* orc_write_obj(const float* pos, const float* nrm, size_t n, const uint32_t* tris, size_t t, int precision,
                    int normals, const char* comments, size_t n_comments, const char* object_name, const float* rgb,
                    const uint64_t* group_first, const char* group_names, size_t n_groups) {
    ExportMesh m = export_mesh(pos, nrm, nullptr, nullptr, nullptr, n, tris, t);
    ObjOptions o;
    o.precision = precision;
    o.normals = normals != 0;
    o.comments = split(comments, n_comments);
    if (object_name) o.object_name = object_name;
    o.vertex_colors = colors(rgb, n);
    std::vector<std::string> names = split(group_names, n_groups);
    for (size_t i = 0; i < n_groups; i++) o.groups.push_back({group_first[i], names[i]});
    return new std::string(write_obj(m, o));
}

void* orc_write_ply(const float* pos, const float* nrm, const uint32_t* atoms, const uint8_t* patches,
                    const uint32_t* vcomp, size_t n, const uint32_t* tris, size_t t, int normals, const char* comments,
                    size_t n_comments, const float* rgb) {
    ExportMesh m = export_mesh(pos, nrm, atoms, patches, vcomp, n, tris, t);
    PlyOptions o;
    o.normals = normals != 0;
    o.comments = split(comments, n_comments);
    o.vertex_colors = colors(rgb, n);
    return new std::string(write_ply(m, o));
}

// The .vert text, then the .face text (vert_len receives the first's length).
void* orc_write_msms(const float* pos, const float* nrm, const uint32_t* atoms, const uint8_t* patches,
                     const uint32_t* vcomp, size_t n, const uint32_t* tris, size_t t, const char* description,
                     size_t spheres, double density, double probe, size_t* vert_len) {
    ExportMesh m = export_mesh(pos, nrm, atoms, patches, vcomp, n, tris, t);
    MsmsHeader h{description, spheres, density, probe};
    std::string v, f;
    write_msms(m, h, v, f);
    *vert_len = v.size();
    return new std::string(v + f);
}

// push_fixed of n values (NUL after each), for the formatter's tests.
void* orc_fixed(const float* x, size_t n, int prec) {
    std::string* s = new std::string();
    for (size_t i = 0; i < n; i++) {
        push_fixed(*s, x[i], prec);
        *s += '\0';
    }
    return s;
}

size_t orc_bytes(void* h, char* out) {
    const std::string& s = *static_cast<std::string*>(h);
    if (out) std::memcpy(out, s.data(), s.size());
    return s.size();
}

void orc_bytes_free(void* h) { delete static_cast<std::string*>(h); }

// Samples (field/kernel.rs Engine::evaluate) of `count` bricks.
void orc_brick_values(void* p, const int32_t* keys, size_t count, float* values) {
    Surface* S = static_cast<Surface*>(p);
    BrickScratch s = S->engine.scratch();
    size_t spb = S->engine.samples_per_brick();
    for (size_t i = 0; i < count; i++) {
        S->engine.evaluate(keys + 3 * i, s);
        std::copy(s.values.begin(), s.values.begin() + spb, values + i * spb);
    }
}

void orc_sub_states(void* p, const int32_t* keys, size_t count, uint8_t* states) {
    Surface* S = static_cast<Surface*>(p);
    for (size_t i = 0; i < count; i++) S->engine.sub_states(keys + 3 * i, states + 8 * i);
}

// Brick classification of given centres with ball radius rho.
void orc_uniform_sign(void* p, const double* centers, size_t count, double rho, uint8_t* out) {
    Surface* S = static_cast<Surface*>(p);
    for (size_t i = 0; i < count; i++) {
        V3 c(centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]);
        out[i] = uniform_sign(S->field, c, rho);
    }
}

// SesField::sample: value; closest point; feature kind and id.
double orc_field_sample(void* p, const double x[3], double closest[3], uint32_t feat[2]) {
    Surface* S = static_cast<Surface*>(p);
    V3 c;
    Feature f;
    double v = S->field.sample(V3(x[0], x[1], x[2]), c, f);
    closest[0] = c.x;
    closest[1] = c.y;
    closest[2] = c.z;
    feat[0] = f.kind;
    feat[1] = f.id;
    return v;
}

// Arrangement arrays in internal order: centres (n*3), expanded (n), radii
// (n), status (n), orig (n).
void orc_arrangement(void* p, double* centers, double* expanded, double* radii, uint8_t* status, uint32_t* orig) {
    const Arrangement& a = static_cast<Surface*>(p)->field.arr;
    for (size_t i = 0; i < a.centers.size(); i++) {
        centers[3 * i] = a.centers[i].x;
        centers[3 * i + 1] = a.centers[i].y;
        centers[3 * i + 2] = a.centers[i].z;
        expanded[i] = a.expanded[i];
        radii[i] = a.radii[i];
        status[i] = a.status[i];
        orig[i] = a.orig[i];
    }
}

// Neighbour and circle-blocker CSR sizes and contents.
size_t orc_csr(void* p, int which, uint32_t* off, uint32_t* items) {
    const Arrangement& a = static_cast<Surface*>(p)->field.arr;
    const Csr& c = which == 0 ? a.neighbors : a.circle_blockers;
    if (off) std::copy(c.off.begin(), c.off.end(), off);
    if (items) std::copy(c.items.begin(), c.items.end(), items);
    return c.items.size();
}

// Circles: i, j (2 u32) and center, axis, e1, radius (10 f64) each.
void orc_circles(void* p, uint32_t* ij, double* geo) {
    const Arrangement& a = static_cast<Surface*>(p)->field.arr;
    for (size_t k = 0; k < a.circles.size(); k++) {
        const Circle& c = a.circles[k];
        ij[2 * k] = c.i;
        ij[2 * k + 1] = c.j;
        double g[10] = {c.center.x, c.center.y, c.center.z, c.axis.x, c.axis.y,
                        c.axis.z,   c.e1.x,     c.e1.y,     c.e1.z,   c.radius};
        for (int q = 0; q < 10; q++) geo[10 * k + q] = g[q];
    }
}

void orc_sas_vertices(void* p, double* pos, uint32_t* atoms) {
    const Arrangement& a = static_cast<Surface*>(p)->field.arr;
    for (size_t k = 0; k < a.vertices.size(); k++) {
        for (int q = 0; q < 3; q++) {
            pos[3 * k + q] = a.vertices[k].position[q];
            atoms[3 * k + q] = a.vertices[k].atoms[q];
        }
    }
}

void orc_mc_table(uint16_t offset[257], uint8_t* tris, size_t cap) {
    const McTable& t = mc_table();
    for (int i = 0; i < 257; i++) offset[i] = t.offset[i];
    for (size_t i = 0; i < t.tris.size() && i < cap; i++)
        for (int k = 0; k < 3; k++) tris[3 * i + k] = t.tris[i][k];
}

// CPU baseline: sample and mesh the bricks first + k * stride (k < count)
// of a stage >= 2 surface with `threads` threads over chunks of
// par::chunk_len bricks; returns wall seconds and the vertices made.
double orc_time_mesh_sample(void* p, size_t first, size_t count, size_t stride, int threads, uint64_t* vertices) {
    Surface* S = static_cast<Surface*>(p);
    std::vector<size_t> idx;
    for (size_t k = 0; k < count && first + k * stride < S->bricks.size(); k++) idx.push_back(first + k * stride);
    if (threads < 1) threads = 1;
    size_t n = idx.size(), chunk = chunk_len(n), nchunks = (n + chunk - 1) / chunk;
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> verts{0};
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; w++)
        pool.emplace_back([&] {
            BrickScratch s = S->engine.scratch();
            for (;;) {
                size_t c = next.fetch_add(1);
                if (c >= nchunks) break;
                BrickMesh part;
                for (size_t q = c * chunk; q < std::min(n, (c + 1) * chunk); q++) {
                    S->engine.evaluate(S->bricks[idx[q]].data(), s);
                    mesh_brick(S->engine, S->bricks[idx[q]].data(), s.values.data(), part);
                }
                verts += part.edges.size();
            }
        });
    for (auto& t : pool) t.join();
    auto t1 = std::chrono::steady_clock::now();
    if (vertices) *vertices = verts.load();
    return std::chrono::duration<double>(t1 - t0).count();
}

// Algorithmic op counts of sampling and refining the bricks first + k *
// stride (k < count), one thread: [f32 ops, evaluated samples, f64 ops of
// eval_local, evaluations, vertices].
// Op counts of one point x: the f32 lane kernel over the sampling lists of
// the ball (x, rho) and the f64 eval_local over the refinement lists of the
// same ball. out = [f32 ops, f64 ops].
void orc_count_point(void* p, const double x[3], double rho, uint64_t out[2]) {
    Surface* S = static_cast<Surface*>(p);
    const Engine& E = S->engine;
    BrickScratch s = E.scratch();
    V3 c(x[0], x[1], x[2]);
    E.gather_ses(c, rho, s);
    V3 l = c - E.lattice.frame_origin();
    float probe = (float)S->field.probe;
    float cut = (float)(S->field.probe + E.band);
    g_ops = OpCounts();
    g_ops.on = true;
    (void)ses_lane(s.g, (float)l.x, (float)l.y, (float)l.z, probe, cut * cut);
    LocalCandidates lc;
    S->field.gather_local(c, rho, E.band, lc);
    (void)S->field.eval_local(c, E.band, lc);
    g_ops.on = false;
    out[0] = g_ops.f32_ops;
    out[1] = g_ops.f64_ops;
}

void orc_count_ops(void* p, size_t first, size_t count, size_t stride, uint64_t out[5]) {
    Surface* S = static_cast<Surface*>(p);
    BrickScratch s = S->engine.scratch();
    BrickMesh part;
    g_ops = OpCounts();
    g_ops.on = true;
    for (size_t k = 0; k < count && first + k * stride < S->bricks.size(); k++) {
        size_t i = first + k * stride;
        S->engine.evaluate(S->bricks[i].data(), s);
        mesh_brick(S->engine, S->bricks[i].data(), s.values.data(), part);
    }
    g_ops.on = false;
    out[0] = g_ops.f32_ops;
    out[1] = g_ops.samples;
    out[2] = g_ops.f64_ops;
    out[3] = g_ops.evals;
    out[4] = part.edges.size();
}

}  // extern "C"
