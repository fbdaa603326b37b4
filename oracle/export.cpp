// Mesh writers of the oracle (test infrastructure only): Wavefront OBJ,
// binary PLY and MSMS .vert/.face, restated from the reference's exporters
// (crates/sesmesh/src/export/{fmt,obj,ply,msms}.rs).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "oracle.hpp"

namespace orc {

namespace {
const uint64_t POW10[10] = {1, 10, 100, 1000, 10000, 100000, 1000000, 10000000, 100000000, 1000000000};

void push_u64(std::string& b, uint64_t n) { b += std::to_string(n); }

// f32::clamp(0.0, 1.0): below 0 gives 0, above 1 gives 1, else the value
// itself (-0.0 and NaN pass through).
float clamp01(float c) { return c < 0.0f ? 0.0f : c > 1.0f ? 1.0f : c; }

std::string clean(const std::string& s) {
    std::string o = s;
    for (char& c : o)
        if (c == '\n' || c == '\r') c = ' ';
    return o;
}
}  // namespace

// export/fmt.rs:30-54: x with `prec` decimals; x * 10^prec is exact in f64
// for an f32, so rounding it half to even is the exact decimal rounding.
void push_fixed(std::string& b, float x, int prec) {
    double scaled = std::nearbyint((double)x * (double)POW10[prec]);  // round half to even
    if (!std::isfinite(scaled) || std::fabs(scaled) >= 1e18) {
        char tmp[512];
        std::snprintf(tmp, sizeof(tmp), "%.*f", prec, (double)x);
        b += tmp;
        return;
    }
    if (std::signbit(x)) b += '-';
    uint64_t m = (uint64_t)std::fabs(scaled);
    push_u64(b, m / POW10[prec]);
    if (prec > 0) {
        b += '.';
        uint64_t f = m % POW10[prec];
        char d[16];
        for (int k = prec - 1; k >= 0; k--) {
            d[k] = (char)('0' + f % 10);
            f /= 10;
        }
        b.append(d, prec);
    }
}

// export/obj.rs:138-224 (write_impl) with one group range per group.
std::string write_obj(const ExportMesh& m, const ObjOptions& o) {
    std::string w;
    w += "# sesmesh " + std::string(SESMESH_VERSION) + "\n";
    for (const auto& c : o.comments) w += "# " + clean(c) + "\n";
    w += "o " + clean(o.object_name) + "\n";
    const bool normals = o.normals && !m.normals.empty();
    auto vec3 = [&](const char* tag, const std::array<float, 3>& v, const std::array<float, 3>* rgb) {
        w += tag;
        for (int k = 0; k < 3; k++) {
            if (k) w += ' ';
            push_fixed(w, v[k], o.precision);
        }
        if (rgb)
            for (int k = 0; k < 3; k++) {
                w += ' ';
                push_fixed(w, clamp01((*rgb)[k]), 4);
            }
        w += '\n';
    };
    for (size_t i = 0; i < m.positions.size(); i++)
        vec3("v ", m.positions[i], o.vertex_colors.empty() ? nullptr : &o.vertex_colors[i]);
    if (normals)
        for (const auto& n : m.normals) vec3("vn ", n, nullptr);
    std::vector<std::pair<size_t, const std::string*>> bounds = {{0, nullptr}};
    for (const auto& g : o.groups) {
        size_t first = std::min(g.first, m.triangles.size());
        if (bounds.back().first == first) bounds.pop_back();
        bounds.push_back({first, &g.second});
    }
    for (size_t k = 0; k < bounds.size(); k++) {
        size_t end = k + 1 < bounds.size() ? bounds[k + 1].first : m.triangles.size();
        if (bounds[k].second) w += "g " + clean(*bounds[k].second) + "\n";
        for (size_t t = bounds[k].first; t < end; t++) {
            w += 'f';
            for (int q = 0; q < 3; q++) {
                uint64_t n = (uint64_t)m.triangles[t][q] + 1;
                w += ' ';
                push_u64(w, n);
                if (normals) {
                    w += "//";
                    push_u64(w, n);
                }
            }
            w += '\n';
        }
    }
    return w;
}

// export/ply.rs:97-182: little-endian binary records.
std::string write_ply(const ExportMesh& m, const PlyOptions& o) {
    const size_t n = m.positions.size();
    const bool normals = o.normals && !m.normals.empty();
    const bool colors = !o.vertex_colors.empty();
    const bool comps = !m.vertex_component.empty();
    std::string h = "ply\nformat binary_little_endian 1.0\n";
    h += "comment sesmesh " + std::string(SESMESH_VERSION) + "\n";
    for (const auto& c : o.comments) h += "comment " + clean(c) + "\n";
    h += "element vertex " + std::to_string(n) + "\n";
    for (const char* p : {"x", "y", "z"}) h += std::string("property float ") + p + "\n";
    if (normals)
        for (const char* p : {"nx", "ny", "nz"}) h += std::string("property float ") + p + "\n";
    if (colors)
        for (const char* p : {"red", "green", "blue"}) h += std::string("property uchar ") + p + "\n";
    h += "property uint atom\nproperty uchar patch\n";
    if (comps) h += "property uint component\n";
    h += "element face " + std::to_string(m.triangles.size()) +
         "\nproperty list uchar uint vertex_indices\nend_header\n";
    std::string w = h;
    auto put = [&](const void* p, size_t k) { w.append((const char*)p, k); };
    for (size_t i = 0; i < n; i++) {
        put(m.positions[i].data(), 12);
        if (normals) put(m.normals[i].data(), 12);
        if (colors)
            for (int k = 0; k < 3; k++) {
                float c = clamp01(o.vertex_colors[i][k]) * 255.0f;
                uint8_t u = (uint8_t)std::round(c);  // f32::round: half away from zero
                put(&u, 1);
            }
        put(&m.vertex_atom[i], 4);
        put(&m.vertex_patch[i], 1);
        if (comps) put(&m.vertex_component[i], 4);
    }
    for (const auto& t : m.triangles) {
        uint8_t three = 3;
        put(&three, 1);
        put(t.data(), 12);
    }
    return w;
}

// export/msms.rs:70-148.
namespace {
uint8_t msms_type(uint8_t patch) { return patch == 1 ? 1 : patch == 2 ? 2 : 3; }
void pad(std::string& w, const std::string& s, size_t width) {
    if (s.size() < width) w.append(width - s.size(), ' ');
    w += s;
}
std::string msms_head(const char* what, size_t count, const MsmsHeader& h) {
    char line[256];
    std::snprintf(line, sizeof(line), "%7zu %7zu %5.2f %5.2f\n", count, h.spheres, h.density, h.probe_radius);
    return "# MSMS solvent excluded surface " + std::string(what) + " for " + clean(h.description) + "\n#" + what +
           " #sphere density probe_r\n" + line;
}
}  // namespace

void write_msms(const ExportMesh& m, const MsmsHeader& h, std::string& vert, std::string& face) {
    auto comp = [&](size_t v) -> uint64_t {
        return v < m.vertex_component.size() ? (uint64_t)m.vertex_component[v] + 1 : 1;
    };
    vert = msms_head("vertices", m.positions.size(), h);
    for (size_t v = 0; v < m.positions.size(); v++) {
        for (int k = 0; k < 6; k++) {
            std::string x;
            push_fixed(x, k < 3 ? m.positions[v][k] : m.normals[v][k - 3], 3);
            pad(vert, x, 9);
            vert += ' ';
        }
        pad(vert, std::to_string(comp(v)), 7);
        vert += ' ';
        pad(vert, std::to_string((uint64_t)m.vertex_atom[v] + 1), 7);
        vert += ' ';
        pad(vert, std::to_string(msms_type(m.vertex_patch[v])), 2);
        vert += '\n';
    }
    face = msms_head("faces", m.triangles.size(), h);
    for (const auto& t : m.triangles) {
        uint8_t ty[3];
        for (int k = 0; k < 3; k++) ty[k] = msms_type(m.vertex_patch[t[k]]);
        uint8_t type = ty[0] == ty[1] && ty[1] == ty[2] ? ty[0] : 1;
        for (int k = 0; k < 3; k++) {
            pad(face, std::to_string((uint64_t)t[k] + 1), 6);
            face += ' ';
        }
        pad(face, std::to_string(type), 2);
        face += ' ';
        pad(face, std::to_string(comp(t[0])), 6);
        face += '\n';
    }
}

}  // namespace orc
