// Oracle: lattice, brick classification, per-sub-block gathering and the f32
// lane kernels (test infrastructure). Restates field/lattice.rs and
// field/kernel.rs of the reference (crates/sesmesh/src/), every surface kind.
#include <algorithm>

#include "oracle.hpp"

namespace orc {

// field/lattice.rs:35-44 (`round` is half away from zero, as C round)
void Lattice::init(double hh, size_t bb, const V3& fc) {
    h = hh;
    b = (int64_t)bb;
    double c[3] = {fc.x, fc.y, fc.z};
    for (int a = 0; a < 3; a++) frame[a] = (int64_t)std::round(c[a] / h);
    h32 = (float)h;
}

// field/lattice.rs:93-100
u64 pack(const int32_t key[3]) {
    u64 k[3];
    for (int a = 0; a < 3; a++) k[a] = (u64)((int64_t)key[a] + KEY_BIAS);
    return k[0] | (k[1] << 21) | (k[2] << 42);
}
void unpack(u64 p, int32_t key[3]) {
    const int s[3] = {0, 21, 42};
    for (int a = 0; a < 3; a++) key[a] = (int32_t)((int64_t)((p >> s[a]) & 0x1fffff) - KEY_BIAS);
}

// field/lattice.rs:105-163 (the chunking does not change the sorted result)
std::string touched_bricks(const SesField& f, const Lattice& l, std::vector<u64>& keys) {
    const Arrangement& a = f.arr;
    keys.clear();
    for (size_t i = 0; i < a.centers.size(); i++) {
        if (a.status[i] == CONTAINED) continue;
        double r = a.expanded[i] + CLASSIFY_EPS;
        const V3& c = a.centers[i];
        int64_t lo[3] = {l.brick_of(c.x - r), l.brick_of(c.y - r), l.brick_of(c.z - r)};
        int64_t hi[3] = {l.brick_of(c.x + r), l.brick_of(c.y + r), l.brick_of(c.z + r)};
        for (int q = 0; q < 3; q++)
            if (std::llabs(lo[q]) >= KEY_BIAS - 1 || std::llabs(hi[q]) >= KEY_BIAS - 1)
                return "coordinates too far from the origin for this grid spacing";
        for (int64_t z = lo[2]; z <= hi[2]; z++)
            for (int64_t y = lo[1]; y <= hi[1]; y++)
                for (int64_t x = lo[0]; x <= hi[0]; x++) {
                    int32_t k[3] = {(int32_t)x, (int32_t)y, (int32_t)z};
                    keys.push_back(pack(k));
                }
        if (keys.size() >= (size_t(1) << 22)) {
            std::sort(keys.begin(), keys.end());
            keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    return "";
}

// field/lattice.rs:190-227
uint8_t uniform_sign(const SesField& f, const V3& center, double rho) {
    const Arrangement& a = f.arr;
    double margin = rho + CLASSIFY_EPS;
    double s_sas = INFINITY, s_vdw = INFINITY;
    a.cells.for_each_near(center, margin + a.max_expanded, [&](u32 k) {
        if (a.status[k] != CONTAINED) {
            double d = center.dist(a.centers[k]);
            s_sas = std::fmin(s_sas, d - a.expanded[k]);
            s_vdw = std::fmin(s_vdw, d - a.radii[k]);
        }
    });
    if (f.kind != K_SES) {  // expanded is the SAS radius, or the vdW radius (probe 0)
        if (s_sas > margin) return SS_OUT;
        if (s_sas < -margin) return SS_IN;
        return SS_NONE;
    }
    if (s_sas >= margin) return SS_OUT;
    if (s_vdw <= -margin) return SS_IN;
    double cutoff = f.probe + margin + CLASSIFY_EPS;
    double v = f.probe - f.ses_eval(center, cutoff).distance;
    if (v > margin) return SS_OUT;
    if (v < -margin) return SS_IN;
    return SS_NONE;
}

// field/kernel.rs:204-261
std::string Engine::init(SesField* f, double spacing, size_t brick) {
    if (!(std::isfinite(spacing) && spacing >= 0.01)) return "grid spacing must be at least 0.01 A";
    if (brick < 2 || brick > 32) return "brick size must be 2..=32 cells";
    field = f;
    const Arrangement& a = f->arr;
    V3 center;
    double half = 0.0;
    if (!a.centers.empty()) {
        V3 lo = a.centers[0], hi = a.centers[0];
        for (const V3& c : a.centers) {
            lo = lo.vmin(c);
            hi = hi.vmax(c);
        }
        V3 d = hi - lo;
        center = (lo + hi) * 0.5;
        half = 0.5 * std::fmax(std::fmax(d.x, d.y), d.z);
    }
    lattice.init(spacing, brick, center);
    double reach = half + 2.0 * a.max_expanded + 2.0 * f->probe + 4.0 * (double)brick * spacing;
    f32_ok = reach < F32_MAX_COORD && reach / spacing < (double)(1u << 23);
    band = std::sqrt(3.0) * spacing * 1.01;
    coord_max = std::fmax(std::fmax(std::fabs(center.x), std::fabs(center.y)), std::fabs(center.z)) + half +
                a.max_expanded + f->probe + 2.0 * band;
    return "";
}

// field/kernel.rs:453-457
void Engine::sub_layout(int64_t& sub, int64_t& nsub) const {
    int64_t b = lattice.b;
    sub = (b % 2 == 0 && b >= 4) ? b / 2 : b;
    nsub = b / sub;
}

BrickScratch Engine::scratch() const {
    size_t padded = (samples_per_brick() + 15) / 16 * 16;
    BrickScratch s;
    s.xs.assign(padded, 0);
    s.ys.assign(padded, 0);
    s.zs.assign(padded, 0);
    s.out.assign(padded, 0);
    s.values.assign(padded, 0);
    s.local_of.assign(field->arr.centers.size(), UINT32_MAX);
    return s;
}

// field/kernel.rs:301-305 and field/lattice.rs:168-183
std::string Engine::surface_bricks(size_t& touched, std::vector<std::array<int32_t, 3>>& bricks) const {
    std::vector<u64> keys;
    std::string err = touched_bricks(*field, lattice, keys);
    if (!err.empty()) return err;
    touched = keys.size();
    double rho = lattice.brick_radius();
    bricks.clear();
    for (u64 k : keys) {
        int32_t key[3];
        unpack(k, key);
        if (uniform_sign(*field, lattice.brick_center(key), rho) == SS_NONE) bricks.push_back({key[0], key[1], key[2]});
    }
    return "";
}

// field/kernel.rs:467-482
void Engine::sub_states(const int32_t key[3], uint8_t st[8]) const {
    int64_t sub, nsub;
    sub_layout(sub, nsub);
    for (int i = 0; i < 8; i++) st[i] = SS_NONE;
    if (nsub != 2) return;
    int64_t o[3];
    lattice.brick_origin(key, o);
    V3 base = lattice.point(o);
    double h = lattice.h;
    double rho = 0.5 * std::sqrt(3.0) * (double)sub * h;
    for (int i = 0; i < 8; i++) {
        double half = (double)sub * 0.5;
        auto off = [&](int bit) { return (double)((i >> bit) & 1) * (double)sub + half; };
        V3 center = base + V3(off(0), off(1), off(2)) * h;
        st[i] = uniform_sign(*field, center, rho);
    }
}

// field/kernel.rs:489-546 (f32 kernels; the f64 fallback of `!f32_ok` is the
// clamped exact value)
void Engine::assign(const int32_t key[3], const uint8_t st[8], BrickScratch& s) const {
    int64_t b = lattice.b, side = b + 1;
    int64_t origin[3];
    lattice.brick_origin(key, origin);
    float fband = (float)band;
    int64_t sub, nsub;
    sub_layout(sub, nsub);
    auto range = [&](int64_t v, int64_t& lo, int64_t& hi) {
        hi = std::min(v / sub, nsub - 1);
        lo = (v > 0 && v % sub == 0) ? std::min(v / sub - 1, nsub - 1) : hi;
    };
    for (auto& g : s.groups) g.clear();
    size_t i = 0;
    for (int64_t z = 0; z < side; z++) {
        int64_t z0, z1;
        range(z, z0, z1);
        for (int64_t y = 0; y < side; y++) {
            int64_t y0, y1;
            range(y, y0, y1);
            for (int64_t x = 0; x < side; x++) {
                int64_t x0, x1;
                range(x, x0, x1);
                int owner = -1;
                bool positive = true;
                for (int64_t sz = z0; sz <= z1 && owner < 0; sz++)
                    for (int64_t sy = y0; sy <= y1 && owner < 0; sy++)
                        for (int64_t sx = x0; sx <= x1; sx++) {
                            int sb = (int)(sx + 2 * sy + 4 * sz);
                            if (st[sb] == SS_NONE) {
                                owner = sb;
                                break;
                            }
                            positive = st[sb] == SS_OUT;
                        }
                if (owner < 0) {
                    s.values[i] = positive ? fband : -fband;
                } else if (f32_ok) {
                    s.groups[owner].push_back((u32)i);
                } else {
                    int64_t idx[3] = {origin[0] + x, origin[1] + y, origin[2] + z};
                    s.values[i] = (float)field->clamped_value(lattice.point(idx), band);
                }
                i++;
            }
        }
    }
}

void Gathered::clear() {
    for (auto* v : {&nx, &ny, &nz, &nr, &nr2, &nr2tol, &sx, &sy, &sz, &sr, &cx, &cy, &cz, &ux, &uy, &uz, &ex, &ey,
                    &ez, &rho, &vx, &vy, &vz})
        v->clear();
    n_inside = 0;
    s_off.assign(1, 0);
    s_blk.clear();
    c_off.assign(1, 0);
    c_blk.clear();
}

// field/kernel.rs:553-598
bool Engine::prepare_sub(const int32_t key[3], size_t sb, BrickScratch& s, size_t& padded) const {
    size_t m = s.groups[sb].size();
    if (m == 0) return false;
    int64_t b = lattice.b, side = b + 1;
    int64_t origin[3];
    lattice.brick_origin(key, origin);
    int64_t sub, nsub;
    sub_layout(sub, nsub);
    double h = lattice.h;
    V3 base = lattice.point(origin);
    double sub_rho = 0.5 * std::sqrt(3.0) * (double)sub * h, half = (double)sub * 0.5;
    V3 center;
    double rho;
    if (nsub == 2) {
        auto off = [&](int bit) { return (double)((sb >> bit) & 1) * (double)sub + half; };
        center = base + V3(off(0), off(1), off(2)) * h;
        rho = sub_rho;
    } else {
        center = lattice.brick_center(key);
        rho = lattice.brick_radius();
    }
    size_t su = (size_t)side;
    for (size_t j = 0; j < m; j++) {
        size_t i = s.groups[sb][j];
        size_t x = i % su, y = (i / su) % su, z = i / (su * su);
        int64_t idx[3] = {origin[0] + (int64_t)x, origin[1] + (int64_t)y, origin[2] + (int64_t)z};
        float l[3];
        lattice.local(idx, l);
        s.xs[j] = l[0];
        s.ys[j] = l[1];
        s.zs[j] = l[2];
    }
    padded = (m + 15) / 16 * 16;
    for (size_t k = m; k < padded; k++) {
        s.xs[k] = s.xs[m - 1];
        s.ys[k] = s.ys[m - 1];
        s.zs[k] = s.zs[m - 1];
    }
    if (field->kind == K_SES)
        gather_ses(center, rho, s);
    else
        gather_spheres(center, rho, band, s);
    return true;
}

// field/kernel.rs:604-643
void Engine::gather_near(const V3& b, double rho, double reach, BrickScratch& s) const {
    const Arrangement& a = field->arr;
    V3 frame = lattice.frame_origin();
    s.g.clear();
    s.touched.clear();
    Gathered& g = s.g;
    for (int pass = 0; pass < 2; pass++) {
        a.cells.for_each_near(b, reach + a.max_expanded, [&](u32 k) {
            double r = a.expanded[k];
            double d = b.dist(a.centers[k]);
            bool contains_sample = d < r + rho;
            if (a.status[k] == CONTAINED || d >= r + reach || contains_sample != (pass == 0)) return;
            s.local_of[k] = (u32)g.nx.size();
            s.touched.push_back(k);
            V3 c = a.centers[k] - frame;
            g.nx.push_back((float)c.x);
            g.ny.push_back((float)c.y);
            g.nz.push_back((float)c.z);
            float r32 = (float)r;
            g.nr.push_back(r32);
            g.nr2.push_back(r32 * r32);
            float rt = std::fmax(r32 - INSIDE_TOL, 0.0f);
            g.nr2tol.push_back(rt * rt);
        });
        if (pass == 0) g.n_inside = g.nx.size();
    }
}

// field/kernel.rs:645-652
void Engine::gather_spheres(const V3& b, double rho, double band_, BrickScratch& s) const {
    gather_near(b, rho, band_ + rho, s);
    for (u32 k : s.touched) s.local_of[k] = UINT32_MAX;
}

// field/kernel.rs:659-751
void Engine::gather_ses(const V3& b, double rho, BrickScratch& s) const {
    const Arrangement& a = field->arr;
    const SesField& f = *field;
    V3 frame = lattice.frame_origin();
    double reach = f.probe + band + rho;
    gather_near(b, rho, reach, s);
    Gathered& g = s.g;
    auto& order = s.order;
    auto sort_key = [](std::vector<std::pair<double, u32>>& v) {
        std::sort(v.begin(), v.end(), [](const std::pair<double, u32>& x, const std::pair<double, u32>& y) {
            if (total_lt(x.first, y.first)) return true;
            if (total_lt(y.first, x.first)) return false;
            return x.second < y.second;
        });
    };
    order.clear();
    for (u32 k : s.touched) {
        if (a.status[k] != EXPOSED) continue;
        double lb = std::fabs(b.dist(a.centers[k]) - a.expanded[k]);
        if (lb <= reach) order.push_back({lb, k});
    }
    sort_key(order);
    for (auto& o : order) {
        u32 k = o.second;
        size_t local = s.local_of[k];
        g.sx.push_back(g.nx[local]);
        g.sy.push_back(g.ny[local]);
        g.sz.push_back(g.nz[local]);
        g.sr.push_back(g.nr[local]);
        size_t nb;
        const u32* row = a.neighbors.row(k, nb);
        for (size_t t = 0; t < nb; t++) {
            u32 l = s.local_of[row[t]];
            if (l != UINT32_MAX) g.s_blk.push_back(l);
        }
        g.s_off.push_back((u32)g.s_blk.size());
    }
    order.clear();
    f.circle_cells.for_each_near(b, reach + f.rho_max, [&](u32 ci) {
        const Circle& c = a.circles[ci];
        V3 y = b - c.center;
        double along = y.dot(c.axis);
        double wl = (y - c.axis * along).norm();
        double dist = std::sqrt(along * along + (wl - c.radius) * (wl - c.radius));
        if (dist <= reach) order.push_back({dist, ci});
    });
    sort_key(order);
    for (auto& o : order) {
        const Circle& c = a.circles[o.second];
        V3 m = c.center - frame;
        g.cx.push_back((float)m.x);
        g.cy.push_back((float)m.y);
        g.cz.push_back((float)m.z);
        g.ux.push_back((float)c.axis.x);
        g.uy.push_back((float)c.axis.y);
        g.uz.push_back((float)c.axis.z);
        g.ex.push_back((float)c.e1.x);
        g.ey.push_back((float)c.e1.y);
        g.ez.push_back((float)c.e1.z);
        g.rho.push_back((float)c.radius);
        size_t nb;
        const u32* row = a.circle_blockers.row(o.second, nb);
        for (size_t t = 0; t < nb; t++) {
            u32 l = s.local_of[row[t]];
            if (l != UINT32_MAX) g.c_blk.push_back(l);
        }
        g.c_off.push_back((u32)g.c_blk.size());
    }
    f.vertex_cells.for_each_near(b, reach, [&](u32 vi) {
        const V3& p = a.vertices[vi].position;
        if (b.dist(p) <= reach) {
            V3 q = p - frame;
            g.vx.push_back((float)q.x);
            g.vy.push_back((float)q.y);
            g.vz.push_back((float)q.z);
        }
    });
    for (u32 k : s.touched) s.local_of[k] = UINT32_MAX;
}

// field/kernel.rs:837-952 for one lane. Lane operations are IEEE f32
// add/sub/mul/div/sqrt/min/compare in the order of the vector code.
float ses_lane(const Gathered& g, float x, float y, float z, float probe, float cutoff2) {
    uint64_t ops = 0;
    bool inside = false;
    for (size_t k = 0; k < g.n_inside; k++) {
        float dx = x - g.nx[k], dy = y - g.ny[k], dz = z - g.nz[k];
        inside |= (dx * dx + dy * dy + dz * dz) < g.nr2[k];
        ops += 9;
    }
    float result = probe;
    if (inside) {
        float best2 = cutoff2;
        for (size_t v = 0; v < g.vx.size(); v++) {
            float dx = x - g.vx[v], dy = y - g.vy[v], dz = z - g.vz[v];
            float d2 = dx * dx + dy * dy + dz * dz;
            best2 = d2 < best2 ? d2 : best2;
            ops += 9;
        }
        for (size_t si = 0; si < g.sx.size(); si++) {
            float cx = g.sx[si], cy = g.sy[si], cz = g.sz[si], r = g.sr[si];
            float dx = x - cx, dy = y - cy, dz = z - cz;
            float dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            float lb = std::fabs(dist - r);
            float lb2 = lb * lb;
            ops += 13;
            bool valid = lb2 < best2 && 1e-6f < dist;
            if (!valid) continue;
            float scale = r / dist;
            float qx = cx + dx * scale, qy = cy + dy * scale, qz = cz + dz * scale;
            ops += 7;
            for (u32 t = g.s_off[si]; t < g.s_off[si + 1] && valid; t++) {
                u32 k = g.s_blk[t];
                float ex = qx - g.nx[k], ey = qy - g.ny[k], ez = qz - g.nz[k];
                valid = (ex * ex + ey * ey + ez * ez) >= g.nr2tol[k];
                ops += 9;
            }
            if (valid) best2 = lb2;
        }
        for (size_t ci = 0; ci < g.cx.size(); ci++) {
            float mx = g.cx[ci], my = g.cy[ci], mz = g.cz[ci];
            float ux = g.ux[ci], uy = g.uy[ci], uz = g.uz[ci];
            float rho = g.rho[ci];
            float yx = x - mx, yy = y - my, yz = z - mz;
            float along = yx * ux + yy * uy + yz * uz;
            float wx = yx - ux * along, wy = yy - uy * along, wz = yz - uz * along;
            float wl = std::sqrt(wx * wx + wy * wy + wz * wz);
            float t = wl - rho;
            float d2 = along * along + t * t;
            ops += 25;
            bool valid = d2 < best2;
            if (!valid) continue;
            bool degenerate = wl < 1e-6f;
            float sc = rho / wl;
            float qx = degenerate ? mx + g.ex[ci] * rho : mx + wx * sc;
            float qy = degenerate ? my + g.ey[ci] * rho : my + wy * sc;
            float qz = degenerate ? mz + g.ez[ci] * rho : mz + wz * sc;
            ops += 7;
            for (u32 tt = g.c_off[ci]; tt < g.c_off[ci + 1] && valid; tt++) {
                u32 k = g.c_blk[tt];
                float ex = qx - g.nx[k], ey = qy - g.ny[k], ez = qz - g.nz[k];
                valid = (ex * ex + ey * ey + ez * ez) >= g.nr2tol[k];
                ops += 9;
            }
            if (valid) best2 = d2;
        }
        result = probe - std::sqrt(best2);
        ops += 2;
    }
    if (g_ops.on) {
        g_ops.f32_ops += ops;
        g_ops.samples++;
    }
    return result;
}

// field/kernel.rs:953-981 for one lane: min(band, min_k(|x - c_k| - R_k))
// over every gathered sphere, IEEE f32.
float sphere_lane(const Gathered& g, float x, float y, float z, float band) {
    float best = band;
    for (size_t k = 0; k < g.nx.size(); k++) {
        float dx = x - g.nx[k], dy = y - g.ny[k], dz = z - g.nz[k];
        float v = std::sqrt(dx * dx + dy * dy + dz * dz) - g.nr[k];
        best = std::fmin(best, v);
    }
    if (g_ops.on) {
        g_ops.f32_ops += 10 * (uint64_t)g.nx.size();
        g_ops.samples++;
    }
    return best;
}

// field/kernel.rs:415-449
void Engine::evaluate(const int32_t key[3], BrickScratch& s) const {
    uint8_t st[8];
    sub_states(key, st);
    assign(key, st, s);
    if (!f32_ok) return;
    float probe = (float)field->probe;
    float cutoff = (float)(field->probe + band);
    float cutoff2 = cutoff * cutoff;
    float fband = (float)band;
    size_t nsb = sub_blocks();
    for (size_t sb = 0; sb < nsb; sb++) {
        size_t padded;
        if (!prepare_sub(key, sb, s, padded)) continue;
        for (size_t j = 0; j < s.groups[sb].size(); j++) {
            s.values[s.groups[sb][j]] = field->kind == K_SES
                                            ? ses_lane(s.g, s.xs[j], s.ys[j], s.zs[j], probe, cutoff2)
                                            : sphere_lane(s.g, s.xs[j], s.ys[j], s.zs[j], fband);
        }
    }
}

}  // namespace orc
