// Seeded synthetic molecules of the measured workloads.
//
// Generator: SplitMix64 (as the reference's test clusters,
// crates/sesmesh-testkit/src/clusters.rs:27-55). A globule is a
// self-avoiding random walk with 1.5 A steps confined to a ball: each step
// tries up to 64 random directions from the current tip and takes the first
// that stays inside the ball and keeps every atom at least 1.25 A away; a
// tip with no such step hands over to a random earlier atom (branching).
// Elements C/N/O/S with probabilities 0.62/0.17/0.19/0.02 and Bondi radii
// 1.70/1.55/1.52/1.80 A (reference radii/tables.rs:25-41).
#include <array>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform(double lo, double hi) { return lo + (hi - lo) * (double)(next() >> 11) / 9007199254740992.0; }
    uint64_t below(uint64_t n) { return next() % n; }
    void unit(double u[3]) {
        for (;;) {
            double v[3] = {uniform(-1, 1), uniform(-1, 1), uniform(-1, 1)};
            double n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (n2 >= 0.01 && n2 <= 1.0) {
                double n = std::sqrt(n2);
                for (int a = 0; a < 3; a++) u[a] = v[a] / n;
                return;
            }
        }
    }
};

struct Hash {
    double cell;
    std::unordered_map<uint64_t, std::vector<uint32_t>> map;
    uint64_t key(int64_t x, int64_t y, int64_t z) const {
        return ((uint64_t)(x + (1 << 20)) << 42) | ((uint64_t)(y + (1 << 20)) << 21) | (uint64_t)(z + (1 << 20));
    }
    void add(const double* p, uint32_t i) {
        map[key((int64_t)std::floor(p[0] / cell), (int64_t)std::floor(p[1] / cell), (int64_t)std::floor(p[2] / cell))]
            .push_back(i);
    }
    bool clear_of(const std::vector<double>& xyz, const double* p, double dmin) const {
        int64_t c[3];
        for (int a = 0; a < 3; a++) c[a] = (int64_t)std::floor(p[a] / cell);
        for (int64_t dz = -1; dz <= 1; dz++)
            for (int64_t dy = -1; dy <= 1; dy++)
                for (int64_t dx = -1; dx <= 1; dx++) {
                    auto it = map.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (it == map.end()) continue;
                    for (uint32_t j : it->second) {
                        double x = xyz[3 * j] - p[0], y = xyz[3 * j + 1] - p[1], z = xyz[3 * j + 2] - p[2];
                        if (x * x + y * y + z * z < dmin * dmin) return false;
                    }
                }
        return true;
    }
};

double element_radius(SplitMix& g) {
    double u = g.uniform(0.0, 1.0);
    return u < 0.62 ? 1.70 : (u < 0.79 ? 1.55 : (u < 0.98 ? 1.52 : 1.80));
}

}  // namespace

extern "C" {

// `n` atoms of a globule of ball radius `radius` into xyzr (4 per atom).
void smg_synth_globule(uint64_t seed, size_t n, double radius, double* xyzr) {
    SplitMix g{seed};
    const double step = 1.5, dmin = 1.25;
    std::vector<double> xyz;
    xyz.reserve(3 * n);
    Hash grid{1.5, {}};
    double p0[3] = {0, 0, 0};
    xyz.insert(xyz.end(), p0, p0 + 3);
    grid.add(p0, 0);
    size_t tip = 0;
    while (xyz.size() / 3 < n) {
        bool placed = false;
        for (int attempt = 0; attempt < 64 && !placed; attempt++) {
            double u[3], p[3];
            g.unit(u);
            for (int a = 0; a < 3; a++) p[a] = xyz[3 * tip + a] + step * u[a];
            if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > radius * radius) continue;
            if (!grid.clear_of(xyz, p, dmin)) continue;
            uint32_t id = (uint32_t)(xyz.size() / 3);
            xyz.insert(xyz.end(), p, p + 3);
            grid.add(p, id);
            tip = id;
            placed = true;
        }
        if (!placed) tip = (size_t)g.below(xyz.size() / 3);
    }
    for (size_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) xyzr[4 * i + a] = xyz[3 * i + a];
        xyzr[4 * i + 3] = element_radius(g);
    }
}

// Two stacked rings of `per_ring` copies of one globule: copy k of ring r
// is the globule shifted by `ring_radius` along x, rotated by 2*pi*k/per_ring
// about z, and shifted by -+separation/2 along z.
void smg_synth_rings(uint64_t seed, size_t n_globule, double radius, int per_ring, double ring_radius,
                     double separation, double* xyzr) {
    std::vector<double> one(4 * n_globule);
    smg_synth_globule(seed, n_globule, radius, one.data());
    const double tau = 6.283185307179586;
    size_t at = 0;
    for (int ring = 0; ring < 2; ring++)
        for (int k = 0; k < per_ring; k++) {
            double th = tau * (double)k / (double)per_ring, c = std::cos(th), s = std::sin(th);
            double z0 = ring == 0 ? -0.5 * separation : 0.5 * separation;
            for (size_t i = 0; i < n_globule; i++, at++) {
                double x = one[4 * i] + ring_radius, y = one[4 * i + 1];
                xyzr[4 * at] = c * x - s * y;
                xyzr[4 * at + 1] = s * x + c * y;
                xyzr[4 * at + 2] = one[4 * i + 2] + z0;
                xyzr[4 * at + 3] = one[4 * i + 3];
            }
        }
}

// The 60 rotations of the icosahedral group, as the closure (breadth
// first, from the identity) of a 5-fold rotation about the vertex axis
// (0, 1, phi) and the 2-fold rotation about z of the icosahedron with
// vertices (0, +-1, +-phi) and their cyclic permutations.
static std::vector<std::array<double, 9>> icosahedral_rotations() {
    auto mul = [](const std::array<double, 9>& a, const std::array<double, 9>& b) {
        std::array<double, 9> c{};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) c[3 * i + j] += a[3 * i + k] * b[3 * k + j];
        return c;
    };
    auto axis_angle = [](double x, double y, double z, double th) {
        double n = std::sqrt(x * x + y * y + z * z);
        x /= n, y /= n, z /= n;
        double c = std::cos(th), s = std::sin(th), t = 1.0 - c;
        return std::array<double, 9>{t * x * x + c,     t * x * y - s * z, t * x * z + s * y,
                                     t * x * y + s * z, t * y * y + c,     t * y * z - s * x,
                                     t * x * z - s * y, t * y * z + s * x, t * z * z + c};
    };
    const double phi = 0.5 * (1.0 + std::sqrt(5.0)), tau = 6.283185307179586;
    std::array<double, 9> gens[2] = {axis_angle(0.0, 1.0, phi, tau / 5.0), axis_angle(0.0, 0.0, 1.0, tau / 2.0)};
    std::vector<std::array<double, 9>> group{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    for (size_t q = 0; q < group.size() && group.size() < 60; q++)
        for (auto& g : gens) {
            std::array<double, 9> r = mul(g, group[q]);
            bool seen = false;
            for (auto& e : group) {
                double d = 0;
                for (int i = 0; i < 9; i++) d = std::fmax(d, std::fabs(e[i] - r[i]));
                if (d < 1e-9) seen = true;
            }
            if (!seen) group.push_back(r);
        }
    return group;
}

// `copies` (at most 60) copies of one globule on a shell: copy k is the
// globule shifted by `shell_radius` along a fixed direction off every
// symmetry axis, then turned by icosahedral rotation k.
void smg_synth_shell(uint64_t seed, size_t n_globule, double radius, int copies, double shell_radius,
                     double* xyzr) {
    std::vector<double> one(4 * n_globule);
    smg_synth_globule(seed, n_globule, radius, one.data());
    std::vector<std::array<double, 9>> rot = icosahedral_rotations();
    double o[3] = {0.31, 0.17, 0.93}, on = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    for (int a = 0; a < 3; a++) o[a] *= shell_radius / on;
    size_t at = 0;
    for (int k = 0; k < copies && k < (int)rot.size(); k++)
        for (size_t i = 0; i < n_globule; i++, at++) {
            double p[3] = {one[4 * i] + o[0], one[4 * i + 1] + o[1], one[4 * i + 2] + o[2]};
            for (int a = 0; a < 3; a++)
                xyzr[4 * at + a] = rot[k][3 * a] * p[0] + rot[k][3 * a + 1] * p[1] + rot[k][3 * a + 2] * p[2];
            xyzr[4 * at + 3] = one[4 * i + 3];
        }
}

// Batch of `count` molecules: atoms log-uniform in [lo, hi], probe uniform
// in [1.2, 1.4], spacing from {0.2, 0.25, 0.5}, per-molecule seeds.
void smg_synth_batch_spec(uint64_t seed, size_t count, uint32_t lo, uint32_t hi, uint32_t* atoms, double* probe,
                          double* spacing, uint64_t* seeds) {
    SplitMix g{seed};
    const double h[3] = {0.2, 0.25, 0.5};
    for (size_t i = 0; i < count; i++) {
        double u = g.uniform(std::log((double)lo), std::log((double)hi + 1.0));
        uint32_t n = (uint32_t)std::floor(std::exp(u));
        atoms[i] = n < lo ? lo : (n > hi ? hi : n);
        probe[i] = g.uniform(1.2, 1.4);
        spacing[i] = h[g.below(3)];
        seeds[i] = g.next();
    }
}

}  // extern "C"
