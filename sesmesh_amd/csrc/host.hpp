// Host-side types of the MI355X mesh stage library.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace smg {

// A uniform cell grid in CSR form (reference crates/sesmesh/src/geom/cells.rs:9-18).
struct Grid {
    double origin[3] = {0, 0, 0};
    double cell = 1;
    uint32_t dims[3] = {1, 1, 1};
    std::vector<uint32_t> starts, items;
};

// The f64 SAS arrangement with the field's cell grids, SoA, internal order.
struct Arrangement {
    int kind = 0;  // 0 SES, 1 SAS, 2 vdW
    double probe = 0, max_expanded = 0, rho_max = 0;
    std::vector<double> center, radius, expanded;  // center: 3 per sphere
    std::vector<uint8_t> status;                    // 0 exposed, 1 buried, 2 contained
    std::vector<uint32_t> orig;
    std::vector<uint32_t> nb_off, nb_items;
    std::vector<uint32_t> circle_ij;  // 2 per circle
    std::vector<double> circle;       // 10 per circle
    std::vector<uint32_t> cb_off, cb_items;
    std::vector<double> vertex;  // 3 per SAS vertex
    std::vector<uint32_t> vertex_atoms;
    Grid sphere_grid, circle_grid, vertex_grid;
};

void build_grid(const double* pts, size_t n, double cell, Grid& g);
// The cell edge, dimensions and origin of the grid of n points with bounds
// lo..hi (n > 0); build_grid then files the points into cells.
void grid_shape(const double lo[3], const double hi[3], size_t n, double cell, Grid& g);
// `kind` 0 (SES) builds the circles and SAS vertices; 1 (SAS) and 2 (vdW,
// built with probe 0) only containment and neighbours, every sphere that is
// not contained being exposed.
std::string build_arrangement(const double* xyzr, size_t n, double probe, int kind, int threads, Arrangement& a);

}  // namespace smg
