// Reference-style driver (the shape of benchmarks/bench_solve_2D.cpp) written
// against the C++ host layer: 2D diracs uniform in the unit square, unit box
// domain, Newton solve to uniform target masses, then the Hessian CSR.
// Prints one JSON line; exits non-zero on failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sdot_mi355x/PowerDiagram.h"

struct Pc {
    enum { dim = 2, allow_ball_cut = 1, allow_translations = 0 };
    using TI = std::size_t;
    using TF = double;
    using CI = std::size_t;
};

// A domain of two items (ConvexPolyhedronAssembly.tcc:7-43): a regular hexagon of
// apothem 0.4 about (0.5, 0.5) from its face points and normals, density 1, and a
// triangle of density 2 inside it (density 3 there).  `measure` checks the assembly's measure
// against the closed forms on the host only; `shapes` also solves on the device.
static int shapes(bool device) {
    using namespace sdot_mi355x;
    ConvexPolyhedronAssembly<Pc> bounds;
    const double pi = 3.14159265358979323846, a = 0.4;
    std::vector<Point2> fp, fn;
    for (int k = 0; k < 6; ++k) {
        const double nx = std::cos(k * pi / 3), ny = std::sin(k * pi / 3);
        fp.push_back(Point2{{0.5 + a * nx, 0.5 + a * ny}});
        fn.push_back(Point2{{nx, ny}});
    }
    bounds.add_convex_polyhedron(fp, fn, 1.0);
    const double tri[6] = {0.3, 0.3, 0.7, 0.35, 0.5, 0.7};
    bounds.add_simplex(tri, 2.0);
    const double exact = 2 * std::sqrt(3.0) * a * a + 2.0 * 0.5 * (0.4 * 0.4 - 0.05 * 0.2);
    const double meas = bounds.measure();
    double mass_sum = 0, max_rel = 0;
    int st = 0, iters = 0;
    if (device) {
        const std::size_t n = 4000;
        std::vector<double> u(2 * n);
        sdot_rng_uniform(3, 2 * n, u.data());
        std::vector<Point2> positions(n);
        for (std::size_t i = 0; i < n; ++i) positions[i] = Point2{{0.22 + 0.56 * u[2 * i], 0.22 + 0.56 * u[2 * i + 1]}};
        std::vector<double> weights(n, 0.0), masses(n, meas / n), v(n);
        SpZGrid<Pc> grid(11);
        grid.update(positions.data(), weights.data(), n);
        get_integrals(v.data(), grid, bounds, positions.data(), weights.data(), n, FunctionEnum::Unit());
        for (double m : v) mass_sum += m;
        OptimalTransportSolver<SpZGrid<Pc>, ConvexPolyhedronAssembly<Pc>> solver(&grid, &bounds);
        solver.stop_rule = 0;
        solver.tol = 1e-6;
        st = solver.solve(positions.data(), weights.data(), n, masses.data());
        iters = solver.stats.newton_iters;
        get_integrals(v.data(), grid, bounds, positions.data(), weights.data(), n, FunctionEnum::Unit());
        for (std::size_t i = 0; i < n; ++i) max_rel = std::fmax(max_rel, std::fabs(v[i] / masses[i] - 1));
    }
    std::printf("{\"measure\": %.15f, \"exact\": %.15f, \"mass_sum\": %.15f, \"status\": %d, \"newton_iters\": %d, "
                "\"max_rel_mass_err\": %.3e}\n",
                meas, exact, mass_sum, st, iters, max_rel);
    bool ok = std::fabs(meas - exact) < 1e-13;
    if (device) ok = ok && std::fabs(mass_sum - exact) < 1e-12 && st == 0 && max_rel < 1e-6;
    return ok ? 0 : 1;
}

int main(int argc, char **argv) {
    using namespace sdot_mi355x;
    if (argc > 1 && std::string(argv[1]) == "measure") return shapes(false);
    if (argc > 1 && std::string(argv[1]) == "shapes") return shapes(true);
    const std::size_t n = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 10000;
    std::vector<double> u(2 * n);
    sdot_rng_uniform(0, 2 * n, u.data());
    std::vector<Point2> positions(n);
    for (std::size_t i = 0; i < n; ++i) positions[i] = Point2{{u[2 * i], u[2 * i + 1]}};
    std::vector<double> weights(n, 0.0), masses(n, 1.0 / n);

    SpZGrid<Pc> grid(11);
    ConvexPolyhedronAssembly<Pc> bounds;
    bounds.add_box({{0, 0}}, {{1, 1}});
    grid.update(positions.data(), weights.data(), n);

    OptimalTransportSolver<SpZGrid<Pc>, ConvexPolyhedronAssembly<Pc>> solver(&grid, &bounds);
    solver.stop_rule = 0;
    solver.tol = 1e-6;
    int st = solver.solve(positions.data(), weights.data(), n, masses.data());

    std::vector<std::size_t> offsets, columns;
    std::vector<double> values, v;
    int err = get_der_integrals_wrt_weights(offsets, columns, values, v, grid, bounds, positions.data(), weights.data(), n,
                                            FunctionEnum::Unit());
    double max_rel = 0, sum = 0;
    for (std::size_t i = 0; i < n; ++i) {
        max_rel = std::fmax(max_rel, std::fabs(v[i] * n - 1));
        sum += v[i];
    }
    // symmetric rows with zero sums (no ball)
    double max_rowsum = 0;
    for (std::size_t i = 0; i < n; ++i) {
        double s = 0;
        for (std::size_t k = offsets[i]; k < offsets[i + 1]; ++k) s += values[k];
        max_rowsum = std::fmax(max_rowsum, std::fabs(s));
    }
    std::vector<double> ball(n);
    get_integrals(ball.data(), grid, bounds, positions.data(), weights.data(), n, FunctionEnum::InBallW05());
    // centroids (get_centroids.h:14): the mass-weighted mean of the centroids is the
    // centre of the unit square
    double cm[2] = {0, 0}, csum = 0;
    get_centroids(grid, bounds, positions.data(), weights.data(), n, [&](const Point2 &c, double m, std::size_t) {
        cm[0] += m * c[0];
        cm[1] += m * c[1];
        csum += m;
    });
    const double cerr = std::fmax(std::fabs(cm[0] / csum - 0.5), std::fabs(cm[1] / csum - 0.5));
    // get_der_centroids_and_integrals_wrt_weight_and_positions: the mass rows of the
    // weight columns are the Hessian of the masses (same offsets up to the 3x layout)
    std::vector<std::size_t> joff, jcol;
    std::vector<double> jval, jv;
    const int jerr = get_der_centroids_and_integrals_wrt_weight_and_positions(joff, jcol, jval, jv, grid, bounds,
                                                                              positions.data(), weights.data(), n,
                                                                              FunctionEnum::Unit());
    double jdiff = 0;
    for (std::size_t i = 0; i < n && jerr == 0; ++i) {
        std::size_t k2 = joff[3 * i + 2];
        for (std::size_t k = offsets[i]; k < offsets[i + 1]; ++k, k2 += 3) {   // (mass i, weight col) entries
            jdiff = std::fmax(jdiff, std::fabs(jval[k2 + 2] - values[k]));
            if (jcol[k2 + 2] != 3 * columns[k] + 2) jdiff = 1e300;
        }
    }
    // display_vtk (display_vtk.h:13-25): one polygon per cell, the count read back from the file
    long vtk_polygons = -1;
    if (argc > 2) {
        VtkOutput<1> vo(VtkOutput<1>::CN{{"num_dirac"}});
        display_vtk(vo, grid, bounds, positions.data(), weights.data(), n);
        vo.save(argv[2]);
        if (std::FILE *f = std::fopen(argv[2], "r")) {
            char line[256];
            while (std::fgets(line, sizeof line, f))
                if (std::sscanf(line, "CELL_TYPES %ld", &vtk_polygons) == 1) break;
            std::fclose(f);
        }
    }
    std::printf("{\"status\": %d, \"der_status\": %d, \"newton_iters\": %d, \"max_rel_mass_err\": %.3e, \"mass_sum\": %.15f, "
                "\"nnz\": %zu, \"max_row_sum\": %.3e, \"ball_mass_0\": %.6e, \"centroid_mean_err\": %.3e, "
                "\"der_pos_status\": %d, \"der_pos_vs_hessian\": %.3e, \"vtk_polygons\": %ld}\n",
                st, err, solver.stats.newton_iters, max_rel, sum, columns.size(), max_rowsum, ball[0], cerr, jerr, jdiff,
                vtk_polygons);
    return (st == 0 && err == 0 && max_rel < 1e-6 && max_rowsum < 1e-9 && cerr < 1e-12 && jerr == 0 && jdiff < 1e-9) ? 0
                                                                                                                 : 1;
}
