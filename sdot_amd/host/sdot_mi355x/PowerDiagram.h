// C++ host layer of the MI355X Laguerre-cell path: the reference's template API
// for this path, restated over the C-ABI (include/sdot_mi355x.h), so code written
// against sdot's SpZGrid / ConvexPolyhedronAssembly / get_der_integrals_wrt_weights /
// OptimalTransportSolver compiles against this header with the namespace changed.
//
// Reference interfaces mirrored (paths relative to the sdot tree):
//   SpZGrid<Pc>(max_diracs_per_cell), update(...)   src/sdot/PowerDiagram/Visitors/SpZGrid.h:37-39
//   ConvexPolyhedronAssembly<Pc>::add_box(...)      src/sdot/Domains/ConvexPolyhedronAssembly.h:26-28
//   ScaledImage<Pc>(min, max, data, sizes, nb_coeffs) src/sdot/Domains/ScaledImage.h:26
//   FunctionEnum::Unit, FunctionEnum::InBallW05     src/sdot/Integration/Unit.h, InBallW05.h
//   get_integrals(res, grid, bounds, pos, w, N, rf) src/sdot/PowerDiagram/get_integrals.h:13
//   get_centroids(grid, bounds, pos, w, N, rf, cb)  src/sdot/PowerDiagram/get_centroids.h:14
//   get_der_integrals_wrt_weights(...) -> int       src/sdot/PowerDiagram/get_der_integrals_wrt_weights.h:16
//   OptimalTransportSolver<Grid,Bounds>::solve(...) src/sdot/Solvers/OptimalTransportSolver.h:18-26
//
// Differences: the grid owns a device handle (one HIP stream); failures other
// than an empty cell throw sdot_mi355x::Error instead of aborting; the Newton
// stop rule can be the reference's max|dw| or the max relative mass error.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/sdot_mi355x.h"

namespace sdot_mi355x {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};

inline void check(sdot_pd *pd, int st, const char *what, bool allow_empty = false) {
    if (st == SDOT_OK || (allow_empty && st == SDOT_EMPTY_CELL)) return;
    throw Error(st, std::string(what) + ": " + sdot_status_string(st) + (pd ? std::string(" (") + sdot_pd_last_error(pd) + ")" : ""));
}

template <int dim>
struct Point {
    double v[dim];
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
using Point2 = Point<2>;
using Point3 = Point<3>;

namespace FunctionEnum {
struct Unit {
    static constexpr int ball = 0;
};
struct InBallW05 {   // clip each cell by the ball of radius sqrt(w_i)
    static constexpr int ball = 1;
};
}  // namespace FunctionEnum

/// Pc: the reference's traits struct (dim; TF = double; TI = index type).
template <class Pc>
class SpZGrid {
public:
    static constexpr int dim = Pc::dim;
    using TF = typename Pc::TF;
    using TI = typename Pc::TI;
    using Pt = Point<dim>;

    explicit SpZGrid(TI max_diracs_per_cell = 11, int precision = SDOT_FP64) : max_diracs_per_cell(max_diracs_per_cell) {
        check(nullptr, sdot_pd_create(&pd, dim, precision), "sdot_pd_create");
        // the device grid targets about half the reference's leaf bound per cell
        sdot_pd_set_tuning(pd, (int)(50 * (max_diracs_per_cell > 2 ? max_diracs_per_cell : 2) / 2), 0);
    }
    ~SpZGrid() { sdot_pd_destroy(pd); }
    SpZGrid(const SpZGrid &) = delete;
    SpZGrid &operator=(const SpZGrid &) = delete;

    /// SpZGrid::update (SpZGrid.tcc:24): uploads diracs; the device grid is rebuilt on the next pass.
    void update(const Pt *positions, const TF *weights, TI nb_diracs, bool positions_have_changed = true,
                bool weights_have_changed = true, bool = false) {
        if (positions_have_changed || n != (int64_t)nb_diracs) {
            check(pd, sdot_pd_set_diracs(pd, reinterpret_cast<const double *>(positions), weights, (int64_t)nb_diracs),
                  "update");
            n = (int64_t)nb_diracs;
        } else if (weights_have_changed) {
            check(pd, sdot_pd_set_weights(pd, weights), "update");
        }
    }

    sdot_pd *handle() const { return pd; }
    int64_t size() const { return n; }
    TI max_diracs_per_cell;

private:
    sdot_pd *pd = nullptr;
    int64_t n = 0;
};

/// ConvexPolyhedronAssembly<Pc> restricted to what the device path evaluates: one box.
template <class Pc>
class ConvexPolyhedronAssembly {
public:
    static constexpr int dim = Pc::dim;
    using Pt = Point<dim>;
    using TF = double;

    void add_box(Pt p0, Pt p1, TF coeff = 1.0, std::int64_t cut_id = -1) { boxes.push_back({p0, p1, coeff, cut_id}); }
    TF measure() const {
        TF r = 0;
        for (const auto &b : boxes) {
            TF v = b.coeff;
            for (int d = 0; d < dim; ++d) v *= b.p1[d] - b.p0[d];
            r += v;
        }
        return r;
    }
    void normalize() {
        TF m = measure();
        if (m)
            for (auto &b : boxes) b.coeff /= m;
    }
    void apply(sdot_pd *pd) const {
        check(pd, sdot_pd_clear_shapes(pd), "clear_shapes");
        for (const auto &b : boxes) check(pd, sdot_pd_add_box_shape(pd, b.p0.v, b.p1.v, b.coeff, b.cut_id), "add_box");
    }

private:
    struct Box {
        Pt p0, p1;
        TF coeff;
        std::int64_t cut_id;
    };
    std::vector<Box> boxes;
};

/// ScaledImage<Pc>: piecewise-constant density on a box, x fastest.
template <class Pc>
class ScaledImage {
public:
    static constexpr int dim = Pc::dim;
    using Pt = Point<dim>;
    using TI = typename Pc::TI;
    ScaledImage(Pt min_pt, Pt max_pt, const double *data, std::array<TI, dim> sizes, TI nb_coeffs = 1)
        : min_pt(min_pt), max_pt(max_pt), nb_coeffs(nb_coeffs) {
        std::size_t tot = 1;
        for (int d = 0; d < dim; ++d) {
            this->sizes[d] = (std::int64_t)sizes[d];
            tot *= sizes[d];
        }
        this->data.assign(data, data + tot * nb_coeffs);
    }
    double measure() const {
        double s = 0;
        for (double v : data) s += v;
        s /= data.size();
        for (int d = 0; d < dim; ++d) s *= max_pt[d] - min_pt[d];
        return s;
    }
    void apply(sdot_pd *pd) const {
        check(pd, sdot_pd_clear_shapes(pd), "clear_shapes");
        check(pd, sdot_pd_set_image_density(pd, min_pt.v, max_pt.v, data.data(), sizes, (std::int64_t)nb_coeffs),
              "set_image_density");
    }

private:
    Pt min_pt, max_pt;
    std::int64_t sizes[3] = {1, 1, 1};
    TI nb_coeffs;
    std::vector<double> data;
};

/// get_integrals (get_integrals.h:13): res[i] = integral of the density over cell i.
template <class TF, class Grid, class Bounds, class Pt, class Func>
void get_integrals(TF *res, Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights, std::size_t nb_diracs,
                   const Func &) {
    grid.update(positions, weights, nb_diracs, false, true);
    bounds.apply(grid.handle());
    check(grid.handle(), sdot_pd_set_ball_clip(grid.handle(), Func::ball), "set_ball_clip");
    check(grid.handle(), sdot_pd_integrals(grid.handle(), res), "get_integrals", true);
}

template <class TF, class Grid, class Bounds, class Pt>
void get_integrals(TF *res, Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights, std::size_t nb_diracs) {
    get_integrals(res, grid, bounds, positions, weights, nb_diracs, FunctionEnum::Unit());
}

/// get_centroids (get_centroids.h:14, rand_ratio = 0): cb(centroid, mass, num_dirac) for
/// every cell, in dirac order; the centroid of an empty cell is 0 (get_centroids.h:58).
template <class Grid, class Bounds, class Pt, class TF, class TI, class Func, class CB>
void get_centroids(Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights, TI nb_diracs, const Func &,
                   const CB &cb, TF rand_ratio = TF(0)) {
    if (rand_ratio != TF(0)) throw Error{SDOT_NOT_SUPPORTED, "get_centroids: rand_ratio (random points) is not offered"};
    constexpr int dim = sizeof(Pt) / sizeof(double);
    grid.update(positions, weights, nb_diracs, false, true);
    bounds.apply(grid.handle());
    check(grid.handle(), sdot_pd_set_ball_clip(grid.handle(), Func::ball), "set_ball_clip");
    std::vector<double> c((std::size_t)nb_diracs * dim), m(nb_diracs);
    check(grid.handle(), sdot_pd_centroids(grid.handle(), c.data(), m.data()), "get_centroids", true);
    for (std::size_t i = 0; i < (std::size_t)nb_diracs; ++i) {
        Pt p;
        for (int d = 0; d < dim; ++d) p[d] = c[dim * i + d];
        cb(p, (TF)m[i], i);
    }
}

template <class Grid, class Bounds, class Pt, class TF, class TI, class CB>
void get_centroids(Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights, TI nb_diracs, const CB &cb) {
    get_centroids(grid, bounds, positions, weights, nb_diracs, FunctionEnum::Unit(), cb);
}

/// get_der_integrals_wrt_weights (get_der_integrals_wrt_weights.h:16): CSR of d mass / d w
/// (rows in dirac order, sorted columns, diagonal present) and the masses; returns 1 on an
/// empty cell (the outputs are filled for the other cells), 0 otherwise.
template <class TI, class TF, class Grid, class Bounds, class Pt, class Func>
int get_der_integrals_wrt_weights(std::vector<TI> &m_offsets, std::vector<TI> &m_columns, std::vector<TF> &m_values,
                                  std::vector<TF> &v_values, Grid &grid, Bounds &bounds, const Pt *positions,
                                  const TF *weights, std::size_t nb_diracs, const Func &, bool stop_if_void = true) {
    grid.update(positions, weights, nb_diracs, false, true);
    sdot_pd *pd = grid.handle();
    bounds.apply(pd);
    check(pd, sdot_pd_set_ball_clip(pd, Func::ball), "set_ball_clip");
    std::int64_t nnz = 0;
    int st = sdot_pd_der_integrals_wrt_weights(pd, nullptr, nullptr, nullptr, nullptr, 0, &nnz);
    check(pd, st, "get_der_integrals_wrt_weights", true);
    std::vector<std::int64_t> off(nb_diracs + 1), col(nnz > 0 ? nnz : 1);
    m_values.resize(nnz > 0 ? nnz : 1);
    v_values.resize(nb_diracs);
    st = sdot_pd_der_integrals_wrt_weights(pd, off.data(), col.data(), m_values.data(), v_values.data(), nnz, &nnz);
    check(pd, st, "get_der_integrals_wrt_weights", true);
    m_offsets.assign(off.begin(), off.end());
    m_columns.assign(col.begin(), col.begin() + nnz);
    m_values.resize(nnz);
    return (st == SDOT_EMPTY_CELL && stop_if_void) ? 1 : 0;
}

template <class TI, class TF, class Grid, class Bounds, class Pt>
int get_der_integrals_wrt_weights(std::vector<TI> &m_offsets, std::vector<TI> &m_columns, std::vector<TF> &m_values,
                                  std::vector<TF> &v_values, Grid &grid, Bounds &bounds, const Pt *positions,
                                  const TF *weights, std::size_t nb_diracs) {
    return get_der_integrals_wrt_weights(m_offsets, m_columns, m_values, v_values, grid, bounds, positions, weights,
                                         nb_diracs, FunctionEnum::Unit());
}

/// OptimalTransportSolver (OptimalTransportSolver.h:18-26): Newton on the weights with
/// the reference's rules; the linear solves run as device FP64 Jacobi-PCG.
template <class Grid, class Bounds>
class OptimalTransportSolver {
public:
    using TF = double;
    OptimalTransportSolver(Grid *grid, Bounds *bounds) : grid(*grid), bounds(*bounds) {}

    /// stop_rule 1 (default, as the reference): max |dw| < tol; 0: max relative mass error < tol
    template <class Pt, class TI, class RadialFunc>
    int solve(const Pt *positions, TF *weights, TI nb_diracs, TF *masses, RadialFunc) {
        grid.update(positions, weights, nb_diracs, true, true);
        sdot_pd *pd = grid.handle();
        bounds.apply(pd);
        check(pd, sdot_pd_set_ball_clip(pd, RadialFunc::ball), "set_ball_clip");
        int st = sdot_ot_solve(pd, masses, tol, stop_rule, (int)max_nb_iter, weights, &stats);
        if (st != SDOT_OK && st != SDOT_NOT_CONVERGED) check(pd, st, "OptimalTransportSolver::solve");
        return st;
    }
    template <class Pt, class TI>
    int solve(const Pt *positions, TF *weights, TI nb_diracs, TF *masses) {
        return solve(positions, weights, nb_diracs, masses, FunctionEnum::Unit());
    }

    std::size_t max_nb_iter = 150;   // OptimalTransportSolver.tcc:16
    double tol = 1e-6;               // OptimalTransportSolver.tcc:111
    int stop_rule = 1;
    sdot_solve_stats stats{};

private:
    Grid &grid;
    Bounds &bounds;
};

}  // namespace sdot_mi355x
