// C++ host layer of the MI355X Laguerre-cell path: the reference's template API
// for this path, restated over the C-ABI (include/sdot_mi355x.h), so code written
// against sdot's SpZGrid / ConvexPolyhedronAssembly / get_der_integrals_wrt_weights /
// OptimalTransportSolver compiles against this header with the namespace changed.
//
// Reference interfaces mirrored (paths relative to the sdot tree):
//   SpZGrid<Pc>(max_diracs_per_cell), update(...), translations
//                                                   src/sdot/PowerDiagram/Visitors/SpZGrid.h:37-53
//   ConvexPolyhedronAssembly<Pc>::add_convex_polyhedron / add_simplex / add_box / measure / normalize
//                                                   src/sdot/Domains/ConvexPolyhedronAssembly.h:26-40
//   ScaledImage<Pc>(min, max, data, sizes, nb_coeffs) src/sdot/Domains/ScaledImage.h:26
//   FunctionEnum::Unit, InBallW05, R2, R2InBallW05, WmR2, PpWmR2, ExpWmR2db<TF>{eps}
//                                                   src/sdot/Integration/*.h
//   get_integrals(res, grid, bounds, pos, w, N, rf) src/sdot/PowerDiagram/get_integrals.h:13
//   get_centroids(grid, bounds, pos, w, N, rf, cb)  src/sdot/PowerDiagram/get_centroids.h:14
//   get_der_integrals_wrt_weights(...) -> int       src/sdot/PowerDiagram/get_der_integrals_wrt_weights.h:16
//   get_der_centroids_and_integrals_wrt_weight_and_positions(...) -> int
//                         src/sdot/PowerDiagram/get_der_centroids_and_integrals_wrt_weight_and_positions.h:16
//   OptimalTransportSolver<Grid,Bounds>::solve(...) src/sdot/Solvers/OptimalTransportSolver.h:18-26
//
// Differences: the grid owns a device handle (one HIP stream); failures other
// than an empty cell throw sdot_mi355x::Error instead of aborting; the Newton
// stop rule can be the reference's max|dw| or the max relative mass error.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdio>
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
// `code`: the sdot_pd_set_radial_func value; `ball`: need_ball_cut
struct Unit {
    static constexpr int ball = 0, code = SDOT_FUNC_UNIT;
};
struct InBallW05 {   // clip each cell by the ball of radius sqrt(w_i)
    static constexpr int ball = 1, code = SDOT_FUNC_IN_BALL_W05;
};
struct R2 {          // |x - x_i|^2 (integrals only)
    static constexpr int ball = 0, code = SDOT_FUNC_R2;
};
struct R2InBallW05 {
    static constexpr int ball = 1, code = SDOT_FUNC_R2_IN_BALL_W05;
};
struct WmR2 {        // w_i - |x - x_i|^2
    static constexpr int ball = 0, code = SDOT_FUNC_WMR2;
};
struct PpWmR2 {      // its positive part (ball cut)
    static constexpr int ball = 1, code = SDOT_FUNC_PP_WMR2;
};
template <class TS>
struct ExpWmR2db {   // exp((w_i - |x - x_i|^2) / eps), no ball cut (ExpWmR2db.h), 2D
    static constexpr int ball = 0, code = SDOT_FUNC_EXP_WMR2_DB;
    TS eps;
};
}  // namespace FunctionEnum

/// the radial function on the handle (with its parameter for ExpWmR2db)
template <class Func>
inline int set_radial_func(sdot_pd *pd, const Func &) {
    return sdot_pd_set_radial_func(pd, Func::code);
}
template <class TS>
inline int set_radial_func(sdot_pd *pd, const FunctionEnum::ExpWmR2db<TS> &f) {
    return sdot_pd_set_radial_func_eps(pd, f.code, (double)f.eps);
}

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

    /// SpZGrid::update (SpZGrid.tcc:24): uploads diracs (and `translations` when they
    /// changed); the device grid is rebuilt on the next pass.
    void update(const Pt *positions, const TF *weights, TI nb_diracs, bool positions_have_changed = true,
                bool weights_have_changed = true, bool = false) {
        std::vector<double> tr;
        for (const Pt &t : translations) tr.insert(tr.end(), t.v, t.v + dim);
        if (tr != applied_translations) {
            check(pd, sdot_pd_set_translations(pd, tr.data(), (int64_t)translations.size()), "translations");
            applied_translations = tr;
        }
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
    /// SpZGrid::translations (SpZGrid.h:53): periodic images cut every cell
    std::vector<Pt> translations;

private:
    sdot_pd *pd = nullptr;
    int64_t n = 0;
    std::vector<double> applied_translations;
};

namespace detail {
// Measure of {x in [lo, hi] : n_q . x <= h_q} (dim 2 or 3): the vertices of the
// arrangement (every pair / triple of constraints, kept when feasible), then in
// 2D the shoelace over the vertices in angular order, in 3D the sum over the
// faces of area x distance to an inner point / 3, each face's vertices in
// angular order in its plane.  Coincident vertices only add empty edges.
inline double polytope_measure(int dim, const double *lo, const double *hi,
                               const std::vector<std::array<double, 4>> &planes) {
    std::vector<std::array<double, 4>> c = planes;
    double L = 0;
    for (int d = 0; d < dim; ++d) {
        std::array<double, 4> a{0, 0, 0, -lo[d]}, b{0, 0, 0, hi[d]};
        a[d] = -1;
        b[d] = 1;
        c.push_back(a);
        c.push_back(b);
        L = std::max(L, hi[d] - lo[d]);
    }
    const int m = (int)c.size();
    std::vector<double> nrm(m);
    for (int i = 0; i < m; ++i) nrm[i] = std::sqrt(c[i][0] * c[i][0] + c[i][1] * c[i][1] + c[i][2] * c[i][2]);
    const double tol = 1e-11 * L;
    std::vector<std::array<double, 3>> vs;
    auto feasible = [&](const std::array<double, 3> &x) {
        for (int i = 0; i < m; ++i)
            if (c[i][0] * x[0] + c[i][1] * x[1] + c[i][2] * x[2] - c[i][3] > tol * nrm[i]) return false;
        return true;
    };
    for (int i = 0; i < m; ++i)
        for (int j = i + 1; j < m; ++j) {
            if (dim == 2) {
                const double det = c[i][0] * c[j][1] - c[i][1] * c[j][0];
                if (!(std::fabs(det) > 1e-12 * nrm[i] * nrm[j])) continue;
                const std::array<double, 3> x{(c[i][3] * c[j][1] - c[j][3] * c[i][1]) / det,
                                              (c[i][0] * c[j][3] - c[j][0] * c[i][3]) / det, 0};
                if (feasible(x)) vs.push_back(x);
                continue;
            }
            for (int k = j + 1; k < m; ++k) {
                const auto &a = c[i], &b = c[j], &e = c[k];
                const double bx = b[1] * e[2] - b[2] * e[1], by = b[2] * e[0] - b[0] * e[2], bz = b[0] * e[1] - b[1] * e[0];
                const double ex = e[1] * a[2] - e[2] * a[1], ey = e[2] * a[0] - e[0] * a[2], ez = e[0] * a[1] - e[1] * a[0];
                const double ax = a[1] * b[2] - a[2] * b[1], ay = a[2] * b[0] - a[0] * b[2], az = a[0] * b[1] - a[1] * b[0];
                const double det = a[0] * bx + a[1] * by + a[2] * bz;
                if (!(std::fabs(det) > 1e-12 * nrm[i] * nrm[j] * nrm[k])) continue;
                const std::array<double, 3> x{(a[3] * bx + b[3] * ex + e[3] * ax) / det,
                                              (a[3] * by + b[3] * ey + e[3] * ay) / det,
                                              (a[3] * bz + b[3] * ez + e[3] * az) / det};
                if (feasible(x)) vs.push_back(x);
            }
        }
    if ((int)vs.size() <= dim) return 0;
    std::array<double, 3> g{0, 0, 0};
    for (const auto &v : vs)
        for (int d = 0; d < 3; ++d) g[d] += v[d] / vs.size();
    // polygon area of the points `ids` about their mean, in the plane of unit normal n
    auto polygon_area = [&](std::vector<std::array<double, 3>> pts, const double *n) {
        std::array<double, 3> f{0, 0, 0};
        for (const auto &v : pts)
            for (int d = 0; d < 3; ++d) f[d] += v[d] / pts.size();
        double u[3] = {0, 0, 0};
        const int ax = std::fabs(n[0]) < 0.6 ? 0 : (std::fabs(n[1]) < 0.6 ? 1 : 2);
        u[ax] = 1;
        const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
        for (int d = 0; d < 3; ++d) u[d] -= un * n[d];
        const double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int d = 0; d < 3; ++d) u[d] /= ul;
        const double w[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
        std::vector<std::pair<double, std::array<double, 2>>> q;
        for (const auto &v : pts) {
            const double r[3] = {v[0] - f[0], v[1] - f[1], v[2] - f[2]};
            const double x = r[0] * u[0] + r[1] * u[1] + r[2] * u[2], y = r[0] * w[0] + r[1] * w[1] + r[2] * w[2];
            q.push_back({std::atan2(y, x), {x, y}});
        }
        std::sort(q.begin(), q.end(), [](const auto &A, const auto &B) { return A.first < B.first; });
        double a2 = 0;
        for (std::size_t k = 0; k < q.size(); ++k) {
            const auto &P = q[k].second, &Q = q[(k + 1) % q.size()].second;
            a2 += P[0] * Q[1] - Q[0] * P[1];
        }
        return 0.5 * std::fabs(a2);
    };
    if (dim == 2) {
        const double n[3] = {0, 0, 1};
        return polygon_area(vs, n);
    }
    double vol = 0;
    for (int i = 0; i < m; ++i) {
        std::vector<std::array<double, 3>> face;
        for (const auto &v : vs)
            if (std::fabs(c[i][0] * v[0] + c[i][1] * v[1] + c[i][2] * v[2] - c[i][3]) <= tol * nrm[i]) face.push_back(v);
        if (face.size() < 3) continue;
        const double n[3] = {c[i][0] / nrm[i], c[i][1] / nrm[i], c[i][2] / nrm[i]};
        const double dist = (c[i][3] - (c[i][0] * g[0] + c[i][1] * g[1] + c[i][2] * g[2])) / nrm[i];
        vol += polygon_area(face, n) * dist / 3;
    }
    return vol;
}
}  // namespace detail

/// ConvexPolyhedronAssembly<Pc>: items added with add_box, add_convex_polyhedron or
/// add_simplex, each with its density coefficient, passed to the device handle as
/// they were added (one item is the starting cell; several are integrated piece by
/// piece).  cut_id must be -1.
template <class Pc>
class ConvexPolyhedronAssembly {
public:
    static constexpr int dim = Pc::dim;
    using Pt = Point<dim>;
    using TF = double;
    using TI = std::int64_t;

    void add_convex_polyhedron(const std::vector<Pt> &positions, const std::vector<Pt> &normals, TF coeff = 1.0,
                               TI cut_id = -1) {
        if (positions.size() != normals.size() || positions.empty())
            throw Error(SDOT_INVALID_ARG, "add_convex_polyhedron: one normal per position");
        Item it{Item::POLY, coeff, cut_id, {}, {}};
        for (std::size_t i = 0; i < positions.size(); ++i)
            for (int d = 0; d < dim; ++d) {
                it.a.push_back(positions[i][d]);
                it.b.push_back(normals[i][d]);
            }
        items.push_back(it);
    }
    void add_simplex(const TF *points, TF coeff = 1.0, TI cut_id = -1) {
        items.push_back({Item::SIMPLEX, coeff, cut_id, std::vector<double>(points, points + dim * (dim + 1)), {}});
    }
    void add_box(Pt p0, Pt p1, TF coeff = 1.0, TI cut_id = -1) {
        items.push_back({Item::BOX, coeff, cut_id, std::vector<double>(p0.v, p0.v + dim),
                         std::vector<double>(p1.v, p1.v + dim)});
    }
    /// sum of coeff x measure over the items (ConvexPolyhedronAssembly.tcc:115-120)
    TF measure() const {
        TF r = 0;
        for (const auto &it : items) r += it.coeff * item_measure(it);
        return r;
    }
    void normalize() {
        TF m = measure();
        if (m)
            for (auto &it : items) it.coeff /= m;
    }
    void apply(sdot_pd *pd) const {
        check(pd, sdot_pd_clear_shapes(pd), "clear_shapes");
        for (const auto &it : items) {
            if (it.kind == Item::BOX)
                check(pd, sdot_pd_add_box_shape(pd, it.a.data(), it.b.data(), it.coeff, it.cut_id), "add_box");
            else if (it.kind == Item::POLY)
                check(pd, sdot_pd_add_convex_polyhedron_shape(pd, it.a.data(), it.b.data(), (int64_t)(it.a.size() / dim),
                                                              it.coeff, it.cut_id),
                      "add_convex_polyhedron");
            else
                check(pd, sdot_pd_add_simplex_shape(pd, it.a.data(), it.coeff, it.cut_id), "add_simplex");
        }
    }

private:
    struct Item {
        enum Kind { BOX, POLY, SIMPLEX } kind;
        TF coeff;
        TI cut_id;
        std::vector<double> a, b;   // box: p0, p1; polyhedron: positions, normals; simplex: points
    };
    static TF item_measure(const Item &it) {
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        std::vector<std::array<double, 4>> planes;
        if (it.kind == Item::BOX) {
            TF v = 1;
            for (int d = 0; d < dim; ++d) v *= it.b[d] - it.a[d];
            return v;
        }
        const std::size_t np = it.a.size() / dim;
        for (int d = 0; d < dim; ++d) {
            lo[d] = hi[d] = it.a[d];
            for (std::size_t i = 1; i < np; ++i) {
                lo[d] = std::min(lo[d], it.a[dim * i + d]);
                hi[d] = std::max(hi[d], it.a[dim * i + d]);
            }
        }
        if (it.kind == Item::POLY) {   // the box widened by its extent, cut by (x - p_i) . n_i <= 0
            for (int d = 0; d < dim; ++d) {
                const double e = hi[d] - lo[d];
                lo[d] -= e;
                hi[d] += e;
            }
            for (std::size_t i = 0; i < np; ++i) {
                std::array<double, 4> pl{0, 0, 0, 0};
                for (int d = 0; d < dim; ++d) {
                    pl[d] = it.b[dim * i + d];
                    pl[3] += it.b[dim * i + d] * it.a[dim * i + d];
                }
                planes.push_back(pl);
            }
            return detail::polytope_measure(dim, lo, hi, planes);
        }
        // simplex: |det(p_k - p_0)| / dim!
        const double *q = it.a.data();
        if (dim == 2)
            return 0.5 * std::fabs((q[2] - q[0]) * (q[5] - q[1]) - (q[3] - q[1]) * (q[4] - q[0]));
        const double u[3] = {q[3] - q[0], q[4] - q[1], q[5] - q[2]}, v[3] = {q[6] - q[0], q[7] - q[1], q[8] - q[2]},
                     w[3] = {q[9] - q[0], q[10] - q[1], q[11] - q[2]};
        return std::fabs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                         u[2] * (v[0] * w[1] - v[1] * w[0])) / 6;
    }
    std::vector<Item> items;
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
                   const Func &func) {
    grid.update(positions, weights, nb_diracs, false, true);
    bounds.apply(grid.handle());
    check(grid.handle(), set_radial_func(grid.handle(), func), "radial function");
    check(grid.handle(), sdot_pd_integrals(grid.handle(), res), "get_integrals", true);
}

template <class TF, class Grid, class Bounds, class Pt>
void get_integrals(TF *res, Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights, std::size_t nb_diracs) {
    get_integrals(res, grid, bounds, positions, weights, nb_diracs, FunctionEnum::Unit());
}

/// VtkOutput<nb_cell_fields, TF> (Display/VtkOutput.h:17-72) as display_vtk uses it: one
/// cell field; display_vtk records the diagram to draw and save(filename) writes it (the
/// cells are built on the device from the diagram's last pass).
template <int nb_cell_fields = 1, class TF = double>
class VtkOutput {
    static_assert(nb_cell_fields == 1, "display_vtk writes one cell field (the dirac index)");

public:
    using CN = std::array<std::string, nb_cell_fields>;
    explicit VtkOutput(const CN &names = {{"num_dirac"}}) : names(names) {}
    void save(const std::string &filename) const {
        if (!pd) throw Error{SDOT_INVALID_ARG, "VtkOutput::save: nothing displayed"};
        check(pd, sdot_pd_display_vtk(pd, filename.c_str(), names[0].c_str()), "display_vtk", true);
    }
    sdot_pd *pd = nullptr;   // the diagram drawn by the last display_vtk
    CN names;
};

/// display_vtk (PowerDiagram/display_vtk.h:13-25): the cells of the diagram, drawn into res.
template <class TF, class Grid, class Bounds, class Pt, class Func>
void display_vtk(VtkOutput<1, TF> &res, Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights,
                 std::size_t nb_diracs, const Func &func) {
    grid.update(positions, weights, nb_diracs, false, true);
    bounds.apply(grid.handle());
    check(grid.handle(), set_radial_func(grid.handle(), func), "radial function");
    res.pd = grid.handle();
}

template <class TF, class Grid, class Bounds, class Pt>
void display_vtk(VtkOutput<1, TF> &res, Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights,
                 std::size_t nb_diracs) {
    display_vtk(res, grid, bounds, positions, weights, nb_diracs, FunctionEnum::Unit());
}

/// get_centroids (get_centroids.h:14, rand_ratio = 0): cb(centroid, mass, num_dirac) for
/// every cell, in dirac order; the centroid of an empty cell is 0 (get_centroids.h:58).
template <class Grid, class Bounds, class Pt, class TF, class TI, class Func, class CB>
void get_centroids(Grid &grid, Bounds &bounds, const Pt *positions, const TF *weights, TI nb_diracs, const Func &func,
                   const CB &cb, TF rand_ratio = TF(0)) {
    if (rand_ratio != TF(0)) throw Error{SDOT_NOT_SUPPORTED, "get_centroids: rand_ratio (random points) is not offered"};
    constexpr int dim = sizeof(Pt) / sizeof(double);
    grid.update(positions, weights, nb_diracs, false, true);
    bounds.apply(grid.handle());
    check(grid.handle(), set_radial_func(grid.handle(), func), "radial function");
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
                                  const TF *weights, std::size_t nb_diracs, const Func &func, bool stop_if_void = true) {
    grid.update(positions, weights, nb_diracs, false, true);
    sdot_pd *pd = grid.handle();
    bounds.apply(pd);
    check(pd, set_radial_func(pd, func), "radial function");
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

/// get_der_centroids_and_integrals_wrt_weight_and_positions
/// (get_der_centroids_and_integrals_wrt_weight_and_positions.h:16; one domain shape,
/// several shapes or constant pixels / voxels):
/// rows nu n + d (centroid coordinates, then the mass of dirac n), columns nu m + e
/// (position coordinates, then the weight of dirac m), nu = dim + 1; v_values = centroids
/// and masses; returns 1 on an empty cell, 0 otherwise.
template <class TI, class TF, class Grid, class Bounds, class Pt, class Func>
int get_der_centroids_and_integrals_wrt_weight_and_positions(std::vector<TI> &m_offsets, std::vector<TI> &m_columns,
                                                             std::vector<TF> &m_values, std::vector<TF> &v_values,
                                                             Grid &grid, Bounds &bounds, const Pt *positions,
                                                             const TF *weights, std::size_t nb_diracs, const Func &func,
                                                             bool stop_if_void = true) {
    grid.update(positions, weights, nb_diracs, false, true);
    sdot_pd *pd = grid.handle();
    bounds.apply(pd);
    check(pd, set_radial_func(pd, func), "radial function");
    std::int64_t nnz = 0;
    int st = sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions(pd, nullptr, nullptr, nullptr, nullptr, 0, &nnz);
    check(pd, st, "get_der_centroids_and_integrals_wrt_weight_and_positions", true);
    constexpr int nu = sizeof(Pt) / sizeof(double) + 1;
    std::vector<std::int64_t> off(nu * nb_diracs + 1), col(nnz > 0 ? nnz : 1);
    m_values.resize(nnz > 0 ? nnz : 1);
    v_values.resize(nu * nb_diracs);
    st = sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions(pd, off.data(), col.data(), m_values.data(),
                                                                      v_values.data(), nnz, &nnz);
    check(pd, st, "get_der_centroids_and_integrals_wrt_weight_and_positions", true);
    m_offsets.assign(off.begin(), off.end());
    m_columns.assign(col.begin(), col.begin() + nnz);
    m_values.resize(nnz);
    return (st == SDOT_EMPTY_CELL && stop_if_void) ? 1 : 0;
}

/// The benchmarks' `.xyw` dirac files (benchmarks/set_up_diracs_2D.h:171-181): whitespace
/// separated "x y w" triples, one dirac each.  Returns false when the file cannot be read.
inline bool read_xyw(const std::string &path, std::vector<Point<2>> &positions, std::vector<double> &weights) {
    std::FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    positions.clear();
    weights.clear();
    double x, y, w;
    while (std::fscanf(f, "%lf %lf %lf", &x, &y, &w) == 3) {
        positions.push_back(Point<2>{{x, y}});
        weights.push_back(w);
    }
    std::fclose(f);
    return true;
}

inline bool write_xyw(const std::string &path, const std::vector<Point<2>> &positions, const std::vector<double> &weights) {
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f || positions.size() != weights.size()) {
        if (f) std::fclose(f);
        return false;
    }
    for (std::size_t i = 0; i < positions.size(); ++i)
        std::fprintf(f, "%.17g %.17g %.17g\n", positions[i][0], positions[i][1], weights[i]);
    return std::fclose(f) == 0;
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
    int solve(const Pt *positions, TF *weights, TI nb_diracs, TF *masses, RadialFunc func) {
        grid.update(positions, weights, nb_diracs, true, true);
        sdot_pd *pd = grid.handle();
        bounds.apply(pd);
        check(pd, set_radial_func(pd, func), "radial function");
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
