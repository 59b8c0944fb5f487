// Per-cell routine of the 2D Laguerre-cell kernel (host+device so that the
// same code can be exercised on the CPU when debugging the kernel).
//
// Per cell (semantics of src/sdot/PowerDiagram/Visitors/SpZGrid.tcc:494-645):
//   * start from the domain box (ConvexPolyhedronAssembly with one box,
//     ConvexPolyhedronAssembly.tcc:64-65), in coordinates local to the dirac;
//   * visit grid cells ring by ring; a ring (and every further one) or a grid
//     cell is skipped when no dirac in it can cut the current polygon -- the
//     power-distance bound of SpZGrid::can_be_evicted (:475-480) taken over the
//     disc of radius R (max vertex distance) that holds the polygon (in ball mode
//     over its intersection with the ball, which is what the result depends on);
//   * cut by the radical half-plane of each remaining candidate
//     (SpZGrid.tcc:506-522; ConvexPolyhedron2::plane_cut :687-922: the new
//     points are interpolated from the inside vertex, the entering edge takes
//     the new id, the leaving edge keeps the id of the last removed edge);
//   * in ball mode, clip by the disc of radius sqrt(w_i) streaming the
//     LINE/ARC items of ConvexPolyhedron2::ball_cut (:534-665) straight into the
//     integrals (area :1791-1842, _arc_area :2361-2380, boundary items
//     :437-478) and the Hessian row (get_der_integrals_wrt_weights.h:52-75).
// T = double (FP64 path) or float (mixed path: FP32 local coordinates; plane
// offsets, distances and the mass/Hessian outputs in FP64).
#pragma once

#include <cfloat>
#include <cmath>
#include <type_traits>

#include "internal.h"

namespace sdot_amd {
namespace c2 {

template <class T>
struct Eps;
template <>
struct Eps<double> {
    static constexpr double region = 1e-9, cand = 1e-9;
};
template <>
struct Eps<float> {
    static constexpr double region = 2e-5, cand = 2e-5;
};

// vertex bitmasks of a polygon of at most CAP vertices, and their bit counts
template <int CAP>
struct Mask {
    using type = typename std::conditional<(CAP <= 32), uint32_t, uint64_t>::type;
};
SDOT_HD inline int popc(uint32_t x) { return __builtin_popcount(x); }
SDOT_HD inline int popc(uint64_t x) { return __builtin_popcountll(x); }
SDOT_HD inline int hibit(uint32_t x) { return 31 - __builtin_clz(x); }
SDOT_HD inline int hibit(uint64_t x) { return 63 - __builtin_clzll(x); }

// Can a region whose points are all at distance >= d = sqrt(D2) from c0 and whose
// diracs have weight <= wmax hold a dirac that cuts a polygon contained in the disc
// of radius R = sqrt(R2) about c0?  Dirac j cuts only if |V|*R > (|V|^2 + w0 - wj)/2,
// so the region is skippable when d >= R and d^2 - 2 R d + w0 - wmax >= tol, or when
// d < R and w0 - wmax - R^2 >= tol.  The first test is taken in squares (both sides
// of d^2 + w0 - wmax - tol >= 2 R d are then non-negative): no square root.
template <class T>
SDOT_HD inline bool region_skippable(double D2, double R2, double w0, double wmax) {
    const double dw = w0 - wmax;
    const double tol = Eps<T>::region * (D2 + R2 + fabs(w0) + fabs(wmax));
    if (D2 >= R2) {
        const double l = D2 + dw - tol;
        return l >= 0 && l * l >= 4 * R2 * D2;
    }
    return dw - R2 >= tol;
}

// Neumaier-compensated sum
template <class T>
struct CSum {
    T s = 0, c = 0;
    SDOT_HD void add(T x) {
        T t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    SDOT_HD double value() const { return (double)s + (double)c; }
};

template <class T>
SDOT_HD inline T my_atan2(T y, T x) { return atan2(y, x); }
template <>
SDOT_HD inline float my_atan2<float>(float y, float x) { return atan2f(y, x); }

constexpr double TWO_PI = 6.283185307179586476925286766559;
constexpr double PI_D = 3.1415926535897932384626433832795;

// ---- ExpWmR2db (ExpWmR2db.h: f = exp((w - |x - c|^2) / eps)), coordinates about c ----
// The reference integrates f over a polygon with a generated table of piecewise radial
// polynomials (ConvexPolyhedron2.tcc:1747-1787, precision 1e-10).  Here: the exact
// integral over the fan of triangles (c, P0, P1).  In polar coordinates about c the
// radial part of a triangle is exact, (eps / 2)(1 - exp(-R(theta)^2 / eps)); with rho the
// distance from c to the edge's line, a = rho / sqrt(eps) and s the offset along the line
// from the foot of the perpendicular over sqrt(eps), the angular part becomes
//   T = (eps / 2) [dtheta - sign(P0 x P1) exp(-a^2) (Lam(s1) - Lam(s0))],
//   Lam(s) = a int_0^s exp(-t^2) / (a^2 + t^2) dt,
// and Lam(s) = atan(s / a) + a int_0^s expm1(-t^2) / (a^2 + t^2) dt splits off the Lorentzian
// exactly; the remaining integrand is bounded and smooth away from t ~ a, integrated by
// 16-point Gauss-Legendre on [0, a], [a, 1] (in log t) and [1, s], s capped at 6.5 (the
// tail is below exp(-42)).  Agreement with adaptive 30-digit quadrature: 5e-13 of the
// bracket.
constexpr double GL16X[8] = {0.095012509837637454, 0.28160355077925892, 0.45801677765722737, 0.61787624440264377,
                             0.755404408355003,    0.86563120238783176, 0.9445750230732326,  0.98940093499164994};
constexpr double GL16W[8] = {0.18945061045506859, 0.18260341504492361, 0.16915651939500262, 0.14959598881657676,
                             0.12462897125553403, 0.095158511682492591, 0.062253523938647706, 0.027152459411754037};

// Lam(s) for s >= 0
SDOT_HD inline double gauss_lorentz(double a, double s) {
    const double S = fmin(s, 6.5);
    if (!(S > 0)) return 0.0;
    if (!(a > 1e-300)) return 0.5 * PI_D;
    const double a2 = a * a;
    auto f = [&](double t) { return expm1(-t * t) / (a2 + t * t); };
    auto gl = [&](double lo, double hi, bool logt) {
        const double m = 0.5 * (lo + hi), h = 0.5 * (hi - lo);
        double r = 0;
        for (int k = 0; k < 8; ++k) {
            double u = m - h * GL16X[k], v = m + h * GL16X[k];
            if (logt) {
                u = exp(u);
                v = exp(v);
                r += GL16W[k] * (f(u) * u + f(v) * v);
            } else {
                r += GL16W[k] * (f(u) + f(v));
            }
        }
        return h * r;
    };
    const double b1 = fmin(a, S), b2 = fmin(fmax(a, 1.0), S);
    double tot = gl(0.0, b1, false);
    if (b2 > b1) tot += gl(log(b1), log(b2), true);
    if (S > b2) tot += gl(b2, S, false);
    return atan(S / a) + a * tot;
}

// int over the triangle (0, P0, P1) of exp(-|x|^2 / eps), signed like P0 x P1
SDOT_HD inline double exp_triangle(double x0, double y0, double x1, double y1, double eps) {
    const double cr = x0 * y1 - y0 * x1;
    if (cr == 0) return 0.0;
    const double dx = x1 - x0, dy = y1 - y0, d = sqrt(dx * dx + dy * dy), se = sqrt(eps);
    const double a = fabs(cr) / (d * se);
    const double s0 = (x0 * dx + y0 * dy) / (d * se), s1 = (x1 * dx + y1 * dy) / (d * se);
    const double dth = atan2(cr, x0 * x1 + y0 * y1);
    double e = 0;
    if (a * a < 45) {
        const double l1 = s1 >= 0 ? gauss_lorentz(a, s1) : -gauss_lorentz(a, -s1);
        const double l0 = s0 >= 0 ? gauss_lorentz(a, s0) : -gauss_lorentz(a, -s0);
        e = (cr > 0 ? 1.0 : -1.0) * exp(-a * a) * (l1 - l0);
    }
    return 0.5 * eps * (dth - e);
}

// int along the segment P0 -> P1 of exp((w - |x|^2) / eps) ds (erf closed form; erfc
// differences when both ends lie on one side of the foot); 0 for a zero-length edge
SDOT_HD inline double exp_edge(double x0, double y0, double x1, double y1, double w, double eps) {
    const double dx = x1 - x0, dy = y1 - y0, d2 = dx * dx + dy * dy;
    if (d2 == 0) return 0.0;
    const double d = sqrt(d2), se = sqrt(eps), cr = x0 * y1 - y0 * x1;
    const double s0 = (x0 * dx + y0 * dy) / (d * se), s1 = (x1 * dx + y1 * dy) / (d * se);
    const double diff = s0 >= 0 ? erfc(s0) - erfc(s1) : (s1 <= 0 ? erfc(-s1) - erfc(-s0) : erf(s1) - erf(s0));
    return 0.5 * sqrt(PI_D * eps) * exp((w - cr * cr / d2) / eps) * diff;
}

// Accumulates the boundary items of the final cell (line pieces with a cut id,
// arcs of the ball) into the integrals and the Hessian row.
// pos_integral C and momentum M of a straight boundary item p0 -> p1 (local coordinates
// from the dirac at (c0x, c0y)) in absolute coordinates, added to C and M: the reference
// does not scale them by the density (ConvexPolyhedron2.tcc:462-472), only the measure.
SDOT_HD inline void edge_moments(double c0x, double c0y, double x0, double y0, double x1, double y1, double len,
                                 double *C, double *M) {
    const double P0[2] = {c0x + x0, c0y + y0}, P1[2] = {c0x + x1, c0y + y1};
    for (int u = 0; u < 2; ++u) C[u] += 0.5 * len * (P0[u] + P1[u]);
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v) M[2 * u + v] += len / 6 * (2 * (P0[u] * P0[v] + P1[u] * P1[v]) + P0[u] * P1[v] + P1[u] * P0[v]);
}

// One straight boundary item of get_der_centroids_and_integrals_wrt_weight_and_positions
// (:124-166) between the dirac (c0x, c0y) and the neighbour (or image) at q, with its
// pos_integral C, momentum M (row-major) and measure.  The neighbour's block goes to ELL
// slot `slot` (if within the row capacity), the cell's own image merges into jd, and jd
// gets the diagonal terms.
SDOT_HD inline void jac_edge(const CellArgs &a, int64_t t, double c0x, double c0y, const double *C, const double *M,
                             double meas, const double *q, double dist, bool own, int slot, double *jd) {
    const double c0[2] = {c0x, c0y};
    double d1[9];
    const double bw = 0.5 * meas / dist;
    d1[8] = -bw;
    for (int e = 0; e < 2; ++e) d1[3 * e + 2] = -0.5 * C[e] / dist;
    for (int d = 0; d < 2; ++d) d1[6 + d] = (meas * q[d] - C[d]) / dist;
    for (int d = 0; d < 2; ++d)
        for (int e = 0; e < 2; ++e) d1[3 * e + d] = (C[e] * q[d] - M[2 * d + e]) / dist;
    jd[8] += bw;
    for (int e = 0; e < 2; ++e) jd[3 * e + 2] += 0.5 * C[e] / dist;
    for (int d = 0; d < 2; ++d) jd[6 + d] += (C[d] - meas * c0[d]) / dist;
    for (int d = 0; d < 2; ++d)
        for (int e = 0; e < 2; ++e) jd[3 * e + d] += (M[2 * d + e] - C[e] * c0[d]) / dist;
    if (own) {
        for (int k = 0; k < 9; ++k) jd[k] += d1[k];
    } else if (slot < a.ell_cap) {
        for (int k = 0; k < 9; ++k) a.jblk[((int64_t)slot * 9 + k) * a.ell_stride + t] = d1[k];
    }
}

// JAC: also the blocks of get_der_centroids_and_integrals_wrt_weight_and_positions
// (get_der_centroids_and_integrals_wrt_weight_and_positions.h:46-166): one 3x3 block
// per neighbour slot (a->jblk, beside the Hessian entry of that slot) and the
// diagonal block jd (a->jdiag); entry [3 d + e] for output d (centroid x, y, mass)
// and unknown e (position x, y, weight), before the mass / centroid correction.
// RF: the final integration function -- 0 Unit, 1 R2 (|x - c|^2, integrals only), 2 WmR2
// (w - |x - c|^2): integration(Constant, R2 / WmR2) and for_each_boundary_item(Constant,
// WmR2), ConvexPolyhedron2.tcc:2044-2132 and :391-430; 3 ExpWmR2db (exp((w - |x - c|^2) /
// eps)): integration :1747-1787 and boundary items :346-389.
template <class T, bool PER = false, bool JAC = false, int RF = 0>
struct Integrator {
    const CellArgs *a;
    int64_t t;
    double c0x, c0y;
    double w0v = 0;      // the dirac's weight (RF)
    double ew = 1;       // exp(w / eps) (RF = 3, ExpWmR2db)
    double rint = 0;     // integral of the radial function over the fan of items (RF)
    double jd[JAC ? 9 : 1] = {};
    T r, r2;
    CSum<T> area2;       // twice the polygonal area (shoelace about the dirac)
    double arcs = 0;     // circular segment areas
    double arcm = 0, perim = 0;
    double der_line = 0;   // sum of 0.5*rho*|G_ij|/|x_i-x_j| (rho included)
    double der_arc = 0;    // sum of 0.5*da: 0.5*arc/sqrt(w_i) before the density factor
    double mx = 0, my = 0; // first moments about the dirac (when a->mom): triangles and ball sectors
    int nrow = 0;
    int nlines = 0, narcs = 0;
    bool row_ovf = false;
    // previous / first item
    int m = 0;
    T fx, fy, px, py;
    int ftype, fid, ptype, pid;

    SDOT_HD void line(T x0, T y0, T x1, T y1, int id) {
        T ex = x1 - x0, ey = y1 - y0;
        double len = (double)sqrt(ex * ex + ey * ey);
        if (RF == 3) {   // ExpWmR2db: the fan triangle, and the edge item (ConvexPolyhedron2.tcc:346-389)
            const double X0 = x0, Y0 = y0, X1 = x1, Y1 = y1;
            rint += ew * exp_triangle(X0, Y0, X1, Y1, a->rf_eps);
            if (len == 0) return;   // a zero-length edge carries no item
            len = exp_edge(X0, Y0, X1, Y1, w0v, a->rf_eps);
        } else if (RF) {
            // triangle (dirac, p0, p1): integral of |x - c|^2 = A (|p0|^2 + |p1|^2 + p0.p1) / 6
            const double X0 = x0, Y0 = y0, X1 = x1, Y1 = y1;
            const double at = 0.5 * (X0 * Y1 - Y0 * X1);
            const double r2 = at * (X0 * X0 + Y0 * Y0 + X1 * X1 + Y1 * Y1 + X0 * X1 + Y0 * Y1) / 6;
            rint += RF == 1 ? r2 : w0v * at - r2;
            if (RF == 1) return;   // R2: no boundary items (TODO in the reference)
            if (len == 0) return;  // WmR2 skips zero-length edges
            len *= w0v - (X0 * (X0 + X1) + X1 * X1 + Y0 * (Y0 + Y1) + Y1 * Y1) / 3;
        }
        perim += len;
        if (a->mom) {   // triangle (dirac, p0, p1): area cr/2, centroid (p0 + p1)/3
            const double cr = (double)x0 * y1 - (double)x1 * y0;
            mx += cr * ((double)x0 + x1) / 6;
            my += cr * ((double)y0 + y1) / 6;
        }
        if (id >= 0) {
            double q[2];
            const int64_t j = cut_dirac<2, PER>(*a, id, q);
            const double vx = q[0] - c0x, vy = q[1] - c0y;
            const double dist = sqrt(vx * vx + vy * vy);
            if (JAC && dist != 0) {
                double C[2] = {0, 0}, M[4] = {0, 0, 0, 0};
                edge_moments(c0x, c0y, x0, y0, x1, y1, len, C, M);
                jac_edge(*a, t, c0x, c0y, C, M, a->coeff * len, q, dist, PER && j == t, nrow, jd);
            }
            if (dist != 0 && (!PER || j != t)) {   // a face with the cell's own image adds nothing
                double b = 0.5 * a->coeff * len / dist;
                if (nrow < a->ell_cap) {
                    a->ell_col[nrow * a->ell_stride + t] = (int)j;
                    a->ell_val[nrow * a->ell_stride + t] = -b;
                    ++nrow;
                } else {
                    row_ovf = true;
                }
                der_line += b;
            }
        }
    }
    SDOT_HD void arc(T x0, T y0, T x1, T y1) {
        T a0 = my_atan2(y0, x0), a1 = my_atan2(y1, x1);
        T da = a1 - a0;
        if (a1 < a0) da = (a1 + T(TWO_PI)) - a0;
        double l = (double)da * (double)r;
        double fac = 1;
        if (RF) {   // sector: integral of |x - c|^2 = R^4 da / 4
            const double R2 = (double)r2, r2s = R2 * R2 * (double)da / 4;
            rint += RF == 1 ? r2s : w0v * R2 * (double)da / 2 - r2s;
            fac = RF == 1 ? 0.0 : w0v - R2;   // WmR2 arc item: length x (w - R^2)
        }
        arcm += l * fac;
        perim += l * fac;
        der_arc += 0.5 * (double)da * fac;
        if (JAC) {   // ball item (get_der_centroids_and_integrals_wrt_weight_and_positions.h:66-114)
            const double R = (double)r, A0 = (double)a0, A1 = (double)a0 + (double)da;
            const double mm = 0.5 * a->coeff * (double)da;
            const double mxx = R * (sin(A1) - sin(A0)), myy = R * (cos(A0) - cos(A1));
            jd[8] += mm;
            if (A0 < A1) {
                jd[6] += mxx;
                jd[7] += myy;
                jd[2] += mm * c0x + 0.5 * R * (sin(A1) - sin(A0));
                jd[5] += mm * c0y + 0.5 * R * (cos(A0) - cos(A1));
                jd[0] += mxx * c0x + R * R / 2 * ((A1 + sin(A1) * cos(A1)) - (A0 + sin(A0) * cos(A0)));
                jd[1] += myy * c0x + R * R / 4 * (cos(2 * A0) - cos(2 * A1));
                jd[3] += mxx * c0y + R * R / 4 * (cos(2 * A0) - cos(2 * A1));
                jd[4] += myy * c0y + R * R / 2 * ((A1 - sin(A1) * cos(A1)) - (A0 - sin(A0) * cos(A0)));
            } else {
                jd[2] += mm * c0x;
                jd[5] += mm * c0y;
                jd[0] += PI_D * R * R;
                jd[4] += PI_D * R * R;
            }
        }
        if (a->mom) {   // ball sector (_centroid_arc, ConvexPolyhedron2.tcc:2296-2330)
            const double mid = (double)a0 + 0.5 * (double)da;
            const double k = 2.0 / 3.0 * (double)r * r * r * sin(0.5 * (double)da);
            mx += k * cos(mid);
            my += k * sin(mid);
        }
        T ex = x1 - x0, ey = y1 - y0;
        T d0 = sqrt(ex * ex + ey * ey) / 2;
        T d1 = sqrt(fmax(T(0), r2 - d0 * d0));
        double seg = (double)r2 * (double)da / 2;
        if (da > T(PI_D)) seg += (double)d1 * (double)d0;
        else seg -= (double)d1 * (double)d0;
        arcs += seg;
    }
    SDOT_HD void close(T x1, T y1) {
        area2.add(px * y1 - x1 * py);
        if (ptype == 0) {
            line(px, py, x1, y1, pid);
            ++nlines;
        } else {
            arc(px, py, x1, y1);
            ++narcs;
        }
    }
    SDOT_HD void push(int type, int id, T x, T y) {
        if (m == 0) { fx = x; fy = y; ftype = type; fid = id; }
        else close(x, y);
        px = x; py = y; ptype = type; pid = id;
        ++m;
    }
    SDOT_HD void finish() {
        if (m > 0) close(fx, fy);
    }
};

// half side of the square that encloses the ball of radius sqrt(w0) with a margin
// Sums of two products with the rounding written out (fma): the compiler's own
// contraction may fuse either product, and may choose differently in two
// instantiations of the cell body; written out, every capacity tier rounds alike, so a
// cell resumed in a larger tier matches one started there bit for bit.
SDOT_HD inline double sq2(double x, double y) { return fma(x, x, y * y); }
SDOT_HD inline double det2(double a, double b, double c, double d) { return fma(a, d, -(b * c)); }   // a d - b c

SDOT_HD inline double ball_square(double w0) { return sqrt(fmax(w0, 0.0)) * (1 + 1e-4); }

// The line a*x + b*y = c (coordinates local to the dirac) that carries an edge:
// the radical line of neighbour id >= 0, box side -1 - id (bottom, right, top,
// left) for ids -1..-4, or a side of the ball's enclosing square for -5..-8.
template <bool PER = false>
SDOT_HD inline void line_of(const CellArgs &a, const ProbView &pv, int id, double c0x, double c0y, double w0,
                            double &la, double &lb, double &lc) {
    if (id >= 0) {
        double q[2];
        const int64_t j = cut_dirac<2, PER>(a, id, q);
        la = q[0] - c0x;
        lb = q[1] - c0y;
        lc = 0.5 * (sq2(la, lb) + w0 - a.sw[j]);
        return;
    }
    if (id <= -100) {   // plane -100 - id of the domain item (n . x <= h)
        const double *pl = a.it_planes + 4 * (-100 - id);
        la = pl[0];
        lb = pl[1];
        lc = fma(-pl[1], c0y, fma(-pl[0], c0x, pl[3]));
        return;
    }
    const double sq = ball_square(w0);
    switch (-1 - id) {
        case 0: la = 0; lb = 1; lc = pv.lo[1] - c0y; break;
        case 1: la = 1; lb = 0; lc = pv.hi[0] - c0x; break;
        case 2: la = 0; lb = 1; lc = pv.hi[1] - c0y; break;
        case 3: la = 1; lb = 0; lc = pv.lo[0] - c0x; break;
        case 4: la = 1; lb = 0; lc = sq; break;
        case 5: la = 0; lb = 1; lc = sq; break;
        case 6: la = -1; lb = 0; lc = sq; break;
        default: la = 0; lb = -1; lc = sq; break;
    }
}

// Builds and integrates cell t.  The polygon lives at X/Y/I with stride S
// (an LDS slot on the device).  Returns 1 when the polygon outgrows CAP
// vertices (nothing is written then), 0 otherwise.
// a polygon with one id per edge (edge k runs from vertex k to vertex k+1)
template <int PC>
struct IdPoly2 {
    double x[PC], y[PC];
    int e[PC];
    int n;
};

// one side of a convex polygon clip (Sutherland-Hodgman keeping edge ids): keeps
// la x + lb y <= lc, the new edge along the clipping line gets `id`; returns false if
// the result would not fit.  A polygon with no vertex strictly inside meets the kept
// side in a segment at most and gives an empty piece: an edge lying on the line then
// belongs to the piece on its other side only (not to both)
template <int PC>
SDOT_HD inline bool clip_line(const IdPoly2<PC> &in, IdPoly2<PC> &out, double la, double lb, double lc, int id) {
    int m = 0;
    bool fits = true, any_in = false;
    auto put = [&](double x, double y, int e) {
        if (m < PC) { out.x[m] = x; out.y[m] = y; out.e[m] = e; ++m; } else fits = false;
    };
    for (int k = 0; k < in.n; ++k) {
        const int k1 = k + 1 == in.n ? 0 : k + 1;
        const double dk = lc - (la * in.x[k] + lb * in.y[k]);
        const double d1 = lc - (la * in.x[k1] + lb * in.y[k1]);
        any_in |= dk > 0;
        if (dk >= 0) {
            put(in.x[k], in.y[k], in.e[k]);
            if (d1 < 0) {
                const double u = dk / (dk - d1);
                put(in.x[k] + u * (in.x[k1] - in.x[k]), in.y[k] + u * (in.y[k1] - in.y[k]), id);
            }
        } else if (d1 >= 0) {
            const double u = d1 / (d1 - dk);
            put(in.x[k1] + u * (in.x[k] - in.x[k1]), in.y[k1] + u * (in.y[k] - in.y[k1]), in.e[k]);
        }
    }
    out.n = any_in ? m : 0;
    return fits;
}

// Resume slots (CellArgs::rs_i / rs_d): the FP64 tier of RS_CAP vertices saves a
// cell that outgrows it (its polygon before the cut that did not fit and where the
// traversal stood), and the FP64 tier of 2 RS_CAP vertices goes on from there instead
// of starting the cell again.  Bitwise the result of starting again: the same cuts on
// the same polygon in the same order.
constexpr int RS_CAP = 8;
template <class T, int CAP>
constexpr bool rs_save() { return sizeof(T) == 8 && CAP == RS_CAP; }
template <class T, int CAP>
constexpr bool rs_load() { return sizeof(T) == 8 && CAP == 2 * RS_CAP; }

// BATCH = false: a single problem (CellArgs::prob is null), whose grid and domain
// then stay uniform values (scalar registers) instead of per-thread copies.
// PER: the cell is also cut by the translated diracs (a.ntrans translations).
// JAC: the position / weight derivative blocks (one domain shape, no pieces).
// RF: the final integration function (Integrator).
template <class T, int CAP, bool BALL, bool IMG = false, bool BATCH = true, bool PER = false, bool JAC = false,
          int RF = 0>
SDOT_HD inline int cell2_body(const CellArgs &a, const int64_t t, T *X, T *Y, int *I, const int S, const int SI,
                             const int64_t wk = -1) {
#define PX(k) X[(k) * S]
#define PY(k) Y[(k) * S]
#define PI_(k) I[(k) * SI]
    constexpr bool SAVE = rs_save<T, CAP>(), LOAD = rs_load<T, CAP>();

    const double c0x = a.spos[2 * t], c0y = a.spos[2 * t + 1], w0 = a.sw[t];
    ProbView pv;
    if (BATCH) prob_view(a, t, pv);
    else prob_view_single(a, pv);
    if (a.pdone && a.pdone[pv.p]) return 0;   // converged problem of a batch: nothing to do
    if (void_seen(a, pv.p)) return 0;         // the pass is a backtrack signal already
    const GridInfo &g = pv.g;
    const double W = pv.W;

    // starting cell: the domain box, vertices p0, (p1.x,p0.y), p1, (p0.x,p1.y), edges
    // bottom, right, top, left (ids -1..-4).  In ball mode it is intersected with a
    // square slightly larger than the disc (sides -5..-8): P & square & disc = P & disc,
    // and every edge left is short, which keeps the chord/circle intersections of the
    // ball clip well conditioned.
    int n = 4;
    // a resumed cell: its saved polygon and the candidate to go on from
    bool res = false;
    int rs_s = -1, rs_r = 0, rs_gy = 0, rs_gx = 0, rs_j = 0;
    if (LOAD && a.rs_hdr && wk >= 0 && wk < a.rs_cap) {
        const int4 h = *reinterpret_cast<const int4 *>(a.rs_hdr + 4 * t);
        const int rn = h.x & 255;
        if (rn > 0) {
            const double *d = a.rs_d + 2 * RS_CAP * wk;
            const int *e = a.rs_i + RS_CAP * wk;
            res = true;
            n = rn;
            rs_s = (h.x >> 8) - 1;
            rs_r = h.y;
            rs_gy = h.z & 0xffff;
            rs_gx = h.z >> 16;
            rs_j = h.w;
            for (int k = 0; k < rn; ++k) {
                PX(k) = T(d[2 * k]);
                PY(k) = T(d[2 * k + 1]);
                PI_(k) = e[k];
            }
        }
    }
    if (!res) {
        double bx0 = pv.lo[0] - c0x, by0 = pv.lo[1] - c0y, bx1 = pv.hi[0] - c0x, by1 = pv.hi[1] - c0y;
        int eb = -1, er = -2, et = -3, el = -4;
        if (BALL) {
            const double sq = ball_square(w0);
            if (-sq > by0) { by0 = -sq; eb = -8; }
            if (sq < bx1) { bx1 = sq; er = -5; }
            if (sq < by1) { by1 = sq; et = -6; }
            if (-sq > bx0) { bx0 = -sq; el = -7; }
            if (!(w0 > 0) || !(bx0 < bx1) || !(by0 < by1)) n = 0;   // no disc, or disjoint from the box
        }
        PX(0) = T(bx0); PY(0) = T(by0); PI_(0) = eb;
        PX(1) = T(bx1); PY(1) = T(by0); PI_(1) = er;
        PX(2) = T(bx1); PY(2) = T(by1); PI_(2) = et;
        PX(3) = T(bx0); PY(3) = T(by1); PI_(3) = el;
    }

    auto radius2 = [&]() {
        double R2 = 0;
        for (int k = 0; k < n; ++k) {
            double x = PX(k), y = PY(k);
            R2 = fmax(R2, sq2(x, y));
        }
        if (BALL) R2 = fmin(R2, fmax(w0, 0.0) * (1 + 1e-6));
        return R2;
    };
    // cut by the half-plane q.V <= H; returns 0 (no change), 1 (cut), 2 (capacity exceeded)
    // exact-ish side of vertex k (FP64, from its two lines) for the line q.v <= h
    auto side64 = [&](int k, double vx, double vy, double hh) -> double {
        double a1, b1, e1, a2, b2, e2;
        line_of<PER>(a, pv, PI_(k == 0 ? n - 1 : k - 1), c0x, c0y, w0, a1, b1, e1);
        line_of<PER>(a, pv, PI_(k), c0x, c0y, w0, a2, b2, e2);
        double x = (double)PX(k), y = (double)PY(k);
        const double det = det2(a1, a2, b1, b2);
        if (fabs(det) > 1e-12 * (fabs(a1 * b2) + fabs(a2 * b1))) {
            x = det2(e1, e2, b1, b2) / det;
            y = det2(a1, a2, e1, e2) / det;
        }
        return hh - (x * vx + y * vy);
    };
    auto cut = [&](const T Vx, const T Vy, const T H, const int cid, double vx64, double vy64, double h64) -> int {
        // classify vertices: d = H - q.V, outside when d < 0.  The loop only builds the
        // bitmask of outside vertices (no per-vertex branches, which cost the threads of a
        // wavefront most of the cut); the runs are read from the mask afterwards.
        auto dist = [&](int k) { return fma(-PX(k), Vx, fma(-PY(k), Vy, H)); };
        using M = typename Mask<CAP>::type;   // one bit per vertex (32-bit words up to 32 vertices)
        constexpr int WB = 8 * sizeof(M);
        M mask = 0;
        if (CAP <= 8) {
            // both sides evaluated (a '&&' would branch around each LDS read; reading the
            // slot past n is in bounds and masked out)
#pragma unroll
            for (int k = 0; k < CAP; ++k) mask |= ((k < n) & (dist(k) < 0)) ? (M(1) << k) : M(0);
        } else {
            // the first 8 as above (most cells have fewer vertices), the rest in a loop
#pragma unroll
            for (int k = 0; k < 8; ++k) mask |= ((k < n) & (dist(k) < 0)) ? (M(1) << k) : M(0);
            for (int k = 8; k < n; ++k) mask |= dist(k) < 0 ? (M(1) << k) : M(0);
        }
        if (!mask) return 0;
        const M full = n >= WB ? ~M(0) : (M(1) << n) - 1;
        // starts: outside vertices whose predecessor is inside; ends: whose successor is inside
        auto runs = [&](M m, M &starts, M &ends) {
            const M prev = ((m << 1) | (m >> (n - 1))) & full, next = ((m >> 1) | (m << (n - 1))) & full;
            starts = m & ~prev;
            ends = m & ~next;
        };
        M starts, ends;
        runs(mask, starts, ends);
        bool exact = false;
        if (popc(starts) > 1) {
            // a convex polygon crosses a line once: rounding noise, redo in FP64 (off the fast path)
            exact = true;
            mask = 0;
            for (int k = 0; k < n; ++k) mask |= T(side64(k, vx64, vy64, h64)) < 0 ? (M(1) << k) : M(0);
            if (!mask) return 0;
            runs(mask, starts, ends);
        }
        const int nout = popc(mask);
        if (nout == n) { n = 0; return 1; }
        // the entering vertex i1 and the leaving vertex i2 (with several runs, as a scan
        // keeping the last transitions would choose: index 0 / n - 1 first)
        const int i1 = (starts & 1) ? 0 : hibit(starts);
        const M ends_lo = ends & (full >> 1);
        const int i2 = ((ends >> (n - 1)) & 1) ? n - 1 : hibit(ends_lo);
        const int nn = n - nout + 2;
        if (nn > CAP) return 2;
        const int i0 = i1 == 0 ? n - 1 : i1 - 1;
        const int i3 = i2 == n - 1 ? 0 : i2 + 1;
        auto dval = [&](int k) { return exact ? T(side64(k, vx64, vy64, h64)) : dist(k); };
        const T dAin = dval(i0), dAout = dval(i1), dBout = dval(i2), dBin = dval(i3);
        const T m0 = dAin / (dAout - dAin);
        const T Ax = fma(-m0, PX(i1) - PX(i0), PX(i0)), Ay = fma(-m0, PY(i1) - PY(i0), PY(i0));
        const T m2 = dBin / (dBout - dBin);
        const T Bx = fma(-m2, PX(i2) - PX(i3), PX(i3)), By = fma(-m2, PY(i2) - PY(i3), PY(i3));
        const int idB = PI_(i2);
        if (i1 <= i2) {
            const int shift = nout - 2;
            if (shift > 0) {
                for (int k = i2 + 1; k < n; ++k) { PX(k - shift) = PX(k); PY(k - shift) = PY(k); PI_(k - shift) = PI_(k); }
            } else if (shift < 0) {
                for (int k = n - 1; k > i2; --k) { PX(k + 1) = PX(k); PY(k + 1) = PY(k); PI_(k + 1) = PI_(k); }
            }
            PX(i1) = Ax; PY(i1) = Ay; PI_(i1) = cid;
            PX(i1 + 1) = Bx; PY(i1 + 1) = By; PI_(i1 + 1) = idB;
        } else {
            const int ni = i1 - i2 - 1;
            for (int k = 0; k < ni; ++k) { PX(1 + k) = PX(i2 + 1 + k); PY(1 + k) = PY(i2 + 1 + k); PI_(1 + k) = PI_(i2 + 1 + k); }
            PX(0) = Bx; PY(0) = By; PI_(0) = idB;
            PX(ni + 1) = Ax; PY(ni + 1) = Ay; PI_(ni + 1) = cid;
        }
        n = nn;
        return 1;
    };
    bool overflow = false;
    bool bad = false;
    // SAVE: where the traversal stands when the cell outgrows the tier (rn = 0: before
    // the traversal, the next tier starts again); the caller keeps the polygon
    auto save = [&](int rn, int sv_s, int sv_r, int sv_gy, int sv_gx, int sv_j) {
        if (a.rs_hdr)
            *reinterpret_cast<int4 *>(a.rs_hdr + 4 * t) = make_int4(rn | (sv_s + 1) << 8, sv_r, sv_gy | sv_gx << 16, sv_j);
    };
    if (!BATCH && a.nitems == 1 && !res)   // one domain item: the starting cell is its box cut by its planes
        for (int q = a.it_off[0]; q < a.it_off[1] && n > 0 && !overflow; ++q) {
            const double *pl = a.it_planes + 4 * q;
            const double hl = fma(-pl[1], c0y, fma(-pl[0], c0x, pl[3]));
            if (cut(T(pl[0]), T(pl[1]), T(hl), -100 - q, pl[0], pl[1], hl) == 2) {
                overflow = true;
                if (SAVE) save(0, 0, 0, 0, 0, 0);
            }
        }
    double R2 = radius2();

    const int maxr = g.ncell[0] > g.ncell[1] ? g.ncell[0] : g.ncell[1];
    // s = -1: the diracs themselves; s >= 0 (PER): the diracs translated by trans[s],
    // found around q = c0 - trans[s] in the grid's frame (SpZGrid.tcc:558-561,610)
    for (int s = res ? rs_s : -1; s < (PER ? a.ntrans : 0) && n > 0 && !overflow; ++s) {
    const double qx = s < 0 ? c0x : c0x - a.trans[2 * s], qy = s < 0 ? c0y : c0y - a.trans[2 * s + 1];
    const int idbase = s < 0 ? 0 : (s + 1) * (int)a.N;
    const int hx = grid_coord(g, 0, qx), hy = grid_coord(g, 1, qy);
    for (int r = res ? rs_r : 0; r <= maxr && n > 0 && !overflow; ++r) {
        if (r > 0 && !res) {
            bool lf = hx - r + 1 > 0, rt = hx + r < g.ncell[0], bt = hy - r + 1 > 0, tp = hy + r < g.ncell[1];
            if (!(lf || rt || bt || tp)) break;
            double dm = DBL_MAX;
            if (lf) dm = fmin(dm, qx - (g.lo[0] + (hx - r + 1) * g.h[0]));
            if (rt) dm = fmin(dm, (g.lo[0] + (hx + r) * g.h[0]) - qx);
            if (bt) dm = fmin(dm, qy - (g.lo[1] + (hy - r + 1) * g.h[1]));
            if (tp) dm = fmin(dm, (g.lo[1] + (hy + r) * g.h[1]) - qy);
            dm = fmax(dm, 0.0);
            if (region_skippable<T>(dm * dm, R2, w0, W)) break;
        }
        // the in-range cells of ring r: full rows at dy = +-r, the two side columns in between
        const int gy0 = hy - r < 0 ? 0 : hy - r, gy1 = hy + r >= g.ncell[1] ? g.ncell[1] - 1 : hy + r;
        for (int gy = res ? rs_gy : gy0; gy <= gy1 && n > 0 && !overflow; ++gy) {
          const bool full_row = gy == hy - r || gy == hy + r;
          const int gxa = full_row ? (hx - r < 0 ? 0 : hx - r) : hx - r;
          const int gxb = full_row ? (hx + r >= g.ncell[0] ? g.ncell[0] - 1 : hx + r) : hx + r;
          const int step = full_row ? 1 : 2 * r;
          // the bounds and the largest weight of the row's next grid cell are loaded while this
          // one is processed (one load latency per grid cell instead of two or three)
          const int gxe = gxb < g.ncell[0] - 1 ? gxb : g.ncell[0] - 1;   // last in-range cell of the row
          uint32_t ncs = 0, nce = 0;
          double nwm = 0;
          bool pre = false;
          for (int gx = res ? rs_gx : gxa; gx <= gxb && n > 0 && !overflow; gx += step) {
            if (gx < 0 || gx >= g.ncell[0]) {
                pre = false;
                continue;
            }
            uint32_t cs = ncs, ce = nce;
            double wm = nwm;
            if (!pre) {
                const uint32_t code = morton2(gx, gy);
                cs = pv.cell_start[code];
                ce = pv.cell_start[code + 1];
                wm = pv.cell_wmax[code];
            }
            {   // unconditional (the row's last cell loads itself again): no branch around the loads
                const int gn = gx + step <= gxe ? gx + step : gx;
                pre = gx + step <= gxe;
                const uint32_t cn = morton2(gn, gy);
                ncs = pv.cell_start[cn];
                nce = pv.cell_start[cn + 1];
                nwm = pv.cell_wmax[cn];
            }
            if (cs == ce) continue;
            if (!res) {
                double bx0 = g.lo[0] + gx * g.h[0], by0 = g.lo[1] + gy * g.h[1];
                double ex = fmax(0.0, fmax(bx0 - qx, qx - (bx0 + g.h[0])));
                double ey = fmax(0.0, fmax(by0 - qy, qy - (by0 + g.h[1])));
                if (region_skippable<T>(ex * ex + ey * ey, R2, w0, wm)) continue;
            }
            const uint32_t j0 = res ? (uint32_t)rs_j : cs;
            res = false;
            // the next candidate's position and weight are loaded while the current one is
            // tested and cut (software pipelining: the loads' latency overlaps the cut)
            double nx = a.spos[2 * (int64_t)j0], ny = a.spos[2 * (int64_t)j0 + 1], nw = a.sw[j0];
            for (uint32_t j = j0; j < ce; ++j) {
                const double px = nx, py = ny, pw = nw;
                {   // unconditional (the last candidate loads itself again): no branch around the loads
                    const uint32_t jn = j + 1 < ce ? j + 1 : j;
                    nx = a.spos[2 * (int64_t)jn];
                    ny = a.spos[2 * (int64_t)jn + 1];
                    nw = a.sw[jn];
                }
                const double vx = px - qx, vy = py - qy;
                const double vv = sq2(vx, vy);
                const double hh = 0.5 * (vv + w0 - pw);
                // the cell's own dirac, or a dirac whose line cannot reach the polygon: one
                // branch for both (no short-circuit)
                if ((((int64_t)j == t) & (s < 0)) | ((hh > 0) & (R2 * vv <= hh * hh * (1 - Eps<T>::cand)))) continue;
                if (vv == 0) { bad = true; continue; }                             // coincident diracs
                const T Vx = T(vx), Vy = T(vy), H = T(hh);
                const int rc = cut(Vx, Vy, H, (int)j + idbase, vx, vy, hh);
                if (rc == 2) {
                    overflow = true;
                    if (SAVE) save(n, s, r, gy, gx, (int)j);
                    break;
                }
                if (rc == 0) continue;
                if (n == 0) break;
                R2 = radius2();
            }
          }
        }
    }
    }
    if (bad) flag_or(a.flags, FLAG_BAD_INPUT);
    if (overflow) return 1;

    // ---- integration (and the ball clip, streamed), in FP64 ----
    // FP64 path: the vertices as clipped.  Mixed path: the FP32 clip fixed
    // which lines bound the cell; each vertex is rebuilt in FP64 as the
    // intersection of its two lines, so the integrals carry no FP32 error.
    Integrator<double, PER, JAC, RF> it;
    it.w0v = w0;
    if (RF == 3) it.ew = exp(w0 / a.rf_eps);
    it.a = &a;
    it.t = t;
    it.c0x = c0x;
    it.c0y = c0y;
    it.r2 = BALL ? w0 : 0.0;
    it.r = BALL ? sqrt(fmax(w0, 0.0)) : 0.0;
    auto vtx = [&](int k, double &x, double &y) {
        x = (double)PX(k);
        y = (double)PY(k);
        if (sizeof(T) < sizeof(double)) {
            double a1, b1, e1, a2, b2, e2;
            line_of<PER>(a, pv, PI_(k == 0 ? n - 1 : k - 1), c0x, c0y, w0, a1, b1, e1);
            line_of<PER>(a, pv, PI_(k), c0x, c0y, w0, a2, b2, e2);
            double det = det2(a1, a2, b1, b2);
            if (fabs(det) > 1e-12 * (fabs(a1 * b2) + fabs(a2 * b1))) {
                x = det2(e1, e2, b1, b2) / det;
                y = det2(a1, a2, e1, e2) / det;
            }
        }
    };
    bool full_disc = false;
    if (IMG) {
        // Pieces, each integrated with its own density; a neighbour's entry sums density
        // x length over its edge's pieces.  Pixels (ScaledImage::for_each_intersection,
        // ScaledImage.tcc:98-149): the cell clipped by every pixel under its bounding box.
        // Several domain items (ConvexPolyhedronAssembly::for_each_intersection,
        // ConvexPolyhedronAssembly.tcc:132-142): the cell clipped by each item.
        constexpr int PC = CAP + 8;
        IdPoly2<PC> cell, pa, pb;
        double edge_rho_len[CAP];
        bool edge_hit[CAP];   // crossed by a piece: only such edges are boundary items of the cell
        double edge_c[JAC ? CAP : 1][2], edge_m[JAC ? CAP : 1][4];   // JAC: unweighted moments, summed over pieces
        double mn[2] = {DBL_MAX, DBL_MAX}, mx[2] = {-DBL_MAX, -DBL_MAX};
        cell.n = n;
        for (int k = 0; k < n; ++k) {
            vtx(k, cell.x[k], cell.y[k]);
            cell.e[k] = PI_(k);
            edge_rho_len[k] = 0;
            edge_hit[k] = false;
            if (JAC) {
                edge_c[k][0] = edge_c[k][1] = 0;
                edge_m[k][0] = edge_m[k][1] = edge_m[k][2] = edge_m[k][3] = 0;
            }
            mn[0] = fmin(mn[0], cell.x[k]); mx[0] = fmax(mx[0], cell.x[k]);
            mn[1] = fmin(mn[1], cell.y[k]); mx[1] = fmax(mx[1], cell.y[k]);
        }
        double mass_sum = 0, mom_x = 0, mom_y = 0, rf_area = 0;
        bool fits = true;
        const bool poly = a.img && a.img_nc == 6;
        const IdPoly2<PC> *src = &cell;
        IdPoly2<PC> *dst = &pa;
        auto clip = [&](double la, double lb, double lc) {
            fits &= clip_line(*src, *dst, la, lb, lc, -9);
            src = dst;
            dst = dst == &pa ? &pb : &pa;
        };
        // integrates the piece in *src with density rho (or the quadratic q)
        auto add_piece = [&](double rho, const double *q) {
            auto q_at = [&](double x, double y) {
                return q[0] + x * (q[1] + q[3] * x + q[4] * y) + y * (q[2] + q[5] * y);
            };
            const IdPoly2<PC> &pc = *src;
            // area moments of the piece (Green's theorem over its edges)
            double m0 = 0, sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0, tri = 0;
            for (int k = 0; k < pc.n; ++k) {
                const int k1 = k + 1 == pc.n ? 0 : k + 1;
                const double x0 = pc.x[k], y0 = pc.y[k], x1 = pc.x[k1], y1 = pc.y[k1];
                const double cr = x0 * y1 - x1 * y0;
                m0 += cr;
                if (RF == 3) tri += exp_triangle(x0, y0, x1, y1, a.rf_eps);   // the piece's fan about the dirac
                if (q || a.mom) {
                    sx += (x0 + x1) * cr;
                    sy += (y0 + y1) * cr;
                }
                if (q || RF) {
                    sxx += (x0 * x0 + x0 * x1 + x1 * x1) * cr;
                    syy += (y0 * y0 + y0 * y1 + y1 * y1) * cr;
                    sxy += (x0 * y1 + 2 * x0 * y0 + 2 * x1 * y1 + x1 * y0) * cr;
                }
                if (pc.e[k] < 0 || RF == 1) continue;   // R2: no boundary items (TODO in the reference)
                const double ex = x1 - x0, ey = y1 - y0;
                // ExpWmR2db: the edge piece's integral of exp((w - |x - c|^2) / eps)
                const double len = RF == 3 ? exp_edge(x0, y0, x1, y1, w0, a.rf_eps) : sqrt(ex * ex + ey * ey);
                // density integrated along the edge piece (Simpson, exact for a quadratic);
                // WmR2: times the mean of w - |x - c|^2 over the edge
                const double dens =
                    (q ? (q_at(x0, y0) + 4 * q_at(0.5 * (x0 + x1), 0.5 * (y0 + y1)) + q_at(x1, y1)) / 6 : rho) *
                    (RF == 2 ? w0 - (x0 * x0 + y0 * y0 + x0 * x1 + y0 * y1 + x1 * x1 + y1 * y1) / 3 : 1.0);
                for (int j = 0; j < n; ++j)
                    if (cell.e[j] == pc.e[k]) {
                        edge_rho_len[j] += dens * len;
                        edge_hit[j] = true;
                        if (JAC) edge_moments(c0x, c0y, x0, y0, x1, y1, len, edge_c[j], edge_m[j]);
                        break;
                    }
            }
            if (RF == 3) {
                mass_sum += rho * exp(w0 / a.rf_eps) * tri;
            } else if (RF) {   // constant pieces: the second moment about the dirac, (w - |x - c|^2) for WmR2
                const double r2 = (sxx + syy) / 12;
                mass_sum += rho * (RF == 1 ? r2 : w0 * 0.5 * m0 - r2);
                rf_area += rho * 0.5 * m0;
            } else {
                mass_sum += q ? q[0] * m0 / 2 + q[1] * sx / 6 + q[2] * sy / 6 + q[3] * sxx / 12 + q[4] * sxy / 24 +
                                    q[5] * syy / 12
                              : rho * 0.5 * m0;
            }
            if (!q) {   // first moments of a constant piece (quadratic pixels: not offered)
                mom_x += rho * sx / 6;
                mom_y += rho * sy / 6;
            }
        };
        // ball mode: each piece is also clipped by the disc of radius sqrt(w0) about the dirac
        // (intersect_with re-applies the cell's ball cut, ConvexPolyhedron2.tcc:2474-2496, with the
        // LINE / ARC rules of ball_cut :534-665); arcs add rho x length to the ball measure and
        // 0.5 rho x angle to the diagonal (get_der_integrals_wrt_weights.h:58-60)
        double arc_sum = 0, der_arc_sum = 0;
        auto add_ball_piece = [&](double rho) {
            const IdPoly2<PC> &pc = *src;
            const double r2 = w0, r = sqrt(fmax(w0, 0.0));
            double area2 = 0, segs = 0, pmx = 0, pmy = 0;
            int m = 0, ptype = 0, pid = 0, ftype = 0, fid = 0;
            double px = 0, py = 0, fx = 0, fy = 0;
            auto item = [&](int type, int id, double x0, double y0, double x1, double y1) {
                area2 += x0 * y1 - x1 * y0;
                if (type == 0) {
                    if (a.mom) {   // triangle (dirac, p0, p1)
                        const double cr = x0 * y1 - x1 * y0;
                        pmx += cr * (x0 + x1) / 6;
                        pmy += cr * (y0 + y1) / 6;
                    }
                    if (id < 0) return;   // domain, pixel and enclosing-square sides carry no row
                    const double ex = x1 - x0, ey = y1 - y0, len = sqrt(ex * ex + ey * ey);
                    for (int j = 0; j < n; ++j)
                        if (cell.e[j] == id) {
                            edge_rho_len[j] += rho * len;
                            edge_hit[j] = true;
                            break;
                        }
                    return;
                }
                const double a0 = atan2(y0, x0), a1 = atan2(y1, x1);
                const double da = a1 < a0 ? (a1 + TWO_PI) - a0 : a1 - a0;
                arc_sum += rho * da * r;
                der_arc_sum += rho * 0.5 * da;
                const double ex = x1 - x0, ey = y1 - y0, h0 = sqrt(ex * ex + ey * ey) / 2, h1 = sqrt(fmax(0.0, r2 - h0 * h0));
                segs += r2 * da / 2 + (da > PI_D ? h1 * h0 : -h1 * h0);
                if (a.mom) {   // ball sector (_centroid_arc, ConvexPolyhedron2.tcc:2296-2330)
                    const double mid = a0 + 0.5 * da, k = 2.0 / 3.0 * r * r * r * sin(0.5 * da);
                    pmx += k * cos(mid);
                    pmy += k * sin(mid);
                }
            };
            auto push = [&](int type, int id, double x, double y) {
                if (m == 0) { fx = x; fy = y; ftype = type; fid = id; }
                else item(ptype, pid, px, py, x, y);
                px = x; py = y; ptype = type; pid = id;
                ++m;
            };
            bool all_inside = true;
            for (int k = 0; k < pc.n; ++k) all_inside &= pc.x[k] * pc.x[k] + pc.y[k] * pc.y[k] - r2 <= 0;
            for (int k0 = 0; k0 < pc.n; ++k0) {
                const int k1 = k0 + 1 == pc.n ? 0 : k0 + 1, id = pc.e[k0];
                const double p0x = pc.x[k0], p0y = pc.y[k0], p1x = pc.x[k1], p1y = pc.y[k1];
                if (all_inside) {
                    push(0, id, p0x, p0y);
                    continue;
                }
                const double d0 = p0x * p0x + p0y * p0y - r2, d1 = p1x * p1x + p1y * p1y - r2;
                auto root = [&](double ax, double ay, double ex, double ey, double c) {   // |A + t E|^2 = r2 nearest 1/2
                    const double aa = ex * ex + ey * ey, b = ax * ex + ay * ey, dd = sqrt(fmax(0.0, b * b - aa * c));
                    const double u = (-b + dd) / aa, v = (-b - dd) / aa;
                    return fmin(1.0, fmax(0.0, fabs(u - 0.5) <= fabs(v - 0.5) ? u : v));
                };
                if (d0 <= 0) {
                    push(0, id, p0x, p0y);
                    if (d1 > 0) {
                        const double tt = root(p0x, p0y, p1x - p0x, p1y - p0y, d0);
                        push(1, 0, p0x + tt * (p1x - p0x), p0y + tt * (p1y - p0y));
                    }
                } else if (d1 <= 0) {
                    const double tt = root(p1x, p1y, p0x - p1x, p0y - p1y, d1);
                    push(0, id, p1x + tt * (p0x - p1x), p1y + tt * (p0y - p1y));
                } else {   // both outside: 0 or 2 crossings
                    const double ex = p1x - p0x, ey = p1y - p0y, aa = ex * ex + ey * ey;
                    if (aa == 0) continue;
                    const double b = p0x * ex + p0y * ey, s = b * b - aa * d0;
                    if (s <= 0) continue;
                    const double dd = sqrt(s);
                    double u = (-b - dd) / aa, v = (-b + dd) / aa;
                    if (u > 0 && u < 1) v = fmax(0.0, fmin(1.0, v));
                    else if (v > 0 && v < 1) u = fmax(0.0, fmin(1.0, u));
                    else continue;
                    const double q0x = p0x + u * ex, q0y = p0y + u * ey, q1x = p0x + v * ex, q1y = p0y + v * ey;
                    if (atan2(q0y, q0x) == atan2(q1y, q1x)) continue;
                    push(0, id, q0x, q0y);
                    push(1, 0, q1x, q1y);
                }
            }
            double area = 0;
            if (m > 0) {
                item(ptype, pid, px, py, fx, fy);
                area = 0.5 * area2 + segs;
                (void)ftype;
                (void)fid;
            } else if (pc.n > 0) {   // no item: the disc lies inside the piece, or outside it
                bool outside = false;
                for (int k0 = 0; k0 < pc.n; ++k0) {
                    const int k1 = k0 + 1 == pc.n ? 0 : k0 + 1;
                    const double ex = pc.x[k1] - pc.x[k0], ey = pc.y[k1] - pc.y[k0];
                    if (ex * pc.y[k0] - ey * pc.x[k0] > 0) outside = true;   // the dirac lies right of the edge
                }
                if (!outside && r > 0) {   // the whole circle (moment 0 about the dirac)
                    area = PI_D * r2;
                    arc_sum += rho * TWO_PI * r;
                    der_arc_sum += rho * PI_D;
                }
            }
            mass_sum += rho * area;
            mom_x += rho * pmx;
            mom_y += rho * pmy;
        };
        auto add = [&](double rho, const double *q) {
            if (BALL) add_ball_piece(rho);
            else add_piece(rho, q);
        };
        if (a.img) {
            const double c0[2] = {c0x, c0y};
            int i0[2] = {0, 0}, i1[2] = {-1, -1};
            double ps[2] = {0, 0};
            for (int d = 0; d < 2 && n > 0; ++d) {
                ps[d] = (a.dom_hi[d] - a.dom_lo[d]) / a.img_n[d];
                const int u0 = (int)floor((mn[d] + c0[d] - a.dom_lo[d]) / ps[d]);
                const int u1 = (int)floor((mx[d] + c0[d] - a.dom_lo[d]) / ps[d]);
                i0[d] = u0 < 0 ? 0 : (u0 >= a.img_n[d] ? a.img_n[d] - 1 : u0);
                i1[d] = u1 < 0 ? 0 : (u1 >= a.img_n[d] ? a.img_n[d] - 1 : u1);
            }
            const int64_t npix = (int64_t)a.img_n[0] * a.img_n[1];
            for (int py = i0[1]; py <= i1[1]; ++py)
                for (int px = i0[0]; px <= i1[0]; ++px) {
                    const int64_t pix = (int64_t)py * a.img_n[0] + px;
                    const double rho = a.img[pix];
                    // quadratic pixel (Polynomial<6>, Polynomial.h:37-50): the coefficients are
                    // given in domain coordinates; q re-expands them about the dirac
                    double q[6] = {rho, 0, 0, 0, 0, 0};
                    if (poly) {
                        double c[6];
                        for (int k = 0; k < 6; ++k) c[k] = a.img[k * npix + pix];
                        q[0] = c[0] + c0x * (c[1] + c[3] * c0x + c[4] * c0y) + c0y * (c[2] + c[5] * c0y);
                        q[1] = c[1] + 2 * c[3] * c0x + c[4] * c0y;
                        q[2] = c[2] + c[4] * c0x + 2 * c[5] * c0y;
                        q[3] = c[3];
                        q[4] = c[4];
                        q[5] = c[5];
                    }
                    // sides of the pixel that lie inside the bounding box cut the cell; the
                    // others cannot, so the piece of an edge pixel keeps the cell's own edges
                    src = &cell;
                    dst = &pa;
                    if (px > i0[0]) clip(-1, 0, -(a.dom_lo[0] + ps[0] * px - c0x));
                    if (px < i1[0]) clip(1, 0, a.dom_lo[0] + ps[0] * (px + 1) - c0x);
                    if (py > i0[1]) clip(0, -1, -(a.dom_lo[1] + ps[1] * py - c0y));
                    if (py < i1[1]) clip(0, 1, a.dom_lo[1] + ps[1] * (py + 1) - c0y);
                    add(rho, poly ? q : nullptr);
                }
        } else {
            for (int k = 0; k < a.nitems && n > 0; ++k) {
                const double *b = a.it_box + 6 * k;
                const double lx = b[0] - c0x, ly = b[1] - c0y, hx = b[3] - c0x, hy = b[4] - c0y;
                if (lx > mx[0] || hx < mn[0] || ly > mx[1] || hy < mn[1]) continue;   // item box misses the cell
                src = &cell;
                dst = &pa;
                if (lx > mn[0]) clip(-1, 0, -lx);
                if (hx < mx[0]) clip(1, 0, hx);
                if (ly > mn[1]) clip(0, -1, -ly);
                if (hy < mx[1]) clip(0, 1, hy);
                for (int q = a.it_off[k]; q < a.it_off[k + 1]; ++q) {
                    const double *pl = a.it_planes + 4 * q;
                    clip(pl[0], pl[1], pl[3] - pl[0] * c0x - pl[1] * c0y);
                }
                add(a.it_coeff[k], nullptr);
            }
        }
        if (!fits) flag_or(a.flags, FLAG_CELL_OVERFLOW);
        double der_line = 0, perim = 0;
        int nrow = 0;
        bool row_ovf = false;
        double jdp[JAC ? 9 : 1] = {};
        for (int k = 0; k < n; ++k) {
            const int id = cell.e[k];
            if (id < 0 || !edge_hit[k]) continue;
            perim += edge_rho_len[k];   // image density: faces shared with neighbours only
            double q[2];
            const int64_t j = cut_dirac<2, PER>(a, id, q);
            const double vx = q[0] - c0x, vy = q[1] - c0y;
            const double dist = sqrt(vx * vx + vy * vy);
            if (JAC && dist != 0)   // the items of this edge's pieces together (the blocks are linear in them)
                jac_edge(a, t, c0x, c0y, edge_c[k], edge_m[k], edge_rho_len[k], q, dist, PER && j == t, nrow, jdp);
            if (dist == 0 || (PER && j == t)) continue;
            const double b = 0.5 * edge_rho_len[k] / dist;
            if (nrow < a.ell_cap) {
                a.ell_col[nrow * a.ell_stride + t] = (int)j;
                a.ell_val[nrow * a.ell_stride + t] = -b;
                ++nrow;
            } else {
                row_ovf = true;
            }
            der_line += b;
        }
        a.mass[t] = mass_sum;
        if (a.mom) {
            a.mom[2 * t] = mom_x;
            a.mom[2 * t + 1] = mom_y;
        }
        if (JAC)
            for (int k = 0; k < 9; ++k) a.jdiag[(int64_t)k * a.ell_stride + t] = jdp[k];
        a.arc[t] = arc_sum;
        a.perim[t] = perim + arc_sum;   // pieces: the faces shared with neighbours and the ball arcs
        // integration_der_wrt_weight: WmR2 the measure, ExpWmR2db the integral / eps (:124-138)
        a.diag[t] = der_line + der_arc_sum + (RF == 2 ? rf_area : RF == 3 ? mass_sum / a.rf_eps : 0.0);
        a.row_len[t] = nrow;
        if (a.ctr) {
            const unsigned long long K = n;
            const int slot = (int)(t % WCTR);
            ctr_add(a.ctr + slot, 4 * K * K + 37 * K);
            ctr_add(a.ctr + WCTR + slot, K);
        }
        if (row_ovf) flag_or(a.flags, FLAG_ROW_OVERFLOW);
        if (mass_sum == 0) {
            flag_or(a.flags, FLAG_EMPTY);
            if (a.pflags) flag_or(a.pflags + pv.p, FLAG_EMPTY);
        }
        return 0;
    }
    if (!BALL) {
        for (int k = 0; k < n; ++k) {
            double x, y;
            vtx(k, x, y);
            it.push(0, PI_(k), x, y);
        }
        it.finish();
    } else if (n > 0) {
        // one item stream with a single push site: edge k0 gives its start (inside the
        // disc) and, where it leaves the disc, the exit point that starts an arc
        // (ball_cut's LINE / ARC rules, ConvexPolyhedron2.tcc:534-665)
        const double r2 = it.r2;
        double x0v, y0v;
        vtx(0, x0v, y0v);
        double p1x = x0v, p1y = y0v;
        int k0 = 0;
        bool pend = false;
        double qx1 = 0, qy1 = 0;
        while (true) {
            int ty = 0, idd = 0;
            double x, y;
            if (pend) {
                ty = 1;
                x = qx1;
                y = qy1;
                pend = false;
            } else {
                if (k0 >= n) break;
                const int k1 = k0 + 1 == n ? 0 : k0 + 1;
                const double p0x = p1x, p0y = p1y;
                if (k1 == 0) { p1x = x0v; p1y = y0v; }
                else vtx(k1, p1x, p1y);
                const double d0 = p0x * p0x + p0y * p0y - r2, d1 = p1x * p1x + p1y * p1y - r2;
                idd = PI_(k0);
                ++k0;
                if (d0 <= 0 && d1 <= 0) {
                    x = p0x;
                    y = p0y;
                } else if (d0 <= 0 || d1 <= 0) {
                    // find_unique_intersection: the root nearest t = 1/2 from the inside end
                    const bool in0 = d0 <= 0;
                    const double ax = in0 ? p0x : p1x, ay = in0 ? p0y : p1y;
                    const double ex = in0 ? p1x - p0x : p0x - p1x, ey = in0 ? p1y - p0y : p0y - p1y;
                    const double aa = ex * ex + ey * ey, b = ax * ex + ay * ey, c = in0 ? d0 : d1;
                    const double dd = sqrt(fmax(0.0, b * b - aa * c));
                    const double u = (-b + dd) / aa, v = (-b - dd) / aa;
                    double tt = fabs(u - 0.5) <= fabs(v - 0.5) ? u : v;
                    tt = fmin(1.0, fmax(0.0, tt));
                    const double qx = ax + tt * ex, qy = ay + tt * ey;
                    if (in0) {
                        x = p0x;
                        y = p0y;
                        qx1 = qx;
                        qy1 = qy;
                        pend = true;
                    } else {
                        x = qx;
                        y = qy;
                    }
                } else {
                    // find_two_cuts: both ends outside, the chord through the disc if any
                    const double ex = p1x - p0x, ey = p1y - p0y;
                    const double aa = ex * ex + ey * ey;
                    if (aa == 0) continue;
                    const double b = p0x * ex + p0y * ey, s2 = b * b - aa * d0;
                    if (s2 <= 0) continue;
                    const double dd = sqrt(s2);
                    double u = (-b - dd) / aa, v = (-b + dd) / aa;
                    if (u > 0 && u < 1) v = fmax(0.0, fmin(1.0, v));
                    else if (v > 0 && v < 1) u = fmax(0.0, fmin(1.0, u));
                    else continue;
                    const double q0x = p0x + u * ex, q0y = p0y + u * ey, q1x = p0x + v * ex, q1y = p0y + v * ey;
                    if (atan2(q0y, q0x) == atan2(q1y, q1x)) continue;
                    x = q0x;
                    y = q0y;
                    qx1 = q1x;
                    qy1 = q1y;
                    pend = true;
                }
            }
            it.push(ty, idd, x, y);
        }
        if (it.m == 0) {
            // no item left: the disc is either inside the polygon or disjoint from it
            bool outside = false;
            double qx = x0v, qy = y0v;
            for (int j0 = 0; j0 < n; ++j0) {
                double px = qx, py = qy;
                if (j0 + 1 < n) vtx(j0 + 1, qx, qy);
                else { qx = x0v; qy = y0v; }
                double ex = qx - px, ey = qy - py;
                if (ex * py - ey * px > 0) outside = true;   // the dirac lies right of the edge
            }
            full_disc = !outside;
        }
        it.finish();
    }
    double area = 0.5 * it.area2.value() + it.arcs;
    if (full_disc) {
        area = PI_D * (double)w0;
        double l = TWO_PI * sqrt(w0);
        it.arcm = l;
        it.perim = l;
        it.der_arc = PI_D;
        if (RF) {   // the whole disc: pi R^4 / 2 (R2), pi R^2 (w - R^2 / 2) (WmR2), arcs of measure 0 (R = sqrt(w))
            it.rint = RF == 1 ? 0.5 * PI_D * (double)w0 * (double)w0 : PI_D * (double)w0 * ((double)w0 - (double)w0 / 2);
            it.arcm = it.perim = it.der_arc = 0;
        }
        if (JAC) {   // the whole circle: a0 = 1 > a1 = 0 in the reference's boundary item
            const double mm = PI_D * a.coeff;
            it.jd[8] += mm;
            it.jd[2] += mm * c0x;
            it.jd[5] += mm * c0y;
            it.jd[0] += PI_D * (double)w0;
            it.jd[4] += PI_D * (double)w0;
        }
    }
    if (JAC && BALL && !full_disc && it.m == 0) {
        // a void cell still has the reference's whole-circle item (sphere radius 0,
        // ConvexPolyhedron2.tcc:548-551 and :440-448): measure 0, and pi w in the
        // centroid / position entries (get_der_centroids_and_integrals_wrt_weight_and_positions.h:113-117)
        it.jd[0] += PI_D * (double)w0;
        it.jd[4] += PI_D * (double)w0;
    }
    if (JAC)
        for (int k = 0; k < 9; ++k) a.jdiag[(int64_t)k * a.ell_stride + t] = it.jd[k];
    const double coeff = a.coeff;
    const double mass = coeff * (RF ? it.rint : area);
    if (RF == 2) it.der_line += coeff * area;   // integration_der_wrt_weight(WmR2) = the measure (:135-138)
    if (RF == 3) it.der_line += mass / a.rf_eps;   // integration_der_wrt_weight(ExpWmR2db) = integral / eps (:124-127)
    a.mass[t] = mass;
    if (a.mom) {
        a.mom[2 * t] = coeff * it.mx;
        a.mom[2 * t + 1] = coeff * it.my;
    }
    a.arc[t] = coeff * it.arcm;
    a.perim[t] = coeff * it.perim;
    a.diag[t] = it.der_line + coeff * it.der_arc;
    a.row_len[t] = it.nrow;
    if (a.ctr) {
        const unsigned long long K = it.nlines, A = it.narcs;
        const int slot = (int)(t % WCTR);
        ctr_add(a.ctr + slot, 4 * K * K + 37 * K + 40 * A);
        ctr_add(a.ctr + WCTR + slot, K);
        ctr_add(a.ctr + 2 * WCTR + slot, A);
    }
    if (it.row_ovf && a.ell_cap > 0) flag_or(a.flags, FLAG_ROW_OVERFLOW);   // ell_cap 0: a pass without rows
    if (mass == 0) {
        flag_or(a.flags, FLAG_EMPTY);
        if (a.pflags) flag_or(a.pflags + pv.p, FLAG_EMPTY);
    }
    return 0;
#undef PX
#undef PY
#undef PI_
}


}  // namespace c2
}  // namespace sdot_amd
