/*
 * sdot_oracle.c -- CPU restatement of the sdot Laguerre-cell path.
 *
 * TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).  The product
 * library never links or calls this file.  See sdot_oracle.h for what each
 * function follows in the reference and how it is pinned.
 */
#define _GNU_SOURCE
#include "sdot_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* 2D polygon: ConvexPolyhedron2 restated                                     */
/* ------------------------------------------------------------------------- */

/* Box ctor: ConvexPolyhedron2.tcc:65-93 (vertices p0,(p1.x,p0.y),p1,(p0.x,p1.y)) */
void or_cp2_box(or_cp2 *cp, double x0, double y0, double x1, double y1, int64_t id) {
    cp->n = 4;
    cp->x[0] = x0; cp->y[0] = y0;
    cp->x[1] = x1; cp->y[1] = y0;
    cp->x[2] = x1; cp->y[2] = y1;
    cp->x[3] = x0; cp->y[3] = y1;
    cp->nx[0] = 0;  cp->ny[0] = -1;
    cp->nx[1] = 1;  cp->ny[1] = 0;
    cp->nx[2] = 0;  cp->ny[2] = 1;
    cp->nx[3] = -1; cp->ny[3] = 0;
    for (int k = 0; k < 4; ++k) { cp->id[k] = id; cp->arc[k] = 0; }
    cp->scx = 0.5 * (x0 + x1);
    cp->scy = 0.5 * (y0 + y1);
    cp->sr = -1;
    cp->sid = -1;
}

int or_cp2_empty(const or_cp2 *cp) { return cp->n == 0 && cp->sr <= 0; } /* ConvexPolyhedron2.h:95 */

/* plane_cut: ConvexPolyhedron2.tcc:687-922.  Keeps p with (o-p).n >= 0.  New
 * points are interpolated from the inside vertex toward the outside one
 * (:752-783); the edge entering the removed run takes the new id and normal,
 * the edge leaving it keeps the id of the last removed edge.  The streaming
 * form below emits the same vertex order as the reference's four cases. */
int or_cp2_plane_cut(or_cp2 *cp, double ox, double oy, double nx, double ny, int64_t id, int normal_is_normalized) {
    int n = cp->n;
    if (n == 0) return 0;
    double d[OR_CP2_CAP];
    int nout = 0;
    for (int k = 0; k < n; ++k) {
        d[k] = (ox - cp->x[k]) * nx + (oy - cp->y[k]) * ny;
        nout += d[k] < 0;
    }
    if (nout == 0) return 0;
    if (nout == n) { cp->n = 0; return 0; }
    if (!normal_is_normalized) {
        double s = 1 / sqrt(nx * nx + ny * ny);
        for (int k = 0; k < n; ++k) d[k] *= s;
        nx *= s; ny *= s;
    }
    if (n + 1 > OR_CP2_CAP) return OR_CAPACITY;
    or_cp2 t;
    int m = 0;
    for (int k = 0; k < n; ++k) {
        int km = (k + n - 1) % n;
        int ok = d[k] < 0, okm = d[km] < 0;
        if (ok && !okm) {            /* inside km -> outside k: new point, new edge id */
            double m0 = d[km] / (d[k] - d[km]);
            t.x[m] = cp->x[km] - m0 * (cp->x[k] - cp->x[km]);
            t.y[m] = cp->y[km] - m0 * (cp->y[k] - cp->y[km]);
            t.nx[m] = nx; t.ny[m] = ny; t.id[m] = id; t.arc[m] = 0;
            ++m;
        }
        if (!ok && okm) {            /* outside km -> inside k: point keeps edge km's id */
            double m1 = d[k] / (d[km] - d[k]);
            t.x[m] = cp->x[k] - m1 * (cp->x[km] - cp->x[k]);
            t.y[m] = cp->y[k] - m1 * (cp->y[km] - cp->y[k]);
            t.nx[m] = cp->nx[km]; t.ny[m] = cp->ny[km]; t.id[m] = cp->id[km]; t.arc[m] = cp->arc[km];
            ++m;
        }
        if (!ok) {
            t.x[m] = cp->x[k]; t.y[m] = cp->y[k];
            t.nx[m] = cp->nx[k]; t.ny[m] = cp->ny[k]; t.id[m] = cp->id[k]; t.arc[m] = cp->arc[k];
            ++m;
        }
    }
    memcpy(cp->x, t.x, m * sizeof(double));
    memcpy(cp->y, t.y, m * sizeof(double));
    memcpy(cp->nx, t.nx, m * sizeof(double));
    memcpy(cp->ny, t.ny, m * sizeof(double));
    memcpy(cp->id, t.id, m * sizeof(int64_t));
    memcpy(cp->arc, t.arc, m);
    cp->n = m;
    return 1;
}

/* ball_cut: ConvexPolyhedron2.tcc:534-665 */
int or_cp2_ball_cut(or_cp2 *cp, double cx, double cy, double r, int64_t id) {
    cp->scx = cx; cp->scy = cy; cp->sr = r; cp->sid = id;
    int n = cp->n;
    if (n == 0) { cp->sr = 0; return 0; }
    double dist[OR_CP2_CAP];
    int all_inside = 1;
    for (int k = 0; k < n; ++k) {
        double px = cp->x[k] - cx, py = cp->y[k] - cy;
        dist[k] = px * px + py * py - r * r;
        all_inside &= dist[k] <= 0;
    }
    if (all_inside) return 0;
    if (2 * n > OR_CP2_CAP) return OR_CAPACITY;

    or_cp2 t;
    int m = 0;
#define PUSH(TYPE, ID, NX, NY, X, Y) do { t.arc[m] = (TYPE); t.id[m] = (ID); t.nx[m] = (NX); t.ny[m] = (NY); t.x[m] = (X); t.y[m] = (Y); ++m; } while (0)
    for (int i0 = 0; i0 < n; ++i0) {
        int i1 = (i0 + 1) % n;
        double p0x = cp->x[i0], p0y = cp->y[i0], p1x = cp->x[i1], p1y = cp->y[i1];
        if (dist[i0] <= 0) {
            if (dist[i1] <= 0) {
                PUSH(0, cp->id[i0], cp->nx[i0], cp->ny[i0], p0x, p0y);
            } else {
                /* find_unique_intersection(p0, p1) :574-584 */
                double ex = p1x - p0x, ey = p1y - p0y;
                double a = ex * ex + ey * ey;
                double b = (p0x - cx) * ex + (p0y - cy) * ey;
                double c = (p0x - cx) * (p0x - cx) + (p0y - cy) * (p0y - cy) - r * r;
                double dd = sqrt(fmax(0.0, b * b - a * c));
                double u = (-b + dd) / a, v = (-b - dd) / a;
                double tt = fabs(u - 0.5) <= fabs(v - 0.5) ? u : v;
                tt = fmin(1.0, fmax(0.0, tt));
                PUSH(0, cp->id[i0], cp->nx[i0], cp->ny[i0], p0x, p0y);
                PUSH(1, id, 0.0, 0.0, p0x + tt * ex, p0y + tt * ey);
            }
        } else {
            if (dist[i1] <= 0) {
                double ex = p0x - p1x, ey = p0y - p1y;   /* unique(p1, p0) */
                double a = ex * ex + ey * ey;
                double b = (p1x - cx) * ex + (p1y - cy) * ey;
                double c = (p1x - cx) * (p1x - cx) + (p1y - cy) * (p1y - cy) - r * r;
                double dd = sqrt(fmax(0.0, b * b - a * c));
                double u = (-b + dd) / a, v = (-b - dd) / a;
                double tt = fabs(u - 0.5) <= fabs(v - 0.5) ? u : v;
                tt = fmin(1.0, fmax(0.0, tt));
                PUSH(0, cp->id[i0], cp->nx[i0], cp->ny[i0], p1x + tt * ex, p1y + tt * ey);
            } else {
                /* find_two_cuts :586-612 */
                double ex = p1x - p0x, ey = p1y - p0y;
                double a = ex * ex + ey * ey;
                if (a == 0) continue;
                double b = (p0x - cx) * ex + (p0y - cy) * ey;
                double c = (p0x - cx) * (p0x - cx) + (p0y - cy) * (p0y - cy) - r * r;
                double s = b * b - a * c;
                if (s <= 0) continue;
                double dd = sqrt(s);
                double u = (-b - dd) / a, v = (-b + dd) / a;
                if (u > 0 && u < 1) v = fmax(0.0, fmin(1.0, v));
                else if (v > 0 && v < 1) u = fmax(0.0, fmin(1.0, u));
                else continue;
                double q0x = p0x + u * ex, q0y = p0y + u * ey;
                double q1x = p0x + v * ex, q1y = p0y + v * ey;
                double a0 = atan2(q0y - cy, q0x - cx), a1 = atan2(q1y - cy, q1x - cx);
                if (a0 == a1) continue;
                PUSH(0, cp->id[i0], cp->nx[i0], cp->ny[i0], q0x, q0y);
                PUSH(1, id, 0.0, 0.0, q1x, q1y);
            }
        }
    }
#undef PUSH
    if (m == 0) {                       /* :646-653 */
        for (int i0 = 0; i0 < n; ++i0) {
            if ((cx - cp->x[i0]) * cp->nx[i0] + (cy - cp->y[i0]) * cp->ny[i0] > 0) {
                cp->sr = 0;
                break;
            }
        }
    }
    memcpy(cp->x, t.x, m * sizeof(double));
    memcpy(cp->y, t.y, m * sizeof(double));
    memcpy(cp->nx, t.nx, m * sizeof(double));
    memcpy(cp->ny, t.ny, m * sizeof(double));
    memcpy(cp->id, t.id, m * sizeof(int64_t));
    memcpy(cp->arc, t.arc, m);
    cp->n = m;
    return 0;
}

static double arc_angle(const or_cp2 *cp, int i0, int i1) {
    double a0 = atan2(cp->y[i0] - cp->scy, cp->x[i0] - cp->scx);
    double a1 = atan2(cp->y[i1] - cp->scy, cp->x[i1] - cp->scx);
    if (a1 < a0) a1 += 2 * M_PI;
    return a1 - a0;
}

/* integration(Constant, Unit): ConvexPolyhedron2.tcc:1791-1842, _arc_area :2361-2380 */
double or_cp2_area(const or_cp2 *cp, double coeff) {
    int n = cp->n;
    if (n == 0) return cp->sr > 0 ? coeff * M_PI * cp->sr * cp->sr : 0.0;
    double res = 0;
    double ax = cp->x[0], ay = cp->y[0];
    for (int i = 2; i < n; ++i) {
        double bx = cp->x[i - 1], by = cp->y[i - 1], cx = cp->x[i], cy = cp->y[i];
        res += ax * (by - cy) + bx * (cy - ay) + cx * (ay - by);
    }
    res *= 0.5;
    for (int i0 = n - 1, i1 = 0; i1 < n; i0 = i1++) {
        if (!cp->arc[i0]) continue;
        double da = arc_angle(cp, i0, i1);
        double ex = cp->x[i0] - cp->x[i1], ey = cp->y[i0] - cp->y[i1];
        double d0 = sqrt(ex * ex + ey * ey) / 2;
        double d1 = sqrt(fmax(0.0, cp->sr * cp->sr - d0 * d0));
        double rs = cp->sr * cp->sr * da / 2;
        if (da > M_PI) rs += d1 * d0;
        else rs -= d1 * d0;
        res += rs;
    }
    return coeff * res;
}

/* integration(Constant, R2) and (Constant, WmR2), ConvexPolyhedron2.tcc:2044-2132: the
 * integral of |x - c|^2 (rf = 1) or of w - |x - c|^2 (rf = 2) over the polygon, c the
 * sphere centre (the dirac), as the fan of boundary items seen from c: a straight edge
 * (P0, P1) gives the triangle (c, P0, P1), whose integral of |x - c|^2 is
 * A (|P0|^2 + |P1|^2 + P0.P1) / 6 with A its signed area; an arc of angle da gives the
 * sector, R^4 da / 4; the whole disc pi R^4 / 2. */
double or_cp2_radial_integral(const or_cp2 *cp, double coeff, int rf, double w) {
    const double R = cp->sr;
    if (cp->n == 0) {
        if (!(R > 0)) return 0.0;
        return coeff * (rf == 1 ? 0.5 * M_PI * R * R * R * R : M_PI * R * R * (w - R * R / 2));
    }
    double area = 0, r2 = 0;
    for (int i0 = cp->n - 1, i1 = 0; i1 < cp->n; i0 = i1++) {
        if (cp->arc[i0]) {
            double da = arc_angle(cp, i0, i1);
            area += R * R * da / 2;
            r2 += R * R * R * R * da / 4;
        } else {
            double px = cp->x[i0] - cp->scx, py = cp->y[i0] - cp->scy, qx = cp->x[i1] - cp->scx, qy = cp->y[i1] - cp->scy;
            double a = 0.5 * (px * qy - py * qx);
            area += a;
            r2 += a * (px * px + py * py + qx * qx + qy * qy + px * qx + py * qy) / 6;
        }
    }
    return coeff * (rf == 1 ? r2 : w * area - r2);
}

/* for_each_boundary_item(Constant, WmR2), ConvexPolyhedron2.tcc:391-430: measure = coeff
 * x the integral of w - |x - c|^2 along the item (edge: length x (w - (|P0|^2 + P0.P1 +
 * |P1|^2) / 3); arc: length x (w - R^2)); zero-length edges are skipped */
int or_cp2_boundary_items_wmr2(const or_cp2 *cp, double coeff, double w, double *measures, int64_t *ids, int cap) {
    const double R = cp->sr;
    if (cp->n == 0) {
        if (R >= 0 && cap > 0) {
            measures[0] = coeff * 2 * M_PI * R * (w - R * R);
            ids[0] = cp->sid;
            return 1;
        }
        return 0;
    }
    int m = 0;
    for (int i1 = 0, i0 = cp->n - 1; i1 < cp->n; i0 = i1++) {
        if (m >= cap) return -1;
        if (cp->arc[i0]) {
            measures[m] = coeff * arc_angle(cp, i0, i1) * R * (w - R * R);
            ids[m++] = cp->id[i0];
        } else {
            double px = cp->x[i0] - cp->scx, py = cp->y[i0] - cp->scy, qx = cp->x[i1] - cp->scx, qy = cp->y[i1] - cp->scy;
            double d2 = (qx - px) * (qx - px) + (qy - py) * (qy - py);
            if (d2 == 0) continue;
            measures[m] = coeff * sqrt(d2) * (w - (px * (px + qx) + qx * qx + py * (py + qy) + qy * qy) / 3);
            ids[m++] = cp->id[i0];
        }
    }
    return m;
}

/* ---- ExpWmR2db (ExpWmR2db.h): f(x) = exp((w - |x - c|^2) / eps), no ball cut ----
 * integration(Constant, ExpWmR2db), ConvexPolyhedron2.tcc:1747-1787: the reference integrates
 * a tabulated piecewise-polynomial fit of the Gaussian (precision 1e-10).  Restated here as
 * the exact integral over the fan of triangles (c, P0, P1) of the edges, each triangle
 * written along its edge in polar coordinates about c with the radial part done exactly:
 *   T = (P0 x P1) int_0^1 (eps / 2) (1 - exp(-q(u) / eps)) / q(u) du,  q(u) = |P0 + u (P1 - P0)|^2,
 * integrated by adaptive Gauss-Legendre (8 points, bisection).  Boundary items (:346-389):
 * coeff int_edge exp((w - |x - c|^2) / eps) ds, by the same adaptive rule. */
static const double GL8X[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
static const double GL8W[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};

typedef struct edge_fn {
    double px, py, dx, dy, eps;
    int kind;   /* 0: (1 - exp(-q/eps)) / q (mass), 1: exp(-q/eps) (edge item) */
} edge_fn;

static double edge_fn_eval(const edge_fn *f, double u) {
    const double x = f->px + u * f->dx, y = f->py + u * f->dy, q = x * x + y * y;
    if (f->kind == 1) return exp(-q / f->eps);
    return q > 1e-300 * f->eps ? -expm1(-q / f->eps) / q : 1.0 / f->eps;
}

static double gl8(const edge_fn *f, double a, double b) {
    const double m = 0.5 * (a + b), h = 0.5 * (b - a);
    double s = 0;
    for (int k = 0; k < 4; ++k) s += GL8W[k] * (edge_fn_eval(f, m - h * GL8X[k]) + edge_fn_eval(f, m + h * GL8X[k]));
    return h * s;
}

/* bisection until the two halves agree with the whole to tol (an absolute bound fixed
 * from the first estimate, so rounding noise cannot drive the recursion), at most 30 levels */
static double gl8_adaptive(const edge_fn *f, double a, double b, double whole, double tol, int depth) {
    const double m = 0.5 * (a + b), l = gl8(f, a, m), r = gl8(f, m, b);
    if (depth >= 30 || fabs(l + r - whole) <= tol) return l + r;
    return gl8_adaptive(f, a, m, l, tol, depth + 1) + gl8_adaptive(f, m, b, r, tol, depth + 1);
}

static double edge_integral(edge_fn *f) {
    const double g = gl8(f, 0, 1);
    return gl8_adaptive(f, 0, 1, g, 1e-14 * fabs(g) + 1e-300, 0);
}

double or_cp2_expwmr2db_integral(const or_cp2 *cp, double coeff, double w, double eps) {
    double res = 0;
    for (int i0 = cp->n - 1, i1 = 0; i1 < cp->n; i0 = i1++) {
        const double px = cp->x[i0] - cp->scx, py = cp->y[i0] - cp->scy, qx = cp->x[i1] - cp->scx, qy = cp->y[i1] - cp->scy;
        const double cr = px * qy - py * qx;
        if (cr == 0) continue;
        edge_fn f = {px, py, qx - px, qy - py, eps, 0};
        res += cr * 0.5 * eps * edge_integral(&f);
    }
    return coeff * exp(w / eps) * res;
}

int or_cp2_boundary_items_expwmr2db(const or_cp2 *cp, double coeff, double w, double eps, double *measures, int64_t *ids,
                                    int cap) {
    int m = 0;
    for (int i1 = 0, i0 = cp->n - 1; i1 < cp->n; i0 = i1++) {
        if (m >= cap) return -1;
        const double px = cp->x[i0] - cp->scx, py = cp->y[i0] - cp->scy, qx = cp->x[i1] - cp->scx, qy = cp->y[i1] - cp->scy;
        const double d2 = (qx - px) * (qx - px) + (qy - py) * (qy - py);
        if (d2 == 0) continue;
        edge_fn f = {px, py, qx - px, qy - py, eps, 1};
        measures[m] = coeff * exp(w / eps) * sqrt(d2) * edge_integral(&f);
        ids[m++] = cp->id[i0];
    }
    return m;
}

/* add_centroid_contrib(Constant, Unit), ConvexPolyhedron2.tcc:1047-1073, _centroid_arc
 * :2296-2330: area and first moments of the polygon, as signed triangles and
 * circular sectors seen from (ox, oy) (the ball centre for arcs), relative to it */
void or_cp2_moments(const or_cp2 *cp, double coeff, double ox, double oy, double *m0, double *mx, double *my) {
    double a = 0, sx = 0, sy = 0;
    if (cp->n == 0) {
        if (cp->sr > 0) {
            a = M_PI * cp->sr * cp->sr;
            sx = a * (cp->scx - ox);
            sy = a * (cp->scy - oy);
        }
    } else {
        for (int i0 = cp->n - 1, i1 = 0; i1 < cp->n; i0 = i1++) {
            double x0 = cp->x[i0] - ox, y0 = cp->y[i0] - oy, x1 = cp->x[i1] - ox, y1 = cp->y[i1] - oy;
            if (cp->arc[i0]) {   /* sector of the ball (centred at the origin of the moments) */
                double a0 = atan2(y0, x0), t = arc_angle(cp, i0, i1), r = cp->sr;
                double mid = a0 + t / 2, k = 2.0 / 3.0 * r * r * r * sin(t / 2);
                a += r * r * t / 2;
                sx += k * cos(mid);
                sy += k * sin(mid);
            } else {
                double cr = x0 * y1 - x1 * y0;
                a += cr / 2;
                sx += cr * (x0 + x1) / 6;
                sy += cr * (y0 + y1) / 6;
            }
        }
    }
    *m0 += coeff * a;
    *mx += coeff * sx;
    *my += coeff * sy;
}

/* integration(Polynomial<6>, Unit), ConvexPolyhedron2.tcc:2181-2225: the integral of
 * c0 + c1 x + c2 y + c3 x^2 + c4 x y + c5 y^2 over the polygon, from its area moments
 * (Green's theorem over the edges, cr = x_k y_k+1 - x_k+1 y_k). */
double or_cp2_poly_area(const or_cp2 *cp, const double *c6) {
    int n = cp->n;
    double A = 0, Sx = 0, Sy = 0, Sxx = 0, Sxy = 0, Syy = 0;
    for (int k = 0; k < n; ++k) {
        int k1 = k + 1 == n ? 0 : k + 1;
        double x0 = cp->x[k], y0 = cp->y[k], x1 = cp->x[k1], y1 = cp->y[k1];
        double cr = x0 * y1 - x1 * y0;
        A += cr;
        Sx += (x0 + x1) * cr;
        Sy += (y0 + y1) * cr;
        Sxx += (x0 * x0 + x0 * x1 + x1 * x1) * cr;
        Syy += (y0 * y0 + y0 * y1 + y1 * y1) * cr;
        Sxy += (x0 * y1 + 2 * x0 * y0 + 2 * x1 * y1 + x1 * y0) * cr;
    }
    return c6[0] * A / 2 + c6[1] * Sx / 6 + c6[2] * Sy / 6 + c6[3] * Sxx / 12 + c6[4] * Sxy / 24 + c6[5] * Syy / 12;
}

static double poly_at(const double *c, double x, double y) {
    return c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y;
}

/* for_each_boundary_item(Polynomial<6>, Unit), ConvexPolyhedron2.tcc:180-214: the
 * integral along a segment, by Simpson's rule (exact for a quadratic) */
double or_seg_poly(double x0, double y0, double x1, double y1, const double *c6) {
    double len = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    return len * (poly_at(c6, x0, y0) + 4 * poly_at(c6, 0.5 * (x0 + x1), 0.5 * (y0 + y1)) + poly_at(c6, x1, y1)) / 6;
}

/* for_each_boundary_item(Constant, Unit): ConvexPolyhedron2.tcc:437-478 */
int or_cp2_boundary_items(const or_cp2 *cp, double coeff, double *measures, int64_t *ids, int cap) {
    int n = cp->n;
    if (n == 0) {
        if (cp->sr >= 0 && cap > 0) {
            measures[0] = coeff * 2 * M_PI * cp->sr;
            ids[0] = cp->sid;
            return 1;
        }
        return 0;
    }
    int m = 0;
    for (int i1 = 0, i0 = n - 1; i1 < n; i0 = i1++) {
        if (m >= cap) return -1;
        ids[m] = cp->id[i0];
        if (cp->arc[i0]) {
            measures[m] = coeff * arc_angle(cp, i0, i1) * cp->sr;
        } else {
            double ex = cp->x[i1] - cp->x[i0], ey = cp->y[i1] - cp->y[i0];
            measures[m] = coeff * sqrt(ex * ex + ey * ey);
        }
        ++m;
    }
    return m;
}

int or_cp2_eval(double x0, double y0, double x1, double y1, int ncuts, const double *cuts,
                int ball, double bcx, double bcy, double br,
                double *area, int *nb_points, double *measures, int64_t *ids, int cap, int *nitems) {
    or_cp2 *cp = (or_cp2 *)malloc(sizeof(or_cp2));
    or_cp2_box(cp, x0, y0, x1, y1, -1);
    int st = OR_OK;
    for (int k = 0; k < ncuts && st == OR_OK; ++k) {
        int r = or_cp2_plane_cut(cp, cuts[4 * k + 0], cuts[4 * k + 1], cuts[4 * k + 2], cuts[4 * k + 3], k, 1);
        if (r == OR_CAPACITY) st = r;
    }
    if (ball && st == OR_OK) st = or_cp2_ball_cut(cp, bcx, bcy, br, -2);
    *area = or_cp2_area(cp, 1.0);
    *nb_points = cp->n;
    *nitems = or_cp2_boundary_items(cp, 1.0, measures, ids, cap);
    free(cp);
    return st;
}

int or_cp2_centroid(double x0, double y0, double x1, double y1, int ncuts, const double *cuts,
                    int ball, double bcx, double bcy, double br, double *centroid, double *area) {
    or_cp2 *cp = (or_cp2 *)malloc(sizeof(or_cp2));
    or_cp2_box(cp, x0, y0, x1, y1, -1);
    int st = OR_OK;
    for (int k = 0; k < ncuts && st == OR_OK; ++k) {
        int r = or_cp2_plane_cut(cp, cuts[4 * k + 0], cuts[4 * k + 1], cuts[4 * k + 2], cuts[4 * k + 3], k, 1);
        if (r == OR_CAPACITY) st = r;
    }
    if (ball && st == OR_OK) st = or_cp2_ball_cut(cp, bcx, bcy, br, -2);
    const double ox = ball ? bcx : 0.0, oy = ball ? bcy : 0.0;
    double m0 = 0, mx = 0, my = 0;
    or_cp2_moments(cp, 1.0, ox, oy, &m0, &mx, &my);
    *area = m0;
    centroid[0] = m0 != 0 ? ox + mx / m0 : 0.0;
    centroid[1] = m0 != 0 ? oy + my / m0 : 0.0;
    free(cp);
    return st;
}

/* ------------------------------------------------------------------------- */
/* 3D polyhedron (ConvexPolyhedron3 semantics)                                */
/* ------------------------------------------------------------------------- */

static void cp3_set_face(or_cp3 *cp, int f, const double v[4][3], double nx, double ny, double nz, int64_t id) {
    cp->nv[f] = 4;
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) cp->v[f][k][d] = v[k][d];
    cp->n[f][0] = nx; cp->n[f][1] = ny; cp->n[f][2] = nz;
    cp->id[f] = id;
}

/* clear(Box): ConvexPolyhedron3.tcc:799-866 -- six faces, all with cut id `id` */
void or_cp3_box(or_cp3 *cp, const double *p0, const double *p1, int64_t id) {
    double x0 = p0[0], y0 = p0[1], z0 = p0[2], x1 = p1[0], y1 = p1[1], z1 = p1[2];
    cp->nf = 6;
    cp->overflow = 0;
    { const double v[4][3] = {{x0,y0,z0},{x0,y1,z0},{x1,y1,z0},{x1,y0,z0}}; cp3_set_face(cp, 0, v, 0, 0, -1, id); }
    { const double v[4][3] = {{x0,y0,z1},{x1,y0,z1},{x1,y1,z1},{x0,y1,z1}}; cp3_set_face(cp, 1, v, 0, 0, +1, id); }
    { const double v[4][3] = {{x0,y0,z0},{x1,y0,z0},{x1,y0,z1},{x0,y0,z1}}; cp3_set_face(cp, 2, v, 0, -1, 0, id); }
    { const double v[4][3] = {{x0,y1,z0},{x0,y1,z1},{x1,y1,z1},{x1,y1,z0}}; cp3_set_face(cp, 3, v, 0, +1, 0, id); }
    { const double v[4][3] = {{x0,y0,z0},{x0,y0,z1},{x0,y1,z1},{x0,y1,z0}}; cp3_set_face(cp, 4, v, -1, 0, 0, id); }
    { const double v[4][3] = {{x1,y0,z0},{x1,y1,z0},{x1,y1,z1},{x1,y0,z1}}; cp3_set_face(cp, 5, v, +1, 0, 0, id); }
}

typedef struct { double p[3]; double ang; } cap_pt;

static int cmp_cap(const void *a, const void *b) {
    double x = ((const cap_pt *)a)->ang, y = ((const cap_pt *)b)->ang;
    return (x > y) - (x < y);
}

/* plane_cut: ConvexPolyhedron3.tcc:614-754 semantics.  Keeps p with
 * (o-p).n >= 0; every face loop is clipped (intersections from the inside
 * vertex toward the outside one, as in 2D), removed faces vanish, and the new
 * face is closed from the crossing points.  Returns 1 on an effective cut. */
int or_cp3_plane_cut(or_cp3 *cp, const double *origin, const double *normal, int64_t id) {
    double nn = sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    double nx = normal[0] / nn, ny = normal[1] / nn, nz = normal[2] / nn;
    double off = origin[0] * nx + origin[1] * ny + origin[2] * nz;
    int any_out = 0, any_in = 0;
    for (int f = 0; f < cp->nf; ++f)
        for (int k = 0; k < cp->nv[f]; ++k) {
            double d = off - (cp->v[f][k][0] * nx + cp->v[f][k][1] * ny + cp->v[f][k][2] * nz);
            if (d < 0) any_out = 1; else any_in = 1;
        }
    if (!any_out) return 0;
    if (!any_in) { cp->nf = 0; return 0; }

    static __thread cap_pt capp[2 * OR_CP3_MAXF + 8];
    int ncap = 0;
    int nf2 = 0;
    double tv[OR_CP3_MAXFV * 2][3];
    for (int f = 0; f < cp->nf; ++f) {
        int nv = cp->nv[f];
        double d[OR_CP3_MAXFV];
        int nout = 0;
        for (int k = 0; k < nv; ++k) {
            d[k] = off - (cp->v[f][k][0] * nx + cp->v[f][k][1] * ny + cp->v[f][k][2] * nz);
            nout += d[k] < 0;
        }
        if (nout == nv) continue;                 /* face removed */
        if (nout == 0) {                          /* face kept as is */
            if (nf2 != f) {
                cp->nv[nf2] = nv;
                memcpy(cp->v[nf2], cp->v[f], sizeof(double) * 3 * nv);
                memcpy(cp->n[nf2], cp->n[f], sizeof(double) * 3);
                cp->id[nf2] = cp->id[f];
            }
            ++nf2;
            continue;
        }
        int m = 0;
        for (int k = 0; k < nv; ++k) {
            int km = (k + nv - 1) % nv;
            int ok = d[k] < 0, okm = d[km] < 0;
            if (ok != okm) {
                int ii = ok ? km : k, io = ok ? k : km;
                double mm = d[ii] / (d[io] - d[ii]);
                for (int c = 0; c < 3; ++c)
                    tv[m][c] = cp->v[f][ii][c] - mm * (cp->v[f][io][c] - cp->v[f][ii][c]);
                if (ncap < 2 * OR_CP3_MAXF + 8) { memcpy(capp[ncap].p, tv[m], sizeof(double) * 3); ++ncap; }
                ++m;
            }
            if (!ok) { memcpy(tv[m], cp->v[f][k], sizeof(double) * 3); ++m; }
        }
        if (m > OR_CP3_MAXFV) { cp->overflow = 1; m = OR_CP3_MAXFV; }
        cp->nv[nf2] = m;
        memcpy(cp->v[nf2], tv, sizeof(double) * 3 * m);
        if (nf2 != f) {
            memcpy(cp->n[nf2], cp->n[f], sizeof(double) * 3);
            cp->id[nf2] = cp->id[f];
        }
        ++nf2;
    }
    cp->nf = nf2;
    if (ncap < 3) return 1;
    /* new face: unique crossing points ordered counter-clockwise about +n */
    double cx = 0, cy = 0, cz = 0;
    for (int k = 0; k < ncap; ++k) { cx += capp[k].p[0]; cy += capp[k].p[1]; cz += capp[k].p[2]; }
    cx /= ncap; cy /= ncap; cz /= ncap;
    double ux, uy, uz;
    if (fabs(nx) < 0.6) { ux = 0; uy = nz; uz = -ny; } else { ux = -nz; uy = 0; uz = nx; }
    double un = sqrt(ux * ux + uy * uy + uz * uz);
    ux /= un; uy /= un; uz /= un;
    double vx = ny * uz - nz * uy, vy = nz * ux - nx * uz, vz = nx * uy - ny * ux;
    for (int k = 0; k < ncap; ++k) {
        double px = capp[k].p[0] - cx, py = capp[k].p[1] - cy, pz = capp[k].p[2] - cz;
        capp[k].ang = atan2(px * vx + py * vy + pz * vz, px * ux + py * uy + pz * uz);
    }
    qsort(capp, ncap, sizeof(cap_pt), cmp_cap);
    if (cp->nf >= OR_CP3_MAXF) { cp->overflow = 1; return 1; }
    int f = cp->nf++;
    int m = 0;
    for (int k = 0; k < ncap; ++k) {
        if (m && capp[k].p[0] == cp->v[f][m - 1][0] && capp[k].p[1] == cp->v[f][m - 1][1] && capp[k].p[2] == cp->v[f][m - 1][2])
            continue;
        if (m >= OR_CP3_MAXFV) { cp->overflow = 1; break; }
        memcpy(cp->v[f][m], capp[k].p, sizeof(double) * 3);
        ++m;
    }
    while (m > 1 && cp->v[f][m - 1][0] == cp->v[f][0][0] && cp->v[f][m - 1][1] == cp->v[f][0][1] && cp->v[f][m - 1][2] == cp->v[f][0][2])
        --m;
    cp->nv[f] = m;
    cp->n[f][0] = nx; cp->n[f][1] = ny; cp->n[f][2] = nz;
    cp->id[f] = id;
    return 1;
}

/* area(Face): ConvexPolyhedron3.tcc:1367-1418 -- fan of cross products . n */
double or_cp3_face_area(const or_cp3 *cp, int f) {
    double s = 0;
    const double (*v)[3] = cp->v[f];
    for (int k = 2; k < cp->nv[f]; ++k) {
        double ax = v[k - 1][0] - v[0][0], ay = v[k - 1][1] - v[0][1], az = v[k - 1][2] - v[0][2];
        double bx = v[k][0] - v[0][0], by = v[k][1] - v[0][1], bz = v[k][2] - v[0][2];
        s += (ay * bz - az * by) * cp->n[f][0] + (az * bx - ax * bz) * cp->n[f][1] + (ax * by - ay * bx) * cp->n[f][2];
    }
    return s / 2;
}

/* integration(Constant, Unit): ConvexPolyhedron3.tcc:1022-1056, taken about `center` */
double or_cp3_volume(const or_cp3 *cp, const double *center) {
    double res = 0;
    for (int f = 0; f < cp->nf; ++f) {
        if (cp->nv[f] < 3) continue;
        double h = (cp->v[f][0][0] - center[0]) * cp->n[f][0] + (cp->v[f][0][1] - center[1]) * cp->n[f][1] +
                   (cp->v[f][0][2] - center[2]) * cp->n[f][2];
        res += h * or_cp3_face_area(cp, f) / 3;
    }
    return res;
}

/* add_centroid_contrib(Constant, Unit), ConvexPolyhedron3.tcc:884-924: volume and first
 * moments as signed tetrahedra (centre, fan triangle of each face), relative to centre */
void or_cp3_moments(const or_cp3 *cp, const double *c, double coeff, double *m0, double *m) {
    for (int f = 0; f < cp->nf; ++f) {
        if (cp->nv[f] < 3) continue;
        const double (*v)[3] = cp->v[f];
        double a[3] = {v[0][0] - c[0], v[0][1] - c[1], v[0][2] - c[2]};
        for (int k = 2; k < cp->nv[f]; ++k) {
            double b[3] = {v[k - 1][0] - c[0], v[k - 1][1] - c[1], v[k - 1][2] - c[2]};
            double d[3] = {v[k][0] - c[0], v[k][1] - c[1], v[k][2] - c[2]};
            double vol = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                          a[2] * (b[0] * d[1] - b[1] * d[0])) / 6;
            *m0 += coeff * vol;
            for (int q = 0; q < 3; ++q) m[q] += coeff * vol * (a[q] + b[q] + d[q]) / 4;
        }
    }
}

int or_cp3_eval(const double *p0, const double *p1, int ncuts, const double *cuts,
                double *volume, int *nb_faces, double *areas, int64_t *ids, int cap) {
    or_cp3 *cp = (or_cp3 *)malloc(sizeof(or_cp3));
    or_cp3_box(cp, p0, p1, -1);
    for (int k = 0; k < ncuts; ++k) or_cp3_plane_cut(cp, cuts + 6 * k, cuts + 6 * k + 3, k);
    double c[3] = {0, 0, 0};
    *volume = or_cp3_volume(cp, c);
    int m = 0;
    for (int f = 0; f < cp->nf && m < cap; ++f) {
        if (cp->nv[f] < 3) continue;
        areas[m] = or_cp3_face_area(cp, f);
        ids[m] = cp->id[f];
        ++m;
    }
    *nb_faces = m;
    int st = cp->overflow ? OR_CAPACITY : OR_OK;
    free(cp);
    return st;
}

/* ------------------------------------------------------------------------- */
/* Neighbour search: uniform grid with the SpZGrid eviction rule              */
/* ------------------------------------------------------------------------- */

typedef struct {
    int dim;
    int64_t ncell[3];
    double lo[3], h[3];
    int64_t *start;    /* ncells + 1 */
    int64_t *items;    /* N dirac indices, grouped per cell, ascending within cell */
    double *wmax;      /* per cell max weight */
    double wglob;
} or_grid;

static int64_t grid_coord(const or_grid *g, int d, double x) {
    int64_t c = (int64_t)floor((x - g->lo[d]) / g->h[d]);
    if (c < 0) c = 0;
    if (c >= g->ncell[d]) c = g->ncell[d] - 1;
    return c;
}

static void grid_build(or_grid *g, int dim, int64_t N, const double *pos, const double *w) {
    g->dim = dim;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) { lo[d] = INFINITY; hi[d] = -INFINITY; }
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < dim; ++d) {
            lo[d] = fmin(lo[d], pos[dim * i + d]);
            hi[d] = fmax(hi[d], pos[dim * i + d]);
        }
    double vol = 1, lmax = 0;
    for (int d = 0; d < dim; ++d) lmax = fmax(lmax, hi[d] - lo[d]);
    for (int d = 0; d < dim; ++d) {
        double l = hi[d] - lo[d];
        double pad = 1e-9 * l + 1e-300;
        if (l + 2 * pad < 1e-3 * lmax) pad = 0.5 * (1e-3 * lmax - l);   /* flat inputs: keep cells isotropic */
        lo[d] -= pad; hi[d] += pad;
        vol *= hi[d] - lo[d];
    }
    double h = pow(vol * 2.0 / (double)(N > 0 ? N : 1), 1.0 / dim);
    int64_t tot = 1;
    for (int d = 0; d < 3; ++d) {
        if (d < dim) {
            int64_t nc = (int64_t)floor((hi[d] - lo[d]) / h);
            if (nc < 1) nc = 1;
            if (nc > 4096) nc = 4096;
            g->ncell[d] = nc;
            g->lo[d] = lo[d];
            g->h[d] = (hi[d] - lo[d]) / nc;
        } else {
            g->ncell[d] = 1; g->lo[d] = 0; g->h[d] = 1;
        }
        tot *= g->ncell[d];
    }
    g->start = (int64_t *)calloc(tot + 1, sizeof(int64_t));
    g->items = (int64_t *)malloc((N > 0 ? N : 1) * sizeof(int64_t));
    g->wmax = (double *)malloc(tot * sizeof(double));
    int64_t *cell = (int64_t *)malloc((N > 0 ? N : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < N; ++i) {
        int64_t c = 0, mul = 1;
        for (int d = 0; d < dim; ++d) { c += mul * grid_coord(g, d, pos[dim * i + d]); mul *= g->ncell[d]; }
        cell[i] = c;
        g->start[c + 1]++;
    }
    for (int64_t c = 0; c < tot; ++c) g->start[c + 1] += g->start[c];
    int64_t *cur = (int64_t *)malloc(tot * sizeof(int64_t));
    memcpy(cur, g->start, tot * sizeof(int64_t));
    for (int64_t i = 0; i < N; ++i) g->items[cur[cell[i]]++] = i;   /* ascending i within cell */
    g->wglob = -INFINITY;
    for (int64_t c = 0; c < tot; ++c) {
        double m = -INFINITY;
        for (int64_t k = g->start[c]; k < g->start[c + 1]; ++k) m = fmax(m, w[g->items[k]]);
        g->wmax[c] = m;
        g->wglob = fmax(g->wglob, m);
    }
    free(cur);
    free(cell);
}

static void grid_free(or_grid *g) { free(g->start); free(g->items); free(g->wmax); }

/* ------------------------------------------------------------------------- */
/* Cell evaluation                                                            */
/* ------------------------------------------------------------------------- */

typedef struct {
    int dim;
    int64_t N;
    const double *pos, *w;
    const or_domain *dom;
    int flags;
    double *masses, *ball_arc, *perimeter;
    double *centroids;     /* N*dim or NULL: mass-weighted centroid of each cell */
    int64_t cap;
    int64_t *row_len, *cols;
    double *vals;
    or_grid grid;
    /* get_der_centroids_and_integrals_wrt_weight_and_positions (2D, one shape): per
     * dirac up to cap (column, 3x3 block) items and v = (centroid, mass) */
    int64_t *jlen, *jcols;
    double *jblocks, *jv;
    int64_t next;          /* shared chunk counter */
    int64_t end;           /* cells [next0, end) are evaluated */
    pthread_mutex_t mtx;
    int status;
} cells_job;

typedef struct {
    int64_t *cols;
    double *vals;
    int64_t len, cap;
    double der0;
    int overflow;
} row_acc;

static void row_add(row_acc *r, int64_t j, double v) {
    for (int64_t k = 1; k < r->len; ++k)
        if (r->cols[k] == j) { r->vals[k] += v; return; }
    if (r->len >= r->cap) { r->overflow = 1; return; }
    r->cols[r->len] = j;
    r->vals[r->len] = v;
    r->len++;
}

static double dist_between(const double *pos, int dim, int64_t i, int64_t j) {
    double s = 0;
    for (int d = 0; d < dim; ++d) { double e = pos[dim * i + d] - pos[dim * j + d]; s += e * e; }
    return sqrt(s);
}

/* the radical cut of cell i by dirac j translated by t (t = NULL: no translation),
 * cut id `id` (SpZGrid.tcc:506-522, translated positions :610) */
static void power_cut2_t(or_cp2 *cp, const double *pos, const double *w, int64_t i, int64_t j, const double *t,
                         int64_t id) {
    double c0x = pos[2 * i], c0y = pos[2 * i + 1];
    double vx = pos[2 * j] + (t ? t[0] : 0.0) - c0x, vy = pos[2 * j + 1] + (t ? t[1] : 0.0) - c0y;
    double n = vx * vx + vy * vy;
    double x = 0.5 + 0.5 * (w[i] - w[j]) / n;
    or_cp2_plane_cut(cp, c0x + x * vx, c0y + x * vy, vx, vy, id, 0);
}

static void power_cut2(or_cp2 *cp, const double *pos, const double *w, int64_t i, int64_t j) {
    power_cut2_t(cp, pos, w, i, j, NULL, j);
}

/* a cut id's dirac and its position: j, or the image of j = id mod N under translation
 * id / N - 1 (get_der_integrals_wrt_weights.h:61-62) */
static int64_t cut_dirac(const cells_job *J, int64_t id, double *p) {
    const int dim = J->dim;
    int64_t j = id % J->N, s = id / J->N - 1;
    for (int d = 0; d < dim; ++d) p[d] = J->pos[dim * j + d] + (s >= 0 ? J->dom->trans[dim * s + d] : 0.0);
    return j;
}

/* Hessian terms of the face of measure `meas` between cell i and cut id `id`
 * (get_der_integrals_wrt_weights.h:56-70): the image of i itself adds -b to M_ii and
 * +b to the diagonal, i.e. nothing */
static void face_term(const cells_job *J, int64_t i, int64_t id, double meas, row_acc *row) {
    double p[3];
    int64_t j = cut_dirac(J, id, p);
    if (j == i) return;
    double dd = 0;
    for (int d = 0; d < J->dim; ++d) {
        double e = J->pos[J->dim * i + d] - p[d];
        dd += e * e;
    }
    dd = sqrt(dd);
    if (dd != 0) {
        double b = 0.5 * meas / dd;
        row_add(row, j, -b);
        row->der0 += b;
    }
}

/* ---- 2D ---- */

/* an image, or several assembly items (for_each_intersection,
 * ConvexPolyhedronAssembly.tcc:132-142): each cell is integrated piece by piece */
static int pieces_mode(const or_domain *dom) { return dom->img != NULL || dom->nitems >= 2; }

/* the per-piece part of the get_der_integrals_wrt_weights callback
 * (get_der_integrals_wrt_weights.h:52-75) for one 2D piece of density coeff */
static void piece2_accumulate(const cells_job *J, int64_t i, const or_cp2 *cp, double coeff, const double *poly,
                              double *mass, double *arcm, double *perim, row_acc *row, double *mom) {
    if (J->centroids && !poly) {
        double m0 = 0;
        or_cp2_moments(cp, coeff, J->pos[2 * i], J->pos[2 * i + 1], &m0, mom, mom + 1);
    }
    static __thread double meas[2 * OR_CP2_CAP + 2];
    static __thread int64_t ids[2 * OR_CP2_CAP + 2];
    int m;
    const int rf = (J->flags >> 4) & 3;   /* 0 Unit, 1 R2, 2 WmR2, 3 ExpWmR2db (the final integration function) */
    if (!poly && rf == 3) {
        const double eps = J->dom->rf_eps, mi = or_cp2_expwmr2db_integral(cp, coeff, J->w[i], eps);
        *mass += mi;
        m = or_cp2_boundary_items_expwmr2db(cp, coeff, J->w[i], eps, meas, ids, 2 * OR_CP2_CAP + 2);
        row->der0 += mi / eps;   /* integration_der_wrt_weight(ExpWmR2db) = integration / eps (:124-127) */
    } else if (!poly && rf) {
        *mass += or_cp2_radial_integral(cp, coeff, rf, J->w[i]);
        if (rf == 1) return;   /* R2 has no boundary items in the reference (TODO there) */
        m = or_cp2_boundary_items_wmr2(cp, coeff, J->w[i], meas, ids, 2 * OR_CP2_CAP + 2);
        row->der0 += or_cp2_area(cp, coeff);   /* integration_der_wrt_weight(WmR2) = the measure (:135-138) */
    } else if (!poly) {
        *mass += or_cp2_area(cp, coeff);
        m = or_cp2_boundary_items(cp, coeff, meas, ids, 2 * OR_CP2_CAP + 2);
    } else {                                     /* straight edges only (no ball with a polynomial) */
        *mass += or_cp2_poly_area(cp, poly);
        m = cp->n;
        for (int k = 0; k < m; ++k) {
            int k1 = k + 1 == m ? 0 : k + 1;
            ids[k] = cp->id[k];
            meas[k] = or_seg_poly(cp->x[k], cp->y[k], cp->x[k1], cp->y[k1], poly);
        }
    }
    double wi = J->w[i];
    for (int k = 0; k < m; ++k) {
        if (!pieces_mode(J->dom) || ids[k] != -1) *perim += meas[k];   /* pieces: their sides are not boundary */
        if (ids[k] == -1) continue;
        if (ids[k] == i) {
            *arcm += meas[k];
            row->der0 += 0.5 * meas[k] / sqrt(wi);
        } else {
            face_term(J, i, ids[k], meas[k], row);
        }
    }
}

/* can_be_evicted, SpZGrid.tcc:475-480: the box cannot cut if, for every vertex p,
 * |c0-p|^2 - w0 <= |clamp(p,box)-p|^2 - wmax */
static int evict2(const or_cp2 *cp, double c0x, double c0y, double w0, double bx0, double by0, double bx1, double by1,
                  double wmax) {
    for (int k = 0; k < cp->n; ++k) {
        double px = cp->x[k], py = cp->y[k];
        double qx = fmin(bx1, fmax(bx0, px)) - px, qy = fmin(by1, fmax(by0, py)) - py;
        double ax = c0x - px, ay = c0y - py;
        if (!(ax * ax + ay * ay - w0 <= qx * qx + qy * qy - wmax)) return 0;
    }
    return 1;
}

/* ConvexPolyhedronAssembly (ConvexPolyhedronAssembly.tcc:7-142): the polygon of item
 * k is its box cut by its planes (add_convex_polyhedron :7-33 cuts a box the same way) */
static void item_poly2(const or_domain *dom, int64_t k, or_cp2 *cp) {
    const double *b = dom->it_box + 6 * k;
    or_cp2_box(cp, b[0], b[1], b[3], b[4], -1);
    for (int64_t q = dom->it_off[k]; q < dom->it_off[k + 1] && cp->n; ++q) {
        const double *pl = dom->it_planes + 4 * q;
        double n2 = pl[0] * pl[0] + pl[1] * pl[1];
        or_cp2_plane_cut(cp, pl[3] * pl[0] / n2, pl[3] * pl[1] / n2, pl[0], pl[1], -1, 0);
    }
}

static void item_poly3(const or_domain *dom, int64_t k, or_cp3 *cp) {
    const double *b = dom->it_box + 6 * k;
    or_cp3_box(cp, b, b + 3, -1);
    for (int64_t q = dom->it_off[k]; q < dom->it_off[k + 1] && cp->nf; ++q) {
        const double *pl = dom->it_planes + 4 * q;
        double n2 = pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2], nn = sqrt(n2);
        double o[3] = {pl[3] * pl[0] / n2, pl[3] * pl[1] / n2, pl[3] * pl[2] / n2};
        double nu[3] = {pl[0] / nn, pl[1] / nn, pl[2] / nn};
        or_cp3_plane_cut(cp, o, nu, -1);
    }
}

/* englobing_convex_polyhedron (:60-92): one item is the item itself, several give a box
 * around them all, widened by its extent on each side */
static void englobing_box(const or_domain *dom, int dim, double *lo, double *hi) {
    for (int d = 0; d < dim; ++d) { lo[d] = INFINITY; hi[d] = -INFINITY; }
    for (int64_t k = 0; k < dom->nitems; ++k)
        for (int d = 0; d < dim; ++d) {
            lo[d] = fmin(lo[d], dom->it_box[6 * k + d]);
            hi[d] = fmax(hi[d], dom->it_box[6 * k + 3 + d]);
        }
    for (int d = 0; d < dim; ++d) {
        double e = hi[d] - lo[d];
        lo[d] -= e;
        hi[d] += e;
    }
}


static void start_cell2(const or_domain *dom, or_cp2 *cp) {
    if (dom->nitems == 1) {
        item_poly2(dom, 0, cp);
    } else if (dom->nitems >= 2) {
        double lo[3], hi[3];
        englobing_box(dom, 2, lo, hi);
        or_cp2_box(cp, lo[0], lo[1], hi[0], hi[1], -1);
    } else {
        or_cp2_box(cp, dom->p0[0], dom->p0[1], dom->p1[0], dom->p1[1], -1);
    }
}

static void start_cell3(const or_domain *dom, or_cp3 *cp) {
    if (dom->nitems == 1) {
        item_poly3(dom, 0, cp);
    } else if (dom->nitems >= 2) {
        double lo[3], hi[3];
        englobing_box(dom, 3, lo, hi);
        or_cp3_box(cp, lo, hi, -1);
    } else {
        or_cp3_box(cp, dom->p0, dom->p1, -1);
    }
}

/* Ring traversal of the grid around the cell of q = c0 - t: the diracs translated by
 * t (none for t = NULL) as candidates, cut ids j + idbase; the stop rule and the box
 * eviction are those of the untranslated traversal taken in the grid's frame. */
static void traverse2(const cells_job *J, int64_t i, or_cp2 *cp, const double *t, int64_t idbase) {
    const double *pos = J->pos, *w = J->w;
    const double tx = t ? t[0] : 0.0, ty = t ? t[1] : 0.0;
    double c0x = pos[2 * i], c0y = pos[2 * i + 1], w0 = w[i];
    const or_grid *g = &J->grid;
    int64_t hx = grid_coord(g, 0, c0x - tx), hy = grid_coord(g, 1, c0y - ty);
    int64_t maxr = g->ncell[0] > g->ncell[1] ? g->ncell[0] : g->ncell[1];
    for (int64_t r = 0; r <= maxr && cp->n; ++r) {
        if (r > 0) {
            /* every dirac in rings >= r lies outside the box of rings < r */
            double rx0 = g->lo[0] + (hx - r + 1) * g->h[0] + tx, rx1 = g->lo[0] + (hx + r) * g->h[0] + tx;
            double ry0 = g->lo[1] + (hy - r + 1) * g->h[1] + ty, ry1 = g->lo[1] + (hy + r) * g->h[1] + ty;
            int covers = hx - r + 1 <= 0 && hy - r + 1 <= 0 && hx + r >= g->ncell[0] && hy + r >= g->ncell[1];
            if (covers) break;
            int stop = 1;
            for (int k = 0; k < cp->n && stop; ++k) {
                double px = cp->x[k], py = cp->y[k];
                double dm = INFINITY;
                if (hx - r + 1 > 0) dm = fmin(dm, px - rx0);
                if (hx + r < g->ncell[0]) dm = fmin(dm, rx1 - px);
                if (hy - r + 1 > 0) dm = fmin(dm, py - ry0);
                if (hy + r < g->ncell[1]) dm = fmin(dm, ry1 - py);
                if (dm < 0) dm = 0;
                double ax = c0x - px, ay = c0y - py;
                if (!(ax * ax + ay * ay - w0 <= dm * dm - g->wglob)) stop = 0;
            }
            if (stop) break;
        }
        for (int64_t gy = hy - r < 0 ? 0 : hy - r; gy <= hy + r && gy < g->ncell[1]; ++gy) {
            int full = llabs(gy - hy) == r;
            for (int64_t gx = hx - r; gx <= hx + r; gx += (full || r == 0) ? 1 : 2 * r) {
                if (gx < 0 || gx >= g->ncell[0]) continue;
                int64_t c = gy * g->ncell[0] + gx;
                if (g->start[c] == g->start[c + 1]) continue;
                double bx0 = g->lo[0] + gx * g->h[0] + tx, by0 = g->lo[1] + gy * g->h[1] + ty;
                if (evict2(cp, c0x, c0y, w0, bx0, by0, bx0 + g->h[0], by0 + g->h[1], g->wmax[c])) continue;
                for (int64_t k = g->start[c]; k < g->start[c + 1]; ++k) {
                    int64_t j = g->items[k];
                    if (j != i || idbase) power_cut2_t(cp, pos, w, i, j, t, j + idbase);
                }
            }
        }
    }
}

static int build_cell2(const cells_job *J, int64_t i, or_cp2 *cp) {
    const or_domain *dom = J->dom;
    const double *pos = J->pos, *w = J->w;
    start_cell2(dom, cp);
    double c0x = pos[2 * i], c0y = pos[2 * i + 1], w0 = w[i];
    if (J->flags & 2) {
        for (int64_t j = 0; j < J->N; ++j)
            if (j != i) power_cut2(cp, pos, w, i, j);
        for (int64_t s = 0; s < dom->ntrans; ++s)
            for (int64_t j = 0; j < J->N; ++j) power_cut2_t(cp, pos, w, i, j, dom->trans + 2 * s, j + (s + 1) * J->N);
    } else {
        traverse2(J, i, cp, NULL, 0);
        for (int64_t s = 0; s < dom->ntrans && cp->n; ++s) traverse2(J, i, cp, dom->trans + 2 * s, (s + 1) * J->N);
    }
    if (J->flags & 1) or_cp2_ball_cut(cp, c0x, c0y, sqrt(w0), i);
    else { cp->scx = c0x; cp->scy = c0y; }
    return or_cp2_empty(cp) ? OR_EMPTY_CELL : OR_OK;
}

/* sort a cell's (column, block) items by column, merge duplicates, apply the mass /
 * centroid correction to the centroid rows (dim rows of nu = dim + 1 entries), store */
static void jac_finish(cells_job *J, int64_t i, int64_t m, int dim, double mass, const double *cent, int overflow) {
    const int nu = dim + 1, B = nu * nu;
    const int64_t cap = J->cap;
    int64_t *cols = J->jcols + i * cap;
    double *blk = J->jblocks + i * cap * 16;
    double tb[16];
    for (int64_t a = 1; a < m; ++a)
        for (int64_t b2 = a; b2 > 0 && cols[b2 - 1] > cols[b2]; --b2) {
            int64_t tc = cols[b2]; cols[b2] = cols[b2 - 1]; cols[b2 - 1] = tc;
            memcpy(tb, blk + 16 * b2, sizeof tb);
            memcpy(blk + 16 * b2, blk + 16 * (b2 - 1), sizeof tb);
            memcpy(blk + 16 * (b2 - 1), tb, sizeof tb);
        }
    int64_t w = 0;
    for (int64_t a = 0; a < m; ++a) {
        if (w > 0 && cols[w - 1] == cols[a]) {
            for (int k = 0; k < B; ++k) blk[16 * (w - 1) + k] += blk[16 * a + k];
            continue;
        }
        if (w != a) { cols[w] = cols[a]; memcpy(blk + 16 * w, blk + 16 * a, 16 * sizeof(double)); }
        ++w;
    }
    double *v = J->jv + nu * i;
    for (int d = 0; d < dim; ++d) v[d] = cent[d];
    v[dim] = mass;
    for (int64_t a = 0; a < w; ++a)
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < nu; ++c)
                blk[16 * a + nu * r + c] = (blk[16 * a + nu * r + c] - blk[16 * a + nu * dim + c] * v[r]) / (mass + (mass == 1));
    J->jlen[i] = w;
    if (overflow) {
        pthread_mutex_lock(&J->mtx);
        J->status = OR_CAPACITY;
        pthread_mutex_unlock(&J->mtx);
    }
}

/* get_der_centroids_and_integrals_wrt_weight_and_positions.h:46-190 for one 2D cell
 * integrated as one piece of density coeff: the boundary items of for_each_boundary_item
 * (ConvexPolyhedron2.tcc:437-477: straight edges carry pos_integral = l (p0 + p1) / 2 and
 * momentum[a][b] = l / 6 (2 (p0a p0b + p1a p1b) + p0a p1b + p1a p0b), not scaled by the
 * density, measure = coeff l; arcs carry their angles), block entries [3 d + e] for
 * output d (centroid x, y, mass) and unknown e (position x, y, weight) */
static void jac2_items(cells_job *J, int64_t i, const or_cp2 *cp, double coeff, double *der0, int64_t *mp,
                       int *overflow) {
    const int64_t cap = J->cap;
    int64_t *cols = J->jcols + i * cap;
    double *blk = J->jblocks + i * cap * 16;
    int64_t m = *mp;
    const double c0[2] = {J->pos[2 * i], J->pos[2 * i + 1]};
    const double R = sqrt(J->w[i]);   /* the ball's radius (:70), also for a void part whose sphere radius is 0 */
    int nitem = cp->n ? cp->n : (cp->sr >= 0 ? 1 : 0);
    for (int k = 0; k < nitem; ++k) {
        int64_t id;
        double measure, a0 = 1, a1 = 0, C[2] = {0, 0}, M[2][2] = {{0, 0}, {0, 0}};
        int is_arc;
        if (cp->n == 0) {                       /* the whole circle (measure 0 for a void part, :444-452) */
            id = cp->sid;
            measure = coeff * 2 * M_PI * cp->sr;
            is_arc = 1;
        } else {
            int i0 = k, i1 = k + 1 == cp->n ? 0 : k + 1;
            const double p0[2] = {cp->x[i0], cp->y[i0]}, p1[2] = {cp->x[i1], cp->y[i1]};
            id = cp->id[i0];
            is_arc = cp->arc[i0];
            if (is_arc) {
                a0 = atan2(p0[1] - cp->scy, p0[0] - cp->scx);
                a1 = atan2(p1[1] - cp->scy, p1[0] - cp->scx);
                if (a1 < a0) a1 += 2 * M_PI;
                measure = coeff * (a1 - a0) * R;
            } else {
                double l = sqrt((p1[0] - p0[0]) * (p1[0] - p0[0]) + (p1[1] - p0[1]) * (p1[1] - p0[1]));
                for (int a = 0; a < 2; ++a) C[a] = 0.5 * l * (p0[a] + p1[a]);
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        M[a][b] = l / 6 * (2 * (p0[a] * p0[b] + p1[a] * p1[b]) + p0[a] * p1[b] + p1[a] * p0[b]);
                measure = coeff * l;
            }
        }
        if (id == -1) continue;
        if (id == i) {                          /* the ball (:66-122) */
            double mm = 0.5 * measure / R;
            der0[8] += mm;
            double mx = 0, my = 0;
            if (a0 < a1) {
                mx = R * (sin(a1) - sin(a0));
                my = R * (cos(a0) - cos(a1));
                der0[6] += mx;
                der0[7] += my;
            }
            double T[2] = {mm * c0[0], mm * c0[1]};
            if (a0 < a1) {
                T[0] += 0.5 * R * (sin(a1) - sin(a0));
                T[1] += 0.5 * R * (cos(a0) - cos(a1));
            }
            der0[2] += T[0];
            der0[5] += T[1];
            if (a0 < a1) {
                der0[0] += mx * c0[0] + R * R / 2 * ((a1 + sin(a1) * cos(a1)) - (a0 + sin(a0) * cos(a0)));
                der0[1] += my * c0[0] + R * R / 4 * (cos(2 * a0) - cos(2 * a1));
                der0[3] += mx * c0[1] + R * R / 4 * (cos(2 * a0) - cos(2 * a1));
                der0[4] += my * c0[1] + R * R / 2 * ((a1 - sin(a1) * cos(a1)) - (a0 - sin(a0) * cos(a0)));
            } else {
                der0[0] += M_PI * R * R;
                der0[4] += M_PI * R * R;
            }
            continue;
        }
        /* a neighbour or the image of one (:124-166) */
        double q[2];
        int64_t j = cut_dirac(J, id, q);
        double dist = sqrt((q[0] - c0[0]) * (q[0] - c0[0]) + (q[1] - c0[1]) * (q[1] - c0[1]));
        double b = 0.5 * measure / dist, d1[9] = {0};
        der0[8] += b;
        d1[8] = -b;
        for (int e = 0; e < 2; ++e) {
            double v = 0.5 * C[e] / dist;
            der0[3 * e + 2] += v;
            d1[3 * e + 2] = -v;
        }
        for (int d = 0; d < 2; ++d) {
            der0[6 + d] += (C[d] - measure * c0[d]) / dist;
            d1[6 + d] = (measure * q[d] - C[d]) / dist;
        }
        for (int d = 0; d < 2; ++d)
            for (int e = 0; e < 2; ++e) {
                der0[3 * e + d] += (M[d][e] - C[e] * c0[d]) / dist;
                d1[3 * e + d] = (C[e] * q[d] - M[d][e]) / dist;
            }
        if (m >= cap) { *overflow = 1; continue; }
        cols[m] = j;
        memcpy(blk + 16 * m, d1, sizeof d1);
        ++m;
    }
    *mp = m;
}

/* one cell: the items of its pieces (one per domain shape or pixel under the cell, each
 * with its density: for_each_intersection, get_der_centroids_and_integrals_wrt_weight_and_positions.h:63),
 * then the diagonal block, sorting, merging and the correction */
static void jac2_close(cells_job *J, int64_t i, const double *der0, int64_t m, int overflow, double mass,
                       const double *cent) {
    const int64_t cap = J->cap;
    if (m >= cap) overflow = 1;
    else {
        J->jcols[i * cap + m] = i;
        memcpy(J->jblocks + (i * cap + m) * 16, der0, 9 * sizeof(double));
        ++m;
    }
    jac_finish(J, i, m, 2, mass, cent, overflow);
}

static void eval_cell2(cells_job *J, int64_t i, or_cp2 *cp, or_cp2 *piece) {
    row_acc row = { J->cols + i * J->cap, J->vals + i * J->cap, 1, J->cap, 0.0, 0 };
    row.cols[0] = i;
    double mass = 0, arcm = 0, perim = 0, mom[2] = {0, 0};
    double jder0[9] = {0};
    int64_t jm = 0;
    int jovf = 0;
    int st = build_cell2(J, i, cp);
    const or_domain *dom = J->dom;
    if (dom->nitems >= 2) {
        if (!or_cp2_empty(cp))
            for (int64_t k = 0; k < dom->nitems; ++k) {
                item_poly2(dom, k, piece);
                /* intersect_with, ConvexPolyhedron2.tcc:2475-2497 */
                for (int q = 0; q < cp->n && piece->n; ++q)
                    or_cp2_plane_cut(piece, cp->x[q], cp->y[q], cp->nx[q], cp->ny[q], cp->id[q], 1);
                if (cp->sr > 0) or_cp2_ball_cut(piece, cp->scx, cp->scy, cp->sr, cp->sid);
                else { piece->scx = cp->scx; piece->scy = cp->scy; piece->sr = cp->sr; piece->sid = cp->sid; }
                piece2_accumulate(J, i, piece, dom->it_coeff[k], NULL, &mass, &arcm, &perim, &row, mom);
                if (J->jblocks) jac2_items(J, i, piece, dom->it_coeff[k], jder0, &jm, &jovf);
            }
    } else if (!dom->img) {
        piece2_accumulate(J, i, cp, dom->nitems == 1 ? dom->it_coeff[0] : dom->coeff, NULL, &mass, &arcm, &perim, &row,
                          mom);
    } else {
        /* ScaledImage::for_each_intersection, ScaledImage.tcc:98-149 */
        double mn[2], mx[2];
        if (cp->n == 0) {
            double r = cp->sr > 0 ? cp->sr : 0;
            mn[0] = cp->scx - r; mx[0] = cp->scx + r; mn[1] = cp->scy - r; mx[1] = cp->scy + r;
        } else {
            /* min_position / max_position (ConvexPolyhedron2.tcc:502-523): an arc
             * contributes the box of its whole disc */
            mn[0] = mn[1] = DBL_MAX; mx[0] = mx[1] = -DBL_MAX;
            for (int k = 0; k < cp->n; ++k) {
                double lx = cp->x[k], ly = cp->y[k], hx = lx, hy = ly;
                if (cp->arc[k]) { lx = cp->scx - cp->sr; ly = cp->scy - cp->sr; hx = cp->scx + cp->sr; hy = cp->scy + cp->sr; }
                mn[0] = fmin(mn[0], lx); mx[0] = fmax(mx[0], hx);
                mn[1] = fmin(mn[1], ly); mx[1] = fmax(mx[1], hy);
            }
        }
        int64_t mi[2], ma[2];
        double ps[2];
        for (int d = 0; d < 2; ++d) {
            double L = dom->p1[d] - dom->p0[d];
            int64_t S = dom->img_sizes[d];
            double a = fmax(mn[d], dom->p0[d]), b = fmin(mx[d], dom->p1[d]);
            int64_t ia = (int64_t)((a - dom->p0[d]) * S / L), ib = (int64_t)((b - dom->p0[d]) * S / L);
            if (ia < 0) ia = 0;
            if (ib < 0) ib = 0;
            mi[d] = (ia > 1 ? ia : 1) - 1;
            ma[d] = ib + 2 < S ? ib + 2 : S;
            ps[d] = L / S;
        }
        if (!or_cp2_empty(cp)) {
            for (int64_t py = mi[1]; py < ma[1]; ++py)
                for (int64_t px = mi[0]; px < ma[0]; ++px) {
                    int64_t pix = py * dom->img_sizes[0] + px, npix = dom->img_sizes[0] * dom->img_sizes[1];
                    double rho = dom->img[pix], c6[6];
                    for (int k = 0; k < 6 && dom->img_nc == 6; ++k) c6[k] = dom->img[k * npix + pix];
                    or_cp2_box(piece, dom->p0[0] + ps[0] * px, dom->p0[1] + ps[1] * py,
                               dom->p0[0] + ps[0] * (px + 1), dom->p0[1] + ps[1] * (py + 1), -1);
                    /* intersect_with, ConvexPolyhedron2.tcc:2475-2497 */
                    for (int k = 0; k < cp->n; ++k)
                        or_cp2_plane_cut(piece, cp->x[k], cp->y[k], cp->nx[k], cp->ny[k], cp->id[k], 1);
                    if (cp->sr > 0) or_cp2_ball_cut(piece, cp->scx, cp->scy, cp->sr, cp->sid);
                    else { piece->scx = cp->scx; piece->scy = cp->scy; piece->sr = cp->sr; piece->sid = cp->sid; }
                    piece2_accumulate(J, i, piece, rho, dom->img_nc == 6 ? c6 : NULL, &mass, &arcm, &perim, &row, mom);
                    if (J->jblocks) jac2_items(J, i, piece, rho, jder0, &jm, &jovf);
                }
        }
    }
    row.vals[0] = row.der0;
    J->masses[i] = mass;
    if (J->centroids) {   /* get_centroids.h:53-58: centroid / (mass + (mass == 0)) */
        for (int d = 0; d < 2; ++d) J->centroids[2 * i + d] = mass != 0 ? J->pos[2 * i + d] + mom[d] / mass : 0.0;
    }
    if (J->jblocks) {     /* constant pieces only (checked by or_der_pos) */
        double cent[2];
        for (int d = 0; d < 2; ++d) cent[d] = mass != 0 ? J->pos[2 * i + d] + mom[d] / mass : 0.0;
        if (!pieces_mode(dom)) jac2_items(J, i, cp, dom->nitems == 1 ? dom->it_coeff[0] : dom->coeff, jder0, &jm, &jovf);
        jac2_close(J, i, jder0, jm, jovf, mass, cent);
    }
    if (J->ball_arc) J->ball_arc[i] = arcm;
    if (J->perimeter) J->perimeter[i] = perim;
    J->row_len[i] = row.len;
    if (row.overflow) st = OR_CAPACITY;
    if (st == OR_OK && mass == 0) st = OR_EMPTY_CELL;      /* get_der_integrals_wrt_weights.h:93-96 */
    if (st != OR_OK) {
        pthread_mutex_lock(&J->mtx);
        if (J->status == OR_OK || st == OR_CAPACITY) J->status = st;
        pthread_mutex_unlock(&J->mtx);
    }
}

/* ---- 3D ---- */

static void power_cut3_t(or_cp3 *cp, const double *pos, const double *w, int64_t i, int64_t j, const double *t,
                         int64_t id) {
    const double *c0 = pos + 3 * i, *c1 = pos + 3 * j;
    double V[3];
    for (int d = 0; d < 3; ++d) V[d] = c1[d] + (t ? t[d] : 0.0) - c0[d];
    double n = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
    double x = 0.5 + 0.5 * (w[i] - w[j]) / n;
    double o[3] = {c0[0] + x * V[0], c0[1] + x * V[1], c0[2] + x * V[2]};
    or_cp3_plane_cut(cp, o, V, id);
}

static void power_cut3(or_cp3 *cp, const double *pos, const double *w, int64_t i, int64_t j) {
    power_cut3_t(cp, pos, w, i, j, NULL, j);
}

static int evict3(const or_cp3 *cp, const double *c0, double w0, const double *b0, const double *b1, double wmax) {
    for (int f = 0; f < cp->nf; ++f)
        for (int k = 0; k < cp->nv[f]; ++k) {
            const double *p = cp->v[f][k];
            double q = 0, a = 0;
            for (int d = 0; d < 3; ++d) {
                double e = fmin(b1[d], fmax(b0[d], p[d])) - p[d];
                q += e * e;
                double g = c0[d] - p[d];
                a += g * g;
            }
            if (!(a - w0 <= q - wmax)) return 0;
        }
    return 1;
}

/* the 3D counterpart of traverse2 */
static void traverse3(const cells_job *J, int64_t i, or_cp3 *cp, const double *t, int64_t idbase) {
    const double *pos = J->pos, *w = J->w;
    const double *c0 = pos + 3 * i;
    const double sh[3] = {t ? t[0] : 0.0, t ? t[1] : 0.0, t ? t[2] : 0.0};
    double w0 = w[i];
    const or_grid *g = &J->grid;
    int64_t h[3];
    for (int d = 0; d < 3; ++d) h[d] = grid_coord(g, d, c0[d] - sh[d]);
    int64_t maxr = g->ncell[0];
    if (g->ncell[1] > maxr) maxr = g->ncell[1];
    if (g->ncell[2] > maxr) maxr = g->ncell[2];
    for (int64_t r = 0; r <= maxr && cp->nf; ++r) {
        if (r > 0) {
            int covers = 1;
            for (int d = 0; d < 3; ++d) covers &= h[d] - r + 1 <= 0 && h[d] + r >= g->ncell[d];
            if (covers) break;
            int stop = 1;
            for (int f = 0; f < cp->nf && stop; ++f)
                for (int k = 0; k < cp->nv[f] && stop; ++k) {
                    const double *p = cp->v[f][k];
                    double dm = INFINITY, a = 0;
                    for (int d = 0; d < 3; ++d) {
                        if (h[d] - r + 1 > 0) dm = fmin(dm, p[d] - (g->lo[d] + (h[d] - r + 1) * g->h[d] + sh[d]));
                        if (h[d] + r < g->ncell[d]) dm = fmin(dm, g->lo[d] + (h[d] + r) * g->h[d] + sh[d] - p[d]);
                        double e = c0[d] - p[d];
                        a += e * e;
                    }
                    if (dm < 0) dm = 0;
                    if (!(a - w0 <= dm * dm - g->wglob)) stop = 0;
                }
            if (stop) break;
        }
        for (int64_t gz = h[2] - r < 0 ? 0 : h[2] - r; gz <= h[2] + r && gz < g->ncell[2]; ++gz) {
            for (int64_t gy = h[1] - r < 0 ? 0 : h[1] - r; gy <= h[1] + r && gy < g->ncell[1]; ++gy) {
                int full = llabs(gz - h[2]) == r || llabs(gy - h[1]) == r;
                for (int64_t gx = h[0] - r; gx <= h[0] + r; gx += (full || r == 0) ? 1 : 2 * r) {
                    if (gx < 0 || gx >= g->ncell[0]) continue;
                    int64_t c = (gz * g->ncell[1] + gy) * g->ncell[0] + gx;
                    if (g->start[c] == g->start[c + 1]) continue;
                    double b0[3] = {g->lo[0] + gx * g->h[0] + sh[0], g->lo[1] + gy * g->h[1] + sh[1],
                                    g->lo[2] + gz * g->h[2] + sh[2]};
                    double b1[3] = {b0[0] + g->h[0], b0[1] + g->h[1], b0[2] + g->h[2]};
                    if (evict3(cp, c0, w0, b0, b1, g->wmax[c])) continue;
                    for (int64_t k = g->start[c]; k < g->start[c + 1]; ++k) {
                        int64_t j = g->items[k];
                        if (j != i || idbase) power_cut3_t(cp, pos, w, i, j, t, j + idbase);
                    }
                }
            }
        }
    }
}

static int build_cell3(const cells_job *J, int64_t i, or_cp3 *cp) {
    const or_domain *dom = J->dom;
    const double *pos = J->pos, *w = J->w;
    start_cell3(dom, cp);
    if (J->flags & 2) {
        for (int64_t j = 0; j < J->N && cp->nf; ++j)
            if (j != i) power_cut3(cp, pos, w, i, j);
        for (int64_t s = 0; s < dom->ntrans; ++s)
            for (int64_t j = 0; j < J->N && cp->nf; ++j)
                power_cut3_t(cp, pos, w, i, j, dom->trans + 3 * s, j + (s + 1) * J->N);
    } else {
        traverse3(J, i, cp, NULL, 0);
        for (int64_t s = 0; s < dom->ntrans && cp->nf; ++s) traverse3(J, i, cp, dom->trans + 3 * s, (s + 1) * J->N);
    }
    if (cp->overflow) return OR_CAPACITY;
    return cp->nf == 0 ? OR_EMPTY_CELL : OR_OK;             /* ConvexPolyhedron3.h:142 */
}

/* integration(Constant, R2), ConvexPolyhedron3.tcc:1120-1123 (Gauss quadrature, exact for
 * a quadratic): the integral of |x - c|^2 over the polyhedron as the tetrahedra (c, v0,
 * vk, vk+1) of each face, each V (sum |a|^2 + sum a.b) / 10 over its corners a, b seen
 * from c, V signed by the face's outward normal */
static double cp3_r2(const or_cp3 *cp, const double *c) {
    double res = 0;
    for (int f = 0; f < cp->nf; ++f) {
        const double (*v)[3] = cp->v[f];
        for (int k = 2; k < cp->nv[f]; ++k) {
            double a[3], b[3], e[3];
            for (int d = 0; d < 3; ++d) { a[d] = v[0][d] - c[d]; b[d] = v[k - 1][d] - c[d]; e[d] = v[k][d] - c[d]; }
            double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, q[3] = {e[0] - a[0], e[1] - a[1], e[2] - a[2]};
            double cr[3] = {u[1] * q[2] - u[2] * q[1], u[2] * q[0] - u[0] * q[2], u[0] * q[1] - u[1] * q[0]};
            double sg = cr[0] * cp->n[f][0] + cr[1] * cp->n[f][1] + cr[2] * cp->n[f][2] < 0 ? -1.0 : 1.0;
            double V = sg * (a[0] * cr[0] + a[1] * cr[1] + a[2] * cr[2]) / 6;
            double s2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + e[0] * e[0] +
                        e[1] * e[1] + e[2] * e[2] + a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[0] * e[0] + a[1] * e[1] +
                        a[2] * e[2] + b[0] * e[0] + b[1] * e[1] + b[2] * e[2];
            res += V * s2 / 10;
        }
    }
    return res;
}

static void piece3_accumulate(const cells_job *J, int64_t i, const or_cp3 *cp, double coeff, double *mass, double *perim,
                              row_acc *row, double *mom) {
    if (((J->flags >> 4) & 3) == 1) {   /* R2: integrals only (no boundary items in the reference) */
        *mass += coeff * cp3_r2(cp, J->pos + 3 * i);
        return;
    }
    *mass += coeff * or_cp3_volume(cp, J->pos + 3 * i);
    if (J->centroids) {
        double m0 = 0;
        or_cp3_moments(cp, J->pos + 3 * i, coeff, &m0, mom);
    }
    for (int f = 0; f < cp->nf; ++f) {
        if (cp->nv[f] < 3) continue;
        double a = coeff * or_cp3_face_area(cp, f);
        int64_t j = cp->id[f];
        if (!pieces_mode(J->dom) || j >= 0) *perim += a;   /* pieces: faces shared with neighbours only */
        if (j == -1) continue;
        face_term(J, i, j, a, row);
    }
}

/* get_der_centroids_and_integrals_wrt_weight_and_positions.h:124-166 for one piece of a
 * 3D cell (no ball: TODO in the reference), density coeff: the faces' boundary items of
 * ConvexPolyhedron3.tcc:226-265 -- pos_integral and momentum over the fan of triangles
 * (v0, vk, vk+1) of each face, not scaled by the density, measure = coeff * area.  The
 * pieces of a cell (for_each_intersection, :63) append their items in turn; jac3_close adds
 * the diagonal block and merges the items of one dirac. */
static void jac3_items(cells_job *J, int64_t i, const or_cp3 *cp, double coeff, double *der0, int64_t *mp,
                       int *overflowp) {
    const int64_t cap = J->cap;
    int64_t *cols = J->jcols + i * cap;
    double *blk = J->jblocks + i * cap * 16;
    int64_t m = *mp;
    const double *c0 = J->pos + 3 * i;
    int overflow = *overflowp;
    for (int f = 0; f < cp->nf; ++f) {
        if (cp->nv[f] < 3) continue;
        int64_t id = cp->id[f];
        if (id < 0) continue;
        const double (*v)[3] = cp->v[f];
        double C[3] = {0, 0, 0}, M[3][3] = {{0}}, area = 0;
        for (int k = 2; k < cp->nv[f]; ++k) {
            const double *A = v[0], *Bp = v[k - 1], *Cp = v[k];
            double e1[3] = {Bp[0] - A[0], Bp[1] - A[1], Bp[2] - A[2]}, e2[3] = {Cp[0] - A[0], Cp[1] - A[1], Cp[2] - A[2]};
            double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            double ta = 0.5 * (cr[0] * cp->n[f][0] + cr[1] * cp->n[f][1] + cr[2] * cp->n[f][2]);
            area += ta;
            for (int a = 0; a < 3; ++a) C[a] += ta * (A[a] + Bp[a] + Cp[a]) / 3;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    M[a][b] += ta / 12 * (A[a] * (2 * A[b] + Bp[b] + Cp[b]) + Bp[a] * (A[b] + 2 * Bp[b] + Cp[b]) +
                                          Cp[a] * (A[b] + Bp[b] + 2 * Cp[b]));
        }
        double q[3];
        int64_t j = cut_dirac(J, id, q);
        double dist = sqrt((q[0] - c0[0]) * (q[0] - c0[0]) + (q[1] - c0[1]) * (q[1] - c0[1]) + (q[2] - c0[2]) * (q[2] - c0[2]));
        double measure = coeff * area, b = 0.5 * measure / dist, d1[16] = {0};
        der0[15] += b;
        d1[15] = -b;
        for (int e = 0; e < 3; ++e) {
            double val = 0.5 * C[e] / dist;
            der0[4 * e + 3] += val;
            d1[4 * e + 3] = -val;
        }
        for (int d = 0; d < 3; ++d) {
            der0[12 + d] += (C[d] - measure * c0[d]) / dist;
            d1[12 + d] = (measure * q[d] - C[d]) / dist;
        }
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) {
                der0[4 * e + d] += (M[d][e] - C[e] * c0[d]) / dist;
                d1[4 * e + d] = (C[e] * q[d] - M[d][e]) / dist;
            }
        if (m >= cap) { overflow = 1; continue; }
        cols[m] = j;
        memcpy(blk + 16 * m, d1, sizeof d1);
        ++m;
    }
    *mp = m;
    *overflowp = overflow;
}

static void jac3_close(cells_job *J, int64_t i, const double *der0, int64_t m, int overflow, double mass,
                       const double *cent) {
    const int64_t cap = J->cap;
    if (m >= cap) overflow = 1;
    else {
        J->jcols[i * cap + m] = i;
        memcpy(J->jblocks + (i * cap + m) * 16, der0, 16 * sizeof(double));
        ++m;
    }
    jac_finish(J, i, m, 3, mass, cent, overflow);
}

static void eval_cell3(cells_job *J, int64_t i, or_cp3 *cp, or_cp3 *piece) {
    row_acc row = { J->cols + i * J->cap, J->vals + i * J->cap, 1, J->cap, 0.0, 0 };
    row.cols[0] = i;
    double mass = 0, perim = 0, mom[3] = {0, 0, 0};
    double jder0[16] = {0};
    int64_t jm = 0;
    int jovf = 0;
    int st = build_cell3(J, i, cp);
    const or_domain *dom = J->dom;
    if (dom->nitems >= 2) {
        if (cp->nf)
            for (int64_t k = 0; k < dom->nitems; ++k) {
                item_poly3(dom, k, piece);
                for (int f = 0; f < cp->nf && piece->nf; ++f) {
                    if (cp->nv[f] < 3) continue;
                    or_cp3_plane_cut(piece, cp->v[f][0], cp->n[f], cp->id[f]);
                }
                if (piece->nf) {
                    piece3_accumulate(J, i, piece, dom->it_coeff[k], &mass, &perim, &row, mom);
                    if (J->jblocks) jac3_items(J, i, piece, dom->it_coeff[k], jder0, &jm, &jovf);
                }
            }
    } else if (!dom->img) {
        const double coeff = dom->nitems == 1 ? dom->it_coeff[0] : dom->coeff;
        piece3_accumulate(J, i, cp, coeff, &mass, &perim, &row, mom);
        if (J->jblocks) jac3_items(J, i, cp, coeff, jder0, &jm, &jovf);
    } else if (cp->nf) {
        double mn[3], mx[3];
        for (int d = 0; d < 3; ++d) { mn[d] = INFINITY; mx[d] = -INFINITY; }
        for (int f = 0; f < cp->nf; ++f)
            for (int k = 0; k < cp->nv[f]; ++k)
                for (int d = 0; d < 3; ++d) { mn[d] = fmin(mn[d], cp->v[f][k][d]); mx[d] = fmax(mx[d], cp->v[f][k][d]); }
        int64_t mi[3], ma[3];
        double ps[3];
        for (int d = 0; d < 3; ++d) {
            double L = dom->p1[d] - dom->p0[d];
            int64_t S = dom->img_sizes[d];
            double a = fmax(mn[d], dom->p0[d]), b = fmin(mx[d], dom->p1[d]);
            int64_t ia = (int64_t)((a - dom->p0[d]) * S / L), ib = (int64_t)((b - dom->p0[d]) * S / L);
            if (ia < 0) ia = 0;
            if (ib < 0) ib = 0;
            mi[d] = (ia > 1 ? ia : 1) - 1;
            ma[d] = ib + 2 < S ? ib + 2 : S;
            ps[d] = L / S;
        }
        for (int64_t pz = mi[2]; pz < ma[2]; ++pz)
            for (int64_t py = mi[1]; py < ma[1]; ++py)
                for (int64_t px = mi[0]; px < ma[0]; ++px) {
                    double rho = dom->img[(pz * dom->img_sizes[1] + py) * dom->img_sizes[0] + px];
                    double q0[3] = {dom->p0[0] + ps[0] * px, dom->p0[1] + ps[1] * py, dom->p0[2] + ps[2] * pz};
                    double q1[3] = {dom->p0[0] + ps[0] * (px + 1), dom->p0[1] + ps[1] * (py + 1), dom->p0[2] + ps[2] * (pz + 1)};
                    or_cp3_box(piece, q0, q1, -1);
                    /* intersect_with, ConvexPolyhedron3.tcc:156-160: cut by every face plane of the cell */
                    for (int f = 0; f < cp->nf && piece->nf; ++f) {
                        if (cp->nv[f] < 3) continue;
                        or_cp3_plane_cut(piece, cp->v[f][0], cp->n[f], cp->id[f]);
                    }
                    if (piece->nf) {
                        piece3_accumulate(J, i, piece, rho, &mass, &perim, &row, mom);
                        if (J->jblocks) jac3_items(J, i, piece, rho, jder0, &jm, &jovf);
                    }
                }
    }
    row.vals[0] = row.der0;
    J->masses[i] = mass;
    if (J->centroids) {
        for (int d = 0; d < 3; ++d) J->centroids[3 * i + d] = mass != 0 ? J->pos[3 * i + d] + mom[d] / mass : 0.0;
    }
    if (J->jblocks) {
        double cent[3];
        for (int d = 0; d < 3; ++d) cent[d] = mass != 0 ? J->pos[3 * i + d] + mom[d] / mass : 0.0;
        jac3_close(J, i, jder0, jm, jovf, mass, cent);
    }
    if (J->ball_arc) J->ball_arc[i] = 0;
    if (J->perimeter) J->perimeter[i] = perim;
    J->row_len[i] = row.len;
    if (row.overflow) st = OR_CAPACITY;
    if (st == OR_OK && mass == 0) st = OR_EMPTY_CELL;
    if (st != OR_OK) {
        pthread_mutex_lock(&J->mtx);
        if (J->status == OR_OK || st == OR_CAPACITY) J->status = st;
        pthread_mutex_unlock(&J->mtx);
    }
}

static void *cells_worker(void *arg) {
    cells_job *J = (cells_job *)arg;
    void *a = NULL, *b = NULL;
    if (J->dim == 2) { a = malloc(sizeof(or_cp2)); b = malloc(sizeof(or_cp2)); }
    else { a = malloc(sizeof(or_cp3)); b = malloc(sizeof(or_cp3)); }
    const int64_t chunk = 64;
    for (;;) {
        int64_t beg = __atomic_fetch_add(&J->next, chunk, __ATOMIC_RELAXED);
        if (beg >= J->end) break;
        int64_t end = beg + chunk < J->end ? beg + chunk : J->end;
        for (int64_t i = beg; i < end; ++i) {
            if (J->dim == 2) eval_cell2(J, i, (or_cp2 *)a, (or_cp2 *)b);
            else eval_cell3(J, i, (or_cp3 *)a, (or_cp3 *)b);
        }
    }
    free(a);
    free(b);
    return NULL;
}

int or_cells(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
             int flags, int nthreads,
             double *masses, double *ball_arc, double *perimeter,
             int64_t cap, int64_t *row_len, int64_t *cols, double *vals) {
    return or_cells_range(dim, N, pos, w, dom, flags, nthreads, 0, N, masses, ball_arc, perimeter, cap, row_len, cols, vals);
}

static int cells_run(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
                     int flags, int nthreads, int64_t i_begin, int64_t i_end,
                     double *masses, double *ball_arc, double *perimeter, double *centroids,
                     int64_t cap, int64_t *row_len, int64_t *cols, double *vals);

int or_cells_range(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
                   int flags, int nthreads, int64_t i_begin, int64_t i_end,
                   double *masses, double *ball_arc, double *perimeter,
                   int64_t cap, int64_t *row_len, int64_t *cols, double *vals) {
    return cells_run(dim, N, pos, w, dom, flags, nthreads, i_begin, i_end, masses, ball_arc, perimeter, NULL, cap,
                     row_len, cols, vals);
}

int or_centroids(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom, int flags,
                 int nthreads, double *centroids, double *masses) {
    if (!centroids || !masses || N < 0) return OR_INVALID;
    if (dom && dom->img && dom->img_nc != 1) return OR_INVALID;   /* quadratic pixels: not restated */
    const int64_t cap = dim == 2 ? 64 : 160;
    int64_t *row_len = (int64_t *)malloc(sizeof(int64_t) * (N + 1));
    int64_t *cols = (int64_t *)malloc(sizeof(int64_t) * (N * cap + 1));
    double *vals = (double *)malloc(sizeof(double) * (N * cap + 1));
    int st = cells_run(dim, N, pos, w, dom, flags, nthreads, 0, N, masses, NULL, NULL, centroids, cap, row_len, cols,
                       vals);
    free(row_len);
    free(cols);
    free(vals);
    return st;
}

static int cells_run_j(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
                       int flags, int nthreads, int64_t i_begin, int64_t i_end,
                       double *masses, double *ball_arc, double *perimeter, double *centroids,
                       int64_t cap, int64_t *row_len, int64_t *cols, double *vals,
                       int64_t *jlen, int64_t *jcols, double *jblocks, double *jv);

static int cells_run(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
                     int flags, int nthreads, int64_t i_begin, int64_t i_end,
                     double *masses, double *ball_arc, double *perimeter, double *centroids,
                     int64_t cap, int64_t *row_len, int64_t *cols, double *vals) {
    return cells_run_j(dim, N, pos, w, dom, flags, nthreads, i_begin, i_end, masses, ball_arc, perimeter, centroids,
                       cap, row_len, cols, vals, NULL, NULL, NULL, NULL);
}

int or_der_pos(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom, int flags, int nthreads,
               int64_t cap, int64_t *jlen, int64_t *jcols, double *jblocks, double *jv) {
    if (!dom || dom->dim != dim || !jlen || !jcols || !jblocks || !jv) return OR_INVALID;
    if (dom->img && dom->img_nc != 1) return OR_INVALID;                /* constant pixels */
    double *masses = (double *)malloc(sizeof(double) * (N + 1));
    double *cent = (double *)malloc(sizeof(double) * (dim * N + 1));
    int64_t *row_len = (int64_t *)malloc(sizeof(int64_t) * (N + 1));
    int64_t *cols = (int64_t *)malloc(sizeof(int64_t) * (N * cap + 1));
    double *vals = (double *)malloc(sizeof(double) * (N * cap + 1));
    int st = cells_run_j(dim, N, pos, w, dom, flags, nthreads, 0, N, masses, NULL, NULL, cent, cap, row_len, cols, vals,
                         jlen, jcols, jblocks, jv);
    free(masses); free(cent); free(row_len); free(cols); free(vals);
    return st;
}

static int cells_run_j(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
                       int flags, int nthreads, int64_t i_begin, int64_t i_end,
                       double *masses, double *ball_arc, double *perimeter, double *centroids,
                       int64_t cap, int64_t *row_len, int64_t *cols, double *vals,
                       int64_t *jlen, int64_t *jcols, double *jblocks, double *jv) {
    if (i_begin < 0 || i_end > N || i_begin > i_end) return OR_INVALID;
    if ((dim != 2 && dim != 3) || N < 0 || cap < 1 || !dom || dom->dim != dim) return OR_INVALID;
    if (dim == 3 && (flags & 1)) return OR_INVALID;  /* ConvexPolyhedron3::ball_cut is TODO (:292-293) */
    if (dom->img && !(dom->img_nc == 1 || (dim == 2 && dom->img_nc == 6 && !(flags & 1))))
        return OR_INVALID;   /* apply_poly (Polynomial.h:27-58): 6 coefficients in 2D only, no ball (TODO there) */
    if (N == 0) return OR_OK;
    cells_job J;
    memset(&J, 0, sizeof(J));
    J.dim = dim; J.N = N; J.pos = pos; J.w = w; J.dom = dom; J.flags = flags;
    J.masses = masses; J.ball_arc = ball_arc; J.perimeter = perimeter; J.centroids = centroids;
    J.cap = cap; J.row_len = row_len; J.cols = cols; J.vals = vals;
    J.jlen = jlen; J.jcols = jcols; J.jblocks = jblocks; J.jv = jv;
    J.next = i_begin;
    J.end = i_end;
    J.status = OR_OK;
    pthread_mutex_init(&J.mtx, NULL);
    if (!(flags & 2)) grid_build(&J.grid, dim, N, pos, w);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, cells_worker, &J);
    cells_worker(&J);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    if (!(flags & 2)) grid_free(&J.grid);
    pthread_mutex_destroy(&J.mtx);
    return J.status;
}

/* ------------------------------------------------------------------------- */
/* CSR assembly, PCG, Newton                                                  */
/* ------------------------------------------------------------------------- */

typedef struct { int64_t c; double v; } cv_pair;

static int cmp_cv(const void *a, const void *b) {
    const cv_pair *x = (const cv_pair *)a, *y = (const cv_pair *)b;
    if (x->c != y->c) return (x->c > y->c) - (x->c < y->c);
    return (x->v > y->v) - (x->v < y->v);
}

int64_t or_rows_to_csr(int64_t N, int64_t cap, const int64_t *row_len, const int64_t *rcols, const double *rvals,
                       int64_t *offsets, int64_t *cols, double *vals) {
    cv_pair *tmp = (cv_pair *)malloc(sizeof(cv_pair) * (cap > 0 ? cap : 1));
    int64_t nnz = 0;
    for (int64_t i = 0; i < N; ++i) {
        offsets[i] = nnz;
        int64_t L = row_len[i];
        for (int64_t k = 0; k < L; ++k) { tmp[k].c = rcols[i * cap + k]; tmp[k].v = rvals[i * cap + k]; }
        qsort(tmp, L, sizeof(cv_pair), cmp_cv);                  /* std::sort of (col,val) pairs */
        for (int64_t k = 0; k < L; ++k) {
            if (k + 1 < L && tmp[k].c == tmp[k + 1].c) { tmp[k + 1].v += tmp[k].v; continue; }
            cols[nnz] = tmp[k].c;
            vals[nnz] = tmp[k].v;
            ++nnz;
        }
    }
    offsets[N] = nnz;
    free(tmp);
    return nnz;
}

static void spmv(int64_t N, const int64_t *off, const int64_t *cols, const double *vals, const double *x, double *y) {
    for (int64_t i = 0; i < N; ++i) {
        double s = 0;
        for (int64_t k = off[i]; k < off[i + 1]; ++k) s += vals[k] * x[cols[k]];
        y[i] = s;
    }
}

int or_pcg(int64_t N, const int64_t *offsets, const int64_t *cols, const double *vals,
           const double *b, double *x, double rtol, int max_iter, int *iters) {
    double *r = (double *)malloc(sizeof(double) * N), *z = (double *)malloc(sizeof(double) * N);
    double *p = (double *)malloc(sizeof(double) * N), *q = (double *)malloc(sizeof(double) * N);
    double *dinv = (double *)malloc(sizeof(double) * N);
    for (int64_t i = 0; i < N; ++i) {
        double d = 0;
        for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k)
            if (cols[k] == i) d += vals[k];
        dinv[i] = d != 0 ? 1 / d : 1;
        x[i] = 0;
        r[i] = b[i];
    }
    double bn = 0;
    for (int64_t i = 0; i < N; ++i) bn += b[i] * b[i];
    bn = sqrt(bn);
    int it = 0, st = OR_OK;
    if (bn == 0) { *iters = 0; goto done; }
    double rz = 0;
    for (int64_t i = 0; i < N; ++i) { z[i] = dinv[i] * r[i]; p[i] = z[i]; rz += r[i] * z[i]; }
    for (it = 0; it < max_iter; ++it) {
        spmv(N, offsets, cols, vals, p, q);
        double pq = 0;
        for (int64_t i = 0; i < N; ++i) pq += p[i] * q[i];
        double alpha = rz / pq;
        double rn = 0;
        for (int64_t i = 0; i < N; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; rn += r[i] * r[i]; }
        if (sqrt(rn) <= rtol * bn) { ++it; break; }
        double rz2 = 0;
        for (int64_t i = 0; i < N; ++i) { z[i] = dinv[i] * r[i]; rz2 += r[i] * z[i]; }
        double beta = rz2 / rz;
        rz = rz2;
        for (int64_t i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
    }
    if (it >= max_iter) st = OR_NOT_CONVERGED;
    *iters = it;
done:
    free(r); free(z); free(p); free(q); free(dinv);
    return st;
}

int or_ot_solve(int dim, int64_t N, const double *pos, double *w, const double *targets,
                const or_domain *dom, int flags, int nthreads, double tol, int stop_rule,
                int max_iter, or_solve_stats *stats) {
    const int64_t cap = dim == 2 ? 64 : 160;
    double *old = (double *)malloc(sizeof(double) * N), *m = (double *)malloc(sizeof(double) * N);
    double *dw = (double *)malloc(sizeof(double) * N), *b = (double *)malloc(sizeof(double) * N);
    int64_t *rl = (int64_t *)malloc(sizeof(int64_t) * N), *rc = (int64_t *)malloc(sizeof(int64_t) * N * cap);
    double *rv = (double *)malloc(sizeof(double) * N * cap);
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (N + 1)), *cols = (int64_t *)malloc(sizeof(int64_t) * N * cap);
    double *vals = (double *)malloc(sizeof(double) * N * cap);
    memcpy(old, w, sizeof(double) * N);
    or_solve_stats st;
    memset(&st, 0, sizeof(st));
    st.max_rel_mass_err = INFINITY;
    int status = OR_NOT_CONVERGED;
    for (int it = 0; it < max_iter; ++it) {
        int err = or_cells(dim, N, pos, w, dom, flags, nthreads, m, NULL, NULL, cap, rl, rc, rv);
        st.newton_iters = it + 1;
        if (err == OR_CAPACITY || err == OR_INVALID) { status = err; break; }
        if (err) {                                   /* OptimalTransportSolver.tcc:72-79 */
            for (int64_t i = 0; i < N; ++i) w[i] = 0.9 * old[i] + 0.1 * w[i];
            st.backtracks++;
            continue;
        }
        memcpy(old, w, sizeof(double) * N);
        double rel = 0;
        for (int64_t i = 0; i < N; ++i) {
            b[i] = m[i] - targets[i];
            rel = fmax(rel, fabs(b[i]) / targets[i]);
        }
        st.max_rel_mass_err = rel;
        if (stop_rule == 0 && rel < tol) { status = OR_OK; break; }
        st.nnz = or_rows_to_csr(N, cap, rl, rc, rv, off, cols, vals);
        if (!(flags & 1)) vals[off[0]] *= 2;          /* :62-63 gauge fix */
        int pit = 0;
        double rtol = fmax(1e-14, fmin(1e-3, 1e-3 * rel));
        or_pcg(N, off, cols, vals, b, dw, rtol, 100000, &pit);
        st.pcg_iters_total += pit;
        double mdw = 0;
        for (int64_t i = 0; i < N; ++i) { mdw = fmax(mdw, fabs(dw[i])); w[i] -= dw[i]; }
        st.max_abs_dw = mdw;
        if (stop_rule == 1 && mdw < tol) { status = OR_OK; break; }
    }
    if (stats) *stats = st;
    free(old); free(m); free(dw); free(b); free(rl); free(rc); free(rv); free(off); free(cols); free(vals);
    return status;
}

/* ------------------------------------------------------------------------- */
/* Inputs: splitmix64 -> xoshiro256** (Blackman & Vigna)                      */
/* ------------------------------------------------------------------------- */

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void or_rng_uniform(uint64_t seed, int64_t n, double *out) {
    uint64_t s[4], sm = seed;
    for (int k = 0; k < 4; ++k) s[k] = splitmix64(&sm);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        out[i] = (double)(r >> 11) * 0x1.0p-53;
    }
}
