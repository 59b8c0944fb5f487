// Cell geometry for display (display_vtk, src/sdot/PowerDiagram/display_vtk.h:13-25, and
// the VtkOutput legacy ASCII writer, src/sdot/Display/VtkOutput.tcc:17-79).
//
// The reference visits every Laguerre cell, intersects it with each piece of the domain
// (an item, a pixel, a voxel: bounds.for_each_intersection) and draws the piece: in 2D one
// polygon per piece whose arcs are cut into 20 segments (ConvexPolyhedron2.tcc:2499-2566:
// a disc with no vertex as 21 points, every polygon closed by repeating its first point),
// in 3D one polygon per face of the piece (ConvexPolyhedron3.tcc:1530-1546), each tagged with
// the dirac's index.
//
// Here each cell is rebuilt on the device from the pass that was just run: its faces are
// exactly its Hessian row (the neighbours j with an edge or face in the final cell) plus the
// domain, so one thread per cell clips a starting polygon by the radical half-planes of its
// row, the piece's box sides and planes, and in ball mode the disc.  3D faces are built one
// plane at a time: the polygon of plane k is a large square in that plane clipped by every
// other half-space of the piece.  A count kernel sizes the output, a fill kernel writes it;
// the host writes the file.  Periodic images are not drawn (a row keeps the dirac of an
// image face, not its translation).
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "internal.h"

namespace sdot_amd {

namespace {

constexpr int GCAP = 64;        // vertices of a polygon being clipped
constexpr int ARC_SEGS = 20;    // segments per arc (ConvexPolyhedron2.tcc:2553)

struct GeomArgs {
    int dim, ball;
    int64_t N;
    const double *spos, *sw, *mass;
    const int *row_len, *ell_col;
    int64_t ell_stride;
    const int *inv_perm;          // dirac -> sorted index
    // pieces: items (box lo[3], hi[3] and planes n.x <= h, absolute coordinates), or the
    // pixels / voxels of an image box
    int nitems;
    const double *it_box, *it_planes;
    const int *it_off;
    int img;
    double img_lo[3], img_hi[3];
    int img_n[3];
    double L;                     // half-size of the starting squares (beyond the domain)
    // phase 0: counts per dirac; phase 1: writes at the scanned offsets
    int64_t *cnt_pts, *cnt_poly;
    const int64_t *off_pts, *off_poly;
    double *pts;
    int *poly_len;
    int64_t *poly_owner;
    unsigned *flags;
};

// keep a x + b y <= c (Sutherland-Hodgman); returns the new vertex count, -1 on overflow
__device__ int clip_poly(const double *x, const double *y, int n, double a, double b, double c, double *ox,
                         double *oy) {
    int m = 0;
    for (int k = 0; k < n; ++k) {
        const int k1 = k + 1 == n ? 0 : k + 1;
        const double d0 = a * x[k] + b * y[k] - c, d1 = a * x[k1] + b * y[k1] - c;
        if (d0 <= 0) {
            if (m >= GCAP) return -1;
            ox[m] = x[k];
            oy[m] = y[k];
            ++m;
        }
        if ((d0 < 0 && d1 > 0) || (d0 > 0 && d1 < 0)) {
            if (m >= GCAP) return -1;
            const double u = d0 / (d0 - d1);
            ox[m] = x[k] + u * (x[k1] - x[k]);
            oy[m] = y[k] + u * (y[k1] - y[k]);
            ++m;
        }
    }
    return m;
}

// writer of one dirac's output: counts in phase 0, stores in phase 1
struct Emit {
    const GeomArgs *g;
    int phase;
    int64_t owner;
    int64_t np = 0, npoly = 0;    // emitted so far
    int64_t cur_start = 0;        // first point of the polygon being emitted
    int64_t po = 0, pp = 0;       // output offsets (phase 1)
    __device__ void begin() { cur_start = np; }
    __device__ void point(double x, double y, double z) {
        if (phase == 1) {
            double *q = g->pts + 3 * (po + np);
            q[0] = x;
            q[1] = y;
            q[2] = z;
        }
        ++np;
    }
    __device__ void end() {
        const int64_t len = np - cur_start;
        if (len < 2) {   // VtkOutput::add_polygon drops polygons of fewer than 2 points
            np = cur_start;
            return;
        }
        if (phase == 1) {
            g->poly_len[pp + npoly] = (int)len;
            g->poly_owner[pp + npoly] = owner;
        }
        ++npoly;
    }
};

// Half-space m of a piece, local coordinates (x relative to the dirac c): n.x <= h.
// Order: the piece's box sides (2 dim), its planes, then the cell's Hessian row.
struct Piece {
    double lo[3], hi[3];
    const double *planes;
    int nplanes;
};

__device__ inline int nb_halfspaces(const GeomArgs &g, const Piece &P, int row) { return 2 * g.dim + P.nplanes + row; }

__device__ void halfspace(const GeomArgs &g, const Piece &P, int64_t t, int m, const double *c, double w0, double *n,
                          double &h) {
    n[0] = n[1] = n[2] = 0;
    if (m < 2 * g.dim) {   // box side: -x_d <= c_d - lo_d, x_d <= hi_d - c_d
        const int d = m >> 1;
        if (m & 1) {
            n[d] = 1;
            h = P.hi[d] - c[d];
        } else {
            n[d] = -1;
            h = c[d] - P.lo[d];
        }
        return;
    }
    m -= 2 * g.dim;
    if (m < P.nplanes) {
        const double *pl = P.planes + 4 * m;
        double s = pl[3];
        for (int d = 0; d < g.dim; ++d) {
            n[d] = pl[d];
            s -= pl[d] * c[d];
        }
        h = s;
        return;
    }
    m -= P.nplanes;
    // radical half-plane of neighbour j (SpZGrid.tcc:506-522): x.v <= (|v|^2 + w_i - w_j) / 2
    const int j = g.ell_col[m * g.ell_stride + t];
    double vv = 0;
    for (int d = 0; d < g.dim; ++d) {
        n[d] = g.spos[g.dim * (int64_t)j + d] - c[d];
        vv += n[d] * n[d];
    }
    h = 0.5 * (vv + w0 - g.sw[j]);
}

// 2D: the piece's polygon (all half-spaces), then its disc clip streamed into the points
__device__ void piece2(const GeomArgs &g, const Piece &P, int64_t t, const double *c, double w0, Emit &E, double *bx,
                       double *by) {
    double x[GCAP], y[GCAP], tx[GCAP], ty[GCAP];
    const double L = g.L;
    int n = 4;
    x[0] = -L; y[0] = -L;
    x[1] = L;  y[1] = -L;
    x[2] = L;  y[2] = L;
    x[3] = -L; y[3] = L;
    const int row = g.row_len[t], M = nb_halfspaces(g, P, row);
    for (int m = 0; m < M && n > 0; ++m) {
        double nn[3], h;
        halfspace(g, P, t, m, c, w0, nn, h);
        const int k = clip_poly(x, y, n, nn[0], nn[1], h, tx, ty);
        if (k < 0) {
            atomicOr(g.flags, FLAG_CELL_OVERFLOW);
            return;
        }
        n = k;
        for (int q = 0; q < n; ++q) {
            x[q] = tx[q];
            y[q] = ty[q];
        }
    }
    if (bx) {   // the piece's bounding box (pixel range of a cell)
        bx[0] = by[0] = DBL_MAX;
        bx[1] = by[1] = -DBL_MAX;
        for (int q = 0; q < n; ++q) {
            bx[0] = fmin(bx[0], x[q]); bx[1] = fmax(bx[1], x[q]);
            by[0] = fmin(by[0], y[q]); by[1] = fmax(by[1], y[q]);
        }
        return;
    }
    if (!g.ball) {
        if (n < 3) return;
        E.begin();
        for (int q = 0; q < n; ++q) E.point(c[0] + x[q], c[1] + y[q], 0);
        E.point(c[0] + x[0], c[1] + y[0], 0);
        E.end();
        return;
    }
    // P & disc of radius r (CCW boundary): straight parts and arcs of ARC_SEGS segments
    const double r2 = fmax(w0, 0.0), r = sqrt(r2);
    if (!(r > 0) || n < 3) return;
    bool any_cross = false, all_in = true;
    for (int q = 0; q < n; ++q) all_in &= x[q] * x[q] + y[q] * y[q] <= r2;
    E.begin();
    if (all_in) {
        for (int q = 0; q < n; ++q) E.point(c[0] + x[q], c[1] + y[q], 0);
        E.point(c[0] + x[0], c[1] + y[0], 0);
        E.end();
        return;
    }
    // walk the edges; an exit point waits for the next entry point to draw its arc
    double fx = 0, fy = 0, ex = 0, ey = 0;
    bool have_first = false, have_exit = false;
    double sx = 0, sy = 0;   // first emitted point (closing)
    bool started = false;
    auto put = [&](double px, double py) {
        if (!started) {
            sx = px;
            sy = py;
            started = true;
        }
        E.point(c[0] + px, c[1] + py, 0);
    };
    auto arc = [&](double ax, double ay, double bx2, double by2) {
        const double a0 = atan2(ay, ax);
        double a1 = atan2(by2, bx2);
        if (a1 < a0) a1 += 2 * M_PI;
        for (int i = 0; i < ARC_SEGS; ++i) {
            const double ai = a0 + (a1 - a0) * i / ARC_SEGS;
            put(r * cos(ai), r * sin(ai));
        }
    };
    for (int q = 0; q < n; ++q) {
        const int q1 = q + 1 == n ? 0 : q + 1;
        const double px = x[q], py = y[q], dx = x[q1] - px, dy = y[q1] - py;
        // |p + s d|^2 = r^2: s^2 (d.d) + 2 s (p.d) + p.p - r^2 = 0
        const double A = dx * dx + dy * dy, B = px * dx + py * dy, C = px * px + py * py - r2;
        const double disc = B * B - A * C;
        if (!(A > 0) || disc <= 0) continue;   // the edge's line misses the disc
        const double sq = sqrt(disc);
        const double s0 = fmax(0.0, (-B - sq) / A), s1 = fmin(1.0, (-B + sq) / A);
        if (!(s0 < s1)) continue;
        any_cross = true;
        const double inx = px + s0 * dx, iny = py + s0 * dy, outx = px + s1 * dx, outy = py + s1 * dy;
        if (s0 > 0) {   // entering the disc on this edge
            if (have_exit) arc(ex, ey, inx, iny);
            else if (!have_first) {
                fx = inx;
                fy = iny;
                have_first = true;
            }
            have_exit = false;
            put(inx, iny);
        } else {        // the edge starts inside: its start is a vertex of the boundary
            put(px, py);
        }
        if (s1 < 1) {   // leaving the disc on this edge: the arc that follows starts here
            ex = outx;
            ey = outy;
            have_exit = true;
        }
    }
    if (!any_cross) {
        // no edge meets the disc: the whole disc when its centre (the dirac) is inside P
        bool centre_in = true;
        for (int q = 0; q < n; ++q) {
            const int q1 = q + 1 == n ? 0 : q + 1;
            centre_in &= x[q] * (y[q1] - y[q]) - y[q] * (x[q1] - x[q]) >= 0;   // left of every edge
        }
        if (!centre_in) return;
        for (int i = 0; i < ARC_SEGS; ++i) {
            const double ai = 2 * M_PI * i / ARC_SEGS;
            E.point(c[0] + r * cos(ai), c[1] + r * sin(ai), 0);
        }
        E.point(c[0] + r, c[1], 0);   // the 21st point: the first one again
        E.end();
        return;
    }
    if (have_exit && have_first) arc(ex, ey, fx, fy);   // back to the first entry
    if (started) put(sx, sy);                          // closed by its first point
    E.end();
}

// 3D: one polygon per plane of the piece that bounds it
__device__ void piece3(const GeomArgs &g, const Piece &P, int64_t t, const double *c, double w0, Emit &E,
                       double *bb) {
    double x[GCAP], y[GCAP], tx[GCAP], ty[GCAP];
    const int row = g.row_len[t], M = nb_halfspaces(g, P, row);
    if (bb)
        for (int d = 0; d < 3; ++d) {
            bb[2 * d] = DBL_MAX;
            bb[2 * d + 1] = -DBL_MAX;
        }
    for (int k = 0; k < M; ++k) {
        double nk[3], hk;
        halfspace(g, P, t, k, c, w0, nk, hk);
        const double nl = sqrt(nk[0] * nk[0] + nk[1] * nk[1] + nk[2] * nk[2]);
        if (!(nl > 0)) continue;
        const double nu[3] = {nk[0] / nl, nk[1] / nl, nk[2] / nl};
        const double o[3] = {nu[0] * hk / nl, nu[1] * hk / nl, nu[2] * hk / nl};
        // u: the axis least aligned with the normal, made orthogonal to it; v = nu x u
        double e[3] = {0, 0, 0};
        const double ax = fabs(nu[0]), ay = fabs(nu[1]), az = fabs(nu[2]);
        e[ax <= ay && ax <= az ? 0 : (ay <= az ? 1 : 2)] = 1;
        const double en = e[0] * nu[0] + e[1] * nu[1] + e[2] * nu[2];
        double u[3] = {e[0] - en * nu[0], e[1] - en * nu[1], e[2] - en * nu[2]};
        const double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        u[0] /= ul; u[1] /= ul; u[2] /= ul;
        const double v[3] = {nu[1] * u[2] - nu[2] * u[1], nu[2] * u[0] - nu[0] * u[2], nu[0] * u[1] - nu[1] * u[0]};
        const double L = g.L;
        int n = 4;
        x[0] = -L; y[0] = -L;
        x[1] = L;  y[1] = -L;
        x[2] = L;  y[2] = L;
        x[3] = -L; y[3] = L;
        bool dup = false;
        for (int m = 0; m < M && n > 0 && !dup; ++m) {
            if (m == k) continue;
            double nm[3], hm;
            halfspace(g, P, t, m, c, w0, nm, hm);
            const double A = nm[0] * u[0] + nm[1] * u[1] + nm[2] * u[2];
            const double B = nm[0] * v[0] + nm[1] * v[1] + nm[2] * v[2];
            const double Cc = hm - (nm[0] * o[0] + nm[1] * o[1] + nm[2] * o[2]);
            const double ml = sqrt(nm[0] * nm[0] + nm[1] * nm[1] + nm[2] * nm[2]);
            if (fabs(A) + fabs(B) <= 1e-13 * ml) {   // a plane parallel to k
                const bool same = nm[0] * nu[0] + nm[1] * nu[1] + nm[2] * nu[2] > 0;
                if (same && fabs(Cc) <= 1e-13 * ml * L) {
                    if (m < k) dup = true;   // the same plane twice: drawn once, for the first
                    continue;
                }
                if (Cc < 0) n = 0;           // plane k lies outside this half-space
                continue;
            }
            const int q = clip_poly(x, y, n, A, B, Cc, tx, ty);
            if (q < 0) {
                atomicOr(g.flags, FLAG_CELL_OVERFLOW);
                return;
            }
            n = q;
            for (int p = 0; p < n; ++p) {
                x[p] = tx[p];
                y[p] = ty[p];
            }
        }
        if (dup || n < 3) continue;
        double area = 0;
        for (int p = 0; p < n; ++p) {
            const int p1 = p + 1 == n ? 0 : p + 1;
            area += x[p] * y[p1] - x[p1] * y[p];
        }
        if (!(area > 1e-24 * L * L)) continue;   // a plane that touches the piece along an edge or a point
        if (bb) {
            for (int p = 0; p < n; ++p)
                for (int d = 0; d < 3; ++d) {
                    const double z = o[d] + x[p] * u[d] + y[p] * v[d];
                    bb[2 * d] = fmin(bb[2 * d], z);
                    bb[2 * d + 1] = fmax(bb[2 * d + 1], z);
                }
            continue;
        }
        E.begin();
        for (int p = 0; p < n; ++p)
            E.point(c[0] + o[0] + x[p] * u[0] + y[p] * v[0], c[1] + o[1] + x[p] * u[1] + y[p] * v[1],
                    c[2] + o[2] + x[p] * u[2] + y[p] * v[2]);
        E.end();
    }
}

// one thread per dirac (dirac order, so the file lists the cells in that order)
__global__ __launch_bounds__(64) void k_cell_geom(GeomArgs g, int phase) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int64_t t = g.inv_perm[i];
    Emit E;
    E.g = &g;
    E.phase = phase;
    E.owner = i;
    if (phase == 1) {
        E.po = g.off_pts[i];
        E.pp = g.off_poly[i];
    }
    if (g.mass[t] != 0) {   // an empty cell draws nothing
        const int dim = g.dim;
        double c[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) c[d] = g.spos[dim * t + d];
        const double w0 = g.sw[t];
        if (g.img) {
            // the pixels (voxels) under the cell's bounding box, each a piece
            Piece whole;
            for (int d = 0; d < 3; ++d) {
                whole.lo[d] = g.img_lo[d];
                whole.hi[d] = g.img_hi[d];
            }
            whole.planes = nullptr;
            whole.nplanes = 0;
            double bb[6];
            if (dim == 2) {
                piece2(g, whole, t, c, w0, E, bb, bb + 2);
                bb[4] = 0;
                bb[5] = 0;
            } else {
                piece3(g, whole, t, c, w0, E, bb);
            }
            if (g.ball && dim == 2) {   // pieces are clipped by the disc too: its box bounds them
                const double r = sqrt(fmax(w0, 0.0));
                bb[0] = fmax(bb[0], -r); bb[1] = fmin(bb[1], r);
                bb[2] = fmax(bb[2], -r); bb[3] = fmin(bb[3], r);
            }
            int i0[3] = {0, 0, 0}, i1[3] = {0, 0, 0};
            double vs[3] = {1, 1, 1};
            bool nonempty = true;
            for (int d = 0; d < dim; ++d) {
                nonempty &= bb[2 * d] <= bb[2 * d + 1];
                vs[d] = (g.img_hi[d] - g.img_lo[d]) / g.img_n[d];
                const int u0 = (int)floor((bb[2 * d] + c[d] - g.img_lo[d]) / vs[d]);
                const int u1 = (int)floor((bb[2 * d + 1] + c[d] - g.img_lo[d]) / vs[d]);
                i0[d] = u0 < 0 ? 0 : (u0 >= g.img_n[d] ? g.img_n[d] - 1 : u0);
                i1[d] = u1 < 0 ? 0 : (u1 >= g.img_n[d] ? g.img_n[d] - 1 : u1);
            }
            if (nonempty)
                for (int iz = i0[2]; iz <= i1[2]; ++iz)
                    for (int iy = i0[1]; iy <= i1[1]; ++iy)
                        for (int ix = i0[0]; ix <= i1[0]; ++ix) {
                            Piece px;
                            const int iv[3] = {ix, iy, iz};
                            for (int d = 0; d < 3; ++d) {
                                px.lo[d] = d < dim ? g.img_lo[d] + vs[d] * iv[d] : 0;
                                px.hi[d] = d < dim ? g.img_lo[d] + vs[d] * (iv[d] + 1) : 1;
                            }
                            px.planes = nullptr;
                            px.nplanes = 0;
                            if (dim == 2) piece2(g, px, t, c, w0, E, nullptr, nullptr);
                            else piece3(g, px, t, c, w0, E, nullptr);
                        }
        } else {
            for (int k = 0; k < g.nitems; ++k) {
                Piece P;
                for (int d = 0; d < 3; ++d) {
                    P.lo[d] = g.it_box[6 * k + d];
                    P.hi[d] = g.it_box[6 * k + 3 + d];
                }
                P.planes = g.it_planes + 4 * (int64_t)g.it_off[k];
                P.nplanes = g.it_off[k + 1] - g.it_off[k];
                if (dim == 2) piece2(g, P, t, c, w0, E, nullptr, nullptr);
                else piece3(g, P, t, c, w0, E, nullptr);
            }
        }
    }
    if (phase == 0) {
        g.cnt_pts[i] = E.np;
        g.cnt_poly[i] = E.npoly;
    }
}

inline unsigned nb64(int64_t n) { return (unsigned)((n + 63) / 64); }

}  // namespace

void write_vtk(sdot_pd *pd, const char *filename, const char *field_name) {
    if (!pd->htrans.empty())
        throw StatusError{SDOT_NOT_SUPPORTED, "display of periodic cells (a Hessian row does not keep the image)"};
    const int64_t N = pd->N;
    const int dim = pd->dim;
    hipStream_t s = pd->stream;
    GeomArgs g{};
    g.dim = dim;
    g.ball = pd->ball;
    g.N = N;
    g.spos = pd->spos.p;
    g.sw = pd->sw.p;
    g.mass = pd->mass.p;
    g.row_len = pd->row_len.p;
    g.ell_col = pd->ell_col.p;
    g.ell_stride = N;
    g.inv_perm = pd->inv_perm.p;
    // pieces of the domain: the image's pixels, or the shapes (a box and its planes each)
    double ext = 0;
    std::vector<double> box, planes;
    std::vector<int> off(1, 0);
    if (pd->has_image) {
        g.img = 1;
        for (int d = 0; d < 3; ++d) {
            g.img_lo[d] = d < dim ? pd->img_lo[d] : 0;
            g.img_hi[d] = d < dim ? pd->img_hi[d] : 1;
            g.img_n[d] = d < dim ? pd->img_n[d] : 1;
            ext = std::max(ext, g.img_hi[d] - g.img_lo[d]);
        }
    } else {
        for (const Shape &sh : pd->shapes) {
            for (int d = 0; d < 3; ++d) box.push_back(d < dim ? sh.p0[d] : 0);
            for (int d = 0; d < 3; ++d) box.push_back(d < dim ? sh.p1[d] : 1);
            for (int d = 0; d < dim; ++d) ext = std::max(ext, sh.p1[d] - sh.p0[d]);
            for (const auto &pl : sh.planes) planes.insert(planes.end(), pl.begin(), pl.end());
            off.push_back((int)(planes.size() / 4));
        }
        if (planes.empty()) planes.assign(4, 0.0);
        g.nitems = (int)pd->shapes.size();
    }
    g.L = 4 * ext + 1;
    DevBuf<double> dbox, dplanes;
    DevBuf<int> doff;
    if (!pd->has_image) {
        dbox.ensure(box.size());
        dplanes.ensure(planes.size());
        doff.ensure(off.size());
        HIPCK(hipMemcpyAsync(dbox.p, box.data(), box.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(dplanes.p, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(doff.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, s));
        g.it_box = dbox.p;
        g.it_planes = dplanes.p;
        g.it_off = doff.p;
    }
    DevBuf<int64_t> cnt_pts, cnt_poly, off_pts, off_poly, tmp, owner;
    DevBuf<double> pts;
    DevBuf<int> plen;
    DevBuf<unsigned> flags;
    auto release = [&] {
        for (auto *b : {&cnt_pts, &cnt_poly, &off_pts, &off_poly, &tmp, &owner}) b->release();
        pts.release();
        plen.release();
        flags.release();
        dbox.release();
        dplanes.release();
        doff.release();
    };
    try {
        cnt_pts.ensure(N);
        cnt_poly.ensure(N);
        off_pts.ensure(N + 1);
        off_poly.ensure(N + 1);
        tmp.ensure(scan_tmp_size(N));
        flags.ensure(1);
        fill_u32(flags.p, 1, 0u, s);
        g.cnt_pts = cnt_pts.p;
        g.cnt_poly = cnt_poly.p;
        g.flags = flags.p;
        hipLaunchKernelGGL(k_cell_geom, dim3(nb64(N)), dim3(64), 0, s, g, 0);
        HIPCK(hipGetLastError());
        exclusive_scan<int64_t>(cnt_pts.p, off_pts.p, N, tmp.p, s);
        exclusive_scan<int64_t>(cnt_poly.p, off_poly.p, N, tmp.p, s);
        int64_t np = 0, npoly = 0;
        unsigned fl = 0;
        HIPCK(hipMemcpyAsync(&np, off_pts.p + N, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(&npoly, off_poly.p + N, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(&fl, flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        if (fl) throw StatusError{SDOT_CAPACITY_OVERFLOW, "a polygon of the display outgrew 64 vertices"};
        pts.ensure(3 * std::max<int64_t>(np, 1));
        plen.ensure(std::max<int64_t>(npoly, 1));
        owner.ensure(std::max<int64_t>(npoly, 1));
        g.off_pts = off_pts.p;
        g.off_poly = off_poly.p;
        g.pts = pts.p;
        g.poly_len = plen.p;
        g.poly_owner = owner.p;
        hipLaunchKernelGGL(k_cell_geom, dim3(nb64(N)), dim3(64), 0, s, g, 1);
        HIPCK(hipGetLastError());
        std::vector<double> hp(3 * np);
        std::vector<int> hl(npoly);
        std::vector<int64_t> ho(npoly);
        HIPCK(hipMemcpyAsync(hp.data(), pts.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(hl.data(), plen.p, hl.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(ho.data(), owner.p, ho.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        release();
        // legacy ASCII VTK, the layout of VtkOutput::save (VtkOutput.tcc:23-79): points, cells
        // (type 7, polygon), one cell field
        FILE *f = fopen(filename, "w");
        if (!f) throw StatusError{SDOT_INVALID_ARG, std::string("cannot open ") + filename};
        fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
        fprintf(f, "POINTS %lld double\n", (long long)np);
        for (int64_t k = 0; k < np; ++k) fprintf(f, "%.17g %.17g %.17g\n", hp[3 * k], hp[3 * k + 1], hp[3 * k + 2]);
        fprintf(f, "CELLS %lld %lld\n", (long long)npoly, (long long)(npoly + np));
        int64_t o = 0;
        for (int64_t k = 0; k < npoly; ++k) {
            fprintf(f, "%d", hl[k]);
            for (int q = 0; q < hl[k]; ++q) fprintf(f, " %lld", (long long)(o++));
            fprintf(f, "\n");
        }
        fprintf(f, "CELL_TYPES %lld\n", (long long)npoly);
        for (int64_t k = 0; k < npoly; ++k) fprintf(f, "7\n");
        fprintf(f, "CELL_DATA %lld\nFIELD FieldData 1\n%s 1 %lld float\n", (long long)npoly, field_name, (long long)npoly);
        for (int64_t k = 0; k < npoly; ++k) fprintf(f, " %lld", (long long)ho[k]);
        fprintf(f, "\n");
        const bool ok = ferror(f) == 0;
        if (fclose(f) != 0 || !ok) throw StatusError{SDOT_INVALID_ARG, std::string("write error on ") + filename};
    } catch (...) {
        release();
        throw;
    }
}

}  // namespace sdot_amd
