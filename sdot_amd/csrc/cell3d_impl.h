// Per-cell routine of the 3D Laguerre-cell pass, one thread per cell, with
// the polyhedron in caller-provided memory (host+device: the wide-capacity
// path for cells that outgrow the wavefront kernel, and a CPU check of the
// algorithm).
//
// The polyhedron is kept in the dual form: a list of planes (the half-spaces
// q.n <= h in coordinates local to the dirac; radical planes of neighbours,
// box sides, voxel sides) and a list of vertices, each the intersection of
// three planes (a,b,c) with det(n_a,n_b,n_c) > 0 (counter-clockwise seen from
// outside).  A cut removes the vertices outside the new plane and creates one
// vertex on every edge (a,b) that joins a removed vertex to a kept one, with
// planes (a,b,new).  This is the semantics of ConvexPolyhedron3::plane_cut
// (src/sdot/Geometry/ConvexPolyhedron3.tcc:614-754: graph walk from an outside
// vertex, split cut edges, close the new face) on a different representation.
//
// Integrals (ConvexPolyhedron3.tcc:1022-1056 volume, :1367-1418 face area):
// face a of vertex T=(a,b,c) continues to the vertex across the dual edge
// (a,b); the face area is |sum (p_T x p_next).n_a| / 2 and the volume
// sum_a dist(c0, plane a) * area_a / 3 (pyramids from the dirac).
// Hessian rows and the voxel pieces of a ScaledImage follow
// get_der_integrals_wrt_weights.h:52-75 and ScaledImage.tcc:98-149.
#pragma once

#include <cfloat>
#include <cmath>

#include "internal.h"

namespace sdot_amd {
namespace c3 {

// polyhedron storage of one cell: element k at [k * S]
template <class T>
struct Poly3Mem {
    T *px, *py, *pz;       // vertex positions (local coordinates)   [VCAP]
    int *tri;              // planes a | b << 8 | c << 16            [VCAP]
    T *d;                  // per-vertex scratch                      [VCAP]
    T *nx, *ny, *nz, *nh;  // planes q.n <= h                         [PCAP]
    int *pid;              // plane id: neighbour (sorted index) or < 0 [PCAP]
    double *acc;           // per-plane accumulators                  [PCAP]
    double *hacc;          // per-plane rho-weighted areas            [PCAP]
    // saved copy of the cell (voxel pieces restore it)
    T *sx, *sy, *sz;
    int *stri;
    int S;
    // dual-edge map: vertex holding the directed edge (x -> y) at [(x * PCAP + y) * S];
    // null: edges are found by a scan of the vertices
    unsigned short *emap = nullptr;
    // per-plane first moments (3) and second moments (xx, xy, xz, yy, yz, zz) of the
    // face, local coordinates, for the position derivatives; null: not wanted
    double *jmom = nullptr;   // [(k * 9 + c) * S]
    // with domain pieces: per cell plane, the face's area and its 9 moments summed over
    // the pieces (each piece integrates its own part of the face)
    double *jsum = nullptr;   // [(k * 10 + c) * S]
};

// voxel / box plane ids: box side k -> -1 - k (k = 0..5: -x,+x,-y,+y,-z,+z);
// voxel side k -> -7 - k.
SDOT_HD inline bool is_neighbour(int id) { return id >= 0; }

template <class T>
struct Cell3 {
    const CellArgs &a;
    Poly3Mem<T> m;
    int VCAP, PCAP;
    int nv = 0, np = 0;
    double c0x, c0y, c0z, w0;
    // current voxel box (local coordinates) for the planes of ids -7..-12
    double vlo[3], vhi[3];

    SDOT_HD Cell3(const CellArgs &a_, const Poly3Mem<T> &m_, int vcap, int pcap) : a(a_), m(m_), VCAP(vcap), PCAP(pcap) {}

#define V_X(k) m.px[(k) * m.S]
#define V_Y(k) m.py[(k) * m.S]
#define V_Z(k) m.pz[(k) * m.S]
#define V_T(k) m.tri[(k) * m.S]
#define V_D(k) m.d[(k) * m.S]
#define P_X(k) m.nx[(k) * m.S]
#define P_Y(k) m.ny[(k) * m.S]
#define P_Z(k) m.nz[(k) * m.S]
#define P_H(k) m.nh[(k) * m.S]
#define P_I(k) m.pid[(k) * m.S]

    // exact (FP64) plane of slot k: n.q = h
    SDOT_HD void plane64(int k, double &nx, double &ny, double &nz, double &h) const {
        const int id = P_I(k);
        if (id >= 0) {
            double q[3];
            const int64_t j = cut_dirac<3>(a, id, q);
            nx = q[0] - c0x;
            ny = q[1] - c0y;
            nz = q[2] - c0z;
            h = 0.5 * (nx * nx + ny * ny + nz * nz + w0 - a.sw[j]);
            return;
        }
        if (id <= -100) {   // plane -100 - id of the domain item (n . x <= h)
            const double *pl = a.it_planes + 4 * (-100 - id);
            nx = pl[0];
            ny = pl[1];
            nz = pl[2];
            h = pl[3] - (pl[0] * c0x + pl[1] * c0y + pl[2] * c0z);
            return;
        }
        const int s = id >= -6 ? -1 - id : -7 - id;
        const int axis = s >> 1, up = s & 1;
        const double c0[3] = {c0x, c0y, c0z};
        double off;
        if (id >= -6) off = (up ? sel3(a.dom_hi, axis) : sel3(a.dom_lo, axis)) - sel3(c0, axis);
        else off = up ? sel3(vhi, axis) : sel3(vlo, axis);
        nx = axis == 0 ? (up ? 1.0 : -1.0) : 0.0;
        ny = axis == 1 ? (up ? 1.0 : -1.0) : 0.0;
        nz = axis == 2 ? (up ? 1.0 : -1.0) : 0.0;
        h = up ? off : -off;
    }

    SDOT_HD void add_plane(int k, double nx, double ny, double nz, double h, int id) {
        P_X(k) = T(nx);
        P_Y(k) = T(ny);
        P_Z(k) = T(nz);
        P_H(k) = T(h);
        P_I(k) = id;
    }

    // starting polyhedron: the box [lo, hi] (absolute), planes -1..-6
    SDOT_HD void init_box() {
        const double c0[3] = {c0x, c0y, c0z};
        for (int s = 0; s < 6; ++s) {
            const int axis = s >> 1, up = s & 1;
            double off = (up ? sel3(a.dom_hi, axis) : sel3(a.dom_lo, axis)) - sel3(c0, axis);
            add_plane(s, axis == 0 ? (up ? 1 : -1) : 0, axis == 1 ? (up ? 1 : -1) : 0, axis == 2 ? (up ? 1 : -1) : 0,
                      up ? off : -off, -1 - s);
        }
        np = 6;
        nv = 8;
        for (int k = 0; k < 8; ++k) {
            const int sx = k & 1, sy = (k >> 1) & 1, sz = (k >> 2) & 1;
            int pa = sx, pb = 2 + sy, pc = 4 + sz;
            // det(n_a, n_b, n_c) = product of the three signs: swap when negative
            const int sign = (sx ? 1 : -1) * (sy ? 1 : -1) * (sz ? 1 : -1);
            if (sign < 0) {
                int t = pb;
                pb = pc;
                pc = t;
            }
            V_T(k) = pa | (pb << 8) | (pc << 16);
            V_X(k) = T((sx ? a.dom_hi[0] : a.dom_lo[0]) - c0x);
            V_Y(k) = T((sy ? a.dom_hi[1] : a.dom_lo[1]) - c0y);
            V_Z(k) = T((sz ? a.dom_hi[2] : a.dom_lo[2]) - c0z);
        }
        write_edges();
    }

    SDOT_HD double radius2() const {
        double R2 = 0;
        for (int k = 0; k < nv; ++k) {
            double x = V_X(k), y = V_Y(k), z = V_Z(k);
            R2 = fmax(R2, x * x + y * y + z * z);
        }
        return R2;
    }

    // record the three dual edges of every vertex (after any change of the vertex list)
    SDOT_HD void write_edges() {
        if (!m.emap) return;
        for (int k = 0; k < nv; ++k) {
            const int t = V_T(k);
            const int p0 = t & 255, p1 = (t >> 8) & 255, p2 = (t >> 16) & 255;
            m.emap[(p0 * PCAP + p1) * m.S] = (unsigned short)k;
            m.emap[(p1 * PCAP + p2) * m.S] = (unsigned short)k;
            m.emap[(p2 * PCAP + p0) * m.S] = (unsigned short)k;
        }
    }

    // vertex holding the directed dual edge (x -> y)
    SDOT_HD int find_edge(int x, int y) const {
        if (m.emap) {   // the map holds every current edge; stale entries are rejected by the check
            const int k = m.emap[(x * PCAP + y) * m.S];
            if (k < nv) {
                const int t = V_T(k);
                const int p0 = t & 255, p1 = (t >> 8) & 255, p2 = (t >> 16) & 255;
                if ((p0 == x && p1 == y) || (p1 == x && p2 == y) || (p2 == x && p0 == y)) return k;
            }
        }
        for (int k = 0; k < nv; ++k) {
            const int t = V_T(k);
            const int p0 = t & 255, p1 = (t >> 8) & 255, p2 = (t >> 16) & 255;
            if ((p0 == x && p1 == y) || (p1 == x && p2 == y) || (p2 == x && p0 == y)) return k;
        }
        return -1;
    }

    // Side of vertex k for the plane n.q <= h, exactly from the planes: with rows
    // [n, h] of the vertex's planes a, b, c and of the new plane,
    // det4 = (h - n.v) det3 and det3 = det(n_a, n_b, n_c) > 0 (orientation).
    SDOT_HD double tie_distance(int k, double nx, double ny, double nz, double h) const {
        const int t = V_T(k);
        double q[4][4];
        plane64(t & 255, q[0][0], q[0][1], q[0][2], q[0][3]);
        plane64((t >> 8) & 255, q[1][0], q[1][1], q[1][2], q[1][3]);
        plane64((t >> 16) & 255, q[2][0], q[2][1], q[2][2], q[2][3]);
        q[3][0] = nx;
        q[3][1] = ny;
        q[3][2] = nz;
        q[3][3] = h;
        const double s0 = q[0][0] * q[1][1] - q[1][0] * q[0][1], s1 = q[0][0] * q[1][2] - q[1][0] * q[0][2];
        const double s2 = q[0][0] * q[1][3] - q[1][0] * q[0][3], s3 = q[0][1] * q[1][2] - q[1][1] * q[0][2];
        const double s4 = q[0][1] * q[1][3] - q[1][1] * q[0][3], s5 = q[0][2] * q[1][3] - q[1][2] * q[0][3];
        const double c5 = q[2][2] * q[3][3] - q[3][2] * q[2][3], c4 = q[2][1] * q[3][3] - q[3][1] * q[2][3];
        const double c3 = q[2][1] * q[3][2] - q[3][1] * q[2][2], c2 = q[2][0] * q[3][3] - q[3][0] * q[2][3];
        const double c1 = q[2][0] * q[3][2] - q[3][0] * q[2][2], c0 = q[2][0] * q[3][1] - q[3][0] * q[2][1];
        const double det4 = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
        const double det3 = q[0][0] * (q[1][1] * q[2][2] - q[1][2] * q[2][1]) - q[0][1] * (q[1][0] * q[2][2] - q[1][2] * q[2][0]) +
                            q[0][2] * (q[1][0] * q[2][1] - q[1][1] * q[2][0]);
        return det3 > 0 ? det4 / det3 : (double)m.d[k * m.S];
    }

    // renumber the planes still bounding a face, freeing the slots of removed ones
    SDOT_HD void compact_planes() {
        int newidx[256];
        for (int k = 0; k < np; ++k) newidx[k] = -1;
        for (int k = 0; k < nv; ++k) {
            const int t = V_T(k);
            newidx[t & 255] = newidx[(t >> 8) & 255] = newidx[(t >> 16) & 255] = 0;
        }
        int w = 0;
        for (int k = 0; k < np; ++k) {
            if (newidx[k] < 0) continue;
            newidx[k] = w;
            if (w != k) {
                P_X(w) = P_X(k);
                P_Y(w) = P_Y(k);
                P_Z(w) = P_Z(k);
                P_H(w) = P_H(k);
                P_I(w) = P_I(k);
            }
            ++w;
        }
        for (int k = 0; k < nv; ++k) {
            const int t = V_T(k);
            V_T(k) = newidx[t & 255] | (newidx[(t >> 8) & 255] << 8) | (newidx[(t >> 16) & 255] << 16);
        }
        np = w;
        write_edges();
    }

    // cut by q.n <= h (plane id `id`): 0 = no change, 1 = cut (nv may become 0),
    // 2 = capacity exceeded, 3 = inconsistent topology
    SDOT_HD int cut(T nx, T ny, T nz, T h, int id, double nx64, double ny64, double nz64, double h64,
                    double tie_tol = 0, bool may_compact = true) {
        int nout = 0;
        for (int k = 0; k < nv; ++k) {
            T dk = fma(-V_X(k), nx, fma(-V_Y(k), ny, fma(-V_Z(k), nz, h)));
            V_D(k) = dk;
            if (fabs((double)dk) <= tie_tol) {   // near-tie: decide from the planes
                const double dd = tie_distance(k, nx64, ny64, nz64, h64);
                dk = T(dd);
                if (dd < 0 && !(dk < 0)) dk = -T(1e-30);
                if (!(dd < 0) && dk < 0) dk = 0;
                V_D(k) = dk;
            }
            nout += dk < 0;
        }
        if (nout == 0) return 0;
        if (nout == nv) {
            nv = 0;
            return 1;
        }
        if (np >= PCAP && may_compact) compact_planes();
        if (np >= PCAP) return 2;
        const int pn = np;
        // new vertices are appended after the current ones, then the list is compacted
        int nn = nv;
        for (int k = 0; k < nv; ++k) {
            if (!(V_D(k) < 0)) continue;
            const int t = V_T(k);
            const int p[3] = {t & 255, (t >> 8) & 255, (t >> 16) & 255};
            for (int e = 0; e < 3; ++e) {
                const int x = p[e], y = p[e == 2 ? 0 : e + 1];
                const int q = find_edge(y, x);
                if (q < 0) return 3;
                if (V_D(q) < 0) continue;
                if (nn >= VCAP) return 2;
                const T din = V_D(q), dout = V_D(k);
                const T mm = din / (dout - din);
                V_X(nn) = V_X(q) - mm * (V_X(k) - V_X(q));
                V_Y(nn) = V_Y(q) - mm * (V_Y(k) - V_Y(q));
                V_Z(nn) = V_Z(q) - mm * (V_Z(k) - V_Z(q));
                V_T(nn) = x | (y << 8) | (pn << 16);
                V_D(nn) = 0;
                ++nn;
            }
        }
        int w = 0;
        for (int k = 0; k < nn; ++k) {
            if (k < nv && V_D(k) < 0) continue;
            if (w != k) {
                V_X(w) = V_X(k);
                V_Y(w) = V_Y(k);
                V_Z(w) = V_Z(k);
                V_T(w) = V_T(k);
            }
            ++w;
        }
        nv = w;
        add_plane(pn, nx64, ny64, nz64, h64, id);
        np = pn + 1;
        write_edges();
        return 1;
    }

    // FP64 position of vertex k: the intersection of its three planes (mixed
    // path); the clipped position when the planes are degenerate or in FP64 mode
    SDOT_HD void vertex64(int k, double &x, double &y, double &z) const {
        x = V_X(k);
        y = V_Y(k);
        z = V_Z(k);
        if (sizeof(T) == sizeof(double)) return;
        const int t = V_T(k);
        double n0[4], n1[4], n2[4];
        plane64(t & 255, n0[0], n0[1], n0[2], n0[3]);
        plane64((t >> 8) & 255, n1[0], n1[1], n1[2], n1[3]);
        plane64((t >> 16) & 255, n2[0], n2[1], n2[2], n2[3]);
        // Cramer's rule with the cross products of the normals
        const double c12x = n1[1] * n2[2] - n1[2] * n2[1], c12y = n1[2] * n2[0] - n1[0] * n2[2], c12z = n1[0] * n2[1] - n1[1] * n2[0];
        const double c20x = n2[1] * n0[2] - n2[2] * n0[1], c20y = n2[2] * n0[0] - n2[0] * n0[2], c20z = n2[0] * n0[1] - n2[1] * n0[0];
        const double c01x = n0[1] * n1[2] - n0[2] * n1[1], c01y = n0[2] * n1[0] - n0[0] * n1[2], c01z = n0[0] * n1[1] - n0[1] * n1[0];
        const double det = n0[0] * c12x + n0[1] * c12y + n0[2] * c12z;
        const double nrm = sqrt((n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]) * (n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]) *
                                (n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]));
        if (!(fabs(det) > 1e-9 * nrm)) return;
        x = (n0[3] * c12x + n1[3] * c20x + n2[3] * c01x) / det;
        y = (n0[3] * c12y + n1[3] * c20y + n2[3] * c01y) / det;
        z = (n0[3] * c12z + n1[3] * c20z + n2[3] * c01z) / det;
    }

    // face areas of the current polyhedron into m.acc[0..np), returns the volume; with
    // `off`, also the centroid offset from the dirac (signed tetrahedra: dirac, foot of
    // the face plane, directed edge; add_centroid_contrib, ConvexPolyhedron3.tcc:884-924)
    SDOT_HD double integrate(double *off = nullptr, double *r2int = nullptr) {
        double tv = 0, tm[3] = {0, 0, 0}, tr2 = 0;
        for (int k = 0; k < np; ++k) m.acc[k * m.S] = 0;
        if (m.jmom)
            for (int k = 0; k < np; ++k)
                for (int c = 0; c < 9; ++c) m.jmom[(k * 9 + c) * m.S] = 0;
        for (int k = 0; k < nv; ++k) {
            const int t = V_T(k);
            const int p[3] = {t & 255, (t >> 8) & 255, (t >> 16) & 255};
            double x0, y0, z0;
            vertex64(k, x0, y0, z0);
            for (int e = 0; e < 3; ++e) {
                const int f = p[e], g = p[e == 2 ? 0 : e + 1];
                const int q = find_edge(g, f);
                if (q < 0) continue;
                double x1, y1, z1;
                vertex64(q, x1, y1, z1);
                double fx, fy, fz, fh;
                plane64(f, fx, fy, fz, fh);
                const double cx = y0 * z1 - z0 * y1, cy = z0 * x1 - x0 * z1, cz = x0 * y1 - y0 * x1;
                const double cn = cx * fx + cy * fy + cz * fz, n2 = fx * fx + fy * fy + fz * fz;
                m.acc[f * m.S] += cn / sqrt(n2);
                if (m.jmom) {   // triangle (foot of the plane F, P, Q), signed area cn / (2 |n|)
                    const double sf = fh / n2, sa = 0.5 * cn / sqrt(n2);
                    const double F[3] = {sf * fx, sf * fy, sf * fz}, P[3] = {x0, y0, z0}, Q[3] = {x1, y1, z1};
                    const double S3[3] = {F[0] + P[0] + Q[0], F[1] + P[1] + Q[1], F[2] + P[2] + Q[2]};
                    for (int c = 0; c < 3; ++c) m.jmom[(f * 9 + c) * m.S] += sa * S3[c] / 3;
                    int c = 3;
                    for (int u = 0; u < 3; ++u)
                        for (int v = u; v < 3; ++v, ++c)
                            m.jmom[(f * 9 + c) * m.S] += sa / 12 * (F[u] * F[v] + P[u] * P[v] + Q[u] * Q[v] + S3[u] * S3[v]);
                }
                if (r2int) {   // tetrahedron (dirac, F, P, Q): V (|F|^2 + |P|^2 + |Q|^2 + F.P + F.Q + P.Q) / 10
                    const double s = fh / n2, v = s * cn / 6;
                    if (!off) tv += v;
                    const double F[3] = {s * fx, s * fy, s * fz};
                    tr2 += v * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2] + x0 * x0 + y0 * y0 + z0 * z0 + x1 * x1 + y1 * y1 +
                                z1 * z1 + F[0] * x0 + F[1] * y0 + F[2] * z0 + F[0] * x1 + F[1] * y1 + F[2] * z1 + x0 * x1 +
                                y0 * y1 + z0 * z1) / 10;
                }
                if (off) {
                    const double s = fh / n2, v = s * cn / 6;
                    tv += v;
                    tm[0] += v * (s * fx + x0 + x1) / 4;
                    tm[1] += v * (s * fy + y0 + y1) / 4;
                    tm[2] += v * (s * fz + z0 + z1) / 4;
                }
            }
        }
        if (off)
            for (int d = 0; d < 3; ++d) off[d] = tv != 0 ? tm[d] / tv : 0.0;
        if (r2int) *r2int = tv < 0 ? -tr2 : tr2;   // the edges ran clockwise about the faces: every term flips
        double vol = 0;
        for (int k = 0; k < np; ++k) {
            if (m.jmom && m.acc[k * m.S] < 0)   // the face's edges ran clockwise: flip its moments
                for (int c = 0; c < 9; ++c) m.jmom[(k * 9 + c) * m.S] = -m.jmom[(k * 9 + c) * m.S];
            const double ar = 0.5 * fabs(m.acc[k * m.S]);
            m.acc[k * m.S] = ar;
            double fx, fy, fz, fh;
            plane64(k, fx, fy, fz, fh);
            vol += fh / sqrt(fx * fx + fy * fy + fz * fz) * ar / 3;
        }
        return vol;
    }

    SDOT_HD void save() {
        for (int k = 0; k < nv; ++k) {
            m.sx[k * m.S] = V_X(k);
            m.sy[k * m.S] = V_Y(k);
            m.sz[k * m.S] = V_Z(k);
            m.stri[k * m.S] = V_T(k);
        }
    }
    SDOT_HD void restore(int nv0, int np0) {
        nv = nv0;
        np = np0;
        for (int k = 0; k < nv; ++k) {
            V_X(k) = m.sx[k * m.S];
            V_Y(k) = m.sy[k * m.S];
            V_Z(k) = m.sz[k * m.S];
            V_T(k) = m.stri[k * m.S];
        }
        write_edges();
    }
};

template <class T>
struct Eps3;
template <>
struct Eps3<double> {
    static constexpr double region = 1e-9, cand = 1e-9, tie = 1e-10;
};
template <>
struct Eps3<float> {
    static constexpr double region = 2e-5, cand = 2e-5, tie = 1e-4;
};

template <class T>
SDOT_HD inline bool region_skippable3(double d, double R2, double w0, double wmax) {
    double R = sqrt(R2);
    double dw = w0 - wmax;
    double tol = Eps3<T>::region * (d * d + R2 + fabs(w0) + fabs(wmax));
    if (d >= R) return d * d - 2 * R * d + dw >= tol;
    return dw - R2 >= tol;
}

// Builds and integrates cell t.  Returns 0, or 2 when the capacities are exceeded.
template <class T>
SDOT_HD inline int cell3_body(const CellArgs &a, const int64_t t, const Poly3Mem<T> &mem, int VCAP, int PCAP) {
    if (void_seen(a, 0)) return 0;   // the pass is a backtrack signal already
    Cell3<T> C(a, mem, VCAP, PCAP);
    C.c0x = a.spos[3 * t];
    C.c0y = a.spos[3 * t + 1];
    C.c0z = a.spos[3 * t + 2];
    C.w0 = a.sw[t];
    const double c0x = C.c0x, c0y = C.c0y, c0z = C.c0z, w0 = C.w0;
    const GridInfo &g = a.g;
    const double W = dbl_unorder(*a.wglob);
    C.init_box();
    bool overflow = false, bad = false;
    if (a.nitems == 1) {   // one domain item: the starting cell is its box cut by its planes
        const double r0 = sqrt(C.radius2());
        for (int q = a.it_off[0]; q < a.it_off[1] && C.nv > 0 && !overflow; ++q) {
            const double *pl = a.it_planes + 4 * q;
            const double hl = pl[3] - (pl[0] * c0x + pl[1] * c0y + pl[2] * c0z);
            const double nn = sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
            if (C.cut(T(pl[0]), T(pl[1]), T(pl[2]), T(hl), -100 - q, pl[0], pl[1], pl[2], hl, Eps3<T>::tie * nn * r0) >= 2)
                overflow = true;
        }
    }
    double R2 = C.radius2();
    int maxr = g.ncell[0];
    if (g.ncell[1] > maxr) maxr = g.ncell[1];
    if (g.ncell[2] > maxr) maxr = g.ncell[2];
    const double c0[3] = {c0x, c0y, c0z};
    // s = -1: the diracs themselves; s >= 0: the diracs translated by trans[s], found
    // around q = c0 - trans[s] in the grid's frame (SpZGrid.tcc:558-561,610)
    for (int s = -1; s < a.ntrans && C.nv > 0 && !overflow; ++s) {
    const double q[3] = {s < 0 ? c0x : c0x - a.trans[3 * s], s < 0 ? c0y : c0y - a.trans[3 * s + 1],
                         s < 0 ? c0z : c0z - a.trans[3 * s + 2]};
    const int idbase = s < 0 ? 0 : (s + 1) * (int)a.N;
    const int h0 = grid_coord(g, 0, q[0]), h1 = grid_coord(g, 1, q[1]), h2 = grid_coord(g, 2, q[2]);
    const int hh[3] = {h0, h1, h2};
    for (int r = 0; r <= maxr && C.nv > 0 && !overflow; ++r) {
        if (r > 0) {
            double dm = DBL_MAX;
            bool any = false;
            for (int d = 0; d < 3; ++d) {
                if (hh[d] - r + 1 > 0) { dm = fmin(dm, q[d] - (g.lo[d] + (hh[d] - r + 1) * g.h[d])); any = true; }
                if (hh[d] + r < g.ncell[d]) { dm = fmin(dm, (g.lo[d] + (hh[d] + r) * g.h[d]) - q[d]); any = true; }
            }
            if (!any) break;
            if (region_skippable3<T>(fmax(dm, 0.0), R2, w0, W)) break;
        }
        const int z0 = h2 - r < 0 ? 0 : h2 - r, z1 = h2 + r >= g.ncell[2] ? g.ncell[2] - 1 : h2 + r;
        const int y0 = h1 - r < 0 ? 0 : h1 - r, y1 = h1 + r >= g.ncell[1] ? g.ncell[1] - 1 : h1 + r;
        for (int gz = z0; gz <= z1 && C.nv > 0 && !overflow; ++gz) {
            for (int gy = y0; gy <= y1 && C.nv > 0 && !overflow; ++gy) {
                const bool full = gz == h2 - r || gz == h2 + r || gy == h1 - r || gy == h1 + r;
                const int step = full || r == 0 ? 1 : 2 * r;
                for (int gx = h0 - r; gx <= h0 + r && C.nv > 0 && !overflow; gx += step) {
                    if (gx < 0 || gx >= g.ncell[0]) continue;
                    const uint32_t code = morton3(gx, gy, gz);
                    const uint32_t cs = a.cell_start[code], ce = a.cell_start[code + 1];
                    if (cs == ce) continue;
                    {
                        const int gg[3] = {gx, gy, gz};
                        double e2 = 0;
                        for (int d = 0; d < 3; ++d) {
                            double b0 = g.lo[d] + gg[d] * g.h[d];
                            double e = fmax(0.0, fmax(b0 - q[d], q[d] - (b0 + g.h[d])));
                            e2 += e * e;
                        }
                        if (region_skippable3<T>(sqrt(e2), R2, w0, a.cell_wmax[code])) continue;
                    }
                    for (uint32_t j = cs; j < ce; ++j) {
                        if ((int64_t)j == t && s < 0) continue;
                        const double vx = a.spos[3 * (int64_t)j] - q[0], vy = a.spos[3 * (int64_t)j + 1] - q[1],
                                     vz = a.spos[3 * (int64_t)j + 2] - q[2];
                        const double vv = vx * vx + vy * vy + vz * vz;
                        const double hv = 0.5 * (vv + w0 - a.sw[j]);
                        if (hv > 0 && R2 * vv <= hv * hv * (1 - Eps3<T>::cand)) continue;
                        if (vv == 0) { bad = true; continue; }
                        const int rc = C.cut(T(vx), T(vy), T(vz), T(hv), (int)j + idbase, vx, vy, vz, hv,
                                             Eps3<T>::tie * sqrt(vv * R2));
                        if (rc >= 2) { overflow = true; break; }
                        if (rc == 1) {
                            if (C.nv == 0) break;
                            R2 = C.radius2();
                        }
                    }
                }
            }
        }
    }
    }
    if (bad) flag_or(a.flags, FLAG_BAD_INPUT);
    if (overflow) return 2;
    if (C.nv > 0) C.compact_planes();   // voxel cuts then start from the face count

    // ---- integrals ----
    double mass = 0, perim = 0, mom[3] = {0, 0, 0}, off[3];
    double *const want = a.mom ? off : nullptr;
    for (int k = 0; k < PCAP; ++k) C.m.hacc[k * C.m.S] = 0;
    const int npc = C.np;
    const bool pieces = a.img || a.nitems >= 2;
    // pieces: the cell planes whose face some piece covers in part; only those faces are
    // boundary items (a face outside every shape or voxel gives no entry)
    unsigned long long hit[4] = {0, 0, 0, 0};   // PCAP <= 256
    if (C.m.jsum && pieces)
        for (int k = 0; k < npc; ++k)
            for (int q = 0; q < 10; ++q) C.m.jsum[(k * 10 + q) * C.m.S] = 0;
    // one piece of density rho (the cell clipped by a domain item or a voxel): its mass (or
    // its R2 integral), first moments, the rho-weighted areas of the cell's faces and, for
    // the position derivatives, those faces' areas and moments (not scaled by rho)
    auto piece_integrals = [&](double rho) {
        if (a.rf == 1) {
            double r2 = 0;
            C.integrate(nullptr, &r2);
            mass += rho * r2;
            return;
        }
        const double vol = C.integrate(want);
        mass += rho * vol;
        if (want)
            for (int d = 0; d < 3; ++d) mom[d] += rho * vol * off[d];
        for (int k = 0; k < npc; ++k) {
            C.m.hacc[k * C.m.S] += rho * C.m.acc[k * C.m.S];
            if (C.m.acc[k * C.m.S] > 0) hit[k >> 6] |= 1ull << (k & 63);
        }
        if (C.m.jsum)
            for (int k = 0; k < npc; ++k) {
                C.m.jsum[(k * 10) * C.m.S] += C.m.acc[k * C.m.S];
                for (int q = 0; q < 9; ++q) C.m.jsum[(k * 10 + 1 + q) * C.m.S] += C.m.jmom[(k * 9 + q) * C.m.S];
            }
    };
    if (C.nv > 0) {
        if (!a.img && a.nitems < 2 && a.rf == 1) {
            // R2 (integration(Constant, R2), ConvexPolyhedron3.tcc:1120-1123): integrals only
            double r2 = 0;
            C.integrate(nullptr, &r2);
            mass = a.coeff * r2;
        } else if (!a.img && a.nitems < 2) {
            const double vol = C.integrate(want);
            mass = a.coeff * vol;
            if (want)
                for (int d = 0; d < 3; ++d) mom[d] = mass * off[d];
            for (int k = 0; k < npc; ++k) C.m.hacc[k * C.m.S] = a.coeff * C.m.acc[k * C.m.S];
        } else if (!a.img) {
            // several domain items (ConvexPolyhedronAssembly.tcc:132-142): the cell cut by
            // each item's box sides and planes, integrated with its density
            double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (int k = 0; k < C.nv; ++k) {
                double x, y, z;
                C.vertex64(k, x, y, z);
                mn[0] = fmin(mn[0], x); mx[0] = fmax(mx[0], x);
                mn[1] = fmin(mn[1], y); mx[1] = fmax(mx[1], y);
                mn[2] = fmin(mn[2], z); mx[2] = fmax(mx[2], z);
            }
            const int nv0 = C.nv, np0 = C.np;
            C.save();
            for (int it = 0; it < a.nitems && !overflow; ++it) {
                const double *b = a.it_box + 6 * it;
                bool empty = false;
                for (int d = 0; d < 3; ++d) {
                    C.vlo[d] = b[d] - c0[d];
                    C.vhi[d] = b[3 + d] - c0[d];
                    if (C.vlo[d] > mx[d] || C.vhi[d] < mn[d]) empty = true;   // the item's box misses the cell
                }
                for (int sd = 0; sd < 6 && !empty; ++sd) {
                    const int axis = sd >> 1, up = sd & 1;
                    const double bound = up ? sel3(C.vhi, axis) : sel3(C.vlo, axis);
                    if (up ? !(bound < sel3(mx, axis)) : !(bound > sel3(mn, axis))) continue;   // cannot cut
                    const double nxx = axis == 0 ? (up ? 1.0 : -1.0) : 0.0, nyy = axis == 1 ? (up ? 1.0 : -1.0) : 0.0,
                                 nzz = axis == 2 ? (up ? 1.0 : -1.0) : 0.0, hq = up ? bound : -bound;
                    const int rc = C.cut(T(nxx), T(nyy), T(nzz), T(hq), -7 - sd, nxx, nyy, nzz, hq,
                                         Eps3<T>::tie * sqrt(R2), false);
                    if (rc >= 2) { overflow = true; break; }
                    if (C.nv == 0) empty = true;
                }
                for (int q = a.it_off[it]; q < a.it_off[it + 1] && !empty && !overflow; ++q) {
                    const double *pl = a.it_planes + 4 * q;
                    const double hl = pl[3] - (pl[0] * c0[0] + pl[1] * c0[1] + pl[2] * c0[2]);
                    const double nn = sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
                    const int rc = C.cut(T(pl[0]), T(pl[1]), T(pl[2]), T(hl), -100 - q, pl[0], pl[1], pl[2], hl,
                                         Eps3<T>::tie * nn * sqrt(R2), false);
                    if (rc >= 2) { overflow = true; break; }
                    if (C.nv == 0) empty = true;
                }
                if (overflow) break;
                if (!empty) piece_integrals(a.it_coeff[it]);
                C.restore(nv0, np0);
            }
            if (overflow) return 2;
        } else {
            // voxel pieces: the cell clipped by each overlapping voxel (ScaledImage.tcc:98-149)
            double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (int k = 0; k < C.nv; ++k) {
                double x, y, z;
                C.vertex64(k, x, y, z);
                mn[0] = fmin(mn[0], x); mx[0] = fmax(mx[0], x);
                mn[1] = fmin(mn[1], y); mx[1] = fmax(mx[1], y);
                mn[2] = fmin(mn[2], z); mx[2] = fmax(mx[2], z);
            }
            int i0[3], i1[3];
            double vs[3];
            for (int d = 0; d < 3; ++d) {
                vs[d] = (a.dom_hi[d] - a.dom_lo[d]) / a.img_n[d];
                const double lo = mn[d] + c0[d] - a.dom_lo[d], hi = mx[d] + c0[d] - a.dom_lo[d];
                int u0 = (int)floor(lo / vs[d]), u1 = (int)floor(hi / vs[d]);
                i0[d] = u0 < 0 ? 0 : (u0 >= a.img_n[d] ? a.img_n[d] - 1 : u0);
                i1[d] = u1 < 0 ? 0 : (u1 >= a.img_n[d] ? a.img_n[d] - 1 : u1);
            }
            const int nv0 = C.nv, np0 = C.np;
            C.save();
            for (int iz = i0[2]; iz <= i1[2]; ++iz)
                for (int iy = i0[1]; iy <= i1[1]; ++iy)
                    for (int ix = i0[0]; ix <= i1[0]; ++ix) {
                        const int iv[3] = {ix, iy, iz};
                        const double rho = a.img[((int64_t)iz * a.img_n[1] + iy) * a.img_n[0] + ix];
                        for (int d = 0; d < 3; ++d) {
                            C.vlo[d] = a.dom_lo[d] + vs[d] * iv[d] - c0[d];
                            C.vhi[d] = a.dom_lo[d] + vs[d] * (iv[d] + 1) - c0[d];
                        }
                        bool empty = false;
                        for (int s = 0; s < 6 && !empty; ++s) {
                            const int axis = s >> 1, up = s & 1;
                            if (up ? (sel3(iv, axis) == sel3(i1, axis)) : (sel3(iv, axis) == sel3(i0, axis))) continue;
                            const double off = up ? sel3(C.vhi, axis) : sel3(C.vlo, axis);
                            const double nxx = axis == 0 ? (up ? 1.0 : -1.0) : 0.0, nyy = axis == 1 ? (up ? 1.0 : -1.0) : 0.0,
                                         nzz = axis == 2 ? (up ? 1.0 : -1.0) : 0.0, hq = up ? off : -off;
                            const int rc = C.cut(T(nxx), T(nyy), T(nzz), T(hq), -7 - s, nxx, nyy, nzz, hq,
                                                 Eps3<T>::tie * sqrt(R2), false);
                            if (rc >= 2) { overflow = true; break; }
                            if (C.nv == 0) empty = true;
                        }
                        if (overflow) break;
                        if (!empty) piece_integrals(rho);
                        C.restore(nv0, np0);
                    }
            if (overflow) return 2;
        }
    }
    // Hessian row and boundary measure
    double der = 0;
    int nrow = 0, nfaces = 0;
    double jd[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool row_ovf = false;
    for (int k = 0; k < npc && a.rf != 1; ++k) {   // R2: no boundary items
        const double ar = C.m.hacc[k * C.m.S];
        const int id = C.m.pid[k * C.m.S];
        if (pieces && !((hit[k >> 6] >> (k & 63)) & 1)) continue;
        if (ar == 0 && C.nv > 0) {
            // a plane that bounds no face (removed by later cuts) is skipped
            bool used = false;
            for (int q = 0; q < C.nv && !used; ++q) {
                const int tt = C.m.tri[q * C.m.S];
                used = (tt & 255) == k || ((tt >> 8) & 255) == k || ((tt >> 16) & 255) == k;
            }
            if (!used) continue;
        }
        if (C.nv == 0) continue;
        ++nfaces;
        // boundary measure: all faces (box domain) / the faces shared with neighbours (image density)
        if ((!a.img && a.nitems < 2) || is_neighbour(id)) perim += ar;
        if (!is_neighbour(id)) continue;
        double qn[3];
        const int64_t jn = cut_dirac<3>(a, id, qn);
        if (C.m.jmom && C.nv > 0) {
            // get_der_centroids_and_integrals_wrt_weight_and_positions.h:124-166 with the
            // face's moments in absolute coordinates (ConvexPolyhedron3.tcc:226-265: not
            // scaled by the density; measure = the rho-weighted area)
            const double c0a[3] = {c0x, c0y, c0z};
            // one shape: the face as integrated; pieces: the sums over the pieces
            auto jm = [&](int q) { return pieces ? C.m.jsum[(k * 10 + 1 + q) * C.m.S] : C.m.jmom[(k * 9 + q) * C.m.S]; };
            const double ag = pieces ? C.m.jsum[(k * 10) * C.m.S] : C.m.acc[k * C.m.S];
            double Cl[3], Ml[3][3], Ca[3], Ma[3][3];
            for (int u = 0; u < 3; ++u) Cl[u] = jm(u);
            {
                int c = 3;
                for (int u = 0; u < 3; ++u)
                    for (int v = u; v < 3; ++v, ++c) Ml[u][v] = Ml[v][u] = jm(c);
            }
            for (int u = 0; u < 3; ++u) Ca[u] = ag * c0a[u] + Cl[u];
            for (int u = 0; u < 3; ++u)
                for (int v = 0; v < 3; ++v) Ma[u][v] = ag * c0a[u] * c0a[v] + c0a[u] * Cl[v] + Cl[u] * c0a[v] + Ml[u][v];
            const double dist = sqrt((qn[0] - c0x) * (qn[0] - c0x) + (qn[1] - c0y) * (qn[1] - c0y) + (qn[2] - c0z) * (qn[2] - c0z));
            if (dist != 0) {
                const double bw = 0.5 * ar / dist;
                double d1[16];
                d1[15] = -bw;
                jd[15] += bw;
                for (int e = 0; e < 3; ++e) {
                    d1[4 * e + 3] = -0.5 * Ca[e] / dist;
                    jd[4 * e + 3] += 0.5 * Ca[e] / dist;
                }
                for (int d = 0; d < 3; ++d) {
                    d1[12 + d] = (ar * qn[d] - Ca[d]) / dist;
                    jd[12 + d] += (Ca[d] - ar * c0a[d]) / dist;
                }
                for (int d = 0; d < 3; ++d)
                    for (int e = 0; e < 3; ++e) {
                        d1[4 * e + d] = (Ca[e] * qn[d] - Ma[d][e]) / dist;
                        jd[4 * e + d] += (Ma[d][e] - Ca[e] * c0a[d]) / dist;
                    }
                if (jn == t) {   // the cell's own image: its block merges into the diagonal
                    for (int q = 0; q < 16; ++q) jd[q] += d1[q];
                } else if (nrow < a.ell_cap) {
                    for (int q = 0; q < 16; ++q) a.jblk[((int64_t)nrow * 16 + q) * a.ell_stride + t] = d1[q];
                }
            }
        }
        if (jn == t) continue;   // a face with the cell's own image adds nothing
        double nx, ny, nz, hv;
        C.plane64(k, nx, ny, nz, hv);
        const double dist = sqrt(nx * nx + ny * ny + nz * nz);
        const double b = 0.5 * ar / dist;
        if (nrow < a.ell_cap) {
            a.ell_col[nrow * a.ell_stride + t] = (int)jn;
            a.ell_val[nrow * a.ell_stride + t] = -b;
            ++nrow;
        } else {
            row_ovf = true;
        }
        der += b;
    }
    a.mass[t] = mass;
    if (a.mom)
        for (int d = 0; d < 3; ++d) a.mom[3 * t + d] = mom[d];
    if (C.m.jmom)
        for (int q = 0; q < 16; ++q) a.jdiag[(int64_t)q * a.ell_stride + t] = jd[q];
    a.arc[t] = 0;
    a.perim[t] = perim;
    a.diag[t] = der;
    a.row_len[t] = nrow;
    if (a.ctr) {
        const unsigned long long K = nfaces;
        const int slot = (int)(t % WCTR);
        ctr_add(a.ctr + slot, K >= 24 ? 12 * K * K + 100 * K - 284 : 12 * K * K + 100 * K > 284 ? 12 * K * K + 100 * K - 284 : 0);
        ctr_add(a.ctr + WCTR + slot, K);
    }
    if (row_ovf) flag_or(a.flags, FLAG_ROW_OVERFLOW);
    if (mass == 0) flag_or(a.flags, FLAG_EMPTY);
    return 0;
}

#undef V_X
#undef V_Y
#undef V_Z
#undef V_T
#undef V_D
#undef P_X
#undef P_Y
#undef P_Z
#undef P_H
#undef P_I

}  // namespace c3
}  // namespace sdot_amd
