/*
 * sdot_oracle.h -- CPU restatement of the sdot Laguerre-cell path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library
 * (sdot_amd/libsdot_mi355x.so) never links, loads or calls it.
 *
 * Every function follows a reference file (paths relative to the sdot tree):
 *   2D polygon        src/sdot/Geometry/ConvexPolyhedron2.tcc
 *   3D polyhedron     src/sdot/Geometry/ConvexPolyhedron3.tcc  (semantics; the
 *                     pointer half-edge graph is replaced by per-face loops)
 *   neighbour search  src/sdot/PowerDiagram/Visitors/SpZGrid.tcc
 *   integrals / CSR   src/sdot/PowerDiagram/get_integrals.h,
 *                     src/sdot/PowerDiagram/get_der_integrals_wrt_weights.h
 *   domains           src/sdot/Domains/ConvexPolyhedronAssembly.tcc,
 *                     src/sdot/Domains/ScaledImage.tcc
 *   Newton            src/sdot/Solvers/OptimalTransportSolver.tcc
 *
 * Parity pins (tests/test_oracle_*.py):
 *   or_cp2_*     reference known values (tests/test_ConvexPolyhedron2.cpp),
 *                closed forms (disc segments, lattices), brute-force N.
 *   or_cp3_*     box volumes, cube lattice closed forms, Euler/volume identities,
 *                agreement of the face-area sum with the divergence theorem.
 *   or_cells     brute-force candidate mode vs grid mode, sum of masses = domain
 *                mass, Hessian symmetry, zero row sums (no ball), central finite
 *                differences of the masses, isolated-dirac ball closed forms.
 *   or_ot_solve  converged masses within tolerance (the OT dual is unique up to
 *                an additive constant; compared modulo the mean).
 *   Eigen's SimplicialCholesky (the reference's linear solver, not vendored,
 *   version unpinned) is replaced by Jacobi-PCG: the Newton trajectory is
 *   "parity unpinned", the fixed point is pinned by the mass equations.
 */
#ifndef SDOT_ORACLE_H
#define SDOT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_EMPTY_CELL = 1,
    OR_NOT_CONVERGED = 2,
    OR_CAPACITY = 3,
    OR_INVALID = 4
};

#define OR_CP2_CAP 256

/* ConvexPolyhedron2 restated: CCW vertices; edge k runs from vertex k to
 * vertex k+1 and carries id[k], normal (nx,ny)[k] and arc[k]
 * (ConvexPolyhedron2.h:168-181). */
typedef struct or_cp2 {
    int n;
    double x[OR_CP2_CAP], y[OR_CP2_CAP];
    double nx[OR_CP2_CAP], ny[OR_CP2_CAP];
    int64_t id[OR_CP2_CAP];
    unsigned char arc[OR_CP2_CAP];
    double scx, scy, sr;  /* sphere centre / radius (-1 = no ball) */
    int64_t sid;
} or_cp2;

void    or_cp2_box(or_cp2 *cp, double x0, double y0, double x1, double y1, int64_t id);
int     or_cp2_plane_cut(or_cp2 *cp, double ox, double oy, double nx, double ny, int64_t id, int normal_is_normalized);
int     or_cp2_ball_cut(or_cp2 *cp, double cx, double cy, double r, int64_t id);
double  or_cp2_area(const or_cp2 *cp, double coeff);
/* area and first moments relative to (ox, oy), which must be the ball centre when there are arcs */
void    or_cp2_moments(const or_cp2 *cp, double coeff, double ox, double oy, double *m0, double *mx, double *my);
double  or_cp2_poly_area(const or_cp2 *cp, const double *c6);   /* integral of the quadratic over a straight-edged polygon */
double  or_seg_poly(double x0, double y0, double x1, double y1, const double *c6);   /* integral along a segment */
int     or_cp2_boundary_items(const or_cp2 *cp, double coeff, double *measures, int64_t *ids, int cap);
int     or_cp2_empty(const or_cp2 *cp);
/* convenience for ctypes tests: box + k plane cuts (o,n as 4 doubles each) + optional ball */
int     or_cp2_eval(double x0, double y0, double x1, double y1, int ncuts, const double *cuts,
                    int ball, double bcx, double bcy, double br,
                    double *area, int *nb_points, double *measures, int64_t *ids, int cap, int *nitems);

/* same polygon as or_cp2_eval: its centroid and area */
int     or_cp2_centroid(double x0, double y0, double x1, double y1, int ncuts, const double *cuts,
                        int ball, double bcx, double bcy, double br, double *centroid, double *area);

/* 3D polyhedron as a list of faces, each a CCW (seen from outside) vertex loop. */
#define OR_CP3_MAXF 96
#define OR_CP3_MAXFV 48
typedef struct or_cp3 {
    int nf;
    int nv[OR_CP3_MAXF];
    double v[OR_CP3_MAXF][OR_CP3_MAXFV][3];
    double n[OR_CP3_MAXF][3];
    int64_t id[OR_CP3_MAXF];
    int overflow;
} or_cp3;

void    or_cp3_box(or_cp3 *cp, const double *p0, const double *p1, int64_t id);
int     or_cp3_plane_cut(or_cp3 *cp, const double *origin, const double *normal, int64_t id);
double  or_cp3_face_area(const or_cp3 *cp, int f);
double  or_cp3_volume(const or_cp3 *cp, const double *center);
/* volume and first moments relative to c */
void    or_cp3_moments(const or_cp3 *cp, const double *c, double coeff, double *m0, double *m);
int     or_cp3_eval(const double *p0, const double *p1, int ncuts, const double *cuts,
                    double *volume, int *nb_faces, double *areas, int64_t *ids, int cap);

/* Domain: a single box with a constant density coefficient (the
 * ConvexPolyhedronAssembly with one add_box item), or a ScaledImage of
 * piecewise-constant pixels/voxels (nb_coeffs = 1, x fastest). */
typedef struct or_domain {
    int dim;
    double p0[3], p1[3];
    double coeff;
    const double *img;       /* NULL: constant box */
    int64_t img_sizes[3];
    int64_t img_nc;          /* coefficients per pixel: 1 (constant) or, in 2D, 6:
                              * c0 + c1 x + c2 y + c3 x^2 + c4 x y + c5 y^2 in domain
                              * coordinates, coefficient k of pixel p at img[k * npix + p] */
    /* ConvexPolyhedronAssembly items (nitems = 0: the box p0, p1 with coeff above).
     * Item k: the box it_box[6k..6k+2] .. it_box[6k+3..6k+5] cut by the planes
     * n . x <= h stored as it_planes[4q..4q+3] = (nx, ny, nz, h) for q in
     * [it_off[k], it_off[k+1]), density it_coeff[k]. */
    int64_t nitems;
    const double *it_box, *it_coeff, *it_planes;
    const int64_t *it_off;
    /* SpZGrid::translations (SpZGrid.h:47-53): every cell is also cut by the diracs
     * translated by trans[dim s .. dim s + dim) for s < ntrans, the image of dirac j
     * under translation s carrying the cut id j + (s + 1) N (SpZGrid.tcc:558-561,610) */
    int64_t ntrans;
    const double *trans;
    /* eps of FunctionEnum::ExpWmR2db (ExpWmR2db.h: exp((w - |x - c|^2) / eps)), the final
     * integration function when the cells flags carry radial function 3 */
    double rf_eps;
} or_domain;

/* Per-cell evaluation (get_der_integrals_wrt_weights.h:16-122 restated).
 *  masses[N]            integral of the density over cell i
 *  ball_arc[N]          measure of the ball part of the cell boundary (may be NULL)
 *  perimeter[N]         measure of the whole cell boundary (may be NULL); with an
 *                       image density, the density-weighted measure of the faces
 *                       shared with other cells (and of the ball part)
 *  row_len[N], cols[N*cap], vals[N*cap]:  unsorted raw row of cell i:
 *        slot 0 = diagonal (col i), then (j, -0.5*rho*|G_ij|/|x_i-x_j|) per face
 *  flags: bit0 ball clip (radius sqrt(w_i)); bit1 brute-force candidates (all j)
 * Returns OR_OK, OR_EMPTY_CELL (some cell empty or zero mass), OR_CAPACITY. */
int or_cells(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
             int flags, int nthreads,
             double *masses, double *ball_arc, double *perimeter,
             int64_t cap, int64_t *row_len, int64_t *cols, double *vals);

/* get_centroids (get_centroids.h:14-60, rand_ratio = 0): per cell the centroid of
 * the density over the cell (0 for an empty cell) and its mass.  Constant density
 * or constant pixels/voxels (quadratic pixels: OR_INVALID). */
int or_centroids(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom, int flags,
                 int nthreads, double *centroids, double *masses);

/* Same, evaluating only the cells [i_begin, i_end) (the grid is built over all
 * N diracs): a bounded sample of a large pass for timing. */
int or_cells_range(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom,
                   int flags, int nthreads, int64_t i_begin, int64_t i_end,
                   double *masses, double *ball_arc, double *perimeter,
                   int64_t cap, int64_t *row_len, int64_t *cols, double *vals);

/* Assemble sorted/merged CSR in dirac order from the raw rows
 * (get_der_integrals_wrt_weights.h:76-88,106-119).  offsets[N+1]; cols/vals
 * need room for sum(row_len).  Returns nnz. */
int64_t or_rows_to_csr(int64_t N, int64_t cap, const int64_t *row_len, const int64_t *rcols, const double *rvals,
                       int64_t *offsets, int64_t *cols, double *vals);

/* Jacobi-PCG on a CSR matrix (replaces EigenSolver.cpp's SimplicialCholesky). */
int or_pcg(int64_t N, const int64_t *offsets, const int64_t *cols, const double *vals,
           const double *b, double *x, double rtol, int max_iter, int *iters);

/* get_der_centroids_and_integrals_wrt_weight_and_positions (one domain shape: a box or
 * one item; translations; the ball (flags & 1) in 2D only, TODO in the reference in
 * 3D).  Per dirac i: jlen[i] items, columns jcols[i*cap ..] ascending (diagonal
 * included, duplicates merged), blocks jblocks[(i*cap + k)*16 + nu*d + e], nu = dim + 1,
 * for output d (centroid coordinates, mass at d = dim) and unknown e (position
 * coordinates, weight at e = dim); jv[nu*i ..] = centroid, mass. */
int or_der_pos(int dim, int64_t N, const double *pos, const double *w, const or_domain *dom, int flags, int nthreads,
               int64_t cap, int64_t *jlen, int64_t *jcols, double *jblocks, double *jv);

typedef struct or_solve_stats {
    int newton_iters;
    int pcg_iters_total;
    int backtracks;
    double max_rel_mass_err;
    double max_abs_dw;
    int64_t nnz;
} or_solve_stats;

/* OptimalTransportSolver::solve restated (OptimalTransportSolver.tcc:27-119).
 * stop_rule 0: max_i |m_i - t_i| / t_i < tol (the workload rule);
 * stop_rule 1: max |dw| < tol (the reference rule). */
int or_ot_solve(int dim, int64_t N, const double *pos, double *w, const double *targets,
                const or_domain *dom, int flags, int nthreads, double tol, int stop_rule,
                int max_iter, or_solve_stats *stats);

/* splitmix64 -> xoshiro256** uniform doubles u = (x >> 11) * 2^-53 */
void or_rng_uniform(uint64_t seed, int64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
