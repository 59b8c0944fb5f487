/*
 * sdot_mi355x.h -- C-ABI of the MI355X (gfx950) Laguerre-cell path.
 *
 * A flat, handle-based API in front of hand-written HIP kernels.  It replaces,
 * for the Laguerre-cell hot path, the reference's three template APIs
 * (paths relative to the sdot tree):
 *   Visitor  SpZGrid::update / for_each_laguerre_cell   src/sdot/PowerDiagram/Visitors/SpZGrid.h:39-48
 *   Bounds   ConvexPolyhedronAssembly::add_box ...      src/sdot/Domains/ConvexPolyhedronAssembly.h:26-41
 *            ScaledImage(min, max, data, sizes, nb_coeffs) src/sdot/Domains/ScaledImage.h:26
 *   Functions get_integrals                             src/sdot/PowerDiagram/get_integrals.h:13
 *            get_der_integrals_wrt_weights              src/sdot/PowerDiagram/get_der_integrals_wrt_weights.h:16
 *            OptimalTransportSolver::solve              src/sdot/Solvers/OptimalTransportSolver.h:18-26
 *
 * Conventions (same as the reference):
 *   positions are interleaved (x0,y0[,z0],x1,...) doubles, Point2/Point3 layout
 *   (src/sdot/Geometry/Point2.h:14); outputs are in dirac order; CSR rows are
 *   sorted by column, duplicates merged, diagonal present; the domain cut id is
 *   -1.  Index types are int64_t (the reference's TI = std::size_t).
 * Errors are status codes; no entry point aborts (the reference's ASSERT/TODO
 * call abort(), src/sdot/Support/Assert.h:13-17).
 * Threading: a handle owns one HIP stream (default priority) and is not
 * internally locked; use one handle per host thread.
 */
#ifndef SDOT_MI355X_H
#define SDOT_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sdot_status {
    SDOT_OK = 0,
    SDOT_EMPTY_CELL = 1,         /* an empty Laguerre cell (reference returns 1) */
    SDOT_NOT_CONVERGED = 2,
    SDOT_CAPACITY_OVERFLOW = 3,
    SDOT_INVALID_ARG = 4,
    SDOT_HIP_ERROR = 5,
    SDOT_OOM = 6,
    SDOT_NOT_SUPPORTED = 7,
    SDOT_PENDING = 8             /* serving loop: the request has not finished yet */
} sdot_status;

typedef enum sdot_precision {
    SDOT_FP64 = 0,               /* clipping, integrals and solve in FP64 */
    SDOT_MIXED = 1               /* FP32 cell-local clipping (which planes bound each cell), FP64
                                    plane offsets; the vertices are then rebuilt in FP64 from their
                                    planes and every integral, Hessian entry, residual and the
                                    Newton/CG solve run in FP64 */
} sdot_precision;

typedef struct sdot_pd sdot_pd;

/* Handle life cycle.  dim is 2 or 3. */
sdot_status sdot_pd_create(sdot_pd **pd, int dim, int precision);
sdot_status sdot_pd_destroy(sdot_pd *pd);
const char *sdot_status_string(int status);
/* last HIP / argument error text of this handle ("" if none) */
const char *sdot_pd_last_error(const sdot_pd *pd);

/* Diracs (host arrays, copied to HBM).  SpZGrid::update(pos, w, N) -- SpZGrid.h:39 */
sdot_status sdot_pd_set_diracs(sdot_pd *pd, const double *positions, const double *weights, int64_t nb_diracs);
/* weights only (positions unchanged: the grid's sort is kept) */
sdot_status sdot_pd_set_weights(sdot_pd *pd, const double *weights);
/* Same with device pointers already in HBM (copied device-to-device on the handle's stream). */
sdot_status sdot_pd_set_diracs_device(sdot_pd *pd, const double *d_positions, const double *d_weights, int64_t nb_diracs);

/* Device-resident inputs: the handle's HBM arrays of positions (N*dim) and
 * weights (N), for producers that write them in place; after writing, call
 * sdot_pd_mark_diracs_changed so the next pass rebuilds the grid. */
sdot_status sdot_pd_dirac_buffers(sdot_pd *pd, double **d_positions, double **d_weights);
sdot_status sdot_pd_mark_diracs_changed(sdot_pd *pd, int positions_changed);

/* Domain ("bounds").  ConvexPolyhedronAssembly::add_box(p0, p1, coeff, cut_id) -- .tcc:46.
 * Shapes accumulate (see "Several shapes" below); cut_id = -1. */
sdot_status sdot_pd_add_box_shape(sdot_pd *pd, const double *p0, const double *p1, double coeff, int64_t cut_id);
/* add_convex_polyhedron(positions, normals, coeff, cut_id) -- ConvexPolyhedronAssembly.tcc:7:
 * the region (x - positions[i]) . normals[i] <= 0 for i < n (n*dim values each), inside
 * the box of the positions widened by its extent on each side.  cut_id = -1. */
sdot_status sdot_pd_add_convex_polyhedron_shape(sdot_pd *pd, const double *positions, const double *normals, int64_t n,
                                                double coeff, int64_t cut_id);
/* add_simplex(points, coeff, cut_id) -- ConvexPolyhedronAssembly.tcc:36: (dim+1)*dim points. */
sdot_status sdot_pd_add_simplex_shape(sdot_pd *pd, const double *points, double coeff, int64_t cut_id);
/* Several shapes: each cell is integrated piece by piece over the shapes
 * (for_each_intersection, ConvexPolyhedronAssembly.tcc:132-142), densities adding up
 * where shapes overlap; the perimeter then counts the faces shared with other cells
 * (and the ball arcs) only.  With the ball clip each piece is also clipped by the
 * cell's disc (intersect_with, ConvexPolyhedron2.tcc:2474-2496); position derivatives
 * of a ball-clipped cell over several pieces return SDOT_NOT_SUPPORTED. */
/* ScaledImage(min_pt, max_pt, data, sizes, nb_coeffs) -- ScaledImage.h:26; sizes x fastest.
 * nb_coeffs = 1: constant per pixel (ball clip allowed: pieces as for several shapes).  nb_coeffs = 6 (2D only, no ball): the density in
 * pixel p is c0 + c1 x + c2 y + c3 x^2 + c4 x y + c5 y^2 in domain coordinates, with
 * c_k = data[k * nb_pixels + p] (apply_poly, Polynomial.h:37-50). */
sdot_status sdot_pd_set_image_density(sdot_pd *pd, const double *min_pt, const double *max_pt, const double *data,
                                      const int64_t *sizes, int64_t nb_coeffs);
sdot_status sdot_pd_clear_shapes(sdot_pd *pd);
/* Radial function: 0 = FunctionEnum::Unit, 1 = FunctionEnum::InBallW05 (clip by the
 * ball of radius sqrt(w_i)) -- src/sdot/Integration/Unit.h, InBallW05.h */
sdot_status sdot_pd_set_ball_clip(sdot_pd *pd, int on);
/* Radial function (the Func argument of get_integrals / get_der_integrals_wrt_weights /
 * OptimalTransportSolver::solve).  The last four in 2D: one domain shape; R2 / WmR2 also
 * over constant pixels or several shapes.  R2 also in 3D, over one shape, several shapes or
 * voxels (ConvexPolyhedron3.tcc:1120-1123); WmR2 2D only (TODO in 3D there):
 *   SDOT_FUNC_UNIT, SDOT_FUNC_IN_BALL_W05 -- Unit.h, InBallW05.h (as sdot_pd_set_ball_clip);
 *   SDOT_FUNC_R2, SDOT_FUNC_R2_IN_BALL_W05 -- |x - x_i|^2 (R2.h, R2InBallW05.h): integrals only
 *     (the reference has no R2 boundary items: ConvexPolyhedron2.tcc:432-435);
 *   SDOT_FUNC_WMR2, SDOT_FUNC_PP_WMR2 -- w_i - |x - x_i|^2 (WmR2.h, PpWmR2.h: its positive part,
 *     i.e. with the ball cut): integrals, Hessian (ConvexPolyhedron2.tcc:391-430, 2044-2088;
 *     d/dw_i adds the cell's measure, :135-138) and the Newton solve.
 * Centroids and position derivatives take Unit / InBallW05 only. */
enum { SDOT_FUNC_UNIT = 0, SDOT_FUNC_IN_BALL_W05 = 1, SDOT_FUNC_R2 = 2, SDOT_FUNC_R2_IN_BALL_W05 = 3,
       SDOT_FUNC_WMR2 = 4, SDOT_FUNC_PP_WMR2 = 5, SDOT_FUNC_EXP_WMR2_DB = 6 };
sdot_status sdot_pd_set_radial_func(sdot_pd *pd, int func);
/* FunctionEnum::ExpWmR2db<TF>{eps} (ExpWmR2db.h) -- func = SDOT_FUNC_EXP_WMR2_DB: the
 * integrand exp((w_i - |x - x_i|^2) / eps), no ball cut, 2D (its 3D integration is TODO in
 * the reference, ConvexPolyhedron3.tcc:1010-1013); one domain shape or constant pieces.
 * Masses are the exact integrals (the reference fits a table to precision 1e-10,
 * ConvexPolyhedron2.tcc:1747-1787), edge items the erf closed form (:346-389), and
 * d m_i / d w_i adds m_i / eps (:124-127).  eps > 0; other funcs ignore eps. */
sdot_status sdot_pd_set_radial_func_eps(sdot_pd *pd, int func, double eps);
/* SpZGrid::translations (SpZGrid.h:47-53, traversal SpZGrid.tcc:558-561,610): every
 * cell is also cut by the diracs translated by each of the n vectors
 * (translations[dim s .. dim s + dim)); a face with the image of dirac j adds to
 * M_ij with the distance to that image, a face with the cell's own image adds
 * nothing (get_der_integrals_wrt_weights.h:58-70).  n = 0 clears them.  At most 64
 * vectors, non-zero and finite; N * (n + 1) < 2^31.  Single problems only (not
 * batches or the serving loop). */
sdot_status sdot_pd_set_translations(sdot_pd *pd, const double *translations, int64_t n);

/* Run the cell pass (grid build + cells + Hessian rows) on the device; results
 * stay resident.  Returns SDOT_EMPTY_CELL when a cell is empty or has zero mass
 * (get_der_integrals_wrt_weights.h:93-96), like the reference's int result. */
sdot_status sdot_pd_compute(sdot_pd *pd);

/* get_integrals(res, ...) -- get_integrals.h:13.  Runs the pass if needed. */
sdot_status sdot_pd_integrals(sdot_pd *pd, double *masses);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid allocation after set-up): the masses of
 * sdot_pd_integrals for n diracs given as device arrays (d_positions 2n interleaved,
 * d_weights n), written to the device array d_masses (n, dirac order), with no device
 * allocation and no host allocation.  The handle gives the settings (2D, precision, one
 * box shape, ball clip on or off; no image, translations or radial function) and its
 * stream, and keeps none of the results (its own pass is left as it was).  Every
 * intermediate array lives in one fixed 40 MiB device workspace per device, held
 * until the process ends: n <= 262144 and a grid of at most 2^20 Morton codes
 * (a dirac bounding box of aspect ratio up to about 4 at the largest n), else
 * SDOT_CAPACITY_OVERFLOW.  Calls on one device share the workspace and are
 * serialised by a lock inside the library; the call returns when the masses are
 * written.  Status as sdot_pd_integrals. */
sdot_status sdot_pd_integrals_noalloc(sdot_pd *pd, const double *d_positions, const double *d_weights,
                                      int64_t nb_diracs, double *d_masses);
/* get_der_centroids_and_integrals_wrt_weight_and_positions(m_offsets, m_columns, m_values,
 * v_values, ...) -- get_der_centroids_and_integrals_wrt_weight_and_positions.h:16 (2D:
 * one domain shape with or without the ball, several shapes or constant pixels without it; 3D: one
 * shape, several shapes or voxels; translations in both; quadratic pixels: SDOT_NOT_SUPPORTED).  With nu = dim + 1: rows
 * nu n + d (d < dim: centroid coordinate d; d = dim: mass of dirac n), columns nu m + e
 * (e < dim: position coordinate e; e = dim: weight of dirac m), sorted;
 * offsets[nu N + 1]; v_values[nu N] = centroids and masses.  As in the reference, the centroid rows are divided by mass + (mass == 1), and
 * the first moments of the faces are not scaled by the density coefficient (exact for
 * density 1).  Two-phase like sdot_pd_der_integrals_wrt_weights (cols = NULL: nnz only).
 * Runs its own pass (FP64 geometry in either precision). */
sdot_status sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions(sdot_pd *pd, int64_t *offsets, int64_t *cols,
                                                                         double *vals, double *v_values, int64_t cap,
                                                                         int64_t *nnz);
/* get_centroids(grid, bounds, positions, weights, N, radial_func, cb) -- get_centroids.h:14
 * (rand_ratio = 0): centroids[N*dim] (dirac order) of the density over each cell,
 * 0 for an empty cell; masses may be NULL.  Runs the pass (with first moments) if
 * needed.  SDOT_NOT_SUPPORTED with quadratic pixels. */
sdot_status sdot_pd_centroids(sdot_pd *pd, double *centroids, double *masses);
/* per-cell boundary measures: ball part (arcs with cut id i) and whole perimeter
 * (area of the boundary in 3D); either may be NULL.  With an image density or
 * several shapes both are density-weighted and the perimeter counts the faces
 * shared with other cells and the ball arcs only (the domain boundary is not counted). */
sdot_status sdot_pd_boundary_measures(sdot_pd *pd, double *ball_arc, double *perimeter);
/* get_der_integrals_wrt_weights(offsets, cols, vals, masses, ...) -- get_der_integrals_wrt_weights.h:16.
 * Two-phase: with cols == NULL only *nnz is written.  offsets has N+1 entries;
 * cols/vals need room for *nnz entries (cap).  masses may be NULL. */
sdot_status sdot_pd_der_integrals_wrt_weights(sdot_pd *pd, int64_t *offsets, int64_t *cols, double *vals,
                                              double *masses, int64_t cap, int64_t *nnz);

/* Builds the dirac-order CSR of the last pass in HBM (no host copy) and returns
 * handle-owned device pointers: offsets[N+1]; row i occupies
 * [offsets[i], offsets[i+1]) with strictly ascending columns, one entry per dirac
 * (faces with a dirac and with its periodic images merged), the layout of
 * get_der_integrals_wrt_weights.h:80-88.  Valid until the next call on the handle. */
sdot_status sdot_pd_csr_device(sdot_pd *pd, const int64_t **d_offsets, const int64_t **d_cols, const double **d_vals);

typedef struct sdot_solve_stats {
    int newton_iters;
    int pcg_iters_total;
    int backtracks;
    int precond_levels;          /* multigrid levels of the last linear solve (0: Jacobi) */
    double max_rel_mass_err;     /* at the returned weights */
    double max_abs_dw;           /* last Newton step */
    int64_t nnz;                 /* Hessian entries of the final pass (row slots + diagonals) */
    double seconds_cells;        /* device time in the cell passes */
    double seconds_solve;        /* device time in the PCG solves */
    double seconds_total;        /* wall time of the call */
} sdot_solve_stats;

/* OptimalTransportSolver::solve(pos, w_inout, N, masses, rf) -- OptimalTransportSolver.tcc:27-119.
 * Newton with the reference's rules (M00 doubled without ball; 0.9/0.1 backtrack on an
 * empty cell).  stop_rule 0: max_i |m_i - t_i| / t_i < tol; 1: max |dw| < tol (reference).
 * weights (dirac order) are read as the start and overwritten with the result.
 * In 3D the passes after the first start each cell from its neighbours in the solve's last
 * complete pass (the same cells; only the rounding of their vertices follows the solve's
 * history): the result depends on the call's inputs only, bitwise, whatever the handle
 * computed before. */
sdot_status sdot_ot_solve(sdot_pd *pd, const double *target_masses, double tol, int stop_rule, int max_iter,
                          double *weights_inout, sdot_solve_stats *stats);

/* Preconditioner of the linear solves (the reference's EigenSolver is a sparse Cholesky,
 * EigenSolver.cpp:28-29; its optional AmgclSolver, AmgclSolver.cpp:19-37, an AMG-preconditioned
 * Krylov solver): aggregation multigrid V-cycle (default) or Jacobi. */
typedef enum sdot_precond { SDOT_PRECOND_AMG = 0, SDOT_PRECOND_JACOBI = 1 } sdot_precond;
sdot_status sdot_pd_set_linear_solver(sdot_pd *pd, int precond);

/* Tuning knobs (0 = automatic): mean diracs per grid cell (x100), threads per block. */
sdot_status sdot_pd_set_tuning(sdot_pd *pd, int diracs_per_cell_x100, int block_size);

typedef struct sdot_timings {
    double ms_grid;              /* bbox + sort + cell tables */
    double ms_cells;             /* cell kernel(s), incl. overflow re-runs */
    double ms_rows;              /* CSR export (0 unless requested) */
    double ms_total;
    int64_t nb_cells_overflow;   /* cells re-run with a larger polygon / polyhedron capacity */
    int64_t nnz_offdiag;         /* not counted here (0): a timing read stays free of device work;
                                    sdot_solve_stats.nnz gives the entries of a solve's final pass */
    int64_t grid_cells;          /* grid cells of the last grid build */
    double ms_cell_kernel;       /* the cell kernels (2D: all capacity tiers; 3D: the main
                                    kernel alone), HIP events around the launches */
    double cell_flops;           /* algorithmic FLOPs of the pass: sum over cells of
                                    4K^2 + 37K + 40A (2D) / 12K^2 + 100K - 284 (3D),
                                    K = faces, A = ball arcs of the final cell */
    int64_t sum_faces;           /* sum of K */
    int64_t sum_arcs;            /* sum of A */
} sdot_timings;
/* HIP-event timings of the last sdot_pd_compute (on the handle's stream). */
sdot_status sdot_pd_last_timings(sdot_pd *pd, sdot_timings *t);

/* ---------------------------------------------------------------------------
 * Batches: K independent 2D problems solved together (one launch per kernel
 * over all problems).  The reference solves one problem per
 * OptimalTransportSolver::solve call (OptimalTransportSolver.tcc:27-119); a batch
 * applies the same Newton rules per problem, each problem with its own grid,
 * domain box, convergence test and iteration count.  A problem's weights do not
 * depend on the other problems of the batch (bitwise).
 * ------------------------------------------------------------------------- */
typedef struct sdot_batch sdot_batch;

typedef struct sdot_batch_stats {
    int newton_iters_max;        /* most Newton passes of any problem */
    int nb_converged;
    int pcg_iters_total;         /* sum over problems and Newton steps */
    int backtracks;              /* sum over problems */
    int graph_replays;           /* launches of the captured PCG graph */
    int tuned_dpc_x100;          /* grid resolution used (diracs per grid cell x100) */
    double max_rel_mass_err;     /* over problems, at the returned weights */
    double seconds_cells;        /* device time of the batched cell passes */
    double seconds_solve;        /* wall time of the batched PCG solves */
    double seconds_total;        /* wall time of the call */
    double ms_pcg_iteration;     /* device time of one captured PCG iteration (all problems) */
    int precond_levels;          /* deepest multigrid hierarchy (0: Jacobi) */
    int precond_fallbacks;       /* problem solves whose hierarchy could not be built (Jacobi used) */
    int tuned_cells_per_wave;    /* cells a wavefront grabs at a time in the persistent pass */
    int tuned_first_tier;        /* capacity tier set of the pass (0 / 1) */
    int tuning_passes;           /* timed passes of the autotuner in this call (0: schedule given or cached) */
    int device_newton;           /* 1: Newton driven from the device (newton_replays graph replays) */
    int newton_replays;
    double seconds_device;       /* device-driven Newton: device time of all replays (passes and
                                    solves are not timed apart; seconds_cells / seconds_solve are 0) */
} sdot_batch_stats;

/* dim must be 2 (3 returns SDOT_NOT_SUPPORTED). */
sdot_status sdot_batch_create(sdot_batch **b, int dim, int precision);
sdot_status sdot_batch_destroy(sdot_batch *b);
const char *sdot_batch_last_error(const sdot_batch *b);
/* Problem p owns diracs [offsets[p], offsets[p+1]) of the concatenated arrays
 * (positions interleaved, weights = start, targets > 0); boxes: (x0, y0, x1, y1)
 * per problem, NULL = unit square for all; ball: 1 = InBallW05 for all problems. */
sdot_status sdot_batch_set_problems(sdot_batch *b, int64_t nb_problems, const int64_t *offsets, const double *positions,
                                    const double *weights, const double *targets, const double *boxes, int ball);
/* diracs_per_cell_x100: grid resolution (0 = autotune on the first solve of a size
 * class, then cached); use_graph: replay the PCG iterations from a captured hipGraph. */
sdot_status sdot_batch_set_tuning(sdot_batch *b, int diracs_per_cell_x100, int use_graph);
/* as sdot_pd_set_linear_solver */
sdot_status sdot_batch_set_linear_solver(sdot_batch *b, int precond);
/* Schedule of the batched passes and of Newton (no counterpart in the reference):
 * cells_per_wave = cells a wavefront grabs at a time in the persistent first-tier
 * pass (a multiple of 64 up to 4096; 0 = autotune); first_tier = capacity tier set
 * (0: FP64 starts at 8-vertex LDS slots, mixed at 16; 1: FP64 at 16, mixed at 8;
 * -1 = autotune); device_newton = 1: the Newton iteration of every problem (pass,
 * residual, decision, accept / backtrack, multigrid setup, cg_per_replay CG
 * iterations, update) is one captured hipGraph replayed until a device-side count of
 * unfinished problems reaches 0; 0: one host-driven Newton iteration at a time. */
sdot_status sdot_batch_set_schedule(sdot_batch *b, int cells_per_wave, int first_tier, int device_newton,
                                    int cg_per_replay);
/* Newton on every problem, starting from the weights given to set_problems; stop_rule
 * as sdot_ot_solve; the autotuning passes (first solve of a size class) use the start
 * weights.  weights_out (N), masses_out
 * (N, may be NULL): dirac order; status_out / iters_out (nb_problems, may be NULL).
 * Returns SDOT_OK when every problem converged, SDOT_NOT_CONVERGED otherwise. */
sdot_status sdot_batch_solve(sdot_batch *b, double tol, int stop_rule, int max_iter, double *weights_out,
                             double *masses_out, int *status_out, int *iters_out, sdot_batch_stats *stats);

/* ---------------------------------------------------------------------------
 * Serving loop: independent OT solve requests (2D and 3D, unit box or a given
 * box, FunctionEnum::Unit, density 1) submitted at any time and answered with
 * their converged weights.  No counterpart in the reference, whose solver runs
 * one problem per call (OptimalTransportSolver.tcc:27-119).
 *  - 2D requests join one continuous batch (the batched solve above) at the next
 *    Newton-iteration boundary and leave it when they converge;
 *  - 3D requests run one per worker, each worker with its own device handle and
 *    stream, concurrently with the 2D batch;
 *  - admission by HBM footprint: a request waits in its lane's FIFO queue until
 *    its estimated device footprint fits the budget beside the resident ones;
 *  - a request's weights do not depend on what else is being served (bitwise),
 *    given the same configuration.
 * ------------------------------------------------------------------------- */
typedef struct sdot_srv sdot_srv;

typedef struct sdot_srv_config {
    int precision;               /* SDOT_FP64 / SDOT_MIXED */
    double hbm_budget_bytes;     /* admission limit for resident requests (0: 80 % of free HBM at creation) */
    int max_batch_problems;      /* 2D requests advanced together (0: 4096) */
    int workers_3d;              /* concurrent 3D solves (0: 2) */
    int diracs_per_cell_x100;    /* 2D grid resolution (0: 250) */
    double tol;                  /* stop tolerance (0: 1e-6) */
    int stop_rule;               /* as sdot_ot_solve */
    int max_iter;                /* Newton iterations per request (0: 150) */
} sdot_srv_config;

typedef struct sdot_srv_result {
    int status;                  /* SDOT_OK / SDOT_NOT_CONVERGED / an error of the solve */
    int newton_iters;
    double max_rel_mass_err;
    double t_submit, t_start, t_done;   /* seconds on the server's clock */
} sdot_srv_result;

typedef struct sdot_srv_stats {
    int64_t queued_2d, queued_3d, running, completed;
    double resident_bytes, peak_resident_bytes, budget_bytes;
    int64_t cells_evaluated;     /* cell evaluations of every pass (2D ticks and 3D passes) */
    int64_t ticks_2d;            /* Newton iterations of the 2D batch */
    double seconds;              /* server clock */
} sdot_srv_stats;

sdot_status sdot_srv_create(sdot_srv **srv, const sdot_srv_config *cfg);
/* Copies the inputs and returns at once.  positions n*dim interleaved; weights
 * (start, may be NULL = 0); targets (n, > 0); box: lo[dim], hi[dim] or NULL = unit box.
 * SDOT_OOM when the request's estimated footprint (with a 3D worker's handle) exceeds
 * the HBM budget, so that it could not run even on an idle device.  A queued request
 * starts when its footprint fits beside the resident ones, or when nothing is running. */
sdot_status sdot_srv_submit(sdot_srv *srv, int dim, int64_t n, const double *positions, const double *weights,
                            const double *targets, const double *box, int64_t *request_id);
/* state: 0 queued, 1 running, 2 finished */
sdot_status sdot_srv_poll(sdot_srv *srv, int64_t request_id, int *state);
/* Waits until the request finishes (timeout_s < 0: no limit); then copies its weights
 * (n, dirac order; may be NULL) and result, and forgets the request.  Returns
 * SDOT_PENDING on a timeout, SDOT_INVALID_ARG for an unknown id, SDOT_OK otherwise
 * (the solve's own status is result->status). */
sdot_status sdot_srv_wait(sdot_srv *srv, int64_t request_id, double timeout_s, double *weights_out,
                          sdot_srv_result *result);
sdot_status sdot_srv_get_stats(sdot_srv *srv, sdot_srv_stats *stats);
const char *sdot_srv_last_error(const sdot_srv *srv);
/* Stops the lanes after their current step; unfinished requests are dropped. */
sdot_status sdot_srv_destroy(sdot_srv *srv);

/* display_vtk(VtkOutput<1, TF> &, grid, bounds, positions, weights, nb_diracs[, radial_func])
 * -- PowerDiagram/display_vtk.h:13-25, written as VtkOutput::save does
 * (Display/VtkOutput.tcc:17-79): the cells of the current diracs and weights to `filename`
 * as a legacy ASCII VTK unstructured grid of polygons (type 7) with one cell field,
 * `field_name` (NULL: "num_dirac"), holding the dirac index.  2D: one polygon per piece of
 * the domain (shape, pixel) a cell meets, ball arcs cut into 20 segments, a whole disc as 21
 * points, every polygon closed by repeating its first point; 3D: one polygon per face of
 * each piece.  Cells are listed in dirac order; a cell of zero mass draws nothing.  Runs
 * the pass if needed; SDOT_NOT_SUPPORTED with periodic images. */
sdot_status sdot_pd_display_vtk(sdot_pd *pd, const char *filename, const char *field_name);

/* splitmix64 -> xoshiro256** uniform doubles (x >> 11) * 2^-53: the synthetic-input generator. */
void sdot_rng_uniform(uint64_t seed, int64_t n, double *out);

/* Measurement helper, no counterpart in the reference: copies nbytes (a positive
 * multiple of 16) between two device buffers of the current device with a plain
 * copy kernel (k_copy16) on the null stream and waits.  Its hardware counters
 * (FETCH_SIZE, WRITE_SIZE) for a known size scale the counters of the other kernels. */
sdot_status sdot_calibration_copy(void *dst, const void *src, int64_t nbytes);

#ifdef __cplusplus
}
#endif
#endif
