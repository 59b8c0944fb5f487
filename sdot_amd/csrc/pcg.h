// Internal: FP64 preconditioned CG for K independent SPD systems stored as one
// block-diagonal matrix (problem p owns rows [off[p], off[p+1])), in the
// Hessian layout of the cell pass: diagonal + column-major ELL rows.
//
// Rows are processed in problem-aligned blocks; every dot product is reduced
// per problem over that problem's blocks in a fixed order, so each problem's
// iterates are bitwise independent of the other problems and run to run.
//
// Preconditioners: Jacobi, or an aggregation multigrid V-cycle (the reference's
// optional linear solver is amgcl's smoothed-aggregation AMG,
// src/sdot/Solvers/AmgclSolver.cpp:19-37).  The aggregates here are runs of
// AGG consecutive rows in grid (Morton) order, hence spatially compact; coarse
// operators are the Galerkin products with the piecewise-constant prolongation;
// one damped-Jacobi sweep before and after the coarse correction; the coarsest
// level of each problem (at most CMIN rows) is solved with its dense inverse.
#pragma once

#include <vector>

#include "internal.h"

namespace sdot_amd {

struct PcgMat {                 // rows in grid order
    const double *diag;
    const int *len;
    const int *col;
    const double *val;
    int64_t stride;
};

struct PcgStats {
    int iters_max = 0;          // most iterations of any problem
    int graph_replays = 0;
    int64_t graph_iterations = 0;   // PCG iterations run from the graph (all problems together)
    double ms_graph = 0;        // device time of the graph launches
    int amg_levels = 0;         // levels of the deepest hierarchy (0: Jacobi)
    int amg_fallbacks = 0;      // problems whose hierarchy overflowed (solved with Jacobi)
    double ms_setup = 0;        // hierarchy build (host wall time)
};

class PcgSolver {
public:
    PcgSolver();
    ~PcgSolver();
    PcgSolver(const PcgSolver &) = delete;
    PcgSolver &operator=(const PcgSolver &) = delete;

    // row offsets of the K problems (host array, K+1 entries); drops captured graphs
    void set_layout(const std::vector<int64_t> &off, hipStream_t s);
    void set_options(bool use_amg, bool use_graph);
    void invalidate();          // the matrix or vector buffers moved: recapture the graph
    // Solves A x = b for every problem p with active[p] != 0 (the others keep x
    // untouched), relative tolerance max(1e-14, min(1e-3, 1e-3 * err[p])) on the
    // residual 2-norm, err = d_err (device, K).  iters[p]: iterations of p.
    void solve(const PcgMat &A, const double *b, double *x, const double *d_err, const std::vector<int> &active,
               int max_iter, std::vector<int> &iters, PcgStats &st);

    bool amg() const { return use_amg_; }
    struct Impl;

private:
    Impl *im;
    bool use_amg_ = true, use_graph_ = true;
};

}  // namespace sdot_amd
