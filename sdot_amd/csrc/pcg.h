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

// per-problem solver state, S[p * S_NS + field] (FP64): the CG scalars, the done
// mark and iteration count, the squared tolerance, and the start mark (1: the
// problem's solve starts at the next setup; set by the host in solve(), by the
// caller's own kernels when the solve is driven from the device)
namespace pcg_state {
enum { S_RZ = 0, S_BB, S_ALPHA, S_BETA, S_RR, S_DONE, S_IT, S_RTOL2, S_START, S_NS };
}

struct PcgMat {                 // rows in grid order
    const double *diag;
    const int *len;
    const int *col;
    const double *val;
    int64_t stride;
    int cap = 0;                // entries per row of the layout (0: unknown)
    // FP32 copy of every entry (null: none), written by the coarsening that reads the rows;
    // the V-cycle's smoothers read it, the CG's own product reads val
    float *val32 = nullptr;
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
    void set_dim(int dim);      // the problems' dimension: picks how the V-cycle walks the rows
    // Solves A x = b for every problem p with active[p] != 0 (the others keep x
    // untouched), relative tolerance max(1e-14, min(1e-3, 1e-3 * err[p])) on the
    // residual 2-norm, err = d_err (device, K).  iters[p]: iterations of p.
    void solve(const PcgMat &A, const double *b, double *x, const double *d_err, const std::vector<int> &active,
               int max_iter, std::vector<int> &iters, PcgStats &st);

    // Device-driven use (every call only enqueues kernels on the layout's stream, so a
    // caller may capture it in a graph): the caller's kernels mark in state() the
    // problems that start a solve (S_START = 1, S_DONE = 0, S_IT = 0, their
    // fallback() entry 0) and clear S_START for every other problem.
    // enqueue_start builds the hierarchy and initialises the iterates of the
    // starting problems (the others keep theirs); enqueue_iterations runs n CG
    // iterations of every problem whose S_DONE is 0.
    void enqueue_start(const PcgMat &A, const double *b, double *x, const double *d_err);
    void enqueue_iterations(const PcgMat &A, double *x, int n, int max_iter);
    double *state();
    int *fallback();

    bool amg() const { return use_amg_; }
    int levels() const;         // multigrid levels of the layout (0: Jacobi)
    struct Impl;

private:
    Impl *im;
    bool use_amg_ = true, use_graph_ = true;
};

}  // namespace sdot_amd
