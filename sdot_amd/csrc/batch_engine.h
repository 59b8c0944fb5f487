// Internal: the engine behind batched solves and the serving loop.  It keeps a
// resident set of independent 2D problems in one set of device arrays (problem
// k owns diracs [off_k, off_k + n_k)), advances all of them by one Newton
// iteration per tick, and can change the resident set between ticks: problems
// that stay keep their device state (weights, last accepted weights) and move
// to their new offsets on the device; new problems are uploaded from the host.
#pragma once

#include <limits>
#include <vector>

#include "internal.h"

namespace sdot_amd {

struct BatchProblem {
    int64_t n = 0;
    const double *hpos = nullptr;   // host positions (caller-owned while resident; the grid is built from them)
    double box[4] = {0, 0, 1, 1};   // x0, y0, x1, y1
    GridInfo g{};                   // grid of the problem at the engine's resolution
    double tie = 0;                 // tie-breaking weight offset scale
    int64_t off = 0;                // offset in the current layout
    int64_t tag = 0;                // caller's id
    // Newton state
    int iters = 0, done = 0, dw_ok = 0, status = SDOT_NOT_CONVERGED, backtracks = 0, pcg_iters = 0;
    double err = std::numeric_limits<double>::infinity();
};

struct NewProblem {
    int64_t n;
    const double *pos, *w0, *tgt;   // host arrays (pos interleaved); pos must stay valid while resident
    double box[4];
    int64_t tag;
};

struct TickParams {
    double tol = 1e-6;
    int stop_rule = 0;              // 0: max relative mass error < tol; 1: max |dw| < tol
    int max_iter = 150;
};

struct DeviceSolveStats {
    int64_t replays = 0;            // graph replays
    double ms_device = 0;           // device time from the first replay to the last
    int pcg_iters = 0, backtracks = 0, widened = 0, fallbacks = 0;
    int64_t cells = 0;              // cells evaluated (problem size x passes)
};

struct TickStats {
    double ms_pass = 0, ms_pcg = 0, ms_pcg_graph = 0;
    int graph_replays = 0, pcg_iters = 0, backtracks = 0, solved = 0, amg_levels = 0, amg_fallbacks = 0;
    int64_t graph_iterations = 0;
    int64_t cells = 0;
};

class BatchEngine {
public:
    BatchEngine(int precision, int ball, int dpc_x100, bool use_graph);
    ~BatchEngine();
    BatchEngine(const BatchEngine &) = delete;
    BatchEngine &operator=(const BatchEngine &) = delete;

    // new resident set: the current problems listed in `keep` (in that order), then `add`
    void relayout(const std::vector<int> &keep, const std::vector<NewProblem> &add);
    // grid resolution (rebuilds every resident grid from the host positions)
    void set_resolution(int dpc_x100);
    int resolution() const { return dpc_x100_; }
    void set_use_graph(bool on);
    void set_preconditioner(bool amg);              // multigrid (default) or Jacobi
    // every problem back to its start weights and a fresh Newton state
    void restart();
    // one Newton iteration of every problem not done; appends the problems that
    // finished in this tick to `finished`; returns the number still running
    int tick(const TickParams &tp, std::vector<int> &finished, TickStats &ts);
    // one cell pass at the current weights (autotuning); device ms
    double pass_only();
    // Newton on every problem from its start weights, driven from the device: one
    // captured graph replayed until every problem is done (cg_per_replay CG
    // iterations per replay); fills probs[k]'s results
    int solve_device(const TickParams &tp, int cg_per_replay, DeviceSolveStats &ds);
    // scheduling of the cell passes: cells a wavefront grabs at a time (multiple of
    // 64) and the capacity tier set (0 default, 1 alternative first slot size)
    void set_schedule(int cells_per_wave, int tierset);
    int cells_per_wave() const;
    int tierset() const;
    int precond_levels() const;                      // multigrid levels (0: Jacobi)

    void read_weights(int k, double *out);          // dirac order of problem k
    void read_weights_all(double *out);
    void read_masses_all(double *out);
    void read_masses(int k, double *out);

    std::vector<BatchProblem> probs;
    int64_t total() const { return N_; }
    size_t device_bytes() const;
    static double bytes_per_dirac();                 // footprint estimate for admission
    sdot_pd *pd = nullptr;
    struct Impl;                                     // device tables and PCG state

private:
    Impl *im;
    int64_t N_ = 0;
    int ball_;
    int dpc_x100_;
    bool use_graph_ = true;
    void rebuild_grid();
    void enqueue_replay(const TickParams &tp, int cg_per_replay);
    void capture_replay(const TickParams &tp, int cg_per_replay);
    void widen_rows();
};

GridInfo problem_grid(const double *pos, int64_t n, int dpc_x100, double *tie);

}  // namespace sdot_amd
