// Newton weight solve on the device: OptimalTransportSolver::solve
// (src/sdot/Solvers/OptimalTransportSolver.tcc:27-119) with the linear solve of
// EigenSolver::solve (src/sdot/Solvers/EigenSolver.cpp:9-41, a sparse Cholesky)
// replaced by FP64 Jacobi-preconditioned CG on the Hessian rows the cell pass
// leaves in HBM (diagonal + column-major ELL rows in grid order).
//
// All reductions are two-level and fixed-order (per-block partials, then one
// block), so a solve is bitwise reproducible run to run.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cfloat>
#include <cstdio>
#include <cstdlib>

#include "internal.h"
#include "pcg.h"

namespace sdot_amd {

namespace {

constexpr int BS = 256;
constexpr int MAXB = 1024;   // reduction blocks

constexpr int S_NS = 8;   // scalar slots kept after the reduction partials

__device__ inline double block_sum(double v, double *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    double r = sm[0];
    __syncthreads();
    return r;
}

__device__ inline double block_max(double v, double *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    double r = sm[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(BS) void k_gather(const double *src, const int *perm, int64_t n, double *dst) {
    int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (t < n) dst[t] = src[perm[t]];
}

// b = m - target, partial max |b|/target
__global__ __launch_bounds__(BS) void k_residual(const double *m, const double *tgt, int64_t n, double *b, double *part) {
    __shared__ double sm[BS];
    double mx = 0;
    for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < n; t += (int64_t)gridDim.x * BS) {
        double v = m[t] - tgt[t];
        b[t] = v;
        mx = fmax(mx, fabs(v) / tgt[t]);
    }
    mx = block_max(mx, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

__global__ __launch_bounds__(BS) void k_max_reduce(const double *part, int nb, double *out) {
    __shared__ double sm[BS];
    double mx = 0;
    for (int k = threadIdx.x; k < nb; k += BS) mx = fmax(mx, part[k]);
    mx = block_max(mx, sm);
    if (threadIdx.x == 0) *out = mx;
}

__global__ void k_gauge(double *diag, const int *inv_perm) { diag[inv_perm[0]] *= 2; }

// w[perm[t]] -= dw[t]; partial max |dw|
__global__ __launch_bounds__(BS) void k_update_w(const double *dw, const int *perm, int64_t n, double *w, double *part) {
    __shared__ double sm[BS];
    double mx = 0;
    for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < n; t += (int64_t)gridDim.x * BS) {
        double d = dw[t];
        w[perm[t]] -= d;
        mx = fmax(mx, fabs(d));
    }
    mx = block_max(mx, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

__global__ __launch_bounds__(BS) void k_backtrack(const double *old, int64_t n, double *w) {
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
        w[i] = 0.9 * old[i] + 0.1 * w[i];
}

inline int red_blocks(int64_t n) { return (int)std::min<int64_t>(MAXB, std::max<int64_t>(1, (n + BS - 1) / BS)); }

}  // namespace

// Solves A x = b (A: diag + ELL rows of the last pass) by Jacobi-PCG to the
// relative residual rtol.  Returns the iteration count.
// M x = b on the handle's Hessian rows; relative tolerance from the mass error at d_err
int pcg_solve(sdot_pd *pd, const double *b, double *x, const double *d_err) {
    if (!pd->pcg) {
        pd->pcg = new PcgSolver();
        pd->pcg->set_dim(pd->dim);
    }
    if (pd->pcg_n != pd->N) {
        pd->pcg->set_layout({0, pd->N}, pd->stream);
        pd->pcg_n = pd->N;
    }
    pd->pcg->set_options(pd->precond == SDOT_PRECOND_AMG, pd->use_graph != 0);
    std::vector<int> its;
    PcgStats ps;
    pd->pcg->solve(PcgMat{pd->diag.p, pd->row_len.p, pd->ell_col.p, pd->ell_val.p, pd->N, pd->ell_cap}, b, x, d_err, {1}, 100000,
                   its, ps);
    pd->last_precond_levels = ps.amg_fallbacks ? 0 : ps.amg_levels;
    return its[0];
}

sdot_status ot_solve(sdot_pd *pd, const double *targets, double tol, int stop_rule, int max_iter, double *w_inout,
                     sdot_solve_stats *stats) {
    auto t_start = std::chrono::steady_clock::now();
    const int64_t n = pd->N;
    hipStream_t s = pd->stream;
    for (int k = 0; k < 5; ++k) pd->sv[k].ensure(n);
    double *tgt = pd->sv[0].p, *old = pd->sv[1].p, *tsort = pd->sv[2].p, *b = pd->sv[3].p, *dw = pd->sv[4].p;
    pd->red.ensure(2 * MAXB + S_NS + 8);
    double *part = pd->red.p, *scal = pd->red.p + 2 * MAXB + S_NS;
    HIPCK(hipMemcpyAsync(tgt, targets, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(pd->w.p, w_inout, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(old, pd->w.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i)
        if (!(targets[i] > 0)) throw StatusError{SDOT_INVALID_ARG, "target masses must be positive"};
    pd->pass_valid = false;
    for (int k = 4; k < 8; ++k)
        if (!pd->ev[k]) HIPCK(hipEventCreate(&pd->ev[k]));

    sdot_solve_stats st{};
    st.max_rel_mass_err = INFINITY;
    sdot_status result = SDOT_NOT_CONVERGED;
    bool tsort_ready = false;
    const int nb = red_blocks(n);
    const unsigned nbg = (unsigned)((n + BS - 1) / BS);
    double ms_cells = 0, ms_solve = 0;
    // SDOT_TRACE=1: one line per Newton iteration on stderr (pass / PCG timings)
    const char *trace_env = getenv("SDOT_TRACE");
    const bool trace = trace_env && trace_env[0] == '1';
    auto wall_ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count(); };
    auto trace_pass = [&](int it, sdot_status ps) {
        if (trace)
            fprintf(stderr, "[sdot] it %d pass %.2f ms (grid %.2f, cell kernel %.2f, all cell kernels %.2f, wide cells %lld) status %d wall %.1f ms\n",
                    it, pd->last.ms_total, pd->last.ms_grid, pd->last.ms_cell_kernel, pd->last.ms_cells,
                    (long long)pd->last.nb_cells_overflow, (int)ps, wall_ms());
    };
    // an empty cell only triggers a backtrack (OptimalTransportSolver.tcc:72-79): the
    // passes stop at the first one (get_der_integrals_wrt_weights' stop_if_void)
    struct StopIfVoid {
        sdot_pd *pd;
        explicit StopIfVoid(sdot_pd *p) : pd(p) { pd->stop_if_void = true; }
        ~StopIfVoid() { pd->stop_if_void = false; }
    } stop_guard(pd);
    // 3D: each pass after the first starts its cells from their neighbours in the
    // previous pass (the rows the PCG has just used); the first pass starts cold, so a
    // solve depends on its inputs only
    struct WarmStart {
        sdot_pd *pd;
        explicit WarmStart(sdot_pd *p) : pd(p) {
            pd->hint_ok = false;
            pd->warm = true;
        }
        ~WarmStart() { pd->warm = false; }
    } warm_guard(pd);
    for (int it = 0; it < max_iter; ++it) {
        sdot_status ps = compute_pass(pd);
        ms_cells += pd->last.ms_total;
        st.newton_iters = it + 1;
        trace_pass(it, ps);
        if (ps == SDOT_EMPTY_CELL) {   // OptimalTransportSolver.tcc:72-79
            hipLaunchKernelGGL(k_backtrack, dim3(nb), dim3(BS), 0, s, old, n, pd->w.p);
            HIPCK(hipGetLastError());
            pd->pass_valid = false;
            st.backtracks++;
            continue;
        }
        HIPCK(hipMemcpyAsync(old, pd->w.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (!tsort_ready) {
            hipLaunchKernelGGL(k_gather, dim3(nbg), dim3(BS), 0, s, tgt, pd->perm.p, n, tsort);
            tsort_ready = true;
        }
        hipLaunchKernelGGL(k_residual, dim3(nb), dim3(BS), 0, s, pd->mass.p, tsort, n, b, part);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(BS), 0, s, part, nb, scal);
        HIPCK(hipGetLastError());
        double rel = 0;
        HIPCK(hipMemcpyAsync(&rel, scal, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        st.max_rel_mass_err = rel;
        if (stop_rule == 0 && rel < tol) {
            result = SDOT_OK;
            break;
        }
        if (!pd->ball) {
            hipLaunchKernelGGL(k_gauge, dim3(1), dim3(1), 0, s, pd->diag.p, pd->inv_perm.p);  // :62-63
            HIPCK(hipGetLastError());
        }
        HIPCK(hipEventRecord(pd->ev[4], s));
        st.pcg_iters_total += pcg_solve(pd, b, dw, scal);   // rtol = max(1e-14, min(1e-3, 1e-3 rel))
        HIPCK(hipEventRecord(pd->ev[5], s));
        hipLaunchKernelGGL(k_update_w, dim3(nb), dim3(BS), 0, s, dw, pd->perm.p, n, pd->w.p, part);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(BS), 0, s, part, nb, scal);
        HIPCK(hipGetLastError());
        double mdw = 0;
        HIPCK(hipMemcpyAsync(&mdw, scal, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        float ms = 0;
        if (hipEventElapsedTime(&ms, pd->ev[4], pd->ev[5]) == hipSuccess) ms_solve += ms;
        pd->pass_valid = false;
        st.max_abs_dw = mdw;
        if (trace)
            fprintf(stderr, "[sdot] it %d rel mass err %.3e, pcg %d iters %.2f ms, max |dw| %.3e, wall %.1f ms\n", it, rel,
                    st.pcg_iters_total, ms, mdw, wall_ms());
        if (stop_rule == 1 && mdw < tol) {
            result = SDOT_OK;
            break;
        }
    }
    // the masses at the returned weights: one more pass when the loop ended on an update
    // (a complete one: the caller reads its per-cell results)
    pd->stop_if_void = false;
    if (!pd->pass_valid) {
        sdot_status ps = compute_pass(pd);
        ms_cells += pd->last.ms_total;
        if (ps == SDOT_OK) {
            if (!tsort_ready) {
                hipLaunchKernelGGL(k_gather, dim3(nbg), dim3(BS), 0, s, tgt, pd->perm.p, n, tsort);
                tsort_ready = true;
            }
            hipLaunchKernelGGL(k_residual, dim3(nb), dim3(BS), 0, s, pd->mass.p, tsort, n, b, part);
            hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(BS), 0, s, part, nb, scal);
            HIPCK(hipGetLastError());
            double rel = 0;
            HIPCK(hipMemcpyAsync(&rel, scal, sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCK(hipStreamSynchronize(s));
            st.max_rel_mass_err = rel;
            if (stop_rule == 0 && rel < tol) result = SDOT_OK;
        }
    }
    HIPCK(hipMemcpyAsync(w_inout, pd->w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    st.nnz = pd->pass_valid ? n + count_offdiag(pd) : 0;   // Hessian entries of the final pass (rows + diagonal)
    st.precond_levels = pd->last_precond_levels;
    st.seconds_cells = ms_cells * 1e-3;
    st.seconds_solve = ms_solve * 1e-3;
    st.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = st;
    return result;
}

}  // namespace sdot_amd
