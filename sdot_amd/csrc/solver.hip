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

#include "internal.h"

namespace sdot_amd {

namespace {

constexpr int BS = 256;
constexpr int MAXB = 1024;   // reduction blocks

struct Mat {
    const double *diag;
    const int *len;
    const int *col;
    const double *val;
    int64_t stride;
    int64_t n;
};

// scalars kept on the device: [0] rz, [1] bb, [2] alpha, [3] beta, [4] rr, [5] done, [6] iters, [7] rtol^2
enum { S_RZ = 0, S_BB, S_ALPHA, S_BETA, S_RR, S_DONE, S_IT, S_RTOL2, S_NS };

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

__device__ inline double row_dot(const Mat &A, int64_t t, const double *x) {
    double s = A.diag[t] * x[t];
    const int L = A.len[t];
    for (int e = 0; e < L; ++e) s += A.val[e * A.stride + t] * x[A.col[e * A.stride + t]];
    return s;
}

__global__ __launch_bounds__(BS) void k_init(Mat A, const double *b, double *x, double *r, double *z, double *p,
                                             double *part) {
    __shared__ double sm[BS];
    double rz = 0, bb = 0;
    for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < A.n; t += (int64_t)gridDim.x * BS) {
        double d = A.diag[t];
        double rv = b[t];
        double zv = d != 0 ? rv / d : rv;
        x[t] = 0;
        r[t] = rv;
        z[t] = zv;
        p[t] = zv;
        rz += rv * zv;
        bb += rv * rv;
    }
    rz = block_sum(rz, sm);
    bb = block_sum(bb, sm);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rz;
        part[MAXB + blockIdx.x] = bb;
    }
}

__global__ __launch_bounds__(BS) void k_init_reduce(const double *part, int nb, double *S, double rtol) {
    __shared__ double sm[BS];
    double a = 0, c = 0;
    for (int k = threadIdx.x; k < nb; k += BS) {
        a += part[k];
        c += part[MAXB + k];
    }
    a = block_sum(a, sm);
    c = block_sum(c, sm);
    if (threadIdx.x == 0) {
        S[S_RZ] = a;
        S[S_BB] = c;
        S[S_RR] = c;
        S[S_RTOL2] = rtol * rtol;
        S[S_IT] = 0;
        S[S_DONE] = (c == 0 || !(a > 0)) ? 1.0 : 0.0;
    }
}

// q = A p, partial p.q
__global__ __launch_bounds__(BS) void k_spmv_dot(Mat A, const double *p, double *q, double *part, const double *S) {
    __shared__ double sm[BS];
    if (S[S_DONE] != 0) return;
    double pq = 0;
    for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < A.n; t += (int64_t)gridDim.x * BS) {
        double v = row_dot(A, t, p);
        q[t] = v;
        pq += p[t] * v;
    }
    pq = block_sum(pq, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

__global__ __launch_bounds__(BS) void k_alpha(const double *part, int nb, double *S) {
    __shared__ double sm[BS];
    if (S[S_DONE] != 0) return;
    double a = 0;
    for (int k = threadIdx.x; k < nb; k += BS) a += part[k];
    a = block_sum(a, sm);
    if (threadIdx.x == 0) S[S_ALPHA] = a != 0 ? S[S_RZ] / a : 0.0;
}

// x += alpha p; r -= alpha q; z = D^-1 r; partial r.z, r.r
__global__ __launch_bounds__(BS) void k_axpy(Mat A, const double *p, const double *q, double *x, double *r, double *z,
                                             double *part, const double *S) {
    __shared__ double sm[BS];
    if (S[S_DONE] != 0) return;
    const double alpha = S[S_ALPHA];
    double rz = 0, rr = 0;
    for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < A.n; t += (int64_t)gridDim.x * BS) {
        x[t] += alpha * p[t];
        double rv = r[t] - alpha * q[t];
        r[t] = rv;
        double d = A.diag[t];
        double zv = d != 0 ? rv / d : rv;
        z[t] = zv;
        rz += rv * zv;
        rr += rv * rv;
    }
    rz = block_sum(rz, sm);
    rr = block_sum(rr, sm);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rz;
        part[MAXB + blockIdx.x] = rr;
    }
}

__global__ __launch_bounds__(BS) void k_beta(const double *part, int nb, double *S, int max_iter) {
    __shared__ double sm[BS];
    if (S[S_DONE] != 0) return;
    double a = 0, c = 0;
    for (int k = threadIdx.x; k < nb; k += BS) {
        a += part[k];
        c += part[MAXB + k];
    }
    a = block_sum(a, sm);
    c = block_sum(c, sm);
    if (threadIdx.x == 0) {
        S[S_BETA] = S[S_RZ] != 0 ? a / S[S_RZ] : 0.0;
        S[S_RZ] = a;
        S[S_RR] = c;
        S[S_IT] += 1;
        if (c <= S[S_RTOL2] * S[S_BB] || S[S_IT] >= max_iter || !(a > 0)) S[S_DONE] = 1;
    }
}

__global__ __launch_bounds__(BS) void k_pupdate(int64_t n, const double *z, double *p, const double *S) {
    if (S[S_DONE] != 0) return;
    const double beta = S[S_BETA];
    for (int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x; t < n; t += (int64_t)gridDim.x * BS) p[t] = z[t] + beta * p[t];
}

// Newton helpers (sorted order t, dirac order i = perm[t])
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
int pcg_solve(sdot_pd *pd, const double *b, double *x, double rtol, int max_iter) {
    const int64_t n = pd->N;
    hipStream_t s = pd->stream;
    for (int k = 5; k < 9; ++k) pd->sv[k].ensure(n);
    double *r = pd->sv[5].p, *z = pd->sv[6].p, *p = pd->sv[7].p, *q = pd->sv[8].p;
    pd->red.ensure(2 * MAXB + S_NS + 8);
    double *part = pd->red.p, *S = pd->red.p + 2 * MAXB;
    Mat A{pd->diag.p, pd->row_len.p, pd->ell_col.p, pd->ell_val.p, n, n};
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_init, dim3(nb), dim3(BS), 0, s, A, b, x, r, z, p, part);
    hipLaunchKernelGGL(k_init_reduce, dim3(1), dim3(BS), 0, s, part, nb, S, rtol);
    HIPCK(hipGetLastError());
    const int chunk = 16;
    double hs[S_NS];
    for (int it = 0; it < max_iter; it += chunk) {
        for (int k = 0; k < chunk; ++k) {
            hipLaunchKernelGGL(k_spmv_dot, dim3(nb), dim3(BS), 0, s, A, p, q, part, S);
            hipLaunchKernelGGL(k_alpha, dim3(1), dim3(BS), 0, s, part, nb, S);
            hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(BS), 0, s, A, p, q, x, r, z, part, S);
            hipLaunchKernelGGL(k_beta, dim3(1), dim3(BS), 0, s, part, nb, S, max_iter);
            hipLaunchKernelGGL(k_pupdate, dim3(nb), dim3(BS), 0, s, n, z, p, S);
        }
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(hs, S, sizeof hs, hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        if (hs[S_DONE] != 0) break;
    }
    return (int)hs[S_IT];
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
    for (int it = 0; it < max_iter; ++it) {
        sdot_status ps = compute_pass(pd);
        ms_cells += pd->last.ms_total;
        st.newton_iters = it + 1;
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
        double rtol = std::max(1e-14, std::min(1e-3, 1e-3 * rel));
        st.pcg_iters_total += pcg_solve(pd, b, dw, rtol, 100000);
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
        if (stop_rule == 1 && mdw < tol) {
            result = SDOT_OK;
            break;
        }
    }
    // the masses at the returned weights: one more pass when the loop ended on an update
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
    st.nnz = 0;
    st.seconds_cells = ms_cells * 1e-3;
    st.seconds_solve = ms_solve * 1e-3;
    st.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = st;
    return result;
}

}  // namespace sdot_amd
