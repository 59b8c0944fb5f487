// Batched preconditioned CG (see pcg.h): problem-aligned row blocks, per-problem
// fixed-order reductions, Jacobi or aggregation-multigrid preconditioning, and
// the iterations replayed from a captured hipGraph.
#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "pcg.h"

namespace sdot_amd {

using namespace pcg_state;

namespace {

constexpr int BB = 256;          // rows per block
constexpr int CHUNK = 8;         // PCG iterations per graph launch
constexpr int AGG = 4;           // rows per aggregate
constexpr int CMIN = 64;         // coarsest level: at most CMIN rows per problem
constexpr int CCAP = 64;         // off-diagonal entries per coarse row
constexpr double OMEGA = 0.8;    // damped-Jacobi smoothing
constexpr double CSCALE = 1.5;   // coarse-correction scaling

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ inline double bsum(double v, double *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = BB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sm[0];
    __syncthreads();
    return r;
}

struct Blocks {
    const int *blk_prob;
    const int64_t *blk_off;
    const int64_t *off;
};
struct BlockRows {
    int p;
    int64_t r0, r1;
};
__device__ inline BlockRows block_rows(const Blocks &B, int64_t b) {
    BlockRows br;
    br.p = B.blk_prob[b];
    br.r0 = B.off[br.p] + (b - B.blk_off[br.p]) * BB;
    br.r1 = B.off[br.p + 1] < br.r0 + BB ? B.off[br.p + 1] : br.r0 + BB;
    return br;
}

__global__ __launch_bounds__(BB) void kp_owner(const int64_t *off, int64_t K, int64_t n, int *idx) {
    const int64_t i = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (i >= n) return;
    int64_t lo = 0, hi = K;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    idx[i] = (int)lo;
}

// ---- multigrid setup ------------------------------------------------------------
// Coarse row I of level l+1 = sum of the AGG fine rows of its aggregate, columns
// mapped to aggregates, duplicates merged in order of first appearance.  A group of
// CLPR lanes builds one row: the entries are taken one after the other (the same order,
// hence the same row, as one thread would build), each looked up in the row so far by
// all lanes at once (a ballot finds it), so a lookup costs CCAP / CLPR steps instead of
// a scan of up to CCAP dependent LDS reads.
constexpr int CLPR = 16;
constexpr int CGRP = 256 / CLPR;   // coarse rows per block

__device__ inline void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(256) void kp_coarsen(PcgMat A, const int64_t *offl, const int64_t *offc,
                                                  const int *cprob, int64_t nc, double *cdiag, int *clen, int *ccol,
                                                  double *cval, int *amg_off, const double *S) {
    __shared__ int sc[CGRP][CCAP];
    __shared__ double sv[CGRP][CCAP];
    const int grp = threadIdx.x / CLPR, sub = threadIdx.x % CLPR;
    const int lane = threadIdx.x & 63, gbase = lane - sub;   // the group's first lane in the wavefront
    const int64_t I = blockIdx.x * (int64_t)CGRP + grp;
    if (I >= nc) return;   // uniform over the group
    int *uc = sc[grp];
    double *uv = sv[grp];
    const int p = cprob[I];
    if (S[p * S_NS + S_START] == 0) return;   // only the problems that start a solve
    const int64_t li = I - offc[p];
    const int64_t f0 = offl[p] + AGG * li;
    const int64_t f1 = f0 + AGG < offl[p + 1] ? f0 + AGG : offl[p + 1];
    double d = 0;
    int u = 0;
    bool ovf = false;
    for (int64_t t = f0; t < f1; ++t) {
        d += A.diag[t];
        const int L = A.len[t];
        for (int e = 0; e < L; ++e) {
            const int c = A.col[e * A.stride + t];
            const double v = A.val[e * A.stride + t];
            if (A.val32 && sub == 0) A.val32[e * A.stride + t] = (float)v;
            const int cc = (int)(offc[p] + (c - offl[p]) / AGG);
            if (cc == I) {
                d += v;
                continue;
            }
            int found = -1;
            for (int k = sub; k < u; k += CLPR)
                if (uc[k] == cc) found = k;
            const unsigned long long gm = (__ballot(found >= 0) >> gbase) & ((1ull << CLPR) - 1);
            if (gm) {
                const int k = __shfl(found, gbase + __ffsll((long long)gm) - 1);
                if (sub == 0) uv[k] += v;
            } else if (u < CCAP) {
                if (sub == 0) {
                    uc[u] = cc;
                    uv[u] = v;
                }
                ++u;
            } else {
                ovf = true;
            }
            wave_lds_order();
        }
    }
    wave_lds_order();
    if (sub == 0) {
        cdiag[I] = d;
        clen[I] = u;
        if (ovf || !(d > 0)) amg_off[p] = 1;
    }
    for (int k = sub; k < u; k += CLPR) {
        ccol[k * nc + I] = uc[k];
        cval[k * nc + I] = uv[k];
    }
}

// dense inverse of each problem's coarsest matrix (at most CMIN rows), in-place
// Gauss-Jordan without pivoting (the matrix is SPD); one block per problem
__global__ __launch_bounds__(CMIN) void kp_coarse_inv(PcgMat A, const int64_t *coff, const int *cn, const int *plev,
                                                      int lev, double *inv, int *amg_off, const double *S) {
    __shared__ double M[CMIN][CMIN + 1];
    __shared__ int bad;
    const int p = blockIdx.x, i = threadIdx.x, n = cn[p];
    if (plev[p] != lev || amg_off[p] || S[p * S_NS + S_START] == 0) return;
    const int64_t o = coff[p];
    for (int j = 0; j < CMIN; ++j) M[i][j] = 0;
    if (i == 0) bad = 0;
    __syncthreads();
    if (i < n) {
        const int64_t g = o + i;
        M[i][i] = A.diag[g];
        const int L = A.len[g];
        for (int e = 0; e < L; ++e) M[i][A.col[e * A.stride + g] - o] += A.val[e * A.stride + g];
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const double piv = M[k][k];
        __syncthreads();
        if (!(piv > 0)) {
            if (i == 0) bad = 1;
            break;
        }
        if (i == 0) {
            const double r = 1.0 / piv;
            M[k][k] = 1.0;
            for (int j = 0; j < n; ++j) M[k][j] *= r;
        }
        __syncthreads();
        double f = 0;
        if (i < n && i != k) {
            f = M[i][k];
            M[i][k] = 0;
        }
        __syncthreads();
        if (i < n && i != k)
            for (int j = 0; j < n; ++j) M[i][j] -= f * M[k][j];
        __syncthreads();
    }
    __syncthreads();
    if (bad) {
        if (i == 0) amg_off[p] = 1;
        return;
    }
    if (i < n)
        for (int j = 0; j < n; ++j) inv[((int64_t)p * CMIN + i) * CMIN + j] = M[i][j];
}

// ---- V-cycle ------------------------------------------------------------------
struct LevelView {
    PcgMat A;
    const int64_t *off;     // K+1 row offsets at this level
    const int *prob;        // row -> problem
    double *r;              // right-hand side at this level (written by the restriction)
    double *xpre;           // pre-smoothed iterate
    double *out;            // corrected and post-smoothed iterate
    int64_t n;
    int lev;
};

// gate 0: the iterations (problems still iterating); gate 1: the start of a solve
// (problems whose solve starts now; the others keep their iterates)
__device__ inline bool gated_out(const double *S, int p, int gate) {
    return gate ? S[p * S_NS + S_START] == 0 : S[p * S_NS + S_DONE] != 0;
}
__device__ inline bool skip(const double *S, const int *amg_off, int p, int gate) {
    return gated_out(S, p, gate) || amg_off[p];
}

// matrix entry k (= e * stride + t) as the V-cycle's smoothers read it: the FP32 copy when
// the level has one (the preconditioner approximates A^-1 anyway; the CG uses A in FP64)
__device__ inline double smooth_val(const PcgMat &A, int64_t k) { return A.val32 ? (double)A.val32[k] : A.val[k]; }

// sum over the LPR consecutive lanes of a row (LPR a power of two, at most 64)
template <int LPR>
__device__ inline double row_sum(double v) {
    for (int o = LPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPR);
    return v;
}


// residual of the pre-smoothed iterate x = w r / D of level l (x written with r by its
// producer: the CG update on level 0, the restriction on the coarse levels):
// res = r - A x, for the rows of problems that have a coarser level
template <int LPR>
__global__ __launch_bounds__(BB) void kp_smooth_res(LevelView F, const int *plev, double *res, const double *S,
                                                    const int *amg_off, int gate) {
    const int64_t g = blockIdx.x * (int64_t)BB + threadIdx.x;
    const int64_t t = g / LPR;
    const int sub = (int)(g % LPR);
    if (t >= F.n) return;
    const int p = F.prob[t];
    if (skip(S, amg_off, p, gate) || plev[p] <= F.lev) return;   // uniform over the row's lanes
    const int L = F.A.len[t];
    double ax = 0;
    for (int e = sub; e < L; e += LPR) {
        const int c = F.A.col[e * F.A.stride + t];
        ax += smooth_val(F.A, e * F.A.stride + t) * F.xpre[c];
    }
    ax = row_sum<LPR>(ax);
    if (sub == 0) res[t] = F.r[t] - (F.A.diag[t] * F.xpre[t] + ax);
}

// restriction: coarse row I = sum of the residuals of its aggregate, in row order
__global__ __launch_bounds__(BB) void kp_restrict(LevelView F, LevelView C, const double *res, const double *S,
                                                  const int *amg_off, int gate) {
    const int64_t I = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (I >= C.n) return;
    const int p = C.prob[I];
    if (skip(S, amg_off, p, gate)) return;
    const int64_t f0 = F.off[p] + AGG * (I - C.off[p]);
    const int64_t f1 = f0 + AGG < F.off[p + 1] ? f0 + AGG : F.off[p + 1];
    double sum = 0;
    for (int64_t t = f0; t < f1; ++t) sum += res[t];
    C.r[I] = sum;
    C.xpre[I] = OMEGA * sum / C.A.diag[I];
}

// the corrected iterate before post-smoothing, y = x + s P e (P: the aggregates' constant
// prolongation, e the coarse level's result), for the rows the post-smoothing reads
__global__ __launch_bounds__(BB) void kp_prolong(LevelView F, LevelView C, const int *plev, double *y, const double *S,
                                                 const int *amg_off, int gate, int jacobi) {
    const int64_t t = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (t >= F.n) return;
    const int p = F.prob[t];
    if (jacobi || skip(S, amg_off, p, gate) || plev[p] == F.lev) return;
    y[t] = F.xpre[t] + CSCALE * C.out[C.off[p] + (t - F.off[p]) / AGG];
}

// coarsest level of each problem: out = inverse * r (level 0 problems write z)
__global__ __launch_bounds__(CMIN) void kp_coarse_solve(const int64_t *cbase, const int *cn, const int *clev,
                                                        const double *inv, const double *vr, double *vout,
                                                        const double *r0, double *z0, const double *S,
                                                        const int *amg_off, int gate) {
    __shared__ double rs[CMIN];
    const int p = blockIdx.x, i = threadIdx.x, n = cn[p];
    if (skip(S, amg_off, p, gate)) return;
    const bool top = clev[p] == 0;
    const double *r = top ? r0 : vr;
    double *out = top ? z0 : vout;
    const int64_t o = cbase[p];
    if (i < n) rs[i] = r[o + i];
    __syncthreads();
    if (i < n) {
        double s = 0;
        const double *row = inv + ((int64_t)p * CMIN + i) * CMIN;
        for (int j = 0; j < n; ++j) s += row[j] * rs[j];
        out[o + i] = s;
    }
}

// the prolonged iterate at row t: precomputed (YIN, kp_prolong) or formed here
template <bool YIN>
__device__ inline double prolonged(const LevelView &F, const LevelView &C, const double *y, int p, int64_t t) {
    if (YIN) return y[t];
    return F.xpre[t] + CSCALE * C.out[C.off[p] + (t - F.off[p]) / AGG];
}

// post-smooth the prolonged iterate y (levels >= 1)
template <int LPR, bool YIN>
__global__ __launch_bounds__(BB) void kp_up(LevelView F, LevelView C, const double *y, const int *plev,
                                            const double *S, const int *amg_off, int gate) {
    const int64_t g = blockIdx.x * (int64_t)BB + threadIdx.x;
    const int64_t t = g / LPR;
    const int sub = (int)(g % LPR);
    if (t >= F.n) return;
    const int p = F.prob[t];
    if (skip(S, amg_off, p, gate) || plev[p] == F.lev) return;   // coarsest of p: set by the coarse solve
    const int L = F.A.len[t];
    double ay = 0;
    for (int e = sub; e < L; e += LPR) {
        const int c = F.A.col[e * F.A.stride + t];
        ay += smooth_val(F.A, e * F.A.stride + t) * prolonged<YIN>(F, C, y, p, c);
    }
    ay = row_sum<LPR>(ay);
    if (sub == 0) {
        const double yt = prolonged<YIN>(F, C, y, p, t), dt = F.A.diag[t];
        F.out[t] = yt + OMEGA * (F.r[t] - (dt * yt + ay)) / dt;
    }
}

// level 0: z = M r (multigrid, the coarse solve of a small problem, or Jacobi),
// per-block partial of r.z; on the first application also p = z
template <bool YIN>
__global__ __launch_bounds__(BB) void kp_up0(LevelView F, LevelView C, const double *y, const int *plev, Blocks B,
                                             double *part, double *pvec, int init, int jacobi, const double *S,
                                             const int *amg_off) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    const int p = br.p;
    if (gated_out(S, p, init)) return;   // uniform per block
    const bool jac = jacobi || amg_off[p];
    const bool coarse = !jac && plev[p] == 0;
    double rz = 0;
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
        double z;
        if (jac) {
            const double d = F.A.diag[t];
            z = d != 0 ? F.r[t] / d : F.r[t];
            F.out[t] = z;
        } else if (coarse) {
            z = F.out[t];
        } else {
            const double yt = prolonged<YIN>(F, C, y, p, t);
            double ay = F.A.diag[t] * yt;
            const int L = F.A.len[t];
            for (int e = 0; e < L; ++e) {
                const int c = F.A.col[e * F.A.stride + t];
                ay += smooth_val(F.A, e * F.A.stride + t) * prolonged<YIN>(F, C, y, p, c);
            }
            z = yt + OMEGA * (F.r[t] - ay) / F.A.diag[t];
            F.out[t] = z;
        }
        if (init) pvec[t] = z;
        rz += F.r[t] * z;
    }
    rz = bsum(rz, sm);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = rz;
}

// ---- CG ---------------------------------------------------------------------------
__global__ __launch_bounds__(BB) void kp_init(Blocks B, const double *b, double *x, double *r, double *part,
                                              const double *S, const double *diag, double *xpre) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    if (gated_out(S, br.p, 1)) return;
    double bb = 0;
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
        const double rv = b[t];
        x[t] = 0;
        r[t] = rv;
        xpre[t] = OMEGA * rv / diag[t];   // level 0's pre-smoothed iterate
        bb += rv * rv;
    }
    bb = bsum(bb, sm);
    if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = bb;
}

__global__ __launch_bounds__(BB) void kp_init_red(const double *part, const int64_t *blk_off, const double *err,
                                                  double *S) {
    __shared__ double sm[BB];
    const int p = blockIdx.x;
    double *sp = S + p * S_NS;
    if (gated_out(S, p, 1)) return;
    double a = 0, c = 0;
    for (int64_t k = blk_off[p] + threadIdx.x; k < blk_off[p + 1]; k += BB) {
        a += part[2 * k];
        c += part[2 * k + 1];
    }
    a = bsum(a, sm);
    c = bsum(c, sm);
    if (threadIdx.x == 0) {
        const double rtol = fmax(1e-14, fmin(1e-3, 1e-3 * err[p]));
        sp[S_RZ] = a;
        sp[S_BB] = c;
        sp[S_RR] = c;
        sp[S_RTOL2] = rtol * rtol;
        sp[S_IT] = 0;
        sp[S_DONE] = (c == 0 || !(a > 0)) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(BB) void kp_spmv_dot(PcgMat A, Blocks B, const double *p, double *q, double *part,
                                                  const double *S) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    if (S[br.p * S_NS + S_DONE] != 0) return;
    double pq = 0;
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
        double v = A.diag[t] * p[t];
        const int L = A.len[t];
        for (int e = 0; e < L; ++e) v += A.val[e * A.stride + t] * p[A.col[e * A.stride + t]];
        q[t] = v;
        pq += p[t] * v;
    }
    pq = bsum(pq, sm);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = pq;
}

__global__ __launch_bounds__(BB) void kp_alpha(const double *part, const int64_t *blk_off, double *S) {
    __shared__ double sm[BB];
    const int p = blockIdx.x;
    double *sp = S + p * S_NS;
    if (sp[S_DONE] != 0) return;
    double a = 0;
    for (int64_t k = blk_off[p] + threadIdx.x; k < blk_off[p + 1]; k += BB) a += part[2 * k];
    a = bsum(a, sm);
    if (threadIdx.x == 0) sp[S_ALPHA] = a != 0 ? sp[S_RZ] / a : 0.0;
}

__global__ __launch_bounds__(BB) void kp_axpy(Blocks B, const double *p, const double *q, double *x, double *r,
                                              double *part, const double *S, const double *diag, double *xpre) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    const double *sp = S + br.p * S_NS;
    if (sp[S_DONE] != 0) return;
    const double alpha = sp[S_ALPHA];
    double rr = 0;
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
        x[t] += alpha * p[t];
        const double rv = r[t] - alpha * q[t];
        r[t] = rv;
        xpre[t] = OMEGA * rv / diag[t];   // level 0's pre-smoothed iterate
        rr += rv * rv;
    }
    rr = bsum(rr, sm);
    if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = rr;
}

__global__ __launch_bounds__(BB) void kp_beta(const double *part, const int64_t *blk_off, double *S, int max_iter) {
    __shared__ double sm[BB];
    const int p = blockIdx.x;
    double *sp = S + p * S_NS;
    if (sp[S_DONE] != 0) return;
    double a = 0, c = 0;
    for (int64_t k = blk_off[p] + threadIdx.x; k < blk_off[p + 1]; k += BB) {
        a += part[2 * k];
        c += part[2 * k + 1];
    }
    a = bsum(a, sm);
    c = bsum(c, sm);
    if (threadIdx.x == 0) {
        sp[S_BETA] = sp[S_RZ] != 0 ? a / sp[S_RZ] : 0.0;
        sp[S_RZ] = a;
        sp[S_RR] = c;
        sp[S_IT] += 1;
        if (c <= sp[S_RTOL2] * sp[S_BB] || sp[S_IT] >= max_iter || !(a > 0)) sp[S_DONE] = 1;
    }
}

__global__ __launch_bounds__(BB) void kp_pupdate(Blocks B, const double *z, double *p, const double *S) {
    const BlockRows br = block_rows(B, blockIdx.x);
    const double *sp = S + br.p * S_NS;
    if (sp[S_DONE] != 0) return;
    const double beta = sp[S_BETA];
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) p[t] = z[t] + beta * p[t];
}

__global__ __launch_bounds__(BB) void kp_count_active(const double *S, int64_t K, int *out) {
    __shared__ double sm[BB];
    double c = 0;
    for (int64_t p = threadIdx.x; p < K; p += BB) c += S[p * S_NS + S_DONE] == 0 ? 1.0 : 0.0;
    c = bsum(c, sm);
    if (threadIdx.x == 0) *out = (int)c;
}

// How the V-cycle walks the rows: lanes per row on the coarse levels (rows of up to
// CCAP entries; one lane per row would wait on one gather after the other) and whether
// the prolonged iterate is formed once per row (a pass, then one gather per entry) or
// per entry (two gathers).  Chosen by the problems' dimension only, so that a problem's
// iterates are the same alone and in a batch.
struct RowPlan {
    int lpr_coarse = 8;
    bool prolong0 = true, prolong_coarse = true;
};

void launch_smooth_res(int lpr, const LevelView &F, const int *plev, double *res, const double *S, const int *amg_off,
                       int gate, hipStream_t s) {
    const unsigned nb = nblocks(F.n * lpr, BB);
    switch (lpr) {
        case 1: hipLaunchKernelGGL(kp_smooth_res<1>, dim3(nb), dim3(BB), 0, s, F, plev, res, S, amg_off, gate); break;
        case 2: hipLaunchKernelGGL(kp_smooth_res<2>, dim3(nb), dim3(BB), 0, s, F, plev, res, S, amg_off, gate); break;
        case 4: hipLaunchKernelGGL(kp_smooth_res<4>, dim3(nb), dim3(BB), 0, s, F, plev, res, S, amg_off, gate); break;
        default: hipLaunchKernelGGL(kp_smooth_res<8>, dim3(nb), dim3(BB), 0, s, F, plev, res, S, amg_off, gate); break;
    }
}

template <bool YIN>
void launch_up_y(int lpr, const LevelView &F, const LevelView &C, const double *y, const int *plev, const double *S,
                 const int *amg_off, int gate, hipStream_t s) {
    const unsigned nb = nblocks(F.n * lpr, BB);
    switch (lpr) {
        case 1: hipLaunchKernelGGL((kp_up<1, YIN>), dim3(nb), dim3(BB), 0, s, F, C, y, plev, S, amg_off, gate); break;
        case 2: hipLaunchKernelGGL((kp_up<2, YIN>), dim3(nb), dim3(BB), 0, s, F, C, y, plev, S, amg_off, gate); break;
        case 4: hipLaunchKernelGGL((kp_up<4, YIN>), dim3(nb), dim3(BB), 0, s, F, C, y, plev, S, amg_off, gate); break;
        default: hipLaunchKernelGGL((kp_up<8, YIN>), dim3(nb), dim3(BB), 0, s, F, C, y, plev, S, amg_off, gate); break;
    }
}

void launch_up(int lpr, bool yin, const LevelView &F, const LevelView &C, const double *y, const int *plev,
               const double *S, const int *amg_off, int gate, hipStream_t s) {
    if (yin) launch_up_y<true>(lpr, F, C, y, plev, S, amg_off, gate, s);
    else launch_up_y<false>(lpr, F, C, y, plev, S, amg_off, gate, s);
}

}  // namespace

struct PcgSolver::Impl {
    RowPlan plan;
    bool v32 = false;                      // FP32 copies of the matrix values for the smoothers
    int v32w = 0;                          // level 0's copy: entries per row (the ELL capacity)
    DevBuf<float> val32_0;                 // level 0's copy, slot-major (slot e of row t at e * N + t)
    hipStream_t s = nullptr;
    int64_t K = 0, N = 0, nblk = 0;
    std::vector<int64_t> off, blk_off;
    DevBuf<int64_t> doff, dblk;
    DevBuf<int> blk_prob, row_prob;
    DevBuf<double> r, z, p, q, red, S, xpre0, res;
    // multigrid: levels 1..L (level 0 is the caller's matrix)
    struct Level {
        int64_t n = 0, base = 0;
        std::vector<int64_t> off;
        DevBuf<int64_t> doff;
        DevBuf<int> prob, len, col;
        DevBuf<double> diag, val;
        DevBuf<float> val32;
    };
    std::vector<Level> lv;                 // lv[0] unused placeholder for level 0
    std::vector<int> plev;                 // depth (coarsest level) of each problem
    DevBuf<int> dplev, dcn, amg_off;
    DevBuf<int64_t> dcoff, dcbase;         // coarsest rows: offset within their level, offset in vr/vout
    DevBuf<double> vr, vxpre, vout, inv;
    int L = 0;                             // deepest level
    hipGraphExec_t graph = nullptr;
    bool graph_amg = false;
    PcgMat graph_A{};
    const double *graph_x = nullptr;
    int graph_max_iter = 0;
    hipEvent_t gev[2] = {};
    void drop_graph() {
        if (!graph) return;
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    Blocks blocks() const { return Blocks{blk_prob.p, dblk.p, doff.p}; }
};

PcgSolver::PcgSolver() : im(new Impl) {
    for (auto &e : im->gev) HIPCK(hipEventCreate(&e));
}

PcgSolver::~PcgSolver() {
    im->drop_graph();
    for (auto &e : im->gev)
        if (e) (void)hipEventDestroy(e);
    for (auto *b : {&im->doff, &im->dblk, &im->dcoff, &im->dcbase}) b->release();
    for (auto *b : {&im->blk_prob, &im->row_prob, &im->dplev, &im->dcn, &im->amg_off}) b->release();
    for (auto *b : {&im->r, &im->z, &im->p, &im->q, &im->red, &im->S, &im->xpre0, &im->res, &im->vr, &im->vxpre, &im->vout, &im->inv})
        b->release();
    for (auto &l : im->lv) {
        l.doff.release();
        l.prob.release();
        l.len.release();
        l.col.release();
        l.diag.release();
        l.val.release();
        l.val32.release();
    }
    im->val32_0.release();
    delete im;
}

void PcgSolver::invalidate() { im->drop_graph(); }

void PcgSolver::set_options(bool use_amg, bool use_graph) {
    if (use_amg != use_amg_ || use_graph != use_graph_) im->drop_graph();
    use_amg_ = use_amg;
    use_graph_ = use_graph;
}

void PcgSolver::set_layout(const std::vector<int64_t> &off, hipStream_t s) {
    Impl &I = *im;
    I.drop_graph();
    I.s = s;
    I.off = off;
    I.K = (int64_t)off.size() - 1;
    I.N = off.back();
    const int64_t K = I.K, N = I.N;
    I.blk_off.assign(K + 1, 0);
    for (int64_t k = 0; k < K; ++k) I.blk_off[k + 1] = I.blk_off[k] + (off[k + 1] - off[k] + BB - 1) / BB;
    I.nblk = I.blk_off[K];
    I.doff.ensure(K + 1);
    I.dblk.ensure(K + 1);
    HIPCK(hipMemcpyAsync(I.doff.p, off.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(I.dblk.p, I.blk_off.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    I.blk_prob.ensure(I.nblk);
    I.row_prob.ensure(N);
    hipLaunchKernelGGL(kp_owner, dim3(nblocks(I.nblk, BB)), dim3(BB), 0, s, I.dblk.p, K, I.nblk, I.blk_prob.p);
    hipLaunchKernelGGL(kp_owner, dim3(nblocks(N, BB)), dim3(BB), 0, s, I.doff.p, K, N, I.row_prob.p);
    HIPCK(hipGetLastError());
    for (auto *b : {&I.r, &I.z, &I.p, &I.q, &I.xpre0, &I.res}) b->ensure(N);
    I.red.ensure(2 * I.nblk + 16);
    I.S.ensure(K * S_NS);
    // multigrid levels: problem sizes shrink by AGG until at most CMIN rows
    std::vector<std::vector<int64_t>> n(1, std::vector<int64_t>(K));
    for (int64_t k = 0; k < K; ++k) n[0][k] = off[k + 1] - off[k];
    I.plev.assign(K, 0);
    for (int l = 0;; ++l) {
        bool any = false;
        std::vector<int64_t> nx(K, 0);
        for (int64_t k = 0; k < K; ++k)
            if (n[l][k] > CMIN) {
                nx[k] = (n[l][k] + AGG - 1) / AGG;
                I.plev[k] = l + 1;
                any = true;
            }
        if (!any) break;
        n.push_back(nx);
    }
    I.L = (int)n.size() - 1;
    if ((int)I.lv.size() < I.L + 1) I.lv.resize(I.L + 1);
    int64_t base = 0;
    for (int l = 1; l <= I.L; ++l) {
        Impl::Level &V = I.lv[l];
        V.off.assign(K + 1, 0);
        for (int64_t k = 0; k < K; ++k) V.off[k + 1] = V.off[k] + n[l][k];
        V.n = V.off[K];
        V.base = base;
        base += V.n;
        V.doff.ensure(K + 1);
        HIPCK(hipMemcpyAsync(V.doff.p, V.off.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
        V.prob.ensure(V.n);
        hipLaunchKernelGGL(kp_owner, dim3(nblocks(V.n, BB)), dim3(BB), 0, s, V.doff.p, K, V.n, V.prob.p);
        HIPCK(hipGetLastError());
        V.diag.ensure(V.n);
        V.len.ensure(V.n);
        V.col.ensure((size_t)CCAP * V.n);
        V.val.ensure((size_t)CCAP * V.n);
        if (I.v32) V.val32.ensure((size_t)CCAP * V.n);
    }
    if (I.v32 && I.L > 0) I.val32_0.ensure((size_t)I.v32w * N);
    for (auto *b : {&I.vr, &I.vxpre, &I.vout}) b->ensure(base);
    // coarsest rows of each problem
    std::vector<int64_t> coff(K), cbase(K);
    std::vector<int> cn(K);
    for (int64_t k = 0; k < K; ++k) {
        const int l = I.plev[k];
        cn[k] = (int)n[l][k];
        coff[k] = l == 0 ? off[k] : I.lv[l].off[k];
        cbase[k] = l == 0 ? off[k] : I.lv[l].base + I.lv[l].off[k];
    }
    I.dplev.ensure(K);
    I.dcn.ensure(K);
    I.amg_off.ensure(K);
    I.dcoff.ensure(K);
    I.dcbase.ensure(K);
    HIPCK(hipMemcpyAsync(I.dplev.p, I.plev.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(I.dcn.p, cn.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(I.dcoff.p, coff.data(), K * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(I.dcbase.p, cbase.data(), K * sizeof(int64_t), hipMemcpyHostToDevice, s));
    I.inv.ensure((size_t)K * CMIN * CMIN);
    HIPCK(hipStreamSynchronize(s));
}

namespace {

LevelView level_view(PcgSolver::Impl &I, const PcgMat &A0, int l, double *r0, double *z0) {
    LevelView v;
    v.lev = l;
    if (l == 0) {
        v.A = A0;
        // the FP32 copy, when it holds the caller's whole layout (ELL stride = row count)
        const bool c32 = I.v32 && A0.cap > 0 && A0.cap <= I.v32w && A0.stride == I.N &&
                         I.val32_0.n >= (size_t)I.v32w * I.N;
        v.A.val32 = c32 ? I.val32_0.p : nullptr;
        v.off = I.doff.p;
        v.prob = I.row_prob.p;
        v.r = r0;
        v.xpre = I.xpre0.p;
        v.out = z0;
        v.n = I.N;
    } else {
        PcgSolver::Impl::Level &V = I.lv[l];
        v.A = PcgMat{V.diag.p, V.len.p, V.col.p, V.val.p, V.n, CCAP, I.v32 ? V.val32.p : nullptr};
        v.off = V.doff.p;
        v.prob = V.prob.p;
        v.r = I.vr.p + V.base;
        v.xpre = I.vxpre.p + V.base;
        v.out = I.vout.p + V.base;
        v.n = V.n;
    }
    return v;
}

// z = M r for every problem not done; rz partials per block; p = z when init
void apply_prec(PcgSolver::Impl &I, const PcgMat &A0, bool amg, bool init) {
    hipStream_t s = I.s;
    const int L = amg ? I.L : 0;
    const int gate = init ? 1 : 0;   // the first application runs for the problems that start
    for (int l = 0; l < L; ++l) {
        LevelView F = level_view(I, A0, l, I.r.p, I.z.p), C = level_view(I, A0, l + 1, I.r.p, I.z.p);
        if (C.n == 0) continue;
        launch_smooth_res(l == 0 ? 1 : I.plan.lpr_coarse, F, I.dplev.p, I.res.p, I.S.p, I.amg_off.p, gate, s);
        hipLaunchKernelGGL(kp_restrict, dim3(nblocks(C.n, BB)), dim3(BB), 0, s, F, C, I.res.p, I.S.p, I.amg_off.p,
                           gate);
    }
    if (amg)
        hipLaunchKernelGGL(kp_coarse_solve, dim3((unsigned)I.K), dim3(CMIN), 0, s, I.dcbase.p, I.dcn.p, I.dplev.p,
                           I.inv.p, I.vr.p, I.vout.p, I.r.p, I.z.p, I.S.p, I.amg_off.p, gate);
    // the restriction's residual buffer is free again: it can hold each level's prolonged iterate
    const RowPlan &rp = I.plan;
    for (int l = L - 1; l >= 1; --l) {
        LevelView F = level_view(I, A0, l, I.r.p, I.z.p), C = level_view(I, A0, l + 1, I.r.p, I.z.p);
        if (F.n == 0) continue;
        if (rp.prolong_coarse)
            hipLaunchKernelGGL(kp_prolong, dim3(nblocks(F.n, BB)), dim3(BB), 0, s, F, C, I.dplev.p, I.res.p, I.S.p,
                               I.amg_off.p, gate, 0);
        launch_up(rp.lpr_coarse, rp.prolong_coarse, F, C, I.res.p, I.dplev.p, I.S.p, I.amg_off.p, gate, s);
    }
    LevelView F = level_view(I, A0, 0, I.r.p, I.z.p);
    LevelView C = L >= 1 ? level_view(I, A0, 1, I.r.p, I.z.p) : F;
    if (L >= 1 && rp.prolong0) {
        hipLaunchKernelGGL(kp_prolong, dim3(nblocks(F.n, BB)), dim3(BB), 0, s, F, C, I.dplev.p, I.res.p, I.S.p,
                           I.amg_off.p, gate, 0);
        hipLaunchKernelGGL(kp_up0<true>, dim3((unsigned)I.nblk), dim3(BB), 0, s, F, C, I.res.p, I.dplev.p, I.blocks(),
                           I.red.p, I.p.p, init ? 1 : 0, amg ? 0 : 1, I.S.p, I.amg_off.p);
    } else {
        hipLaunchKernelGGL(kp_up0<false>, dim3((unsigned)I.nblk), dim3(BB), 0, s, F, C, I.res.p, I.dplev.p,
                           I.blocks(), I.red.p, I.p.p, init ? 1 : 0, amg ? 0 : 1, I.S.p, I.amg_off.p);
    }
}

void cg_iteration(PcgSolver::Impl &I, const PcgMat &A, double *x, bool amg, int max_iter) {
    hipStream_t s = I.s;
    const unsigned nb = (unsigned)I.nblk, K = (unsigned)I.K;
    const Blocks bl = I.blocks();
    hipLaunchKernelGGL(kp_spmv_dot, dim3(nb), dim3(BB), 0, s, A, bl, I.p.p, I.q.p, I.red.p, I.S.p);
    hipLaunchKernelGGL(kp_alpha, dim3(K), dim3(BB), 0, s, I.red.p, I.dblk.p, I.S.p);
    hipLaunchKernelGGL(kp_axpy, dim3(nb), dim3(BB), 0, s, bl, I.p.p, I.q.p, x, I.r.p, I.red.p, I.S.p, A.diag,
                       I.xpre0.p);
    apply_prec(I, A, amg, false);
    hipLaunchKernelGGL(kp_beta, dim3(K), dim3(BB), 0, s, I.red.p, I.dblk.p, I.S.p, max_iter);
    hipLaunchKernelGGL(kp_pupdate, dim3(nb), dim3(BB), 0, s, bl, I.z.p, I.p.p, I.S.p);
}

}  // namespace

// hierarchy of the starting problems' matrices, then their first iterates
void PcgSolver::enqueue_start(const PcgMat &A, const double *b, double *x, const double *d_err) {
    Impl &I = *im;
    hipStream_t s = I.s;
    const int64_t K = I.K;
    const bool amg = use_amg_ && I.L > 0;
    if (amg) {
        for (int l = 0; l < I.L; ++l) {
            LevelView F = level_view(I, A, l, nullptr, nullptr);
            Impl::Level &V = I.lv[l + 1];
            if (V.n > 0)
                hipLaunchKernelGGL(kp_coarsen, dim3(nblocks(V.n, CGRP)), dim3(256), 0, s, F.A, F.off, V.doff.p,
                                   V.prob.p, V.n, V.diag.p, V.len.p, V.col.p, V.val.p, I.amg_off.p, I.S.p);
        }
        // dense inverses: each problem's coarsest matrix lives at its own level
        for (int l = 0; l <= I.L; ++l) {
            LevelView F = level_view(I, A, l, nullptr, nullptr);
            hipLaunchKernelGGL(kp_coarse_inv, dim3((unsigned)K), dim3(CMIN), 0, s, F.A, I.dcoff.p, I.dcn.p, I.dplev.p, l,
                               I.inv.p, I.amg_off.p, I.S.p);
        }
    }
    hipLaunchKernelGGL(kp_init, dim3((unsigned)I.nblk), dim3(BB), 0, s, I.blocks(), b, x, I.r.p, I.red.p, I.S.p, A.diag,
                       I.xpre0.p);
    apply_prec(I, A, amg, true);
    hipLaunchKernelGGL(kp_init_red, dim3((unsigned)K), dim3(BB), 0, s, I.red.p, I.dblk.p, d_err, I.S.p);
    HIPCK(hipGetLastError());
}

void PcgSolver::enqueue_iterations(const PcgMat &A, double *x, int n, int max_iter) {
    const bool amg = use_amg_ && im->L > 0;
    for (int k = 0; k < n; ++k) cg_iteration(*im, A, x, amg, max_iter);
    HIPCK(hipGetLastError());
}

double *PcgSolver::state() { return im->S.p; }

void PcgSolver::set_dim(int dim) {
    // measured on the 3D 1M solve and the 512-problem 2D batch: 3D coarse rows (aggregates
    // of 4 neighbouring cells, up to 64 entries) want 8 lanes and the prolonged iterate
    // formed once per row (CG 0.27 -> 0.20 s per solve); 2D rows are short (about 6 entries
    // on level 0), where one lane per row and the iterate formed per entry are fastest
    RowPlan &rp = im->plan;
    if (dim == 3) {
        rp.lpr_coarse = 8;
        rp.prolong0 = rp.prolong_coarse = true;
    } else {
        rp.lpr_coarse = 1;
        rp.prolong0 = rp.prolong_coarse = false;
    }
    // the smoothers read FP32 copies of the values; level 0's copy is slot-major like the
    // ELL arrays, so a row moves only its used slots (its allocation is the row capacity:
    // 128 bytes per row in 2D, 512 in 3D)
    im->v32 = true;
    im->v32w = dim == 2 ? ELL_CAP2 : ELL_CAP3;
    im->drop_graph();
}
int PcgSolver::levels() const { return use_amg_ && im->L > 0 ? im->L + 1 : 0; }
int *PcgSolver::fallback() { return im->amg_off.p; }

void PcgSolver::solve(const PcgMat &A, const double *b, double *x, const double *d_err, const std::vector<int> &active,
                      int max_iter, std::vector<int> &iters, PcgStats &st) {
    Impl &I = *im;
    hipStream_t s = I.s;
    const int64_t K = I.K;
    std::vector<double> S(K * S_NS, 0.0);
    for (int64_t k = 0; k < K; ++k) {
        S[k * S_NS + S_DONE] = active[k] ? 0.0 : 1.0;
        S[k * S_NS + S_START] = active[k] ? 1.0 : 0.0;
    }
    HIPCK(hipMemcpyAsync(I.S.p, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, s));
    fill_u32(reinterpret_cast<uint32_t *>(I.amg_off.p), K, 0u, s);
    const bool amg = use_amg_ && I.L > 0;
    const auto t0 = std::chrono::steady_clock::now();
    enqueue_start(A, b, x, d_err);
    st.ms_setup += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st.amg_levels = std::max(st.amg_levels, amg ? I.L + 1 : 0);
    const bool same_args = I.graph_A.diag == A.diag && I.graph_A.len == A.len && I.graph_A.col == A.col &&
                           I.graph_A.val == A.val && I.graph_A.stride == A.stride && I.graph_x == x &&
                           I.graph_max_iter == max_iter && I.graph_amg == amg;
    if (use_graph_ && (!I.graph || !same_args)) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());   // no device free while recording
        I.drop_graph();
        hipGraph_t g = nullptr;
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < CHUNK; ++k) cg_iteration(I, A, x, amg, max_iter);
        const hipError_t ec = hipStreamEndCapture(s, &g);
        HIPCK(ec);
        const hipError_t e = hipGraphInstantiate(&I.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCK(e);
        I.graph_amg = amg;
        I.graph_A = A;
        I.graph_x = x;
        I.graph_max_iter = max_iter;
    }
    int *dcount = reinterpret_cast<int *>(I.red.p + 2 * I.nblk + 8);
    int act = 1;
    for (int it = 0; it < max_iter && act > 0; it += CHUNK) {
        if (use_graph_) {
            HIPCK(hipEventRecord(I.gev[0], s));
            HIPCK(hipGraphLaunch(I.graph, s));
            HIPCK(hipEventRecord(I.gev[1], s));
            ++st.graph_replays;
            st.graph_iterations += CHUNK;
        } else {
            for (int k = 0; k < CHUNK; ++k) cg_iteration(I, A, x, amg, max_iter);
        }
        hipLaunchKernelGGL(kp_count_active, dim3(1), dim3(BB), 0, s, I.S.p, K, dcount);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(&act, dcount, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        float ms = 0;
        if (use_graph_ && hipEventElapsedTime(&ms, I.gev[0], I.gev[1]) == hipSuccess) st.ms_graph += ms;
        // SDOT_TRACE=1: one line per graph replay (its device time, the problems still iterating)
        static const bool trace = getenv("SDOT_TRACE") && getenv("SDOT_TRACE")[0] == '1';
        if (trace && use_graph_)
            fprintf(stderr, "[sdot] pcg replay %d (%d iterations): %.3f ms, %d of %lld problems still iterating\n",
                    it / CHUNK, CHUNK, ms, act, (long long)K);
    }
    HIPCK(hipMemcpyAsync(S.data(), I.S.p, S.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    std::vector<int> off_flags(K);
    HIPCK(hipMemcpyAsync(off_flags.data(), I.amg_off.p, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    iters.assign(K, 0);
    for (int64_t k = 0; k < K; ++k) {
        iters[k] = (int)S[k * S_NS + S_IT];
        st.iters_max = std::max(st.iters_max, iters[k]);
        if (amg && active[k] && off_flags[k]) ++st.amg_fallbacks;
    }
}

}  // namespace sdot_amd
