// Batched solves: many independent 2D problems (sizes may differ) in one set of
// device arrays, advanced together by one launch per kernel.
//
//  * grid: each problem gets its own uniform grid (its own bounding box and cell
//    size) in a shared code space; the sort key is code_base[p] + Morton code, so
//    one counting sort groups the diracs by problem, then by grid cell, and
//    problem p keeps the index range [off[p], off[p+1]) in grid order;
//  * cells: the 2D cell kernel reads each cell's problem view (grid, domain box,
//    max weight) and skips the problems that have converged;
//  * Newton / PCG: Jacobi-PCG on problem-aligned blocks of rows; every reduction
//    is per problem, over that problem's blocks in a fixed order, and the
//    tie-breaking weight offsets use the index inside the problem, so a problem's
//    result does not depend on the other problems of the batch (bitwise);
//  * the PCG iterations are captured once in a hipGraph (a fixed number of
//    iterations per launch; converged problems turn into no-ops on the device)
//    and replayed until every problem has converged;
//  * the first solve of a size class picks the grid resolution (diracs per grid
//    cell) by timing one pass per candidate; the choice is cached per size class.
// The Newton rules are those of OptimalTransportSolver::solve
// (src/sdot/Solvers/OptimalTransportSolver.tcc:27-119), applied per problem.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <new>

#include "internal.h"

using namespace sdot_amd;

struct sdot_batch {
    sdot_pd *pd = nullptr;      // holds the concatenated dirac arrays, the pass outputs and the stream
    int ball = 0;
    int64_t K = 0;
    int use_graph = 1;
    int dpc_x100 = 0;           // 0: autotune on the first solve of a size class
    std::vector<int64_t> off;   // problem offsets (K+1)
    std::vector<double> hpos;   // host copy of the positions (per-problem bounding boxes)
    std::vector<GridInfo> grids;
    std::vector<int64_t> code_base;
    std::vector<int64_t> blk_off;
    int64_t nblk = 0;           // problem-aligned row blocks
    int grid_dpc = -1;          // resolution of the current grid (-1: none)
    DevBuf<int> prob, blk_prob, pdone, act;
    DevBuf<int64_t> dcode_base, doff, dblk;
    DevBuf<GridInfo> dgrids;
    DevBuf<double> pdom, tie, w0, tgt, tsort, old, b, red, S, perr;
    DevBuf<double> x, r, z, p, q;
    DevBuf<unsigned> pflags;
    DevBuf<unsigned long long> wglob;
    hipGraphExec_t graph = nullptr;
    hipEvent_t gev[2] = {};
    std::string err;
};

namespace {

constexpr int BB = 256;         // rows per block
constexpr int CHUNK = 8;        // PCG iterations per graph launch
enum { S_RZ = 0, S_BB, S_ALPHA, S_BETA, S_RR, S_DONE, S_IT, S_RTOL2, S_NS };
enum { ACT_NONE = 0, ACT_ACCEPT = 1, ACT_BACKTRACK = 2 };

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
__device__ inline double bmax(double v, double *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = BB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    const double r = sm[0];
    __syncthreads();
    return r;
}

struct BMat {                   // ELL Hessian rows in grid order (column-major slots)
    const double *diag;
    const int *len;
    const int *col;
    const double *val;
    int64_t stride;
};

struct Blocks {                 // problem-aligned row blocks
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

// ---- grid -----------------------------------------------------------------
__global__ __launch_bounds__(BB) void kb_keys(const double *pos, const int *prob, int64_t n, const GridInfo *grids,
                                              const int64_t *code_base, uint32_t *keys, uint32_t *rank, uint32_t *cnt) {
    const int64_t i = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (i >= n) return;
    const int p = prob[i];
    const GridInfo g = grids[p];
    const uint32_t code =
        (uint32_t)code_base[p] + morton2(grid_coord(g, 0, pos[2 * i]), grid_coord(g, 1, pos[2 * i + 1]));
    keys[i] = code;
    rank[i] = atomicAdd(&cnt[code], 1u);
}

// sorted weights plus the tie-breaking offset of each problem (hash of the index
// inside the problem, scaled by 1e-12 of the problem's squared mean spacing)
__global__ __launch_bounds__(BB) void kb_gather_w(const double *w, const int *perm, const int *prob, const int64_t *off,
                                                  const double *tie, int64_t n, double *sw, unsigned *flags) {
    const int64_t t = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (t >= n) return;
    const int i = perm[t];
    const int p = prob[i];
    const double v = w[i];
    if (!isfinite(v)) atomicOr(flags, FLAG_BAD_INPUT);
    uint64_t z = (uint64_t)(i - off[p]) * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    sw[t] = v + tie[p] * ((double)(z >> 11) * 0x1.0p-53);
}

__global__ __launch_bounds__(BB) void kb_cell_wmax(const uint32_t *start, int64_t ncodes, const double *sw,
                                                   double *cell_wmax) {
    const int64_t c = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (c >= ncodes) return;
    double m = -DBL_MAX;
    for (uint32_t k = start[c]; k < start[c + 1]; ++k) m = fmax(m, sw[k]);
    cell_wmax[c] = m;
}

// per-problem max weight (ordered bits), one block per problem
__global__ __launch_bounds__(BB) void kb_wmax(const double *sw, const int64_t *off, unsigned long long *wglob) {
    __shared__ double sm[BB];
    const int p = blockIdx.x;
    double m = -DBL_MAX;
    for (int64_t t = off[p] + threadIdx.x; t < off[p + 1]; t += BB) m = fmax(m, sw[t]);
    m = bmax(m, sm);
    if (threadIdx.x == 0) wglob[p] = dbl_order(m);
}

__global__ void kb_gather(const double *src, const int *perm, int64_t n, double *dst) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[perm[t]];
}

// ---- Newton control ------------------------------------------------------------
// b = m - target (grid order) and the per-block max of |b| / target
__global__ __launch_bounds__(BB) void kb_residual(Blocks B, const double *m, const double *tsort, double *b,
                                                  double *part, const int *pdone) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    double mx = 0;
    if (!pdone[br.p])
        for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
            const double v = m[t] - tsort[t];
            b[t] = v;
            mx = fmax(mx, fabs(v) / tsort[t]);
        }
    mx = bmax(mx, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

// per-problem max over the problem's blocks
__global__ __launch_bounds__(BB) void kb_pmax(const double *part, const int64_t *blk_off, double *out) {
    __shared__ double sm[BB];
    const int p = blockIdx.x;
    double mx = 0;
    for (int64_t k = blk_off[p] + threadIdx.x; k < blk_off[p + 1]; k += BB) mx = fmax(mx, part[k]);
    mx = bmax(mx, sm);
    if (threadIdx.x == 0) out[p] = mx;
}

// accept (old = w) or backtrack (w = 0.9 old + 0.1 w, OptimalTransportSolver.tcc:72-79), dirac order
__global__ __launch_bounds__(BB) void kb_accept(const int *prob, int64_t n, const int *act, double *w, double *old) {
    const int64_t i = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (i >= n) return;
    const int a = act[prob[i]];
    if (a == ACT_BACKTRACK) w[i] = 0.9 * old[i] + 0.1 * w[i];
    else if (a == ACT_ACCEPT) old[i] = w[i];
}

// gauge fix of every problem that solves now: M00 *= 2 on the problem's first dirac (:62-63)
__global__ void kb_gauge(double *diag, const int *inv_perm, const int64_t *off, const int *act, int64_t K) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= K || act[p] != ACT_ACCEPT) return;
    diag[inv_perm[off[p]]] *= 2;
}

// w[perm[t]] -= x[t] for the problems that solved now; per-block max |x|
__global__ __launch_bounds__(BB) void kb_update(Blocks B, const double *x, const int *perm, double *w, const int *act,
                                                double *part) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    double mx = 0;
    if (act[br.p] == ACT_ACCEPT)
        for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
            w[perm[t]] -= x[t];
            mx = fmax(mx, fabs(x[t]));
        }
    mx = bmax(mx, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

__global__ void kb_export(const double *sorted, const int *inv_perm, int64_t n, double *out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = sorted[inv_perm[i]];
}

// ---- PCG on problem-aligned blocks --------------------------------------------
// S[p * S_NS + k]: per-problem scalars; S_DONE != 0 freezes the problem.
__global__ __launch_bounds__(BB) void kb_init(BMat A, Blocks B, const double *b, double *x, double *r, double *z,
                                              double *p, double *part, const double *S) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    double rz = 0, bb = 0;
    if (S[br.p * S_NS + S_DONE] == 0)
        for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
            const double d = A.diag[t], rv = b[t], zv = d != 0 ? rv / d : rv;
            x[t] = 0;
            r[t] = rv;
            z[t] = zv;
            p[t] = zv;
            rz += rv * zv;
            bb += rv * rv;
        }
    rz = bsum(rz, sm);
    bb = bsum(bb, sm);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = rz;
        part[2 * blockIdx.x + 1] = bb;
    }
}

// rtol per problem: max(1e-14, min(1e-3, 1e-3 * relative mass error))
__global__ __launch_bounds__(BB) void kb_init_red(const double *part, const int64_t *blk_off, const double *perr,
                                                  double *S) {
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
        const double rtol = fmax(1e-14, fmin(1e-3, 1e-3 * perr[p]));
        sp[S_RZ] = a;
        sp[S_BB] = c;
        sp[S_RR] = c;
        sp[S_RTOL2] = rtol * rtol;
        sp[S_IT] = 0;
        sp[S_DONE] = (c == 0 || !(a > 0)) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(BB) void kb_spmv_dot(BMat A, Blocks B, const double *p, double *q, double *part,
                                                  const double *S) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    if (S[br.p * S_NS + S_DONE] != 0) return;   // uniform per block
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

__global__ __launch_bounds__(BB) void kb_alpha(const double *part, const int64_t *blk_off, double *S) {
    __shared__ double sm[BB];
    const int p = blockIdx.x;
    double *sp = S + p * S_NS;
    if (sp[S_DONE] != 0) return;
    double a = 0;
    for (int64_t k = blk_off[p] + threadIdx.x; k < blk_off[p + 1]; k += BB) a += part[2 * k];
    a = bsum(a, sm);
    if (threadIdx.x == 0) sp[S_ALPHA] = a != 0 ? sp[S_RZ] / a : 0.0;
}

__global__ __launch_bounds__(BB) void kb_axpy(BMat A, Blocks B, const double *p, const double *q, double *x, double *r,
                                              double *z, double *part, const double *S) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    const double *sp = S + br.p * S_NS;
    if (sp[S_DONE] != 0) return;
    const double alpha = sp[S_ALPHA];
    double rz = 0, rr = 0;
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
        x[t] += alpha * p[t];
        const double rv = r[t] - alpha * q[t];
        r[t] = rv;
        const double d = A.diag[t], zv = d != 0 ? rv / d : rv;
        z[t] = zv;
        rz += rv * zv;
        rr += rv * rv;
    }
    rz = bsum(rz, sm);
    rr = bsum(rr, sm);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = rz;
        part[2 * blockIdx.x + 1] = rr;
    }
}

__global__ __launch_bounds__(BB) void kb_beta(const double *part, const int64_t *blk_off, double *S, int max_iter) {
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

__global__ __launch_bounds__(BB) void kb_pupdate(Blocks B, const double *z, double *p, const double *S) {
    const BlockRows br = block_rows(B, blockIdx.x);
    const double *sp = S + br.p * S_NS;
    if (sp[S_DONE] != 0) return;
    const double beta = sp[S_BETA];
    for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) p[t] = z[t] + beta * p[t];
}

__global__ __launch_bounds__(BB) void kb_count_active(const double *S, int64_t K, int *out) {
    __shared__ double sm[BB];
    double c = 0;
    for (int64_t p = threadIdx.x; p < K; p += BB) c += S[p * S_NS + S_DONE] == 0 ? 1.0 : 0.0;
    c = bsum(c, sm);
    if (threadIdx.x == 0) *out = (int)c;
}

std::mutex tune_mutex;
std::map<int, int> tune_cache;  // size class -> diracs per grid cell x100
constexpr int TUNE_CANDIDATES[] = {150, 250, 400};

void drop_graph(sdot_batch *B) {
    if (B->graph) (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

Blocks blocks(sdot_batch *B) { return Blocks{B->blk_prob.p, B->dblk.p, B->doff.p}; }

// per-problem grids in one code space, the counting sort, sorted positions
void batch_grid(sdot_batch *B, int dpc_x100) {
    sdot_pd *pd = B->pd;
    const int64_t K = B->K, N = pd->N;
    hipStream_t s = pd->stream;
    B->grids.resize(K);
    B->code_base.assign(K + 1, 0);
    std::vector<double> tie(K);
    for (int64_t p = 0; p < K; ++p) {
        double lo[2] = {DBL_MAX, DBL_MAX}, hi[2] = {-DBL_MAX, -DBL_MAX};
        for (int64_t i = B->off[p]; i < B->off[p + 1]; ++i)
            for (int d = 0; d < 2; ++d) {
                lo[d] = std::min(lo[d], B->hpos[2 * i + d]);
                hi[d] = std::max(hi[d], B->hpos[2 * i + d]);
            }
        const int64_t np = B->off[p + 1] - B->off[p];
        const GridInfo g = grid_params(2, lo, hi, np, dpc_x100 / 100.0);
        B->grids[p] = g;
        B->code_base[p + 1] = B->code_base[p] + g.ncodes;
        // same scale as the single-problem path: 1e-12 of the squared mean spacing
        tie[p] = 1e-12 * (g.h[0] * g.ncell[0]) * (g.h[1] * g.ncell[1]) / (double)np;
    }
    const int64_t nc = B->code_base[K];
    if (nc >= (int64_t)0xFFFFFFFF) throw StatusError{SDOT_CAPACITY_OVERFLOW, "batch grid code space exceeds 32 bits"};
    B->dgrids.ensure(K);
    B->dcode_base.ensure(K + 1);
    B->tie.ensure(K);
    HIPCK(hipMemcpyAsync(B->dgrids.p, B->grids.data(), K * sizeof(GridInfo), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(B->dcode_base.p, B->code_base.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(B->tie.p, tie.data(), K * sizeof(double), hipMemcpyHostToDevice, s));
    pd->keys.ensure(N);
    pd->rank.ensure(N);
    pd->cell_start.ensure(nc + 1);
    pd->scan_tmp.ensure(nc + scan_tmp_size(nc));
    pd->cell_wmax.ensure(nc);
    pd->perm.ensure(N);
    pd->inv_perm.ensure(N);
    pd->spos.ensure(2 * N);
    pd->sw.ensure(N);
    fill_u32(pd->scan_tmp.p, nc, 0u, s);
    hipLaunchKernelGGL(kb_keys, dim3(nblocks(N, BB)), dim3(BB), 0, s, pd->pos.p, B->prob.p, N, B->dgrids.p,
                       B->dcode_base.p, pd->keys.p, pd->rank.p, pd->scan_tmp.p);
    HIPCK(hipGetLastError());
    sort_by_keys(pd, nc);
    // the handle's own grid descriptor covers the whole code space (the kernels read the per-problem grids)
    pd->g = B->grids[0];
    pd->g.ncodes = nc;
    // targets in grid order
    hipLaunchKernelGGL(kb_gather, dim3(nblocks(N, 256)), dim3(256), 0, s, B->tgt.p, pd->perm.p, N, B->tsort.p);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(s));
    pd->last.grid_cells = nc;
    B->grid_dpc = dpc_x100;
}

// one cell pass over the problems not marked done; returns the device time (ms)
double batch_pass(sdot_batch *B) {
    sdot_pd *pd = B->pd;
    const int64_t K = B->K, N = pd->N, nc = pd->g.ncodes;
    hipStream_t s = pd->stream;
    HIPCK(hipEventRecord(pd->ev[0], s));
    fill_u32(pd->flags.p, 2, 0u, s);
    fill_u32(B->pflags.p, K, 0u, s);
    hipLaunchKernelGGL(kb_gather_w, dim3(nblocks(N, BB)), dim3(BB), 0, s, pd->w.p, pd->perm.p, B->prob.p, B->doff.p,
                       B->tie.p, N, pd->sw.p, pd->flags.p);
    hipLaunchKernelGGL(kb_cell_wmax, dim3(nblocks(nc, BB)), dim3(BB), 0, s, pd->cell_start.p, nc, pd->sw.p,
                       pd->cell_wmax.p);
    hipLaunchKernelGGL(kb_wmax, dim3((unsigned)K), dim3(BB), 0, s, pd->sw.p, B->doff.p, B->wglob.p);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(pd->ev[1], s));
    CellArgs a{};
    a.N = N;
    a.spos = pd->spos.p;
    a.sw = pd->sw.p;
    a.cell_start = pd->cell_start.p;
    a.cell_wmax = pd->cell_wmax.p;
    a.wglob = B->wglob.p;
    a.g = pd->g;
    a.coeff = 1;
    a.mass = pd->mass.p;
    a.arc = pd->arc.p;
    a.perim = pd->perim.p;
    a.diag = pd->diag.p;
    a.row_len = pd->row_len.p;
    a.ell_col = pd->ell_col.p;
    a.ell_val = pd->ell_val.p;
    a.ell_cap = pd->ell_cap;
    a.ell_stride = N;
    a.flags = pd->flags.p;
    a.ovf_list = pd->ovf_list.p;
    a.ovf_count = pd->flags.p + 1;
    a.prob = B->prob.p;
    a.grids = B->dgrids.p;
    a.code_base = B->dcode_base.p;
    a.pdom = B->pdom.p;
    a.pflags = B->pflags.p;
    a.pdone = B->pdone.p;
    run_cells2(pd, a);   // synchronises the stream
    HIPCK(hipEventRecord(pd->ev[2], s));
    unsigned fl = 0;
    HIPCK(hipMemcpyAsync(&fl, pd->flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (fl & FLAG_BAD_INPUT) throw StatusError{SDOT_INVALID_ARG, "non-finite weight or coincident diracs in the batch"};
    if (fl & FLAG_ROW_OVERFLOW) {
        if (pd->ell_cap >= ELL_CAP2_WIDE) throw StatusError{SDOT_CAPACITY_OVERFLOW, "Hessian row longer than the row capacity"};
        // rerun with wide rows; the captured PCG graph read the old row arrays
        pd->ell_cap = ELL_CAP2_WIDE;
        pd->ell_col.ensure((size_t)pd->ell_cap * N);
        pd->ell_val.ensure((size_t)pd->ell_cap * N);
        drop_graph(B);
        return batch_pass(B);
    }
    float ms = 0;
    if (hipEventElapsedTime(&ms, pd->ev[0], pd->ev[2]) != hipSuccess) (void)hipGetLastError();
    return ms;
}

void pcg_iteration(sdot_batch *B, const BMat &A, int max_iter) {
    hipStream_t s = B->pd->stream;
    const Blocks bl = blocks(B);
    const unsigned nb = (unsigned)B->nblk, K = (unsigned)B->K;
    hipLaunchKernelGGL(kb_spmv_dot, dim3(nb), dim3(BB), 0, s, A, bl, B->p.p, B->q.p, B->red.p, B->S.p);
    hipLaunchKernelGGL(kb_alpha, dim3(K), dim3(BB), 0, s, B->red.p, B->dblk.p, B->S.p);
    hipLaunchKernelGGL(kb_axpy, dim3(nb), dim3(BB), 0, s, A, bl, B->p.p, B->q.p, B->x.p, B->r.p, B->z.p, B->red.p,
                       B->S.p);
    hipLaunchKernelGGL(kb_beta, dim3(K), dim3(BB), 0, s, B->red.p, B->dblk.p, B->S.p, max_iter);
    hipLaunchKernelGGL(kb_pupdate, dim3(nb), dim3(BB), 0, s, bl, B->z.p, B->p.p, B->S.p);
}

// solves M x = b for the problems with S_DONE == 0; returns the summed iteration counts
int batch_pcg(sdot_batch *B, int max_iter, sdot_batch_stats &st, double &graph_ms) {
    sdot_pd *pd = B->pd;
    hipStream_t s = pd->stream;
    const int64_t K = B->K;
    const BMat A{pd->diag.p, pd->row_len.p, pd->ell_col.p, pd->ell_val.p, pd->N};
    hipLaunchKernelGGL(kb_init, dim3((unsigned)B->nblk), dim3(BB), 0, s, A, blocks(B), B->b.p, B->x.p, B->r.p, B->z.p,
                       B->p.p, B->red.p, B->S.p);
    hipLaunchKernelGGL(kb_init_red, dim3((unsigned)K), dim3(BB), 0, s, B->red.p, B->dblk.p, B->perr.p, B->S.p);
    HIPCK(hipGetLastError());
    if (B->use_graph && !B->graph) {
        hipGraph_t g = nullptr;
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < CHUNK; ++k) pcg_iteration(B, A, max_iter);
        HIPCK(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCK(e);
    }
    int *dcount = reinterpret_cast<int *>(B->red.p + 2 * B->nblk + 8);
    int active = 1;
    for (int it = 0; it < max_iter && active > 0; it += CHUNK) {
        if (B->graph) {
            HIPCK(hipEventRecord(B->gev[0], s));
            HIPCK(hipGraphLaunch(B->graph, s));
            HIPCK(hipEventRecord(B->gev[1], s));
            ++st.graph_replays;
        } else {
            for (int k = 0; k < CHUNK; ++k) pcg_iteration(B, A, max_iter);
        }
        hipLaunchKernelGGL(kb_count_active, dim3(1), dim3(BB), 0, s, B->S.p, K, dcount);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(&active, dcount, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        float ms = 0;
        if (B->graph && hipEventElapsedTime(&ms, B->gev[0], B->gev[1]) == hipSuccess) graph_ms += ms;
    }
    std::vector<double> S(K * S_NS);
    HIPCK(hipMemcpyAsync(S.data(), B->S.p, S.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    int tot = 0;
    for (int64_t p = 0; p < K; ++p) tot += (int)S[p * S_NS + S_IT];
    return tot;
}

// residuals of the last pass -> per-problem max relative mass error (host)
void batch_errors(sdot_batch *B, std::vector<double> &e) {
    sdot_pd *pd = B->pd;
    hipStream_t s = pd->stream;
    hipLaunchKernelGGL(kb_residual, dim3((unsigned)B->nblk), dim3(BB), 0, s, blocks(B), pd->mass.p, B->tsort.p, B->b.p,
                       B->red.p, B->pdone.p);
    hipLaunchKernelGGL(kb_pmax, dim3((unsigned)B->K), dim3(BB), 0, s, B->red.p, B->dblk.p, B->perr.p);
    HIPCK(hipGetLastError());
    e.resize(B->K);
    HIPCK(hipMemcpyAsync(e.data(), B->perr.p, B->K * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
}

template <class F>
sdot_status batch_guard(sdot_batch *B, F &&f) {
    try {
        return f();
    } catch (const HipError &e) {
        B->err = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.err);
        return SDOT_HIP_ERROR;
    } catch (const OomError &) {
        B->err = "device allocation failed";
        return SDOT_OOM;
    } catch (const StatusError &e) {
        B->err = e.msg;
        return e.st;
    } catch (const std::string &e) {
        B->err = e;
        return SDOT_CAPACITY_OVERFLOW;
    } catch (const std::bad_alloc &) {
        B->err = "host allocation failed";
        return SDOT_OOM;
    }
}

}  // namespace

extern "C" {

sdot_status sdot_batch_create(sdot_batch **out, int dim, int precision) {
    if (!out) return SDOT_INVALID_ARG;
    *out = nullptr;
    if (dim == 3) return SDOT_NOT_SUPPORTED;
    if (dim != 2 || (precision != SDOT_FP64 && precision != SDOT_MIXED)) return SDOT_INVALID_ARG;
    sdot_batch *B = new (std::nothrow) sdot_batch;
    if (!B) return SDOT_OOM;
    sdot_status st = sdot_pd_create(&B->pd, dim, precision);
    if (st == SDOT_OK)
        st = batch_guard(B, [&] {
            for (auto &e : B->gev) HIPCK(hipEventCreate(&e));
            for (int k = 0; k < 4; ++k)
                if (!B->pd->ev[k]) HIPCK(hipEventCreate(&B->pd->ev[k]));
            return SDOT_OK;
        });
    if (st != SDOT_OK) {
        sdot_batch_destroy(B);
        return st;
    }
    *out = B;
    return SDOT_OK;
}

sdot_status sdot_batch_destroy(sdot_batch *B) {
    if (!B) return SDOT_INVALID_ARG;
    drop_graph(B);
    for (auto &e : B->gev)
        if (e) (void)hipEventDestroy(e);
    for (auto *b : {&B->prob, &B->blk_prob, &B->pdone, &B->act}) b->release();
    for (auto *b : {&B->dcode_base, &B->doff, &B->dblk}) b->release();
    for (auto *b : {&B->pdom, &B->tie, &B->w0, &B->tgt, &B->tsort, &B->old, &B->b, &B->red, &B->S, &B->perr, &B->x, &B->r, &B->z,
                    &B->p, &B->q})
        b->release();
    B->dgrids.release();
    B->pflags.release();
    B->wglob.release();
    if (B->pd) sdot_pd_destroy(B->pd);
    delete B;
    return SDOT_OK;
}

const char *sdot_batch_last_error(const sdot_batch *B) {
    if (!B) return "null handle";
    return B->err.c_str();
}

sdot_status sdot_batch_set_tuning(sdot_batch *B, int dpc_x100, int use_graph) {
    if (!B || dpc_x100 < 0 || (dpc_x100 > 0 && dpc_x100 < 25) || dpc_x100 > 10000) return SDOT_INVALID_ARG;
    B->dpc_x100 = dpc_x100;
    B->use_graph = use_graph ? 1 : 0;
    B->grid_dpc = -1;
    drop_graph(B);
    return SDOT_OK;
}

sdot_status sdot_batch_set_problems(sdot_batch *B, int64_t K, const int64_t *offsets, const double *positions,
                                    const double *weights, const double *targets, const double *boxes, int ball) {
    if (!B || K < 1 || K > 0x7FFFFFFF || !offsets || !positions || !weights || !targets || offsets[0] != 0)
        return SDOT_INVALID_ARG;
    for (int64_t p = 0; p < K; ++p)
        if (offsets[p + 1] <= offsets[p]) return SDOT_INVALID_ARG;
    const int64_t N = offsets[K];
    if (N > 0x7FFFFFFF) return SDOT_INVALID_ARG;
    for (int64_t i = 0; i < N; ++i)
        if (!(targets[i] > 0) || !std::isfinite(positions[2 * i]) || !std::isfinite(positions[2 * i + 1]))
            return SDOT_INVALID_ARG;
    if (boxes)
        for (int64_t p = 0; p < K; ++p)
            if (!(boxes[4 * p + 2] > boxes[4 * p]) || !(boxes[4 * p + 3] > boxes[4 * p + 1])) return SDOT_INVALID_ARG;
    return batch_guard(B, [&] {
        sdot_pd *pd = B->pd;
        hipStream_t s = pd->stream;
        drop_graph(B);   // buffers may move
        B->K = K;
        B->ball = ball ? 1 : 0;
        B->grid_dpc = -1;
        B->off.assign(offsets, offsets + K + 1);
        B->hpos.assign(positions, positions + 2 * N);
        pd->N = N;
        pd->ball = B->ball;
        pd->pos.ensure(2 * N);
        pd->w.ensure(N);
        HIPCK(hipMemcpyAsync(pd->pos.p, positions, 2 * N * sizeof(double), hipMemcpyHostToDevice, s));
        B->w0.ensure(N);
        HIPCK(hipMemcpyAsync(B->w0.p, weights, N * sizeof(double), hipMemcpyHostToDevice, s));
        std::vector<int> prob(N);
        for (int64_t p = 0; p < K; ++p) std::fill(prob.begin() + offsets[p], prob.begin() + offsets[p + 1], (int)p);
        B->prob.ensure(N);
        HIPCK(hipMemcpyAsync(B->prob.p, prob.data(), N * sizeof(int), hipMemcpyHostToDevice, s));
        std::vector<double> dom(6 * K);
        for (int64_t p = 0; p < K; ++p) {
            dom[6 * p + 0] = boxes ? boxes[4 * p + 0] : 0.0;
            dom[6 * p + 1] = boxes ? boxes[4 * p + 1] : 0.0;
            dom[6 * p + 2] = 0.0;
            dom[6 * p + 3] = boxes ? boxes[4 * p + 2] : 1.0;
            dom[6 * p + 4] = boxes ? boxes[4 * p + 3] : 1.0;
            dom[6 * p + 5] = 1.0;
        }
        B->pdom.ensure(6 * K);
        HIPCK(hipMemcpyAsync(B->pdom.p, dom.data(), 6 * K * sizeof(double), hipMemcpyHostToDevice, s));
        B->doff.ensure(K + 1);
        HIPCK(hipMemcpyAsync(B->doff.p, offsets, (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
        B->blk_off.assign(K + 1, 0);
        for (int64_t p = 0; p < K; ++p) B->blk_off[p + 1] = B->blk_off[p] + (offsets[p + 1] - offsets[p] + BB - 1) / BB;
        B->nblk = B->blk_off[K];
        std::vector<int> bp(B->nblk);
        for (int64_t p = 0; p < K; ++p) std::fill(bp.begin() + B->blk_off[p], bp.begin() + B->blk_off[p + 1], (int)p);
        B->blk_prob.ensure(B->nblk);
        B->dblk.ensure(K + 1);
        HIPCK(hipMemcpyAsync(B->blk_prob.p, bp.data(), B->nblk * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(B->dblk.p, B->blk_off.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
        B->tgt.ensure(N);
        HIPCK(hipMemcpyAsync(B->tgt.p, targets, N * sizeof(double), hipMemcpyHostToDevice, s));
        for (auto *b : {&B->tsort, &B->old, &B->b, &B->x, &B->r, &B->z, &B->p, &B->q}) b->ensure(N);
        B->red.ensure(2 * B->nblk + 16);
        B->S.ensure(K * S_NS);
        B->perr.ensure(K);
        B->pflags.ensure(K);
        B->pdone.ensure(K);
        B->act.ensure(K);
        B->wglob.ensure(K);
        pd->flags.ensure(2);
        if (pd->ell_cap < ELL_CAP2) pd->ell_cap = ELL_CAP2;
        pd->mass.ensure(N);
        pd->arc.ensure(N);
        pd->perim.ensure(N);
        pd->diag.ensure(N);
        pd->row_len.ensure(N);
        pd->ell_col.ensure((size_t)pd->ell_cap * N);
        pd->ell_val.ensure((size_t)pd->ell_cap * N);
        pd->ovf_list.ensure(2 * (size_t)N + 2);
        HIPCK(hipStreamSynchronize(s));
        return SDOT_OK;
    });
}

sdot_status sdot_batch_solve(sdot_batch *B, double tol, int stop_rule, int max_iter, double *weights_out,
                             double *masses_out, int *status_out, int *iters_out, sdot_batch_stats *stats) {
    if (!B || !weights_out || !(tol > 0) || max_iter < 1 || B->K < 1 || (stop_rule != 0 && stop_rule != 1))
        return SDOT_INVALID_ARG;
    return batch_guard(B, [&] {
        sdot_pd *pd = B->pd;
        const int64_t K = B->K, N = pd->N;
        hipStream_t s = pd->stream;
        const auto t0 = std::chrono::steady_clock::now();
        sdot_batch_stats st{};
        fill_u32(reinterpret_cast<uint32_t *>(B->pdone.p), K, 0u, s);
        // every solve starts from the weights given to set_problems
        HIPCK(hipMemcpyAsync(pd->w.p, B->w0.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCK(hipMemcpyAsync(B->old.p, B->w0.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));

        // grid resolution: given, cached for this size class, or tuned now on the start weights
        int dpc = B->dpc_x100;
        const int size_class = ((int)std::lround(std::log2((double)N / (double)K)) * 2 + B->ball) * 2 +
                               (pd->precision == SDOT_MIXED);
        if (!dpc) {
            std::lock_guard<std::mutex> lk(tune_mutex);
            auto it = tune_cache.find(size_class);
            if (it != tune_cache.end()) dpc = it->second;
        }
        if (!dpc) {
            double best = DBL_MAX;
            for (int cand : TUNE_CANDIDATES) {
                batch_grid(B, cand);
                (void)batch_pass(B);   // first touch
                const double ms = batch_pass(B);
                if (ms < best) {
                    best = ms;
                    dpc = cand;
                }
            }
            std::lock_guard<std::mutex> lk(tune_mutex);
            tune_cache[size_class] = dpc;
        }
        st.tuned_dpc_x100 = dpc;
        if (B->grid_dpc != dpc) batch_grid(B, dpc);

        std::vector<int> done(K, 0), iters(K, 0), pst(K, SDOT_NOT_CONVERGED), dw_ok(K, 0), act(K);
        std::vector<unsigned> pfl(K);
        std::vector<double> err(K, INFINITY), e, S(K * S_NS), mdw;
        double ms_cells = 0, ms_solve = 0, graph_ms = 0;
        bool evaluated = false;   // the last pass saw the current weights of every active problem
        for (int it = 0; it <= max_iter; ++it) {
            ms_cells += batch_pass(B);
            evaluated = true;
            HIPCK(hipMemcpyAsync(pfl.data(), B->pflags.p, K * sizeof(unsigned), hipMemcpyDeviceToHost, s));
            batch_errors(B, e);
            // per-problem decisions
            int nsolve = 0;
            for (int64_t p = 0; p < K; ++p) {
                act[p] = ACT_NONE;
                S[p * S_NS + S_DONE] = 1;
                if (done[p]) continue;
                const bool empty = (pfl[p] & FLAG_EMPTY) != 0;
                if (!empty) err[p] = e[p];
                if (dw_ok[p] || (stop_rule == 0 && !empty && e[p] < tol)) {   // converged at these weights
                    done[p] = 1;
                    pst[p] = SDOT_OK;
                    continue;
                }
                if (it == max_iter) continue;   // this pass only evaluates the last update
                iters[p] = it + 1;
                if (empty) {
                    act[p] = ACT_BACKTRACK;
                    st.backtracks++;
                    continue;
                }
                act[p] = ACT_ACCEPT;
                S[p * S_NS + S_DONE] = 0;
                ++nsolve;
            }
            HIPCK(hipMemcpyAsync(B->act.p, act.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(kb_accept, dim3(nblocks(N, BB)), dim3(BB), 0, s, B->prob.p, N, B->act.p, pd->w.p,
                               B->old.p);
            HIPCK(hipMemcpyAsync(B->pdone.p, done.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
            HIPCK(hipGetLastError());
            if (std::all_of(done.begin(), done.end(), [](int d) { return d != 0; }) || it == max_iter) break;
            if (nsolve == 0) continue;   // only backtracks this round
            HIPCK(hipMemcpyAsync(B->S.p, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, s));
            if (!B->ball)
                hipLaunchKernelGGL(kb_gauge, dim3(nblocks(K, 256)), dim3(256), 0, s, pd->diag.p, pd->inv_perm.p,
                                   B->doff.p, B->act.p, K);
            const auto ts = std::chrono::steady_clock::now();
            st.pcg_iters_total += batch_pcg(B, 100000, st, graph_ms);
            hipLaunchKernelGGL(kb_update, dim3((unsigned)B->nblk), dim3(BB), 0, s, blocks(B), B->x.p, pd->perm.p,
                               pd->w.p, B->act.p, B->red.p);
            hipLaunchKernelGGL(kb_pmax, dim3((unsigned)K), dim3(BB), 0, s, B->red.p, B->dblk.p, B->perr.p);
            HIPCK(hipGetLastError());
            mdw.resize(K);
            HIPCK(hipMemcpyAsync(mdw.data(), B->perr.p, K * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCK(hipStreamSynchronize(s));
            ms_solve += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
            evaluated = false;
            if (stop_rule == 1)
                for (int64_t p = 0; p < K; ++p)
                    if (act[p] == ACT_ACCEPT && mdw[p] < tol) dw_ok[p] = 1;   // done once the next pass evaluates it
        }
        (void)evaluated;
        HIPCK(hipMemcpyAsync(weights_out, pd->w.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
        if (masses_out) {
            hipLaunchKernelGGL(kb_export, dim3(nblocks(N, 256)), dim3(256), 0, s, pd->mass.p, pd->inv_perm.p, N, B->q.p);
            HIPCK(hipGetLastError());
            HIPCK(hipMemcpyAsync(masses_out, B->q.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        HIPCK(hipStreamSynchronize(s));
        int nconv = 0, maxit = 0;
        double maxerr = 0;
        for (int64_t p = 0; p < K; ++p) {
            nconv += done[p];
            maxit = std::max(maxit, iters[p]);
            maxerr = std::max(maxerr, err[p]);
            if (status_out) status_out[p] = pst[p];
            if (iters_out) iters_out[p] = iters[p];
        }
        st.newton_iters_max = maxit;
        st.nb_converged = nconv;
        st.max_rel_mass_err = maxerr;
        st.seconds_cells = ms_cells * 1e-3;
        st.seconds_solve = ms_solve * 1e-3;
        st.ms_pcg_iteration = st.graph_replays ? graph_ms / (st.graph_replays * (double)CHUNK) : 0.0;
        st.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
        return nconv == K ? SDOT_OK : SDOT_NOT_CONVERGED;
    });
}

}  // extern "C"
