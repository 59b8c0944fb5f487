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

#include "batch_engine.h"
#include "pcg.h"

using namespace sdot_amd;
using namespace sdot_amd::pcg_state;

namespace {

constexpr int BB = 256;         // rows per block
enum { ACT_NONE = 0, ACT_ACCEPT = 1, ACT_BACKTRACK = 2 };
// device-driven Newton: where each problem stands (0 must be "needs a pass": the
// pass and residual kernels take this array as their skip mask, like the done marks)
enum { ST_PASS = 0, ST_CG = 1, ST_DONE = 2 };

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
// b = m - target (grid order) and the per-block max of |b| / target, for the problems
// whose skip mark is 0 (done marks, or the device Newton state ST_PASS)
__global__ __launch_bounds__(BB) void kb_residual(Blocks B, const double *m, const double *tsort, double *b,
                                                  double *part, const int *pskip) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    double mx = 0;
    if (!pskip[br.p])
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

// ---- persistent pass work list ----------------------------------------------------
// The problems whose skip mark is 0, in problem order, with their cells numbered
// consecutively (wl_off[j]: first number of the j-th listed problem; wl_off[count] =
// cells in all; wl_meta = {count, cells}); clears the 8 grab counters.  One block:
// block-wide inclusive scans over chunks of BB problems.
__global__ __launch_bounds__(BB) void kb_worklist(const int *pskip, const int64_t *off, int64_t K, int64_t *wl_off,
                                                  int *wl_prob, int64_t *wl_meta, unsigned long long *grab) {
    __shared__ int sn[BB];
    __shared__ int64_t sc[BB];
    if (threadIdx.x < 8) grab[threadIdx.x] = 0;
    int carry_n = 0;
    int64_t carry_c = 0;
    for (int64_t b = 0; b < K; b += BB) {
        const int64_t p = b + threadIdx.x;
        const int need = p < K && pskip[p] == 0;
        const int64_t cells = need ? off[p + 1] - off[p] : 0;
        sn[threadIdx.x] = need;
        sc[threadIdx.x] = cells;
        __syncthreads();
        for (int d = 1; d < BB; d <<= 1) {
            const int vn = threadIdx.x >= d ? sn[threadIdx.x - d] : 0;
            const int64_t vc = threadIdx.x >= d ? sc[threadIdx.x - d] : 0;
            __syncthreads();
            sn[threadIdx.x] += vn;
            sc[threadIdx.x] += vc;
            __syncthreads();
        }
        if (need) {
            const int j = carry_n + sn[threadIdx.x] - 1;
            wl_prob[j] = (int)p;
            wl_off[j] = carry_c + sc[threadIdx.x] - cells;
        }
        carry_n += sn[BB - 1];
        carry_c += sc[BB - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        wl_off[carry_n] = carry_c;
        wl_meta[0] = carry_n;
        wl_meta[1] = carry_c;
    }
}

// ---- device-driven Newton (the per-problem rules of OptimalTransportSolver.tcc:60-112) ----
struct NewtonDev {
    int *state, *iters, *backtracks, *pcg_iters, *dw_ok, *status, *passes;
    int *fallbacks;       // solves whose multigrid hierarchy could not be built (Jacobi used)
    double *err;          // relative mass error of the last pass with no empty cell
    double *mdw;          // max |dw| of the last update
    int *stat;            // [0] problems not done, [1] sticky pass failure flags
};

__global__ __launch_bounds__(BB) void kb_newton_init(NewtonDev nd, int64_t K, int *pstop, double *S) {
    const int64_t p = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (p == 0) {
        nd.stat[0] = (int)K;
        nd.stat[1] = 0;
    }
    if (p >= K) return;
    S[p * S_NS + S_DONE] = 1;   // no solve running
    S[p * S_NS + S_START] = 0;
    nd.state[p] = ST_PASS;
    nd.iters[p] = nd.backtracks[p] = nd.pcg_iters[p] = nd.dw_ok[p] = nd.passes[p] = nd.fallbacks[p] = 0;
    nd.status[p] = SDOT_NOT_CONVERGED;
    nd.err[p] = INFINITY;
    nd.mdw[p] = INFINITY;
    pstop[p] = 0;   // complete passes until the first decision
}

// after a pass: the decision of every problem that had one (tick() on the host makes
// the same one); a pass that failed (flags) decides nothing and is recorded in stat[1]
__global__ __launch_bounds__(BB) void kb_decide(NewtonDev nd, int64_t K, const unsigned *pflags, const double *perr,
                                                const unsigned *flags, int *act, double *S, int *amg_off, double tol,
                                                int stop_rule, int max_iter) {
    const int64_t p = blockIdx.x * (int64_t)BB + threadIdx.x;
    const unsigned bad = flags[0] & (FLAG_BAD_INPUT | FLAG_CELL_OVERFLOW | FLAG_ROW_OVERFLOW);
    if (p == 0 && bad) nd.stat[1] |= (int)bad;
    if (p >= K) return;
    double *sp = S + p * S_NS;
    int a = ACT_NONE;
    sp[S_START] = 0;
    if (nd.state[p] == ST_PASS && !bad) {
        const bool empty = (pflags[p] & FLAG_EMPTY) != 0;
        const double e = perr[p];
        if (!empty) nd.err[p] = e;
        nd.passes[p] += 1;
        if (nd.dw_ok[p] || (stop_rule == 0 && !empty && e < tol)) {   // converged at these weights
            nd.state[p] = ST_DONE;
            nd.status[p] = SDOT_OK;
        } else if (nd.iters[p] >= max_iter) {   // this pass only evaluated the last update
            nd.state[p] = ST_DONE;
            nd.status[p] = SDOT_NOT_CONVERGED;
        } else {
            nd.iters[p] += 1;
            if (empty) {   // :72-79
                a = ACT_BACKTRACK;
                nd.backtracks[p] += 1;
            } else {       // the solve starts in this replay
                a = ACT_ACCEPT;
                nd.state[p] = ST_CG;
                sp[S_START] = 1;
                sp[S_DONE] = 0;
                sp[S_IT] = 0;
                amg_off[p] = 0;
            }
        }
    }
    act[p] = a;
}

// w[perm[t]] -= x[t] for the problems whose solve ended (state ST_CG, S_DONE); per-block max |x|
__global__ __launch_bounds__(BB) void kb_update_dev(Blocks B, const double *x, const int *perm, double *w,
                                                    const int *state, const double *S, double *part) {
    __shared__ double sm[BB];
    const BlockRows br = block_rows(B, blockIdx.x);
    double mx = 0;
    if (state[br.p] == ST_CG && S[br.p * S_NS + S_DONE] != 0)
        for (int64_t t = br.r0 + threadIdx.x; t < br.r1; t += BB) {
            w[perm[t]] -= x[t];
            mx = fmax(mx, fabs(x[t]));
        }
    mx = bmax(mx, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

// the problems whose solve ended go back to a pass; stop_if_void marks of the next
// pass; stat[0] = problems not done.  One block.
__global__ __launch_bounds__(BB) void kb_finish(NewtonDev nd, int64_t K, const double *S, const int *amg_off,
                                                int *pstop, double tol, int stop_rule, int max_iter) {
    __shared__ double sm[BB];
    double cnt = 0;
    for (int64_t p = threadIdx.x; p < K; p += BB) {
        if (nd.state[p] == ST_CG && S[p * S_NS + S_DONE] != 0) {
            nd.pcg_iters[p] += (int)S[p * S_NS + S_IT];
            nd.fallbacks[p] += amg_off[p] != 0;
            if (stop_rule == 1 && nd.mdw[p] < tol) nd.dw_ok[p] = 1;   // done once the next pass evaluates it
            nd.state[p] = ST_PASS;
        }
        // an empty cell in the next pass means a backtrack, whose pass results are not read
        pstop[p] = nd.state[p] == ST_PASS && !nd.dw_ok[p] && nd.iters[p] < max_iter;
        cnt += nd.state[p] != ST_DONE ? 1.0 : 0.0;
    }
    cnt = bsum(cnt, sm);
    if (threadIdx.x == 0) nd.stat[0] = (int)cnt;
}

// ---- PCG on problem-aligned blocks --------------------------------------------
// relayout: kept problems move from their old offset to the new one (dirac order)
__global__ __launch_bounds__(BB) void kb_move(const int *newprob, int64_t n, const int64_t *src_off,
                                              const int64_t *dst_off, const double *pos, const double *w,
                                              const double *old, const double *w0, const double *tgt, double *pos2,
                                              double *w2, double *old2, double *w02, double *tgt2) {
    const int64_t i = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (i >= n) return;
    const int p = newprob[i];
    const int64_t so = src_off[p];
    if (so < 0) return;   // uploaded from the host
    const int64_t j = so + (i - dst_off[p]);
    pos2[2 * i] = pos[2 * j];
    pos2[2 * i + 1] = pos[2 * j + 1];
    w2[i] = w[j];
    old2[i] = old[j];
    w02[i] = w0[j];
    tgt2[i] = tgt[j];
}

// idx[i] = the k with off[k] <= i < off[k+1] (binary search over K+1 offsets)
__global__ __launch_bounds__(BB) void kb_owner(const int64_t *off, int64_t K, int64_t n, int *idx) {
    const int64_t i = blockIdx.x * (int64_t)BB + threadIdx.x;
    if (i >= n) return;
    int64_t lo = 0, hi = K;   // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    idx[i] = (int)lo;
}

std::mutex tune_mutex;
struct Tuned {
    int dpc_x100, cells_per_wave, tierset;
};
std::map<int, Tuned> tune_cache;  // size class -> pass schedule
constexpr int TUNE_DPC[] = {150, 250, 400};
constexpr int TUNE_CPW[] = {64, 128, 256};
constexpr int TUNE_TIERS[] = {0, 1};

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

namespace sdot_amd {

GridInfo problem_grid(const double *pos, int64_t n, int dpc_x100, double *tie) {
    double lo[2] = {DBL_MAX, DBL_MAX}, hi[2] = {-DBL_MAX, -DBL_MAX};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 2; ++d) {
            lo[d] = std::min(lo[d], pos[2 * i + d]);
            hi[d] = std::max(hi[d], pos[2 * i + d]);
        }
    const GridInfo g = grid_params(2, lo, hi, n, dpc_x100 / 100.0);
    // same scale as the single-problem path: 1e-12 of the squared mean spacing
    if (tie) *tie = 1e-12 * (g.h[0] * g.ncell[0]) * (g.h[1] * g.ncell[1]) / (double)n;
    return g;
}

struct BatchEngine::Impl {
    DevBuf<int> prob, blk_prob, pdone, pstop, act;
    DevBuf<int64_t> dcode_base, doff, dblk, src_off;
    DevBuf<GridInfo> dgrids;
    DevBuf<double> pdom, tie, w0, tgt, tsort, old, b, red, perr;
    DevBuf<double> x, q;
    DevBuf<double> pos2, w2, old2, w02, tgt2;   // relayout targets (swapped in)
    DevBuf<unsigned> pflags;
    DevBuf<unsigned long long> wglob;
    std::vector<int64_t> blk_off;
    int64_t nblk = 0, ncodes = 0;
    PcgSolver pcg;
    // persistent pass work list
    DevBuf<int64_t> wl_off, wl_meta;
    DevBuf<int> wl_prob;
    DevBuf<unsigned long long> grab;
    int cpw = 128;                       // cells a wavefront grabs at a time
    // device-driven Newton: per-problem state, the captured replay and its key
    DevBuf<int> nst, nit, nbt, npcg, ndw, nstatus, npass, nfb, nstat;
    DevBuf<double> nerr, nmdw;
    hipGraphExec_t rgraph = nullptr;
    struct Key {
        const void *ell_col = nullptr, *ell_val = nullptr, *w = nullptr, *pcg_state = nullptr;
        int64_t N = -1, K = -1;
        int ell_cap = 0, cpw = 0, tierset = -1, cg_per_replay = 0, amg = -1, stop_rule = -1, max_iter = -1;
        double tol = 0;
        bool operator==(const Key &o) const {
            return ell_col == o.ell_col && ell_val == o.ell_val && w == o.w && pcg_state == o.pcg_state && N == o.N &&
                   K == o.K && ell_cap == o.ell_cap && cpw == o.cpw && tierset == o.tierset &&
                   cg_per_replay == o.cg_per_replay && amg == o.amg && stop_rule == o.stop_rule &&
                   max_iter == o.max_iter && tol == o.tol;
        }
    } rkey;
    int *hstat = nullptr;                // pinned: two slots of {not done, failure flags}
    hipEvent_t sev[2] = {};
    NewtonDev newton() {
        return NewtonDev{nst.p, nit.p, nbt.p, npcg.p, ndw.p, nstatus.p, npass.p, nfb.p, nerr.p, nmdw.p, nstat.p};
    }
    void drop_replay() {
        if (!rgraph) return;
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        (void)hipGraphExecDestroy(rgraph);
        rgraph = nullptr;
        rkey = Key{};
    }
};

BatchEngine::BatchEngine(int precision, int ball, int dpc_x100, bool use_graph)
    : im(new Impl), ball_(ball ? 1 : 0), dpc_x100_(dpc_x100 > 0 ? dpc_x100 : 250), use_graph_(use_graph) {
    im->pcg.set_options(true, use_graph);
    im->pcg.set_dim(2);
    const sdot_status st = sdot_pd_create(&pd, 2, precision);
    if (st != SDOT_OK) {
        delete im;
        throw StatusError{st, "device handle creation failed"};
    }
    pd->ball = ball_;
    for (int k = 0; k < 4; ++k)
        if (!pd->ev[k]) HIPCK(hipEventCreate(&pd->ev[k]));
    pd->flags.ensure(3);
}

BatchEngine::~BatchEngine() {
    im->drop_replay();
    for (auto &e : im->sev)
        if (e) (void)hipEventDestroy(e);
    if (im->hstat) (void)hipHostFree(im->hstat);
    for (auto *b : {&im->nst, &im->nit, &im->nbt, &im->npcg, &im->ndw, &im->nstatus, &im->npass, &im->nfb, &im->nstat,
                    &im->wl_prob})
        b->release();
    for (auto *b : {&im->nerr, &im->nmdw}) b->release();
    for (auto *b : {&im->wl_off, &im->wl_meta}) b->release();
    im->grab.release();
    for (auto *b : {&im->prob, &im->blk_prob, &im->pdone, &im->pstop, &im->act}) b->release();
    for (auto *b : {&im->dcode_base, &im->doff, &im->dblk, &im->src_off}) b->release();
    for (auto *b : {&im->pdom, &im->tie, &im->w0, &im->tgt, &im->tsort, &im->old, &im->b, &im->red, &im->perr,
                    &im->x, &im->q, &im->pos2, &im->w2, &im->old2, &im->w02, &im->tgt2})
        b->release();
    im->dgrids.release();
    im->pflags.release();
    im->wglob.release();
    delete im;
    if (pd) sdot_pd_destroy(pd);
}

double BatchEngine::bytes_per_dirac() {
    // dirac arrays (pos, w, w0, old, tgt, spare copies for a relayout), sort keys and
    // permutations, sorted copies, pass outputs, ELL rows at the wide capacity, PCG vectors,
    // FP64 resume slots (a 16-byte header per cell, 160 bytes for a quarter of the cells),
    // FP32 copies of the matrix values for the multigrid smoothers (32 entries per row on
    // level 0, 64 per coarse row)
    return 16 + 8 * 4 + 8 * 5 + 8 + 8 + 24 + 32 + 4 + ELL_CAP2_WIDE * 12.0 + 8 * 7 + 8 + 16 + 56 + 128 + 86;
}

size_t BatchEngine::device_bytes() const { return (size_t)(bytes_per_dirac() * (double)N_); }

void BatchEngine::set_use_graph(bool on) {
    use_graph_ = on;
    im->pcg.set_options(im->pcg.amg(), on);
}

void BatchEngine::set_preconditioner(bool amg) { im->pcg.set_options(amg, use_graph_); }

void BatchEngine::relayout(const std::vector<int> &keep, const std::vector<NewProblem> &add) {
    hipStream_t s = pd->stream;
    std::vector<BatchProblem> np;
    np.reserve(keep.size() + add.size());
    std::vector<int64_t> src;
    int64_t off = 0;
    for (int k : keep) {
        BatchProblem b = probs.at(k);
        src.push_back(b.off);
        b.off = off;
        off += b.n;
        np.push_back(b);
    }
    for (const NewProblem &a : add) {
        BatchProblem b;
        b.n = a.n;
        b.hpos = a.pos;
        for (int d = 0; d < 4; ++d) b.box[d] = a.box[d];
        b.g = problem_grid(a.pos, a.n, dpc_x100_, &b.tie);
        b.off = off;
        b.tag = a.tag;
        off += a.n;
        src.push_back(-1);
        np.push_back(b);
    }
    const int64_t N = off, K = (int64_t)np.size();
    if (N > 0x7FFFFFFF) throw StatusError{SDOT_INVALID_ARG, "more than 2^31-1 diracs resident"};
    if (K == 0) {
        probs.clear();
        N_ = 0;
        pd->N = 0;
        return;
    }
    // per-problem tables
    std::vector<int64_t> offs(K + 1);
    for (int64_t k = 0; k < K; ++k) offs[k] = np[k].off;
    offs[K] = N;
    im->doff.ensure(K + 1);
    im->src_off.ensure(K);
    HIPCK(hipMemcpyAsync(im->doff.p, offs.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(im->src_off.p, src.data(), K * sizeof(int64_t), hipMemcpyHostToDevice, s));
    im->prob.ensure(N);
    hipLaunchKernelGGL(kb_owner, dim3(nblocks(N, BB)), dim3(BB), 0, s, im->doff.p, K, N, im->prob.p);
    HIPCK(hipGetLastError());
    // dirac arrays: move the kept problems, upload the new ones, swap
    for (auto *b : {&im->w2, &im->old2, &im->w02, &im->tgt2}) b->ensure(N);
    im->pos2.ensure(2 * N);
    if (!keep.empty())
        hipLaunchKernelGGL(kb_move, dim3(nblocks(N, BB)), dim3(BB), 0, s, im->prob.p, N, im->src_off.p, im->doff.p,
                           pd->pos.p, pd->w.p, im->old.p, im->w0.p, im->tgt.p, im->pos2.p, im->w2.p, im->old2.p,
                           im->w02.p, im->tgt2.p);
    HIPCK(hipGetLastError());
    for (size_t a = 0; a < add.size(); ++a) {
        const BatchProblem &b = np[keep.size() + a];
        const NewProblem &q = add[a];
        HIPCK(hipMemcpyAsync(im->pos2.p + 2 * b.off, q.pos, 2 * b.n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(im->w02.p + b.off, q.w0, b.n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(im->w2.p + b.off, q.w0, b.n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(im->old2.p + b.off, q.w0, b.n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(im->tgt2.p + b.off, q.tgt, b.n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    std::swap(pd->pos, im->pos2);
    std::swap(pd->w, im->w2);
    std::swap(im->old, im->old2);
    std::swap(im->w0, im->w02);
    std::swap(im->tgt, im->tgt2);
    // problem-aligned blocks
    im->blk_off.assign(K + 1, 0);
    for (int64_t k = 0; k < K; ++k) im->blk_off[k + 1] = im->blk_off[k] + (np[k].n + BB - 1) / BB;
    im->nblk = im->blk_off[K];
    im->dblk.ensure(K + 1);
    HIPCK(hipMemcpyAsync(im->dblk.p, im->blk_off.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    im->blk_prob.ensure(im->nblk);
    hipLaunchKernelGGL(kb_owner, dim3(nblocks(im->nblk, BB)), dim3(BB), 0, s, im->dblk.p, K, im->nblk, im->blk_prob.p);
    HIPCK(hipGetLastError());
    std::vector<double> dom(6 * K);
    for (int64_t k = 0; k < K; ++k) {
        dom[6 * k + 0] = np[k].box[0];
        dom[6 * k + 1] = np[k].box[1];
        dom[6 * k + 2] = 0.0;
        dom[6 * k + 3] = np[k].box[2];
        dom[6 * k + 4] = np[k].box[3];
        dom[6 * k + 5] = 1.0;
    }
    im->pdom.ensure(6 * K);
    HIPCK(hipMemcpyAsync(im->pdom.p, dom.data(), 6 * K * sizeof(double), hipMemcpyHostToDevice, s));
    // per-dirac work arrays
    for (auto *b : {&im->tsort, &im->b, &im->x, &im->q}) b->ensure(N);
    im->red.ensure(2 * im->nblk + 16);
    im->pcg.set_layout(offs, s);
    im->perr.ensure(K);
    im->pflags.ensure(K);
    im->pdone.ensure(K);
    im->pstop.ensure(K);
    im->act.ensure(K);
    im->wglob.ensure(K);
    im->wl_off.ensure(K + 1);
    im->wl_prob.ensure(K);
    im->wl_meta.ensure(2);
    im->grab.ensure(8);
    for (auto *b : {&im->nst, &im->nit, &im->nbt, &im->npcg, &im->ndw, &im->nstatus, &im->npass, &im->nfb})
        b->ensure(K);
    im->nstat.ensure(2);
    im->nerr.ensure(K);
    im->nmdw.ensure(K);
    im->drop_replay();
    if (pd->ell_cap < ELL_CAP2) pd->ell_cap = ELL_CAP2;
    pd->N = N;
    pd->mass.ensure(N);
    pd->arc.ensure(N);
    pd->perim.ensure(N);
    pd->diag.ensure(N);
    pd->row_len.ensure(N);
    pd->ell_col.ensure((size_t)pd->ell_cap * N);
    pd->ell_val.ensure((size_t)pd->ell_cap * N);
    pd->ovf_list.ensure(2 * (size_t)N + 2);
    if (pd->precision != SDOT_MIXED) ensure_resume(pd, N);
    probs.swap(np);
    N_ = N;
    std::vector<int> done(K);
    for (int64_t k = 0; k < K; ++k) done[k] = probs[k].done;
    HIPCK(hipMemcpyAsync(im->pdone.p, done.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
    fill_u32(reinterpret_cast<uint32_t *>(im->pstop.p), K, 0u, s);   // complete passes until a tick decides
    rebuild_grid();
}

void BatchEngine::set_resolution(int dpc_x100) {
    if (dpc_x100 == dpc_x100_) return;
    dpc_x100_ = dpc_x100;
    for (auto &b : probs) b.g = problem_grid(b.hpos, b.n, dpc_x100_, &b.tie);
    if (!probs.empty()) rebuild_grid();
}

// per-problem grids in one code space, the counting sort, sorted positions, targets in grid order
void BatchEngine::rebuild_grid() {
    hipStream_t s = pd->stream;
    im->drop_replay();   // the grid arrays may move
    const int64_t K = (int64_t)probs.size(), N = N_;
    std::vector<GridInfo> grids(K);
    std::vector<int64_t> cb(K + 1, 0);
    std::vector<double> tie(K);
    for (int64_t k = 0; k < K; ++k) {
        grids[k] = probs[k].g;
        tie[k] = probs[k].tie;
        cb[k + 1] = cb[k] + grids[k].ncodes;
    }
    const int64_t nc = cb[K];
    if (nc >= (int64_t)0xFFFFFFFF) throw StatusError{SDOT_CAPACITY_OVERFLOW, "batch grid code space exceeds 32 bits"};
    im->ncodes = nc;
    im->dgrids.ensure(K);
    im->dcode_base.ensure(K + 1);
    im->tie.ensure(K);
    HIPCK(hipMemcpyAsync(im->dgrids.p, grids.data(), K * sizeof(GridInfo), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(im->dcode_base.p, cb.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(im->tie.p, tie.data(), K * sizeof(double), hipMemcpyHostToDevice, s));
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
    hipLaunchKernelGGL(kb_keys, dim3(nblocks(N, BB)), dim3(BB), 0, s, pd->pos.p, im->prob.p, N, im->dgrids.p,
                       im->dcode_base.p, pd->keys.p, pd->rank.p, pd->scan_tmp.p);
    HIPCK(hipGetLastError());
    sort_by_keys(pd, nc);
    // the handle's own grid descriptor spans the whole code space (the kernels read the per-problem grids)
    pd->g = grids[0];
    pd->g.ncodes = nc;
    hipLaunchKernelGGL(kb_gather, dim3(nblocks(N, 256)), dim3(256), 0, s, im->tgt.p, pd->perm.p, N, im->tsort.p);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(s));
    pd->last.grid_cells = nc;
}

void BatchEngine::restart() {
    hipStream_t s = pd->stream;
    const int64_t K = (int64_t)probs.size();
    if (K == 0) return;
    HIPCK(hipMemcpyAsync(pd->w.p, im->w0.p, N_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCK(hipMemcpyAsync(im->old.p, im->w0.p, N_ * sizeof(double), hipMemcpyDeviceToDevice, s));
    fill_u32(reinterpret_cast<uint32_t *>(im->pdone.p), K, 0u, s);
    fill_u32(reinterpret_cast<uint32_t *>(im->pstop.p), K, 0u, s);
    for (auto &b : probs) {
        b.iters = b.done = b.dw_ok = b.backtracks = b.pcg_iters = 0;
        b.status = SDOT_NOT_CONVERGED;
        b.err = INFINITY;
    }
    HIPCK(hipStreamSynchronize(s));
}

namespace {

// One cell pass, enqueued only (capturable): sorted weights and bounds, the work list
// of the problems whose skip mark is 0, and the capacity tiers of the 2D kernel, the
// first one persistent over the work list.
void enqueue_pass(sdot_pd *pd, BatchEngine::Impl *im, int64_t K, int64_t N, const int *pskip) {
    hipStream_t s = pd->stream;
    const int64_t nc = im->ncodes;
    fill_u32(pd->flags.p, 3, 0u, s);
    fill_u32(im->pflags.p, K, 0u, s);
    hipLaunchKernelGGL(kb_gather_w, dim3(nblocks(N, BB)), dim3(BB), 0, s, pd->w.p, pd->perm.p, im->prob.p, im->doff.p,
                       im->tie.p, N, pd->sw.p, pd->flags.p);
    hipLaunchKernelGGL(kb_cell_wmax, dim3(nblocks(nc, BB)), dim3(BB), 0, s, pd->cell_start.p, nc, pd->sw.p,
                       pd->cell_wmax.p);
    hipLaunchKernelGGL(kb_wmax, dim3((unsigned)K), dim3(BB), 0, s, pd->sw.p, im->doff.p, im->wglob.p);
    hipLaunchKernelGGL(kb_worklist, dim3(1), dim3(BB), 0, s, pskip, im->doff.p, K, im->wl_off.p, im->wl_prob.p,
                       im->wl_meta.p, im->grab.p);
    HIPCK(hipGetLastError());
    CellArgs a{};
    a.N = N;
    a.spos = pd->spos.p;
    a.sw = pd->sw.p;
    a.cell_start = pd->cell_start.p;
    a.cell_wmax = pd->cell_wmax.p;
    a.wglob = im->wglob.p;
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
    if (pd->precision != SDOT_MIXED) use_resume(pd, a);
    a.prob = im->prob.p;
    a.grids = im->dgrids.p;
    a.code_base = im->dcode_base.p;
    a.pdom = im->pdom.p;
    a.pflags = im->pflags.p;
    a.pdone = nullptr;   // the work list holds only the problems that need the pass
    a.pstop = im->pstop.p;
    a.stop_if_void = 1;
    a.wl_off = im->wl_off.p;
    a.wl_prob = im->wl_prob.p;
    a.wl_meta = im->wl_meta.p;
    a.pb_off = im->doff.p;
    a.grab = im->grab.p;
    a.cpw = im->cpw;
    run_cells2(pd, a);
}

// one cell pass over the problems not marked done; device ms
double engine_pass(sdot_pd *pd, BatchEngine::Impl *im, int64_t K, int64_t N, bool &regrow) {
    hipStream_t s = pd->stream;
    regrow = false;
    HIPCK(hipEventRecord(pd->ev[0], s));
    enqueue_pass(pd, im, K, N, im->pdone.p);
    HIPCK(hipEventRecord(pd->ev[2], s));
    unsigned fl = 0;
    HIPCK(hipMemcpyAsync(&fl, pd->flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (fl & FLAG_BAD_INPUT) throw StatusError{SDOT_INVALID_ARG, "non-finite weight or coincident diracs in the batch"};
    if (fl & FLAG_CELL_OVERFLOW) throw StatusError{SDOT_CAPACITY_OVERFLOW, "cell polygon exceeds 64 vertices"};
    if (fl & FLAG_ROW_OVERFLOW) {
        if (pd->ell_cap >= ELL_CAP2_WIDE) throw StatusError{SDOT_CAPACITY_OVERFLOW, "Hessian row longer than the row capacity"};
        // rerun with wide rows (kept from now on); a captured PCG graph read the old row arrays
        pd->ell_cap = ELL_CAP2_WIDE;
        pd->ell_col.ensure((size_t)pd->ell_cap * N);
        pd->ell_val.ensure((size_t)pd->ell_cap * N);
        im->pcg.invalidate();
        im->drop_replay();
        regrow = true;
        return 0;
    }
    float ms = 0;
    if (hipEventElapsedTime(&ms, pd->ev[0], pd->ev[2]) != hipSuccess) (void)hipGetLastError();
    return ms;
}

Blocks blocks_of(BatchEngine::Impl *im) { return Blocks{im->blk_prob.p, im->dblk.p, im->doff.p}; }

}  // namespace

double BatchEngine::pass_only() {
    bool regrow = false;
    double ms = engine_pass(pd, im, (int64_t)probs.size(), N_, regrow);
    if (regrow) ms = engine_pass(pd, im, (int64_t)probs.size(), N_, regrow);
    return ms;
}

int BatchEngine::tick(const TickParams &tp, std::vector<int> &finished, TickStats &ts) {
    sdot_pd *pd_ = pd;
    hipStream_t s = pd_->stream;
    const int64_t K = (int64_t)probs.size(), N = N_;
    if (K == 0) return 0;
    bool regrow = false;
    double ms = engine_pass(pd_, im, K, N, regrow);
    if (regrow) ms = engine_pass(pd_, im, K, N, regrow);
    ts.ms_pass += ms;
    std::vector<unsigned> pfl(K);
    HIPCK(hipMemcpyAsync(pfl.data(), im->pflags.p, K * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(kb_residual, dim3((unsigned)im->nblk), dim3(BB), 0, s, blocks_of(im), pd_->mass.p, im->tsort.p,
                       im->b.p, im->red.p, im->pdone.p);
    hipLaunchKernelGGL(kb_pmax, dim3((unsigned)K), dim3(BB), 0, s, im->red.p, im->dblk.p, im->perr.p);
    HIPCK(hipGetLastError());
    std::vector<double> e(K);
    HIPCK(hipMemcpyAsync(e.data(), im->perr.p, K * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    // per-problem decisions (OptimalTransportSolver.tcc:60-112)
    std::vector<int> act(K, ACT_NONE), done(K);
    int nsolve = 0, running = 0;
    for (int64_t k = 0; k < K; ++k) {
        BatchProblem &b = probs[k];
        if (!b.done) {
            ts.cells += b.n;
            const bool empty = (pfl[k] & FLAG_EMPTY) != 0;
            if (!empty) b.err = e[k];
            if (b.dw_ok || (tp.stop_rule == 0 && !empty && e[k] < tp.tol)) {   // converged at these weights
                b.done = 1;
                b.status = SDOT_OK;
                finished.push_back((int)k);
            } else if (b.iters >= tp.max_iter) {   // this pass only evaluated the last update
                b.done = 1;
                b.status = SDOT_NOT_CONVERGED;
                finished.push_back((int)k);
            } else {
                ++b.iters;
                ++running;
                if (empty) {   // :72-79
                    act[k] = ACT_BACKTRACK;
                    ++b.backtracks;
                    ++ts.backtracks;
                } else {
                    act[k] = ACT_ACCEPT;
                    ++nsolve;
                }
            }
        }
        done[k] = b.done;
    }
    HIPCK(hipMemcpyAsync(im->act.p, act.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kb_accept, dim3(nblocks(N, BB)), dim3(BB), 0, s, im->prob.p, N, im->act.p, pd_->w.p, im->old.p);
    HIPCK(hipMemcpyAsync(im->pdone.p, done.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCK(hipGetLastError());
    if (nsolve > 0) {
        if (!ball_)   // :62-63
            hipLaunchKernelGGL(kb_gauge, dim3(nblocks(K, 256)), dim3(256), 0, s, pd_->diag.p, pd_->inv_perm.p, im->doff.p,
                               im->act.p, K);
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<int> active(K), its;
        for (int64_t k = 0; k < K; ++k) active[k] = act[k] == ACT_ACCEPT;
        PcgStats ps;
        im->pcg.solve(PcgMat{pd_->diag.p, pd_->row_len.p, pd_->ell_col.p, pd_->ell_val.p, N, pd_->ell_cap}, im->b.p, im->x.p,
                      im->perr.p, active, 100000, its, ps);
        hipLaunchKernelGGL(kb_update, dim3((unsigned)im->nblk), dim3(BB), 0, s, blocks_of(im), im->x.p, pd_->perm.p,
                           pd_->w.p, im->act.p, im->red.p);
        hipLaunchKernelGGL(kb_pmax, dim3((unsigned)K), dim3(BB), 0, s, im->red.p, im->dblk.p, im->perr.p);
        HIPCK(hipGetLastError());
        std::vector<double> mdw(K);
        HIPCK(hipMemcpyAsync(mdw.data(), im->perr.p, K * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        ts.ms_pcg += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ts.ms_pcg_graph += ps.ms_graph;
        ts.graph_replays += ps.graph_replays;
        ts.graph_iterations += ps.graph_iterations;
        ts.amg_levels = std::max(ts.amg_levels, ps.amg_levels);
        ts.amg_fallbacks += ps.amg_fallbacks;
        ts.solved += nsolve;
        for (int64_t k = 0; k < K; ++k) {
            if (act[k] != ACT_ACCEPT) continue;
            probs[k].pcg_iters += its[k];
            ts.pcg_iters += its[k];
            if (tp.stop_rule == 1 && mdw[k] < tp.tol) probs[k].dw_ok = 1;   // done once the next pass evaluates it
        }
    }
    // stop_if_void for the next pass: only problems that cannot finish on it (an empty
    // cell there means a backtrack, whose pass results are not read)
    std::vector<int> stop(K);
    for (int64_t k = 0; k < K; ++k) {
        const BatchProblem &b = probs[k];
        stop[k] = !b.done && !b.dw_ok && b.iters < tp.max_iter;
    }
    HIPCK(hipMemcpyAsync(im->pstop.p, stop.data(), K * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCK(hipStreamSynchronize(s));
    return running;
}

// ---- device-driven Newton ------------------------------------------------------------
// One replay advances every problem by what its state asks for: a problem that needs a
// pass gets it, then its decision (kb_decide: converged, out of iterations, backtrack,
// or accept, which starts its linear solve); every problem in a solve runs
// cg_per_replay more CG iterations; a problem whose solve ended takes its update and
// needs a pass again.  The replay is one captured graph; the host only replays it and
// reads two counters (problems not done, pass failures), one replay behind the device.
// Each problem follows the same steps as under tick(), so its iterates are the same.

void BatchEngine::set_schedule(int cells_per_wave, int tierset) {
    if (cells_per_wave < 64 || cells_per_wave % 64 || cells_per_wave > 4096 || (tierset != 0 && tierset != 1))
        throw StatusError{SDOT_INVALID_ARG, "cells per wavefront: a multiple of 64 up to 4096; tier set 0 or 1"};
    im->cpw = cells_per_wave;
    pd->tierset = tierset;
}

int BatchEngine::cells_per_wave() const { return im->cpw; }
int BatchEngine::precond_levels() const { return im->pcg.levels(); }
int BatchEngine::tierset() const { return pd->tierset; }

// the kernels of one replay (enqueued only)
void BatchEngine::enqueue_replay(const TickParams &tp, int cg_per_replay) {
    hipStream_t s = pd->stream;
    const int64_t K = (int64_t)probs.size(), N = N_;
    const NewtonDev nd = im->newton();
    const PcgMat A{pd->diag.p, pd->row_len.p, pd->ell_col.p, pd->ell_val.p, N, pd->ell_cap};
    enqueue_pass(pd, im, K, N, nd.state);
    hipLaunchKernelGGL(kb_residual, dim3((unsigned)im->nblk), dim3(BB), 0, s, blocks_of(im), pd->mass.p, im->tsort.p,
                       im->b.p, im->red.p, nd.state);
    hipLaunchKernelGGL(kb_pmax, dim3((unsigned)K), dim3(BB), 0, s, im->red.p, im->dblk.p, im->perr.p);
    hipLaunchKernelGGL(kb_decide, dim3(nblocks(K, BB)), dim3(BB), 0, s, nd, K, im->pflags.p, im->perr.p, pd->flags.p,
                       im->act.p, im->pcg.state(), im->pcg.fallback(), tp.tol, tp.stop_rule, tp.max_iter);
    hipLaunchKernelGGL(kb_accept, dim3(nblocks(N, BB)), dim3(BB), 0, s, im->prob.p, N, im->act.p, pd->w.p, im->old.p);
    if (!ball_)   // :62-63
        hipLaunchKernelGGL(kb_gauge, dim3(nblocks(K, 256)), dim3(256), 0, s, pd->diag.p, pd->inv_perm.p, im->doff.p,
                           im->act.p, K);
    im->pcg.enqueue_start(A, im->b.p, im->x.p, im->perr.p);
    im->pcg.enqueue_iterations(A, im->x.p, cg_per_replay, 100000);
    hipLaunchKernelGGL(kb_update_dev, dim3((unsigned)im->nblk), dim3(BB), 0, s, blocks_of(im), im->x.p, pd->perm.p,
                       pd->w.p, nd.state, im->pcg.state(), im->red.p);
    hipLaunchKernelGGL(kb_pmax, dim3((unsigned)K), dim3(BB), 0, s, im->red.p, im->dblk.p, nd.mdw);
    hipLaunchKernelGGL(kb_finish, dim3(1), dim3(BB), 0, s, nd, K, im->pcg.state(), im->pcg.fallback(), im->pstop.p,
                       tp.tol, tp.stop_rule, tp.max_iter);
    HIPCK(hipGetLastError());
}

void BatchEngine::capture_replay(const TickParams &tp, int cg_per_replay) {
    hipStream_t s = pd->stream;
    const int64_t K = (int64_t)probs.size(), N = N_;
    Impl::Key key;
    key.ell_col = pd->ell_col.p;
    key.ell_val = pd->ell_val.p;
    key.w = pd->w.p;
    key.pcg_state = im->pcg.state();
    key.N = N;
    key.K = K;
    key.ell_cap = pd->ell_cap;
    key.cpw = im->cpw;
    key.tierset = pd->tierset;
    key.cg_per_replay = cg_per_replay;
    key.amg = im->pcg.amg() ? 1 : 0;
    key.stop_rule = tp.stop_rule;
    key.max_iter = tp.max_iter;
    key.tol = tp.tol;
    if (im->rgraph && im->rkey == key) return;
    std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());   // no device free while recording
    im->drop_replay();
    hipGraph_t g = nullptr;
    pd->capturing = true;
    HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_replay(tp, cg_per_replay);
    } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        pd->capturing = false;
        throw;
    }
    pd->capturing = false;
    HIPCK(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&im->rgraph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) im->rgraph = nullptr;
    HIPCK(e);
    im->rkey = key;
}

// Hessian rows widened in place: the layout [e * N + t] keeps the entries of the
// problems that are inside a solve at the same places
void BatchEngine::widen_rows() {
    hipStream_t s = pd->stream;
    if (pd->ell_cap >= ELL_CAP2_WIDE) throw StatusError{SDOT_CAPACITY_OVERFLOW, "Hessian row longer than the row capacity"};
    const size_t old_n = (size_t)pd->ell_cap * N_, new_n = (size_t)ELL_CAP2_WIDE * N_;
    DevBuf<int> c2;
    DevBuf<double> v2;
    c2.ensure(new_n);
    v2.ensure(new_n);
    HIPCK(hipMemcpyAsync(c2.p, pd->ell_col.p, old_n * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIPCK(hipMemcpyAsync(v2.p, pd->ell_val.p, old_n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCK(hipStreamSynchronize(s));
    std::swap(pd->ell_col, c2);
    std::swap(pd->ell_val, v2);
    c2.release();
    v2.release();
    pd->ell_cap = ELL_CAP2_WIDE;
    im->pcg.invalidate();
    im->drop_replay();
}

int BatchEngine::solve_device(const TickParams &tp, int cg_per_replay, DeviceSolveStats &ds) {
    hipStream_t s = pd->stream;
    const int64_t K = (int64_t)probs.size();
    if (K == 0) return 0;
    if (!im->hstat) {
        void *h = nullptr;
        HIPCK(hipHostMalloc(&h, 4 * sizeof(int), hipHostMallocDefault));
        im->hstat = static_cast<int *>(h);
        for (auto &e : im->sev) HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const NewtonDev nd = im->newton();
    hipLaunchKernelGGL(kb_newton_init, dim3(nblocks(K, BB)), dim3(BB), 0, s, nd, K, im->pstop.p, im->pcg.state());
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(pd->ev[0], s));
    // every problem ends: at most max_iter + 1 passes, each solve at most 100000 CG iterations
    const int64_t bound = (int64_t)(tp.max_iter + 2) * (100000 / cg_per_replay + 2);
    int64_t r = 0;
    int running = (int)K;
    bool pending = false;   // a status copy in flight for the previous replay
    auto launch = [&](int64_t k) {
        if (use_graph_) HIPCK(hipGraphLaunch(im->rgraph, s));
        else enqueue_replay(tp, cg_per_replay);
        HIPCK(hipMemcpyAsync(im->hstat + 2 * (k & 1), nd.stat, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCK(hipEventRecord(im->sev[k & 1], s));
    };
    if (use_graph_) capture_replay(tp, cg_per_replay);
    while (running > 0) {
        if (r >= bound) throw StatusError{SDOT_NOT_CONVERGED, "device Newton did not end"};
        launch(r);
        ++r;
        if (!pending) {
            pending = true;
            continue;
        }
        // read the replay before the one just launched
        HIPCK(hipEventSynchronize(im->sev[(r - 2) & 1]));
        const int *st = im->hstat + 2 * ((r - 2) & 1);
        const int fail = st[1];
        running = st[0];
        if (fail) {
            HIPCK(hipStreamSynchronize(s));
            if (fail & FLAG_BAD_INPUT) throw StatusError{SDOT_INVALID_ARG, "non-finite weight or coincident diracs in the batch"};
            if (fail & FLAG_CELL_OVERFLOW) throw StatusError{SDOT_CAPACITY_OVERFLOW, "cell polygon exceeds 64 vertices"};
            // a Hessian row outgrew its capacity: the problems of that pass decided
            // nothing, the others went on; widen and go on
            widen_rows();
            ++ds.widened;
            fill_u32(reinterpret_cast<uint32_t *>(nd.stat + 1), 1, 0u, s);
            if (use_graph_) capture_replay(tp, cg_per_replay);
            pending = false;
            running = 1;
        }
    }
    HIPCK(hipEventRecord(pd->ev[1], s));
    HIPCK(hipStreamSynchronize(s));
    float ms = 0;
    if (hipEventElapsedTime(&ms, pd->ev[0], pd->ev[1]) != hipSuccess) (void)hipGetLastError();
    ds.replays += r;
    ds.ms_device += ms;
    // per-problem results
    std::vector<int> it(K), bt(K), pc(K), stt(K), np(K), fb(K);
    std::vector<double> er(K);
    HIPCK(hipMemcpyAsync(it.data(), nd.iters, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(bt.data(), nd.backtracks, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(pc.data(), nd.pcg_iters, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(stt.data(), nd.status, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(np.data(), nd.passes, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(fb.data(), nd.fallbacks, K * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(er.data(), nd.err, K * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    for (int64_t k = 0; k < K; ++k) {
        BatchProblem &b = probs[k];
        b.iters = it[k];
        b.backtracks = bt[k];
        b.pcg_iters = pc[k];
        b.status = stt[k];
        b.err = er[k];
        b.done = 1;
        ds.backtracks += bt[k];
        ds.pcg_iters += pc[k];
        ds.cells += (int64_t)np[k] * b.n;
        ds.fallbacks += fb[k];
    }
    return 0;
}

void BatchEngine::read_weights(int k, double *out) {
    const BatchProblem &b = probs.at(k);
    HIPCK(hipMemcpyAsync(out, pd->w.p + b.off, b.n * sizeof(double), hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
}

void BatchEngine::read_weights_all(double *out) {
    HIPCK(hipMemcpyAsync(out, pd->w.p, N_ * sizeof(double), hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
}

void BatchEngine::read_masses_all(double *out) {
    hipLaunchKernelGGL(kb_export, dim3(nblocks(N_, 256)), dim3(256), 0, pd->stream, pd->mass.p, pd->inv_perm.p, N_,
                       im->q.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, im->q.p, N_ * sizeof(double), hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
}

void BatchEngine::read_masses(int k, double *out) {
    const BatchProblem &b = probs.at(k);
    hipLaunchKernelGGL(kb_export, dim3(nblocks(N_, 256)), dim3(256), 0, pd->stream, pd->mass.p, pd->inv_perm.p, N_,
                       im->q.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, im->q.p + b.off, b.n * sizeof(double), hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
}

}  // namespace sdot_amd

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
struct sdot_batch {
    BatchEngine *eng = nullptr;
    int precision = SDOT_FP64;
    int dpc_x100 = 0;             // 0: autotune on the first solve of a size class
    int use_graph = 1;
    int cells_per_wave = 0;       // 0: autotune
    int tierset = -1;             // -1: autotune
    int device_newton = 1;        // Newton driven from the device (one replayed graph)
    int cg_per_replay = 16;
    int precond = SDOT_PRECOND_AMG;
    std::vector<double> hpos;     // host copy of the positions (grids are built from it)
    std::string err;
};

namespace {

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
    B->precision = precision;
    *out = B;
    return SDOT_OK;
}

sdot_status sdot_batch_destroy(sdot_batch *B) {
    if (!B) return SDOT_INVALID_ARG;
    try {
        delete B->eng;
    } catch (...) {
    }
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
    if (B->eng) B->eng->set_use_graph(B->use_graph != 0);
    return SDOT_OK;
}

sdot_status sdot_batch_set_linear_solver(sdot_batch *B, int precond) {
    if (!B || (precond != SDOT_PRECOND_AMG && precond != SDOT_PRECOND_JACOBI)) return SDOT_INVALID_ARG;
    B->precond = precond;
    if (B->eng) B->eng->set_preconditioner(precond == SDOT_PRECOND_AMG);
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
        if (!B->eng || B->eng->pd->ball != (ball ? 1 : 0)) {
            delete B->eng;
            B->eng = nullptr;
            B->eng = new BatchEngine(B->precision, ball, B->dpc_x100 ? B->dpc_x100 : 250, B->use_graph != 0);
            B->eng->set_preconditioner(B->precond == SDOT_PRECOND_AMG);
        }
        B->hpos.assign(positions, positions + 2 * N);
        std::vector<NewProblem> add(K);
        for (int64_t p = 0; p < K; ++p) {
            NewProblem &a = add[p];
            a.n = offsets[p + 1] - offsets[p];
            a.pos = B->hpos.data() + 2 * offsets[p];
            a.w0 = weights + offsets[p];
            a.tgt = targets + offsets[p];
            for (int d = 0; d < 4; ++d) a.box[d] = boxes ? boxes[4 * p + d] : (d < 2 ? 0.0 : 1.0);
            a.tag = p;
        }
        B->eng->relayout({}, add);
        return SDOT_OK;
    });
}

sdot_status sdot_batch_set_schedule(sdot_batch *B, int cells_per_wave, int first_tier, int device_newton,
                                    int cg_per_replay) {
    if (!B || (cells_per_wave != 0 && (cells_per_wave < 64 || cells_per_wave % 64 || cells_per_wave > 4096)) ||
        first_tier < -1 || first_tier > 1 || cg_per_replay < 1 || cg_per_replay > 256)
        return SDOT_INVALID_ARG;
    B->cells_per_wave = cells_per_wave;
    B->tierset = first_tier;
    B->device_newton = device_newton ? 1 : 0;
    B->cg_per_replay = cg_per_replay;
    return SDOT_OK;
}

sdot_status sdot_batch_solve(sdot_batch *B, double tol, int stop_rule, int max_iter, double *weights_out,
                             double *masses_out, int *status_out, int *iters_out, sdot_batch_stats *stats) {
    if (!B || !weights_out || !(tol > 0) || max_iter < 1 || (stop_rule != 0 && stop_rule != 1)) return SDOT_INVALID_ARG;
    if (!B->eng || B->eng->probs.empty()) {
        B->err = "no problems set";
        return SDOT_INVALID_ARG;
    }
    return batch_guard(B, [&] {
        BatchEngine &E = *B->eng;
        const int64_t K = (int64_t)E.probs.size(), N = E.total();
        const auto t0 = std::chrono::steady_clock::now();
        sdot_batch_stats st{};
        // pass schedule (grid resolution, cells per wavefront grab, capacity tiers): given,
        // cached for this size class, or tuned now on the start weights by coordinate
        // descent over the candidates (one pass per candidate after a first touch)
        Tuned tu{B->dpc_x100, B->cells_per_wave, B->tierset};
        const bool autotune = !tu.dpc_x100 || !tu.cells_per_wave || tu.tierset < 0;
        const int size_class = ((int)std::lround(std::log2((double)N / (double)K)) * 2 + E.pd->ball) * 2 +
                               (B->precision == SDOT_MIXED);
        bool cached = false;
        if (autotune) {
            std::lock_guard<std::mutex> lk(tune_mutex);
            auto it = tune_cache.find(size_class);
            if (it != tune_cache.end()) {
                if (!tu.dpc_x100) tu.dpc_x100 = it->second.dpc_x100;
                if (!tu.cells_per_wave) tu.cells_per_wave = it->second.cells_per_wave;
                if (tu.tierset < 0) tu.tierset = it->second.tierset;
                cached = true;
            }
        }
        E.restart();
        if (autotune && !cached) {
            Tuned best{tu.dpc_x100 ? tu.dpc_x100 : 250, tu.cells_per_wave ? tu.cells_per_wave : 128,
                       tu.tierset >= 0 ? tu.tierset : 0};
            auto time_of = [&](const Tuned &c) {
                E.set_resolution(c.dpc_x100);
                E.set_schedule(c.cells_per_wave, c.tierset);
                (void)E.pass_only();   // first touch
                return E.pass_only();
            };
            double best_ms = DBL_MAX;
            auto sweep = [&](auto field, const int *cand, int nc, bool fixed) {
                if (fixed) return;
                Tuned keep = best;
                for (int q = 0; q < nc; ++q) {
                    Tuned c = keep;
                    c.*field = cand[q];
                    const double ms = time_of(c);
                    ++st.tuning_passes;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
            };
            sweep(&Tuned::dpc_x100, TUNE_DPC, 3, tu.dpc_x100 != 0);
            sweep(&Tuned::cells_per_wave, TUNE_CPW, 3, tu.cells_per_wave != 0);
            sweep(&Tuned::tierset, TUNE_TIERS, 2, tu.tierset >= 0);
            tu = best;
            std::lock_guard<std::mutex> lk(tune_mutex);
            tune_cache[size_class] = tu;
        }
        st.tuned_dpc_x100 = tu.dpc_x100;
        st.tuned_cells_per_wave = tu.cells_per_wave;
        st.tuned_first_tier = tu.tierset;
        E.set_resolution(tu.dpc_x100);
        E.set_schedule(tu.cells_per_wave, tu.tierset);
        E.restart();
        TickParams tp;
        tp.tol = tol;
        tp.stop_rule = stop_rule;
        tp.max_iter = max_iter;
        if (B->device_newton) {
            DeviceSolveStats ds;
            E.solve_device(tp, B->cg_per_replay, ds);
            st.device_newton = 1;
            st.newton_replays = (int)ds.replays;
            st.pcg_iters_total = ds.pcg_iters;
            st.backtracks = ds.backtracks;
            st.seconds_device = ds.ms_device * 1e-3;
            st.precond_levels = E.precond_levels();
            st.precond_fallbacks = ds.fallbacks;
        } else {
            TickStats ts;
            std::vector<int> fin;
            for (int guard = 0; guard <= max_iter + 1; ++guard)
                if (E.tick(tp, fin, ts) == 0) break;
            st.pcg_iters_total = ts.pcg_iters;
            st.backtracks = ts.backtracks;
            st.graph_replays = ts.graph_replays;
            st.seconds_cells = ts.ms_pass * 1e-3;
            st.seconds_solve = ts.ms_pcg * 1e-3;
            st.ms_pcg_iteration = ts.graph_iterations ? ts.ms_pcg_graph / (double)ts.graph_iterations : 0.0;
            st.precond_levels = ts.amg_levels;
            st.precond_fallbacks = ts.amg_fallbacks;
        }
        E.read_weights_all(weights_out);
        if (masses_out) E.read_masses_all(masses_out);
        int nconv = 0, maxit = 0;
        double maxerr = 0;
        for (int64_t k = 0; k < K; ++k) {
            const BatchProblem &b = E.probs[k];
            nconv += b.status == SDOT_OK;
            maxit = std::max(maxit, b.iters);
            maxerr = std::max(maxerr, b.err);
            if (status_out) status_out[k] = b.status;
            if (iters_out) iters_out[k] = b.iters;
        }
        st.newton_iters_max = maxit;
        st.nb_converged = nconv;
        st.max_rel_mass_err = maxerr;
        st.seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
        return nconv == K ? SDOT_OK : SDOT_NOT_CONVERGED;
    });
}

}  // extern "C"
