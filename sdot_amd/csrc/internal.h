// Internal declarations of the MI355X Laguerre-cell library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>
#include <array>

#include "../../include/sdot_mi355x.h"

namespace sdot_amd {

// ---------------------------------------------------------------------------
// error handling: every HIP call goes through HIPCK, which records the error
// text on the handle and returns SDOT_HIP_ERROR from the entry point.
// ---------------------------------------------------------------------------
struct HipError {
    hipError_t err;
    const char *what;
    int line;
};

#define HIPCK(call)                                                        \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess) throw ::sdot_amd::HipError{e_, #call, __LINE__}; \
    } while (0)

struct OomError {};

struct StatusError {
    sdot_status st;
    std::string msg;
};

// Device allocation and stream capture exclude each other across host threads:
// hipFree synchronises the device, which must not happen while another thread
// records a graph (the capture holds the stream the synchronisation waits on).
std::recursive_mutex &device_alloc_mutex();

// grow-only device buffer; a borrowed one points into memory it does not own
// (a caller's array or a slice of a fixed workspace) and never allocates or frees
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool borrowed = false;
    void borrow(T *q, size_t m) {
        p = q;
        n = m;
        borrowed = true;
    }
    void ensure(size_t m) {
        if (m <= n && p) return;
        if (borrowed) throw StatusError{SDOT_CAPACITY_OVERFLOW, "fixed workspace too small for this input"};
        std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
        if (p) HIPCK(hipFree(p));
        p = nullptr;
        n = 0;
        size_t want = m > 0 ? m : 1;
        if (hipMalloc((void **)&p, want * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw OomError{};
        }
        n = want;
    }
    void release() {
        if (p && !borrowed) {
            std::lock_guard<std::recursive_mutex> lk(device_alloc_mutex());
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        borrowed = false;
    }
};

// Uniform grid over the dirac bounding box; cells are addressed by the Morton
// code of their integer coordinates, so a sort by code groups each cell's
// diracs contiguously and keeps nearby cells close in memory.
struct GridInfo {
    int dim;
    int bits;             // bits per axis of the Morton code
    int ncell[3];
    double lo[3];
    double h[3];
    double inv_h[3];
    int64_t ncodes;       // 2^(dim*bits)
};

__host__ __device__ inline uint32_t spread2(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__host__ __device__ inline uint32_t spread3(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__host__ __device__ inline uint32_t morton2(int x, int y) { return spread2(x) | (spread2(y) << 1); }
__host__ __device__ inline uint32_t morton3(int x, int y, int z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}
__host__ __device__ inline int grid_coord(const GridInfo &g, int d, double x) {
    // clamped before the conversion: points far outside the grid (translated
    // centres) stay in int range, NaN goes to cell 0
    const double f = floor((x - g.lo[d]) * g.inv_h[d]);
    if (!(f >= 0)) return 0;
    if (f >= g.ncell[d]) return g.ncell[d] - 1;
    return (int)f;
}

// blockIdx -> work tile mapping that gives each of the 8 XCDs (dispatched
// round-robin) one contiguous range of tiles, so neighbouring tiles share an L2.
__device__ inline int64_t xcd_tile(int64_t b, int64_t nb) {
    int64_t q = nb / 8, rem = nb % 8, x = b % 8, i = b / 8;
    return x < rem ? x * (q + 1) + i : rem * (q + 1) + (x - rem) * q + i;
}

// order-preserving map double -> uint64 (for atomicMax over doubles)
__host__ __device__ inline unsigned long long dbl_order(double v) {
    unsigned long long b;
    memcpy(&b, &v, 8);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double dbl_unorder(unsigned long long b) {
    b = (b & 0x8000000000000000ull) ? (b & 0x7FFFFFFFFFFFFFFFull) : ~b;
    double v;
    memcpy(&v, &b, 8);
    return v;
}

enum : unsigned { FLAG_EMPTY = 1u, FLAG_ROW_OVERFLOW = 2u, FLAG_BAD_INPUT = 4u, FLAG_CELL_OVERFLOW = 8u };

// Hessian row capacity (ELL slots per row): the first pass uses the narrow one;
// a row overflow re-runs the pass with the wide one, which the handle then keeps
constexpr int ELL_CAP2 = 32, ELL_CAP2_WIDE = 64, ELL_CAP3 = 128, ELL_CAP3_WIDE = 255;
constexpr int HCAP3 = 32;   // warm-start neighbours kept per 3D cell

#define SDOT_HD __host__ __device__

// element i (0..2) of a 3-array with constant indices only: a per-lane index
// into an array held in a kernel argument would make the compiler copy the
// whole argument block to scratch memory
template <class A>
SDOT_HD inline auto sel3(const A &v, int i) -> decltype(v[0]) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

SDOT_HD inline void ctr_add(unsigned long long *p, unsigned long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(p, v);
#else
    *p += v;
#endif
}

// flag update usable from host+device code (the device side is atomic)
SDOT_HD inline void flag_or(unsigned *p, unsigned v) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    *p |= v;
#endif
}

// arguments of the cell kernels
struct CellArgs {
    int64_t N;
    const double *spos;          // sorted positions, interleaved
    const double *sw;            // sorted weights
    const uint32_t *cell_start;  // by Morton code, ncodes + 1 entries
    const double *cell_wmax;     // by Morton code
    const unsigned long long *wglob;  // ordered bits of the global max weight
    // max-weight pyramid over the Morton code space: level k (0 = cell_wmax) holds
    // ncodes >> (dim k) nodes at wpyr + pyr_off(k); null when not built
    const double *wpyr;
    GridInfo g;
    double dom_lo[3], dom_hi[3];
    double coeff;
    // piecewise-constant image density (3D voxel / 2D pixel), null when constant
    // ConvexPolyhedronAssembly items when the domain is not a single box (nitems = 0):
    // item k is the box it_box[6k..6k+5] (lo, hi) cut by the planes it_planes[4q..4q+3]
    // = (nx, ny, nz, h), n . x <= h, q in [it_off[k], it_off[k+1]), density it_coeff[k].
    // One item is the starting cell (planes with cut ids -100 - q); several items are
    // integrated piece by piece from a box around them (dom_lo, dom_hi).
    int nitems;
    const double *it_box, *it_coeff, *it_planes;
    const int *it_off;
    // SpZGrid::translations: each cell is also cut by the diracs translated by
    // trans[dim s .. dim s + dim), s < ntrans; the image of sorted dirac j under s has
    // the cut id j + (s + 1) N (single problems only; N (ntrans + 1) < 2^31)
    int ntrans;
    const double *trans;
    // final integration function (2D): 0 Unit, 1 R2, 2 WmR2, 3 ExpWmR2db (exp((w - |x - c|^2) / rf_eps))
    int rf;
    double rf_eps;
    // position / weight derivative blocks (9 doubles): per neighbour slot e at
    // jblk[(e * 9 + k) * ell_stride + t], the diagonal at jdiag[k * ell_stride + t]
    double *jblk, *jdiag;
    const double *img;
    int img_n[3];
    int img_nc;   // coefficients per pixel: 1, or 6 (2D quadratic, coefficient-major)
    // outputs, sorted order
    double *mass, *arc, *perim, *diag;
    double *mom;   // density-weighted first moments about the dirac, dim per cell (null: not wanted)
    int *row_len;
    int *ell_col;
    double *ell_val;
    int ell_cap;
    int64_t ell_stride;
    unsigned *flags;
    // stop_if_void (for_each_laguerre_cell, SpZGrid.h:44): once a cell has come out
    // empty (FLAG_EMPTY in flags, or in pflags for a batch problem) the pass result
    // is only a backtrack signal, so the remaining cells are skipped
    int stop_if_void;
    int *ovf_list;
    unsigned *ovf_count;
    // resume slots of the FP64 8-vertex tier (null: a cell that outgrows it starts
    // again in the next tier).  Sorted cell t that outgrows it leaves where its
    // traversal stood in rs_hdr[4 t .. 4 t + 4) (vertex count | (image + 1) << 8, or 0
    // to start again; ring; grid row | column << 16; candidate), and, at position
    // k < rs_cap of the overflow list, its polygon: x and y of vertex f < 8 at
    // rs_d[16 k + 2 f], rs_d[16 k + 2 f + 1], its edge id at rs_i[8 k + f]
    int *rs_hdr;
    int *rs_i;
    double *rs_d;
    int64_t rs_cap;
    // warm start of the 3D wavefront kernel (null: none): sorted cell t's neighbours in the
    // last complete pass over the same sorted diracs, hint_col[HCAP3 t + e], e < HCAP3
    // (-1: none), which the cell cuts first
    const int *hint_col;
    const int *work;             // optional list of sorted indices to process
    int64_t work_n;              // its length (3D; host-known)
    const unsigned *work_count;  // its length in device memory (2D tiers; no host round trip)
    unsigned long long *ctr;     // work counters: [k] = sum F, [WCTR + k] = sum K, [2*WCTR + k] = sum A
    // batches of independent problems (all null for a single problem): problem of
    // each sorted index, per-problem grid, code offset into cell_start/cell_wmax,
    // domain box (lo[dim], hi[dim] per problem, stride 6), flags and done marks
    const int *prob;
    const GridInfo *grids;
    const int64_t *code_base;
    const double *pdom;
    unsigned *pflags;
    const int *pdone;
    const int *pstop;   // batch: problems whose pass may stop at an empty cell (null: none)
    // persistent work list of a batch pass (first capacity tier; null: every cell,
    // one thread each): the cells of the problems wl_prob[j], j < wl_meta[0], problem
    // wl_prob[j]'s cells numbered from wl_off[j] (wl_meta[1] in all, wl_off has
    // wl_meta[0] + 1 entries); pb_off: the problems' row offsets.  The numbering is
    // cut into 8 segments; a wavefront grabs cpw consecutive cells at a time with
    // grab[x] (cells taken from segment x), starting in the segment of its XCD
    // (workgroups go round-robin over the 8 XCDs) and then helping the others.
    const int64_t *wl_off;
    const int *wl_prob;
    const int64_t *wl_meta;
    const int64_t *pb_off;
    unsigned long long *grab;
    int cpw;
};

// true when the pass may stop: stop_if_void is set and an empty cell has been
// seen (in the problem's own flags for a batch)
SDOT_HD inline bool void_seen(const CellArgs &a, int p) {
    if (!a.stop_if_void) return false;
    if (a.pflags && !(a.pstop && a.pstop[p])) return false;
    const unsigned *f = a.pflags ? a.pflags + p : a.flags;
    return (__atomic_load_n(f, __ATOMIC_RELAXED) & FLAG_EMPTY) != 0;
}

// The dirac of cut id `id` >= 0 and its position p: sorted dirac id, or for id >= N
// the image of dirac id mod N under translation id / N - 1 (get_der_integrals_wrt_weights.h:61-62)
template <int DIM, bool PER = true>
SDOT_HD inline int64_t cut_dirac(const CellArgs &a, int64_t id, double *p) {
    int64_t j = id;
    if (PER && j >= a.N) {
        const int64_t s = j / a.N - 1;
        j -= (s + 1) * a.N;
        for (int d = 0; d < DIM; ++d) p[d] = a.spos[DIM * j + d] + a.trans[DIM * s + d];
    } else {
        for (int d = 0; d < DIM; ++d) p[d] = a.spos[DIM * j + d];
    }
    return j;
}

// the per-problem view of CellArgs for sorted index t
struct ProbView {
    int p;
    GridInfo g;
    const uint32_t *cell_start;
    const double *cell_wmax;
    double W;
    double lo[3], hi[3];
};
SDOT_HD inline void prob_view_single(const CellArgs &a, ProbView &v) {
    v.p = 0;
    v.g = a.g;
    v.cell_start = a.cell_start;
    v.cell_wmax = a.cell_wmax;
    v.W = dbl_unorder(a.wglob[0]);
    for (int d = 0; d < 3; ++d) {
        v.lo[d] = a.dom_lo[d];
        v.hi[d] = a.dom_hi[d];
    }
}
SDOT_HD inline void prob_view(const CellArgs &a, int64_t t, ProbView &v) {
    if (!a.prob) {
        prob_view_single(a, v);
        return;
    }
    const int p = a.prob[t];
    v.p = p;
    v.g = a.grids[p];
    v.cell_start = a.cell_start + a.code_base[p];
    v.cell_wmax = a.cell_wmax + a.code_base[p];
    v.W = dbl_unorder(a.wglob[p]);
    for (int d = 0; d < 3; ++d) {
        v.lo[d] = a.pdom[6 * p + d];
        v.hi[d] = a.pdom[6 * p + 3 + d];
    }
}
// offset of level k of the max-weight pyramid (level 0 has ncodes nodes)
SDOT_HD inline int64_t pyr_off(int64_t ncodes, int dim, int k) {
    int64_t o = 0;
    for (int j = 0; j < k; ++j) o += ncodes >> (dim * j);
    return o;
}
SDOT_HD inline uint32_t compact3(uint32_t x) {
    x &= 0x09249249u;
    x = (x | (x >> 2)) & 0x030C30C3u;
    x = (x | (x >> 4)) & 0x0300F00Fu;
    x = (x | (x >> 8)) & 0x030000FFu;
    x = (x | (x >> 16)) & 0x000003FFu;
    return x;
}
constexpr int WCTR = 1024;       // counters are spread over WCTR slots to avoid contention
constexpr int BBOX_BLOCKS = 256; // blocks of the bounding-box reduction (2 dim partial values each)

// ---------------------------------------------------------------------------
// the handle
// ---------------------------------------------------------------------------
// one ConvexPolyhedronAssembly item: the box p0..p1 cut by the planes n . x <= h
// ({nx, ny, nz, h}), with a constant density
struct Shape {
    double p0[3], p1[3];
    double q0[3], q1[3];   // the box the cells are clipped to: p0, p1 with redundant faces moved out
    double coeff;
    int64_t cut_id;
    std::vector<std::array<double, 4>> planes;
};

class PcgSolver;
}  // namespace sdot_amd

struct sdot_pd {
    int dim = 2;
    int precision = SDOT_FP64;
    int ball = 0;
    int dpc_x100 = 0;  // tuning: diracs per grid cell x100 (0 = auto)
    int block = 0;     // tuning: threads per block (0 = auto)
    hipStream_t stream = nullptr;
    int device = 0;    // the device current at creation (the stream's)
    std::string err;

    // domain
    std::vector<sdot_amd::Shape> shapes;
    bool has_image = false;
    double img_lo[3] = {0, 0, 0}, img_hi[3] = {1, 1, 1};
    int img_n[3] = {1, 1, 1};
    int img_nc = 1;
    sdot_amd::DevBuf<double> img;

    // diracs (original order)
    int64_t N = 0;
    sdot_amd::DevBuf<double> pos, w;
    bool grid_valid = false;    // positions unchanged since the last sort
    // hint_col holds the neighbours of a complete 3D pass over the current sorted order
    // (warm start of the next passes of a Newton solve, which sets `warm`)
    bool hint_ok = false;
    bool warm = false;
    sdot_amd::DevBuf<int> hint_col;
    bool pass_valid = false;    // results match the current diracs/weights/settings
    bool stop_if_void = false;  // passes of a Newton solve: an empty cell ends the pass early

    // grid
    sdot_amd::GridInfo g{};
    sdot_amd::DevBuf<uint32_t> keys, rank, cell_start, scan_tmp;
    sdot_amd::DevBuf<int> perm, inv_perm;
    sdot_amd::DevBuf<double> spos, sw, cell_wmax, bbox_part, wpyr;
    sdot_amd::DevBuf<unsigned long long> wglob;

    // pass outputs (sorted order)
    sdot_amd::DevBuf<double> mass, arc, perim, diag, ell_val;
    sdot_amd::DevBuf<double> mom;   // first moments (centroids), dim per cell
    sdot_amd::DevBuf<double> it_box, it_coeff, it_planes;   // domain items on the device
    std::vector<double> htrans;                             // translations (dim per vector)
    sdot_amd::DevBuf<double> trans;
    bool want_jac = false;                                  // next pass writes the derivative blocks
    int rf = 0;                                             // final integration function: 0 Unit, 1 R2, 2 WmR2, 3 ExpWmR2db
    double rf_eps = 1.0;                                    // eps of ExpWmR2db
    sdot_amd::DevBuf<double> jblk, jdiag;
    sdot_amd::DevBuf<int> it_off;
    bool want_mom = false;          // the next pass computes the first moments
    bool pass_has_mom = false;      // the current pass did
    sdot_amd::DevBuf<int> row_len, ell_col, ovf_list;
    sdot_amd::DevBuf<int> rs_hdr, rs_i;   // resume slots of the FP64 8-vertex 2D tier
    sdot_amd::DevBuf<double> rs_d;
    sdot_amd::DevBuf<unsigned> flags;  // [0] flags, [1] overflow count, [2] second-tier overflow count
    int ell_cap = 0;
    unsigned last_flags = 0;
    sdot_amd::DevBuf<unsigned long long> wctr;
    sdot_amd::DevBuf<char> scratch3;   // per-thread polyhedron slots of the 3D thread kernel
    sdot_amd::DevBuf<sdot_amd::CellArgs> dargs;  // kernel arguments of the 3D kernels, in HBM
    sdot_amd::CellArgs hargs{};

    // CSR export scratch
    sdot_amd::DevBuf<int64_t> csr_off, csr_col, csr_cnt, csr_tmp, csr_off2, csr_col2;
    sdot_amd::DevBuf<double> csr_val, out_tmp, csr_val2;

    // solver scratch
    sdot_amd::DevBuf<double> sv[10];
    sdot_amd::DevBuf<double> red;
    sdot_amd::PcgSolver *pcg = nullptr;   // created on the first solve
    int64_t pcg_n = -1;                   // layout size of pcg
    int precond = SDOT_PRECOND_AMG;
    int use_graph = 1;                    // replay the PCG iterations from a captured graph
    int last_precond_levels = 0;          // multigrid levels of the last solve (0: Jacobi)

    // batches: capacity tiers of the 2D pass (0: the default sequence; 1: FP64 starts at
    // 16-vertex slots, mixed adds an 8-vertex first tier); the stream is being captured
    // into a graph (no event records)
    int tierset = 0;
    bool capturing = false;

    // timing
    hipEvent_t ev[8] = {};
    sdot_timings last{};
};

namespace sdot_amd {
// runs an entry point's body: HIP failures, invalid input and capacity errors
// become status codes, with the text on the handle
template <class F>
sdot_status guard(sdot_pd *pd, F &&f) {
    try {
        return f();
    } catch (const HipError &e) {
        if (pd) {
            char buf[512];
            snprintf(buf, sizeof buf, "HIP error %d (%s) in %s (line %d)", (int)e.err, hipGetErrorString(e.err), e.what,
                     e.line);
            pd->err = buf;
        }
        return SDOT_HIP_ERROR;
    } catch (const OomError &) {
        if (pd) pd->err = "device allocation failed";
        return SDOT_OOM;
    } catch (const StatusError &e) {
        if (pd) pd->err = e.msg;
        return e.st;
    } catch (const std::string &e) {
        if (pd) pd->err = e;
        return SDOT_INVALID_ARG;
    } catch (const std::bad_alloc &) {
        if (pd) pd->err = "host allocation failed";
        return SDOT_OOM;
    }
}

// grid.hip
void build_grid(sdot_pd *pd);         // bbox, keys, sort, sorted positions (positions changed)
void refresh_weights(sdot_pd *pd);    // sorted weights, per-cell and global max weight
void sort_by_keys(sdot_pd *pd, int64_t ncodes);  // counts (scan_tmp) + keys/rank -> cell_start, perm, sorted positions
GridInfo grid_params(int dim, const double *lo, const double *hi, int64_t N, double dpc);
template <class T>
void exclusive_scan(const T *in, T *out, int64_t n, T *tmp, hipStream_t s);  // out has n+1 entries
size_t scan_tmp_size(int64_t n);
// cells.hip
void run_cells(sdot_pd *pd);          // cell kernels for the current dim / precision / ball
int64_t build_csr_device(sdot_pd *pd);
// display.hip: the cells of the last pass as a legacy ASCII VTK file (pass valid)
void write_vtk(sdot_pd *pd, const char *filename, const char *field_name);  // dirac-order CSR in HBM (csr_off/csr_col/csr_val); returns slot bound
void run_cells2(sdot_pd *pd, CellArgs &a);
// copies the first HCAP3 neighbours of every row of the pass just completed to hint_col
void snapshot_hints(sdot_pd *pd);
// resume slots of the FP64 2D tiers for N cells: ensure_resume allocates them (not
// under graph capture), use_resume points the pass's arguments at them when they fit
void ensure_resume(sdot_pd *pd, int64_t N);
void use_resume(const sdot_pd *pd, CellArgs &a);
void run_cells3(sdot_pd *pd, CellArgs &a);
void fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s);
void export_csr(sdot_pd *pd, int64_t *h_off, int64_t *h_col, double *h_val, int64_t cap, int64_t *nnz);
// the position / weight derivative CSR of the last (Jacobian) pass, 2D
void export_jac(sdot_pd *pd, int64_t *h_off, int64_t *h_col, double *h_val, double *h_v, int64_t cap, int64_t *nnz);
void export_vec(sdot_pd *pd, const double *d_sorted, double *h_out);  // sorted -> dirac order, to host
void export_vec_device(sdot_pd *pd, const double *d_sorted, double *d_out);  // sorted -> dirac order, in HBM
int64_t count_offdiag(sdot_pd *pd);   // off-diagonal Hessian entries (ELL slots) of the last pass
void export_centroids(sdot_pd *pd, double *h_out);   // position + moment / mass (0 when empty), dirac order
// capi.hip
sdot_status compute_pass(sdot_pd *pd);  // grid (if needed) + weights + cell pass; flags -> status
// solver.hip
sdot_status ot_solve(sdot_pd *pd, const double *targets, double tol, int stop_rule, int max_iter, double *w_inout,
                     sdot_solve_stats *stats);
}  // namespace sdot_amd
