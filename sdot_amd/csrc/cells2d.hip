// 2D Laguerre-cell kernel: one thread per cell (sorted order, so a wavefront
// covers a compact patch of the grid and its neighbour reads hit the same
// cache lines), with the cell polygon held in that thread's LDS slot.
//
// Per cell (semantics of src/sdot/PowerDiagram/Visitors/SpZGrid.tcc:494-645):
//   * start from the domain box (ConvexPolyhedronAssembly with one box,
//     ConvexPolyhedronAssembly.tcc:64-65), in coordinates local to the dirac;
//   * visit grid cells ring by ring; a ring (and every further one) or a grid
//     cell is skipped when no dirac in it can cut the current polygon -- the
//     exact power-distance bound of SpZGrid::can_be_evicted (:475-480) taken
//     over the disc of radius R (max vertex distance) that holds the polygon;
//   * cut by the radical half-plane of each remaining candidate
//     (SpZGrid.tcc:506-522; ConvexPolyhedron2::plane_cut :687-922);
//   * in ball mode, clip by the disc of radius sqrt(w_i) streaming the
//     LINE/ARC items of ConvexPolyhedron2::ball_cut (:534-665) straight into the
//     integrals (area :1791-1842, _arc_area :2361-2380, boundary items
//     :437-478) and the Hessian row (get_der_integrals_wrt_weights.h:52-75).
// T = double (FP64 path) or float (mixed path: FP32 local coordinates; plane
// offsets, distances and the mass/Hessian outputs in FP64).
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "cell2d_impl.h"

namespace sdot_amd {

namespace {

using namespace c2;

// Persistent work-list traversal of a batch pass (CellArgs::wl_prob): the state of
// one wavefront, uniform over its lanes.  next() moves to the next 64 cells of the
// current grab, grabbing cpw more cells when it is used up (first from the
// wavefront's own segment, then from the others); false when every segment is done.
struct WaveGrab {
    int64_t cb = 0, ce = 0, j = -64, seg = 0, total = 0;
    int sx = 0, x0 = 0;
    __device__ bool next(const CellArgs &a, int lane) {
        j += 64;
        while (cb + j >= ce) {
            if (sx >= 8) return false;
            const int x = (x0 + sx) & 7;
            const int64_t s0 = x * seg, s1 = s0 + seg < total ? s0 + seg : total;
            unsigned long long got = 0;
            if (lane == 0) got = atomicAdd(a.grab + x, (unsigned long long)a.cpw);
            got = __shfl(got, 0);
            const int64_t base = s0 + (int64_t)got;
            if (base >= s1) {
                ++sx;
                continue;
            }
            cb = base;
            ce = base + a.cpw < s1 ? base + a.cpw : s1;
            j = 0;
        }
        return true;
    }
    // sorted index of the lane's cell in the current 64 (-1: none)
    __device__ int64_t cell(const CellArgs &a, int lane) const {
        const int64_t k = cb + j + lane;
        if (k >= ce) return -1;
        int64_t lo = 0, hi = a.wl_meta[0];   // wl_off[lo] <= k < wl_off[hi]
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (a.wl_off[mid] <= k) lo = mid;
            else hi = mid;
        }
        return a.pb_off[a.wl_prob[lo]] + (k - a.wl_off[lo]);
    }
};

// One thread per cell, polygon in the thread's LDS slot (CAP vertices).  Without
// a work list: every cell, in grid order.  With one: the cells listed, count in
// device memory, grid-stride.  With the persistent list of a batch (BATCH,
// wl_prob): a grid of resident wavefronts that grab runs of cells.  A cell that
// outgrows CAP goes to ovf_list for the next tier; in the last tier (no ovf_list)
// it raises FLAG_CELL_OVERFLOW.
// WAVES: the wavefronts per SIMD the register allocation must allow.
template <class T, int CAP, int BS, bool BALL, bool IMG, int WAVES, bool BATCH, bool PER, bool JAC = false, int RF = 0>
__global__ __launch_bounds__(BS, WAVES) void k_cells2(CellArgs a) {
    __shared__ T sxy[2 * CAP * BS];   // x and y of a vertex side by side (one paired LDS access)
    __shared__ int sid[CAP * BS];
    // one call site of the cell body (so it is inlined): without a work list the
    // loop runs once
    int64_t k, count, step;
    const bool grab = BATCH && a.wl_prob != nullptr && a.work == nullptr;   // first tier of a batch pass
    WaveGrab wg;
    const int lane = threadIdx.x & 63;
    if (grab) {
        k = 0;
        count = 1;
        step = 0;
        wg.total = a.wl_meta[1];
        wg.seg = (wg.total + 7) / 8;
        wg.x0 = blockIdx.x & 7;
    } else if (a.work) {
        k = blockIdx.x * (int64_t)BS + threadIdx.x;
        count = *a.work_count;
        step = (int64_t)gridDim.x * BS;
    } else {
        k = xcd_tile(blockIdx.x, gridDim.x) * BS + threadIdx.x;
        count = a.N;
        step = a.N;
    }
    for (; k < count; k += step) {
        int64_t t;
        if (grab) {
            if (!wg.next(a, lane)) break;   // uniform over the wavefront
            t = wg.cell(a, lane);
        } else {
            t = a.work ? a.work[k] : k;
        }
        const bool ovf = t >= 0 && cell2_body<T, CAP, BALL, IMG, BATCH, PER, JAC, RF>(
                                       a, t, sxy + 2 * threadIdx.x, sxy + 2 * threadIdx.x + 1, sid + threadIdx.x,
                                       2 * BS, BS, a.work ? k : -1);
        if (a.ovf_list) {
            // the wavefront's overflowing cells go to the list together, in lane (hence
            // grid) order: the next tier's wavefronts then read runs of nearby cells
            const unsigned long long m = __ballot(ovf);
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                unsigned base = 0;
                if (lane == leader) base = atomicAdd(a.ovf_count, (unsigned)__popcll(m));
                base = __shfl(base, leader);
                const unsigned q = base + __popcll(m & ((1ull << lane) - 1));
                if (ovf) {
                    a.ovf_list[q] = (int)t;
                    if (rs_save<T, CAP>() && a.rs_hdr && (int64_t)q < a.rs_cap) {
                        // the polygon before the cut that did not fit is still in the slot
                        const T *xy = sxy + 2 * threadIdx.x;
                        const int *id = sid + threadIdx.x;
                        double *d = a.rs_d + 2 * RS_CAP * (int64_t)q;
                        int *e = a.rs_i + RS_CAP * (int64_t)q;
                        for (int f = 0; f < CAP; ++f) {
                            d[2 * f] = (double)xy[f * 2 * BS];
                            d[2 * f + 1] = (double)xy[f * 2 * BS + 1];
                            e[f] = id[f * BS];
                        }
                    }
                }
            }
        } else if (ovf) {
            flag_or(a.flags, FLAG_CELL_OVERFLOW);
        }
    }
}

// ---- dirac-order export ----
// sum of the row lengths (off-diagonal Hessian entries of the pass)
__global__ __launch_bounds__(256) void k_sum_rows(const int *row_len, int64_t n, unsigned long long *out) {
    __shared__ unsigned long long sm[256];
    unsigned long long acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) acc += row_len[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, sm[0]);
}

__global__ void k_export_vec(const double *v, const int *inv_perm, int64_t n, double *out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = v[inv_perm[i]];
}

__global__ void k_csr_count(const int *row_len, const int *inv_perm, int64_t n, int64_t *cnt) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = row_len[inv_perm[i]] + 1;
}

// one row per thread, threads in grid order (the ELL slots [e * stride + t] of a
// wavefront are contiguous): the row (columns mapped to dirac indices) and the diagonal,
// sorted by column with images of one dirac merged, written as the dirac-order row
// i = perm[t] at its CSR offset (row_len + 1 places; merged duplicates leave padding,
// marked by column -1).  Rows of up to R entries are sorted in registers (insertion with
// static indices, no LDS: high occupancy for the gathers), as (column, value) pairs for
// R <= 16, as packed (column, slot) keys above (values read when written); longer rows
// are written unsorted and sorted in place in global memory.
template <int R>
__global__ __launch_bounds__(256) void k_csr_fill(const int *row_len, const int *ell_col, const double *ell_val,
                                                  int64_t stride, const double *diag, const int *perm, int64_t n,
                                                  const int64_t *off, int64_t *cols, double *vals, int64_t *mlen) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = perm[t];
    const int L = row_len[t];
    const int64_t o = off[i];
    const int m = L + 1;
    if (R <= 16 && m <= R) {
        // columns and values in registers: the value loads stay coalesced over the wavefront
        int c[R];
        double v[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            c[k] = INT_MAX;   // dirac indices are below 2^31 - 1
            v[k] = 0;
        }
        auto insert = [&](int cc, double vv) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int oc = c[k];
                const double ov = v[k];
                const bool sw = oc > cc;
                c[k] = sw ? cc : oc;
                v[k] = sw ? vv : ov;
                cc = sw ? oc : cc;
                vv = sw ? ov : vv;
            }
        };
        insert((int)i, diag[t]);
        // slots four at a time: their column, permutation and value loads are issued
        // together (slots past the row re-read slot e0 and insert the key INT_MAX, which
        // leaves the row as it is), so a row waits on two load latencies per four entries
        for (int e0 = 0; e0 < L; e0 += 4) {
            int cl[4];
            double vl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t e = e0 + u < L ? e0 + u : e0;
                cl[u] = ell_col[e * stride + t];
                vl[u] = ell_val[e * stride + t];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) cl[u] = perm[cl[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) insert(e0 + u < L ? cl[u] : INT_MAX, vl[u]);
        }
        int w = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k < m) {
                if (k + 1 < R && k + 1 < m && c[k] == c[k + 1]) {
                    v[k + 1] += v[k];
                } else {
                    cols[o + w] = c[k];
                    vals[o + w] = v[k];
                    ++w;
                }
            }
        }
        for (int k = w; k < m; ++k) {
            cols[o + k] = -1;
            vals[o + k] = 0;
        }
        if (mlen) mlen[i] = w;
        return;
    }
    if (R > 16 && m <= R) {
        // keys column << 8 | ELL slot (slot 255: the diagonal; ELL rows hold at most 255)
        uint64_t c[R];
#pragma unroll
        for (int k = 0; k < R; ++k) c[k] = ~0ull;
        auto insert = [&](uint64_t key) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint64_t ok = c[k];
                const bool sw = ok > key;
                c[k] = sw ? key : ok;
                key = sw ? ok : key;
            }
        };
        insert((uint64_t)i << 8 | 255u);
        for (int e0 = 0; e0 < L; e0 += 4) {   // as above; the key ~0 leaves the row as it is
            int cl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) cl[u] = ell_col[(int64_t)(e0 + u < L ? e0 + u : e0) * stride + t];
#pragma unroll
            for (int u = 0; u < 4; ++u) cl[u] = perm[cl[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                insert(e0 + u < L ? (uint64_t)cl[u] << 8 | (unsigned)(e0 + u) : ~0ull);
        }
        int w = 0;
        int64_t pc = -1;
        double acc = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k < m) {
                const int64_t col = (int64_t)(c[k] >> 8);
                const int e = (int)(c[k] & 255u);
                const double val = e == 255 ? diag[t] : ell_val[e * stride + t];
                if (col == pc) {
                    acc += val;
                } else {
                    if (pc >= 0) {
                        cols[o + w] = pc;
                        vals[o + w] = acc;
                        ++w;
                    }
                    pc = col;
                    acc = val;
                }
            }
        }
        cols[o + w] = pc;
        vals[o + w] = acc;
        ++w;
        for (int k = w; k < m; ++k) {
            cols[o + k] = -1;
            vals[o + k] = 0;
        }
        if (mlen) mlen[i] = w;
        return;
    }
    // long row: in place
    int64_t *c = cols + o;
    double *v = vals + o;
    c[0] = i;
    v[0] = diag[t];
    for (int e = 0; e < L; ++e) {
        const int64_t cc = perm[ell_col[e * stride + t]];
        const double vv = ell_val[e * stride + t];
        int k = e + 1;
        while (k > 0 && c[k - 1] > cc) {
            c[k] = c[k - 1];
            v[k] = v[k - 1];
            --k;
        }
        c[k] = cc;
        v[k] = vv;
    }
    int w = 0;
    for (int k = 0; k < m; ++k) {
        if (w > 0 && c[w - 1] == c[k]) {
            v[w - 1] += v[k];
            continue;
        }
        c[w] = c[k];
        v[w] = v[k];
        ++w;
    }
    for (int k = w; k < m; ++k) {
        c[k] = -1;
        v[k] = 0;
    }
    if (mlen) mlen[i] = w;
}

// merged rows moved to their final offsets (rows whose duplicates merged leave
// trailing padding in the fill's layout), one row per thread
__global__ __launch_bounds__(256) void k_csr_squeeze(const int64_t *off, const int64_t *cols, const double *vals,
                                                     const int64_t *off2, int64_t n, int64_t *cols2, double *vals2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = off[i], o2 = off2[i], m = off2[i + 1] - o2;
    for (int64_t k = 0; k < m; ++k) {
        cols2[o2 + k] = cols[o + k];
        vals2[o2 + k] = vals[o + k];
    }
}

// workgroups of one kernel that the device holds at once (the grid of a persistent launch)
int resident_grid(const void *kernel, int block) {
    int dev = 0, cus = 0, per_cu = 0;
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
    return std::max(1, cus * std::max(1, per_cu));
}

template <class T, int CAP, int BS, bool BALL, bool IMG, int WAVES, bool BATCH, bool PER = false, bool JAC = false,
          int RF = 0>
void launch2(const CellArgs &a, unsigned nb, hipStream_t s) {
    if (nb == 0) return;
    auto kern = k_cells2<T, CAP, BS, BALL, IMG, WAVES, BATCH, PER, JAC, RF>;
    if (BATCH && a.wl_prob && !a.work) {
        static const int resident = resident_grid(reinterpret_cast<const void *>(kern), BS);   // per instantiation
        nb = (unsigned)resident;
    }
    hipLaunchKernelGGL(kern, dim3(nb), dim3(BS), 0, s, a);
    HIPCK(hipGetLastError());
}

// batches carry no image and no translations; ball clipping over pieces (pixels or
// several domain items) clips each piece by the disc
template <class T, int CAP, int WAVES, int BS = 64>
void dispatch2(bool ball, bool img, const CellArgs &a, unsigned nb, hipStream_t s) {
    if constexpr (std::is_same<T, float>::value && CAP == 8) {   // a batch-only tier
        if (ball) launch2<T, CAP, BS, true, false, WAVES, true>(a, nb, s);
        else launch2<T, CAP, BS, false, false, WAVES, true>(a, nb, s);
        return;
    }
    if (a.prob) {
        if (ball) launch2<T, CAP, BS, true, false, WAVES, true>(a, nb, s);
        else launch2<T, CAP, BS, false, false, WAVES, true>(a, nb, s);
    } else if (a.ntrans > 0) {
        if (ball && img) launch2<T, CAP, BS, true, true, WAVES, false, true>(a, nb, s);
        else if (ball) launch2<T, CAP, BS, true, false, WAVES, false, true>(a, nb, s);
        else if (img) launch2<T, CAP, BS, false, true, WAVES, false, true>(a, nb, s);
        else launch2<T, CAP, BS, false, false, WAVES, false, true>(a, nb, s);
    } else {
        if (ball && img) launch2<T, CAP, BS, true, true, WAVES, false>(a, nb, s);
        else if (ball) launch2<T, CAP, BS, true, false, WAVES, false>(a, nb, s);
        else if (img) launch2<T, CAP, BS, false, true, WAVES, false>(a, nb, s);
        else launch2<T, CAP, BS, false, false, WAVES, false>(a, nb, s);
    }
}

}  // namespace

namespace {
// the first N / 4 + 1024 cells that outgrow the FP64 8-vertex tier get a resume slot
// (about 9 % of the cells of a 1M ball-clipped pass outgrow it); the others start again
int64_t resume_cap(int64_t N) { return std::min<int64_t>(N, N / 4 + 1024); }
}  // namespace

void ensure_resume(sdot_pd *pd, int64_t N) {
    const int64_t C = resume_cap(N);
    pd->rs_hdr.ensure(4 * (size_t)N);
    pd->rs_i.ensure((size_t)RS_CAP * C);
    pd->rs_d.ensure((size_t)(2 * RS_CAP) * C);
}

void use_resume(const sdot_pd *pd, CellArgs &a) {
    const int64_t C = resume_cap(a.N);
    if (pd->rs_hdr.p && pd->rs_i.p && pd->rs_d.p && pd->rs_hdr.n >= 4 * (size_t)a.N &&
        pd->rs_i.n >= (size_t)RS_CAP * C && pd->rs_d.n >= (size_t)(2 * RS_CAP) * C) {
        a.rs_hdr = pd->rs_hdr.p;
        a.rs_i = pd->rs_i.p;
        a.rs_d = pd->rs_d.p;
        a.rs_cap = C;
    } else {
        a.rs_hdr = nullptr;
        a.rs_i = nullptr;
        a.rs_d = nullptr;
        a.rs_cap = 0;
    }
}

// Capacity tiers, all on the stream with no host round trip: every cell with a
// small polygon slot in LDS, the cells that outgrew it again with a larger one,
// and so on up to 64 vertices (a cell beyond that raises FLAG_CELL_OVERFLOW).  A
// 2D power cell averages about 6 vertices.  Mixed: 16 vertices (FP32 slots,
// three wavefronts per SIMD), then 64.  FP64: 8 vertices (four wavefronts per
// SIMD; a few percent of the cells outgrow it while they are being clipped),
// then 16, then 64.
void run_cells2(sdot_pd *pd, CellArgs &a) {
    hipStream_t s = pd->stream;
    const bool ball = pd->ball != 0;
    const bool mixed = pd->precision == SDOT_MIXED;
    const bool img = a.img != nullptr || a.nitems >= 2;   // pieces: pixels or several domain items
    const int64_t N = a.N;
    const unsigned nb_all = (unsigned)((N + 63) / 64);
    int *lists[2] = {pd->ovf_list.p, pd->ovf_list.p + N + 1};
    int tier = 0;
    // tier `tier` reads the previous tier's list (none for the first) and appends to
    // its own unless it is the last
    auto next = [&](bool last) {
        a.work = tier == 0 ? nullptr : lists[(tier - 1) & 1];
        a.work_count = tier == 0 ? nullptr : pd->flags.p + 1 + ((tier - 1) & 1);
        a.ovf_list = last ? nullptr : lists[tier & 1];
        a.ovf_count = last ? nullptr : pd->flags.p + 1 + (tier & 1);
        ++tier;
        return tier == 1 ? nb_all : std::min(nb_all, 4096u);
    };
    if (a.rf) {
        // R2 / WmR2 / ExpWmR2db integration functions: one FP64 tier of 64 vertices (pieces:
        // constant pixels or several shapes; ExpWmR2db has no ball cut)
        unsigned nb = next(true);
        auto go = [&](auto rfc) {
            constexpr int RF = decltype(rfc)::value;
            constexpr bool BL = RF != 3;   // instantiations with the ball clip
            if (a.ntrans > 0) {
                if (BL && ball) launch2<double, 64, 64, BL, false, 1, false, true, false, RF>(a, nb, s);
                else if (img) launch2<double, 64, 64, false, true, 1, false, true, false, RF>(a, nb, s);
                else launch2<double, 64, 64, false, false, 1, false, true, false, RF>(a, nb, s);
            } else {
                if (BL && ball) launch2<double, 64, 64, BL, false, 1, false, false, false, RF>(a, nb, s);
                else if (img) launch2<double, 64, 64, false, true, 1, false, false, false, RF>(a, nb, s);
                else launch2<double, 64, 64, false, false, 1, false, false, false, RF>(a, nb, s);
            }
        };
        if (a.rf == 1) go(std::integral_constant<int, 1>());
        else if (a.rf == 2) go(std::integral_constant<int, 2>());
        else go(std::integral_constant<int, 3>());
    } else if (a.jdiag) {
        // position / weight derivatives: one FP64 tier of 64 vertices (pieces: constant
        // pixels or several shapes, no ball)
        unsigned nb = next(true);
        if (a.ntrans > 0) {
            if (ball) launch2<double, 64, 64, true, false, 1, false, true, true>(a, nb, s);
            else if (img) launch2<double, 64, 64, false, true, 1, false, true, true>(a, nb, s);
            else launch2<double, 64, 64, false, false, 1, false, true, true>(a, nb, s);
        } else {
            if (ball) launch2<double, 64, 64, true, false, 1, false, false, true>(a, nb, s);
            else if (img) launch2<double, 64, 64, false, true, 1, false, false, true>(a, nb, s);
            else launch2<double, 64, 64, false, false, 1, false, false, true>(a, nb, s);
        }
    } else if (mixed) {
        unsigned nb = next(false);
        if (a.prob && pd->tierset == 1) {   // batch option: 8-vertex FP32 slots first
            dispatch2<float, 8, 4>(ball, img, a, nb, s);
            nb = next(false);
        }
        if (a.work) dispatch2<float, 16, 3, 32>(ball, img, a, nb, s);   // a second tier: 32 cells a wavefront
        else dispatch2<float, 16, 3>(ball, img, a, nb, s);
        nb = next(true);
        dispatch2<float, 64, 1>(ball, img, a, nb, s);
    } else {
        unsigned nb = next(false);
        if (!(a.prob && pd->tierset == 1)) {   // batch option: start at 16-vertex slots
            dispatch2<double, 8, 4>(ball, img, a, nb, s);
            nb = next(false);
        }
        // the cells that outgrew the first tier run 32 to a wavefront (the wavefront's time
        // is its slowest lanes' chain); a batch's persistent first tier keeps 64
        if (a.work) dispatch2<double, 16, 2, 32>(ball, img, a, nb, s);
        else dispatch2<double, 16, 2>(ball, img, a, nb, s);
        nb = next(true);
        dispatch2<double, 64, 1>(ball, img, a, nb, s);
    }
    if (pd->ev[3] && !pd->capturing) HIPCK(hipEventRecord(pd->ev[3], s));
    a.work = nullptr;
    a.work_count = nullptr;
}

namespace {
// get_centroids.h:53-58: centroid / (mass + (mass == 0)); the moments are about the dirac
__global__ void k_export_centroids(const double *spos, const double *mass, const double *mom, const int *inv_perm,
                                   int64_t n, int dim, double *out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t t = inv_perm[i];
    const double m = mass[t];
    for (int d = 0; d < dim; ++d) out[dim * i + d] = m != 0 ? spos[dim * t + d] + mom[dim * t + d] / m : 0.0;
}
}  // namespace

namespace {
// Export of get_der_centroids_and_integrals_wrt_weight_and_positions
// (get_der_centroids_and_integrals_wrt_weight_and_positions.h:167-226): per sorted row t
// (dirac i = perm[t]) the items -- neighbour slots and the diagonal -- sorted by dirac
// column, duplicates (images of one dirac) merged.
constexpr int JMAX = 256;   // up to 255 ELL slots + the diagonal

__device__ int jac_items(const int *row_len, const int *ell_col, int64_t stride, const int *perm, int64_t t, int *c,
                         int *slot) {
    const int L = row_len[t];
    int m = 0;
    c[m] = perm[t];
    slot[m++] = -1;
    for (int e = 0; e < L && m < JMAX; ++e) {
        c[m] = perm[ell_col[e * stride + t]];
        slot[m++] = e;
    }
    for (int a = 1; a < m; ++a)
        for (int b = a; b > 0 && c[b - 1] > c[b]; --b) {
            const int tc = c[b], ts = slot[b];
            c[b] = c[b - 1];
            slot[b] = slot[b - 1];
            c[b - 1] = tc;
            slot[b - 1] = ts;
        }
    return m;
}

// cnt[i] = nb x the number of distinct columns of dirac i (nb = (dim + 1)^2)
__global__ void k_jac_count(const int *row_len, const int *ell_col, int64_t stride, const int *perm, int64_t n, int nb,
                            int64_t *cnt) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    int c[JMAX], slot[JMAX];
    const int m = jac_items(row_len, ell_col, stride, perm, t, c, slot);
    int u = 0;
    for (int a = 0; a < m; ++a) u += a == 0 || c[a] != c[a - 1];
    cnt[perm[t]] = nb * (int64_t)u;
}

// rows NU i + d (centroid coordinates, then the mass), columns NU m + e (position
// coordinates, then the weight), NU = DIM + 1; the centroid rows get the mass /
// centroid correction (:172-177: (b - b_mass * centroid) / (mass + (mass == 1)), the
// reference's own divisor).  Blocks are NU x NU at NU * NU doubles per slot.
template <int DIM>
__global__ void k_jac_fill(const int *row_len, const int *ell_col, const double *jblk, const double *jdiag, int64_t stride,
                           const double *spos, const double *mass, const double *mom, const int *perm, int64_t n,
                           const int64_t *base, int64_t *cols, double *vals) {
    constexpr int NU = DIM + 1, NB = NU * NU;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    int c[JMAX], slot[JMAX];
    const int m = jac_items(row_len, ell_col, stride, perm, t, c, slot);
    int u = 0;
    for (int a = 0; a < m; ++a) u += a == 0 || c[a] != c[a - 1];
    const double ms = mass[t];
    double v[DIM];
    for (int d = 0; d < DIM; ++d) v[d] = ms != 0 ? spos[DIM * t + d] + mom[DIM * t + d] / ms : 0.0;
    const double den = ms + (ms == 1 ? 1.0 : 0.0);
    const int64_t b0 = base[perm[t]];
    int k = 0;
    for (int a = 0; a < m;) {
        const int col = c[a];
        double B[NB];
        for (int q = 0; q < NB; ++q) B[q] = 0;
        for (; a < m && c[a] == col; ++a)
            for (int q = 0; q < NB; ++q)
                B[q] += slot[a] < 0 ? jdiag[q * stride + t] : jblk[((int64_t)slot[a] * NB + q) * stride + t];
        for (int r = 0; r < DIM; ++r)
            for (int cc = 0; cc < NU; ++cc) B[NU * r + cc] = (B[NU * r + cc] - B[NU * DIM + cc] * v[r]) / den;
        for (int d = 0; d < NU; ++d) {
            const int64_t p = b0 + (int64_t)d * NU * u + NU * k;
            for (int e = 0; e < NU; ++e) {
                cols[p + e] = NU * (int64_t)col + e;
                vals[p + e] = B[NU * d + e];
            }
        }
        ++k;
    }
}
}  // namespace

void export_jac(sdot_pd *pd, int64_t *h_off, int64_t *h_col, double *h_val, double *h_v, int64_t cap, int64_t *nnz) {
    const int64_t N = pd->N;
    const int dim = pd->dim, nu = dim + 1;
    hipStream_t s = pd->stream;
    pd->csr_cnt.ensure(N);
    pd->csr_off.ensure(N + 1);
    pd->csr_tmp.ensure(scan_tmp_size(N));
    const unsigned nb = (unsigned)((N + 127) / 128);
    hipLaunchKernelGGL(k_jac_count, dim3(nb), dim3(128), 0, s, pd->row_len.p, pd->ell_col.p, (int64_t)N, pd->perm.p, N,
                       nu * nu, pd->csr_cnt.p);
    HIPCK(hipGetLastError());
    exclusive_scan<int64_t>(pd->csr_cnt.p, pd->csr_off.p, N, pd->csr_tmp.p, s);
    std::vector<int64_t> base(N + 1);
    HIPCK(hipMemcpyAsync(base.data(), pd->csr_off.p, (N + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    const int64_t total = base[N];
    *nnz = total;
    if (h_v) {
        std::vector<double> m(N), c(dim * N);
        export_vec(pd, pd->mass.p, m.data());
        export_centroids(pd, c.data());
        for (int64_t i = 0; i < N; ++i) {
            for (int d = 0; d < dim; ++d) h_v[nu * i + d] = c[dim * i + d];
            h_v[nu * i + dim] = m[i];
        }
    }
    if (!h_col) return;
    if (total > cap) throw std::string("Jacobian capacity too small");
    for (int64_t i = 0; i < N; ++i) {
        const int64_t row = (base[i + 1] - base[i]) / nu;
        for (int d = 0; d < nu; ++d) h_off[nu * i + d] = base[i] + d * row;
    }
    h_off[nu * N] = total;
    pd->csr_col.ensure(total > 0 ? total : 1);
    pd->csr_val.ensure(total > 0 ? total : 1);
    if (dim == 2)
        hipLaunchKernelGGL(k_jac_fill<2>, dim3(nb), dim3(128), 0, s, pd->row_len.p, pd->ell_col.p, pd->jblk.p,
                           pd->jdiag.p, (int64_t)N, pd->spos.p, pd->mass.p, pd->mom.p, pd->perm.p, N, pd->csr_off.p,
                           pd->csr_col.p, pd->csr_val.p);
    else
        hipLaunchKernelGGL(k_jac_fill<3>, dim3(nb), dim3(128), 0, s, pd->row_len.p, pd->ell_col.p, pd->jblk.p,
                           pd->jdiag.p, (int64_t)N, pd->spos.p, pd->mass.p, pd->mom.p, pd->perm.p, N, pd->csr_off.p,
                           pd->csr_col.p, pd->csr_val.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(h_col, pd->csr_col.p, total * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(h_val, pd->csr_val.p, total * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
}

void export_centroids(sdot_pd *pd, double *h_out) {
    const int64_t N = pd->N;
    const int dim = pd->dim;
    if (N == 0) return;
    pd->out_tmp.ensure((size_t)dim * N);
    hipLaunchKernelGGL(k_export_centroids, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, pd->stream, pd->spos.p,
                       pd->mass.p, pd->mom.p, pd->inv_perm.p, N, dim, pd->out_tmp.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(h_out, pd->out_tmp.p, (size_t)dim * N * sizeof(double), hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
}

int64_t count_offdiag(sdot_pd *pd) {
    const int64_t N = pd->N;
    if (N <= 0 || !pd->row_len.p) return 0;
    pd->wctr.ensure(3 * WCTR);   // its first slot is free between passes
    unsigned long long *acc = pd->wctr.p;
    fill_u32(reinterpret_cast<uint32_t *>(acc), 2, 0u, pd->stream);
    const unsigned nb = (unsigned)std::min<int64_t>(1024, (N + 255) / 256);
    hipLaunchKernelGGL(k_sum_rows, dim3(nb), dim3(256), 0, pd->stream, pd->row_len.p, N, acc);
    HIPCK(hipGetLastError());
    unsigned long long h = 0;
    HIPCK(hipMemcpyAsync(&h, acc, sizeof h, hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
    return (int64_t)h;
}

void export_vec_device(sdot_pd *pd, const double *d_sorted, double *d_out) {
    const int64_t N = pd->N;
    hipLaunchKernelGGL(k_export_vec, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, pd->stream, d_sorted, pd->inv_perm.p,
                       N, d_out);
    HIPCK(hipGetLastError());
}

void export_vec(sdot_pd *pd, const double *d_sorted, double *h_out) {
    const int64_t N = pd->N;
    pd->out_tmp.ensure(N);
    hipLaunchKernelGGL(k_export_vec, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, pd->stream, d_sorted, pd->inv_perm.p,
                       N, pd->out_tmp.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(h_out, pd->out_tmp.p, N * sizeof(double), hipMemcpyDeviceToHost, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
}

int64_t build_csr_device(sdot_pd *pd) {
    const int64_t N = pd->N;
    hipStream_t s = pd->stream;
    pd->csr_cnt.ensure(N);
    pd->csr_off.ensure(N + 1);
    pd->csr_tmp.ensure(scan_tmp_size(N));
    const unsigned nb = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_csr_count, dim3(nb), dim3(256), 0, s, pd->row_len.p, pd->inv_perm.p, N, pd->csr_cnt.p);
    HIPCK(hipGetLastError());
    exclusive_scan<int64_t>(pd->csr_cnt.p, pd->csr_off.p, N, pd->csr_tmp.p, s);
    // upper bound of the entries: every ELL slot of every row plus the diagonal
    const int64_t bound = (int64_t)N * (pd->ell_cap + 1);
    pd->csr_col.ensure(bound);
    pd->csr_val.ensure(bound);
    // Only periodic images put one dirac twice in a row (a face with j and one with an
    // image of j; without translations each neighbour bounds at most one face): then
    // the fill records each row's merged length and the rows are squeezed to their
    // final offsets, so the returned rows never carry padding.
    const bool dup = pd->htrans.size() > 0;
    int64_t *mlen = dup ? pd->csr_cnt.p : nullptr;   // the counts are consumed by the scan above
    // register rows of 16 entries in 2D (about 7 on average), 32 in 3D (about 16)
    if (pd->dim == 2)
        hipLaunchKernelGGL((k_csr_fill<16>), dim3(nb), dim3(256), 0, s, pd->row_len.p, pd->ell_col.p, pd->ell_val.p,
                           (int64_t)N, pd->diag.p, pd->perm.p, N, pd->csr_off.p, pd->csr_col.p, pd->csr_val.p, mlen);
    else
        hipLaunchKernelGGL((k_csr_fill<32>), dim3(nb), dim3(256), 0, s, pd->row_len.p, pd->ell_col.p, pd->ell_val.p,
                           (int64_t)N, pd->diag.p, pd->perm.p, N, pd->csr_off.p, pd->csr_col.p, pd->csr_val.p, mlen);
    HIPCK(hipGetLastError());
    if (dup) {
        pd->csr_off2.ensure(N + 1);
        exclusive_scan<int64_t>(pd->csr_cnt.p, pd->csr_off2.p, N, pd->csr_tmp.p, s);
        pd->csr_col2.ensure(bound);
        pd->csr_val2.ensure(bound);
        hipLaunchKernelGGL(k_csr_squeeze, dim3(nb), dim3(256), 0, s, pd->csr_off.p, pd->csr_col.p, pd->csr_val.p,
                           pd->csr_off2.p, N, pd->csr_col2.p, pd->csr_val2.p);
        HIPCK(hipGetLastError());
        std::swap(pd->csr_off, pd->csr_off2);
        std::swap(pd->csr_col, pd->csr_col2);
        std::swap(pd->csr_val, pd->csr_val2);
    }
    return bound;
}

void export_csr(sdot_pd *pd, int64_t *h_off, int64_t *h_col, double *h_val, int64_t cap, int64_t *nnz) {
    const int64_t N = pd->N;
    hipStream_t s = pd->stream;
    build_csr_device(pd);
    int64_t total = 0;
    HIPCK(hipMemcpyAsync(&total, pd->csr_off.p + N, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    *nnz = total;
    if (!h_col) return;
    if (total > cap) throw std::string("CSR capacity too small");
    HIPCK(hipMemcpyAsync(h_off, pd->csr_off.p, (N + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(h_col, pd->csr_col.p, total * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(h_val, pd->csr_val.p, total * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
}

}  // namespace sdot_amd
