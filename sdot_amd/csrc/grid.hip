// Device grid: bounding box, Morton keys, counting (single-digit radix) sort of
// the diracs into grid cells, sorted positions/weights and per-cell maximum
// weights -- the device counterpart of SpZGrid::update / update_box
// (src/sdot/PowerDiagram/Visitors/SpZGrid.tcc:24-187: bounding boxes :65-71,
// per-box max weight :123-143), laid out as a flat grid instead of a tree.
#include <cfloat>
#include <cmath>

#include "internal.h"

namespace sdot_amd {

namespace {

constexpr int RED_BLOCKS = BBOX_BLOCKS;
constexpr int BS = 256;

template <int DIM>
__global__ __launch_bounds__(BS) void k_bbox(const double *pos, int64_t n, double *part, unsigned *flags) {
    __shared__ double smin[DIM][BS], smax[DIM][BS];
    double mn[DIM], mx[DIM];
    for (int d = 0; d < DIM; ++d) { mn[d] = DBL_MAX; mx[d] = -DBL_MAX; }
    bool bad = false;
    for (int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
        for (int d = 0; d < DIM; ++d) {
            double x = pos[i * DIM + d];
            if (!isfinite(x)) bad = true;
            mn[d] = fmin(mn[d], x);
            mx[d] = fmax(mx[d], x);
        }
    }
    if (bad) atomicOr(flags, FLAG_BAD_INPUT);
    for (int d = 0; d < DIM; ++d) { smin[d][threadIdx.x] = mn[d]; smax[d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < DIM; ++d) {
                smin[d][threadIdx.x] = fmin(smin[d][threadIdx.x], smin[d][threadIdx.x + s]);
                smax[d][threadIdx.x] = fmax(smax[d][threadIdx.x], smax[d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < DIM; ++d) {
            part[blockIdx.x * 2 * DIM + d] = smin[d][0];
            part[blockIdx.x * 2 * DIM + DIM + d] = smax[d][0];
        }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

template <int DIM>
__global__ __launch_bounds__(BS) void k_keys(const double *pos, int64_t n, GridInfo g, uint32_t *keys, uint32_t *rank,
                                             uint32_t *cnt) {
    int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (i >= n) return;
    uint32_t code;
    if (DIM == 2)
        code = morton2(grid_coord(g, 0, pos[2 * i]), grid_coord(g, 1, pos[2 * i + 1]));
    else
        code = morton3(grid_coord(g, 0, pos[3 * i]), grid_coord(g, 1, pos[3 * i + 1]), grid_coord(g, 2, pos[3 * i + 2]));
    keys[i] = code;
    rank[i] = atomicAdd(&cnt[code], 1u);
}

__global__ __launch_bounds__(BS) void k_scatter(const uint32_t *keys, const uint32_t *rank, const uint32_t *start,
                                                int64_t n, int *perm) {
    int64_t i = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (i >= n) return;
    perm[start[keys[i]] + rank[i]] = (int)i;
}

// deterministic order inside each grid cell: ascending dirac index
__global__ __launch_bounds__(BS) void k_binsort(const uint32_t *start, int64_t ncodes, int *perm) {
    int64_t c = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (c >= ncodes) return;
    uint32_t s = start[c], e = start[c + 1];
    for (uint32_t k = s + 1; k < e; ++k) {
        int v = perm[k];
        uint32_t m = k;
        while (m > s && perm[m - 1] > v) {
            perm[m] = perm[m - 1];
            --m;
        }
        perm[m] = v;
    }
}

template <int DIM>
__global__ __launch_bounds__(BS) void k_gather_pos(const double *pos, const int *perm, int64_t n, double *spos,
                                                   int *inv_perm) {
    int64_t p = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (p >= n) return;
    int i = perm[p];
    for (int d = 0; d < DIM; ++d) spos[p * DIM + d] = pos[(int64_t)i * DIM + d];
    inv_perm[i] = (int)p;
}

// Sorted weights.  Each weight gets a deterministic offset of at most `tie`
// (about 1e-12 of the squared dirac spacing, a hash of the dirac index): the
// diagram stays a consistent power diagram, moved by ~1e-12 relative, and
// exact ties (lattices: four or more planes through one point, three planes
// through one line) no longer reach the FP64 clipping.
__device__ inline double tie_unit(int i) {   // in [0, 1)
    uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// pyramid level k from level k-1: node n = max of its 2^dim children (Morton order)
__global__ __launch_bounds__(BS) void k_pyr_level(const double *fine, int64_t nnodes, int children, double *coarse) {
    const int64_t n = blockIdx.x * (int64_t)BS + threadIdx.x;
    if (n >= nnodes) return;
    double m = -DBL_MAX;
    for (int c = 0; c < children; ++c) m = fmax(m, fine[n * children + c]);
    coarse[n] = m;
}

// One thread per grid cell: its sorted weights (gathered through the permutation,
// with their tie offsets), their maximum, and the maximum over all cells (one atomic
// per block).  The cell's diracs are taken four at a time so that their permutation
// and weight loads are issued together.
__global__ __launch_bounds__(BS) void k_weights(const uint32_t *start, int64_t ncodes, const double *w, const int *perm,
                                                double *sw, double *cell_wmax, unsigned long long *wglob,
                                                unsigned *flags, double tie) {
    __shared__ double sm[BS / 64];
    const int64_t c = blockIdx.x * (int64_t)BS + threadIdx.x;
    double m = -DBL_MAX;
    bool bad = false;
    if (c < ncodes) {
        const uint32_t s = start[c], e = start[c + 1];
        for (uint32_t k0 = s; k0 < e; k0 += 4) {
            int i[4];
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) i[u] = perm[k0 + u < e ? k0 + u : k0];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = w[i[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k0 + u < e) {
                    bad |= !isfinite(v[u]);
                    const double x = v[u] + tie * tie_unit(i[u]);
                    sw[k0 + u] = x;
                    m = fmax(m, x);
                }
            }
        }
        cell_wmax[c] = m;
    }
    if (bad) atomicOr(flags, FLAG_BAD_INPUT);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = sm[0];
        for (int q = 1; q < BS / 64; ++q) b = fmax(b, sm[q]);
        if (b > -DBL_MAX) atomicMax(wglob, dbl_order(b));
    }
}

// ---- exclusive scan (out[0] = 0, out[k+1] = in[0] + ... + in[k]) ----
constexpr int SCAN_BS = 256, SCAN_IPT = 8, SCAN_E = SCAN_BS * SCAN_IPT;

template <class T>
__global__ __launch_bounds__(SCAN_BS) void k_scan_block(const T *in, T *out, int64_t n, T *block_sums, int write_total) {
    __shared__ T sm[SCAN_BS];
    int64_t base = blockIdx.x * (int64_t)SCAN_E + threadIdx.x * SCAN_IPT;
    T v[SCAN_IPT];
    T tot = 0;
    for (int k = 0; k < SCAN_IPT; ++k) {
        v[k] = base + k < n ? in[base + k] : T(0);
        tot += v[k];
    }
    sm[threadIdx.x] = tot;
    __syncthreads();
    for (int s = 1; s < SCAN_BS; s <<= 1) {
        T a = threadIdx.x >= s ? sm[threadIdx.x - s] : T(0);
        __syncthreads();
        sm[threadIdx.x] += a;
        __syncthreads();
    }
    T run = sm[threadIdx.x] - tot;  // exclusive prefix of this thread inside the block
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == SCAN_BS - 1) {
        if (block_sums) block_sums[blockIdx.x] = sm[SCAN_BS - 1];
        if (write_total) out[n] = sm[SCAN_BS - 1];
    }
}

template <class T>
__global__ __launch_bounds__(SCAN_BS) void k_scan_add(T *out, int64_t n, const T *offs) {
    int64_t i = blockIdx.x * (int64_t)SCAN_BS + threadIdx.x;
    if (i < n) out[i] += offs[i / SCAN_E];
    if (i == 0) out[n] = offs[(n + SCAN_E - 1) / SCAN_E];
}

}  // namespace

void fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s) {
    unsigned nb = (unsigned)((n + 255) / 256);
    if (nb > 1024) nb = 1024;
    if (nb == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(256), 0, s, p, n, v);
    HIPCK(hipGetLastError());
}

size_t scan_tmp_size(int64_t n) {
    size_t tot = 0;
    while (n > SCAN_E) {
        int64_t nb = (n + SCAN_E - 1) / SCAN_E;
        tot += 2 * (nb + 1);
        n = nb;
    }
    return tot + 2;
}

template <class T>
void exclusive_scan(const T *in, T *out, int64_t n, T *tmp, hipStream_t s) {
    if (n <= SCAN_E) {
        hipLaunchKernelGGL(k_scan_block<T>, dim3(1), dim3(SCAN_BS), 0, s, in, out, n, (T *)nullptr, 1);
        HIPCK(hipGetLastError());
        return;
    }
    int64_t nb = (n + SCAN_E - 1) / SCAN_E;
    T *sums = tmp, *sums_x = tmp + (nb + 1), *rest = tmp + 2 * (nb + 1);
    hipLaunchKernelGGL(k_scan_block<T>, dim3((unsigned)nb), dim3(SCAN_BS), 0, s, in, out, n, sums, 0);
    HIPCK(hipGetLastError());
    exclusive_scan<T>(sums, sums_x, nb, rest, s);
    hipLaunchKernelGGL(k_scan_add<T>, dim3((unsigned)((n + SCAN_BS - 1) / SCAN_BS)), dim3(SCAN_BS), 0, s, out, n, sums_x);
    HIPCK(hipGetLastError());
}
template void exclusive_scan<uint32_t>(const uint32_t *, uint32_t *, int64_t, uint32_t *, hipStream_t);
template void exclusive_scan<int64_t>(const int64_t *, int64_t *, int64_t, int64_t *, hipStream_t);

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void build_grid(sdot_pd *pd) {
    const int dim = pd->dim;
    const int64_t N = pd->N;
    hipStream_t s = pd->stream;
    pd->flags.ensure(3);
    pd->bbox_part.ensure(RED_BLOCKS * 2 * 3);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, pd->flags.p, (int64_t)3, 0u);
    if (dim == 2)
        hipLaunchKernelGGL(k_bbox<2>, dim3(RED_BLOCKS), dim3(BS), 0, s, pd->pos.p, N, pd->bbox_part.p, pd->flags.p);
    else
        hipLaunchKernelGGL(k_bbox<3>, dim3(RED_BLOCKS), dim3(BS), 0, s, pd->pos.p, N, pd->bbox_part.p, pd->flags.p);
    HIPCK(hipGetLastError());
    double part[RED_BLOCKS * 2 * 3];   // on the stack: the grid build allocates nothing on the host
    unsigned fl = 0;
    HIPCK(hipMemcpyAsync(part, pd->bbox_part.p, RED_BLOCKS * 2 * dim * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(&fl, pd->flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (fl & FLAG_BAD_INPUT) throw std::string("non-finite dirac position");
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = 0; b < RED_BLOCKS; ++b)
        for (int d = 0; d < dim; ++d) {
            lo[d] = fmin(lo[d], part[b * 2 * dim + d]);
            hi[d] = fmax(hi[d], part[b * 2 * dim + dim + d]);
        }
    double dpc = pd->dpc_x100 > 0 ? pd->dpc_x100 / 100.0 : (dim == 2 ? 3.0 : 2.0);
    const GridInfo g = grid_params(dim, lo, hi, N, dpc);
    pd->g = g;

    pd->keys.ensure(N);
    pd->rank.ensure(N);
    pd->cell_start.ensure(g.ncodes + 1);
    pd->scan_tmp.ensure(g.ncodes + scan_tmp_size(g.ncodes));
    pd->perm.ensure(N);
    pd->inv_perm.ensure(N);
    pd->spos.ensure(N * dim);
    uint32_t *cnt = pd->scan_tmp.p;
    uint32_t *tmp = pd->scan_tmp.p + g.ncodes;
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, cnt, g.ncodes, 0u);
    if (dim == 2)
        hipLaunchKernelGGL(k_keys<2>, dim3(nblk(N, BS)), dim3(BS), 0, s, pd->pos.p, N, g, pd->keys.p, pd->rank.p, cnt);
    else
        hipLaunchKernelGGL(k_keys<3>, dim3(nblk(N, BS)), dim3(BS), 0, s, pd->pos.p, N, g, pd->keys.p, pd->rank.p, cnt);
    HIPCK(hipGetLastError());
    (void)tmp;
    sort_by_keys(pd, g.ncodes);
}

void sort_by_keys(sdot_pd *pd, int64_t ncodes) {
    const int dim = pd->dim;
    const int64_t N = pd->N;
    hipStream_t s = pd->stream;
    uint32_t *cnt = pd->scan_tmp.p;
    uint32_t *tmp = pd->scan_tmp.p + ncodes;
    exclusive_scan<uint32_t>(cnt, pd->cell_start.p, ncodes, tmp, s);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(N, BS)), dim3(BS), 0, s, pd->keys.p, pd->rank.p, pd->cell_start.p, N, pd->perm.p);
    hipLaunchKernelGGL(k_binsort, dim3(nblk(ncodes, BS)), dim3(BS), 0, s, pd->cell_start.p, ncodes, pd->perm.p);
    if (dim == 2)
        hipLaunchKernelGGL(k_gather_pos<2>, dim3(nblk(N, BS)), dim3(BS), 0, s, pd->pos.p, pd->perm.p, N, pd->spos.p, pd->inv_perm.p);
    else
        hipLaunchKernelGGL(k_gather_pos<3>, dim3(nblk(N, BS)), dim3(BS), 0, s, pd->pos.p, pd->perm.p, N, pd->spos.p, pd->inv_perm.p);
    HIPCK(hipGetLastError());
    pd->grid_valid = true;
    pd->hint_ok = false;   // a new sorted order: the rows no longer name this order's diracs
}

GridInfo grid_params(int dim, const double *lo_in, const double *hi_in, int64_t N, double dpc) {
    double lo[3], hi[3];
    for (int d = 0; d < dim; ++d) {
        lo[d] = lo_in[d];
        hi[d] = hi_in[d];
    }
    // grid parameters: about `dpc` diracs per cell
    GridInfo g{};
    g.dim = dim;
    double vol = 1;
    double L[3];
    // extents are floored at 1e-3 of the largest one so that diracs on a line
    // or a plane still get cells of comparable size along every axis
    double lmax = 0;
    for (int d = 0; d < dim; ++d) lmax = fmax(lmax, hi[d] - lo[d]);
    if (!(lmax > 0)) {   // every dirac at one point (one dirac): a unit extent around it
        for (int d = 0; d < dim; ++d) {
            lo[d] -= 0.5;
            hi[d] += 0.5;
        }
        lmax = 1;
    }
    for (int d = 0; d < dim; ++d) {
        double l = hi[d] - lo[d];
        double scale = fmax(fabs(lo[d]), fabs(hi[d]));
        double pad = 1e-7 * l + 1e-12 * scale + 1e-300;
        if (l + 2 * pad < 1e-3 * lmax) pad = 0.5 * (1e-3 * lmax - l);
        lo[d] -= pad;
        hi[d] += pad;
        L[d] = hi[d] - lo[d];
        vol *= L[d];
    }
    double h = pow(vol * dpc / (double)(N > 0 ? N : 1), 1.0 / dim);
    if (!(vol > 1e-290)) {   // extents so small that their product underflows: the same cell size from logarithms
        double lh = log(dpc / (double)(N > 0 ? N : 1));
        for (int d = 0; d < dim; ++d) lh += log(L[d]);
        h = exp(lh / dim);
    }
    int maxc = dim == 2 ? 32768 : 1024;
    int mx = 1;
    for (int d = 0; d < dim; ++d) {
        double nc = ceil(L[d] / h);
        int c = nc < 1 ? 1 : (nc > maxc ? maxc : (int)nc);
        g.ncell[d] = c;
        g.lo[d] = lo[d];
        g.h[d] = L[d] / c;
        g.inv_h[d] = c / L[d];
        mx = c > mx ? c : mx;
    }
    for (int d = dim; d < 3; ++d) { g.ncell[d] = 1; g.lo[d] = 0; g.h[d] = 1; g.inv_h[d] = 1; }
    int bits = 0;
    while ((1 << bits) < mx) ++bits;
    g.bits = bits;
    g.ncodes = (int64_t)1 << (dim * bits);

    return g;
}

void refresh_weights(sdot_pd *pd) {
    const int64_t N = pd->N;
    hipStream_t s = pd->stream;
    const GridInfo &g = pd->g;
    pd->sw.ensure(N);
    pd->cell_wmax.ensure(g.ncodes);
    pd->wglob.ensure(1);
    // tie-breaking offset scale: 1e-12 of the squared mean spacing of the grid
    double vol = 1;
    for (int d = 0; d < pd->dim; ++d) vol *= g.h[d] * g.ncell[d];
    const double spacing2 = pow(vol / (double)(N > 0 ? N : 1), 2.0 / pd->dim);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, pd->wglob.p, dbl_order(-DBL_MAX));
    hipLaunchKernelGGL(k_weights, dim3(nblk(g.ncodes, BS)), dim3(BS), 0, s, pd->cell_start.p, g.ncodes, pd->w.p,
                       pd->perm.p, pd->sw.p, pd->cell_wmax.p, pd->wglob.p, pd->flags.p,
                       N > 1 ? 1e-12 * spacing2 : 0.0);   // a lone dirac has no ties to break
    HIPCK(hipGetLastError());
    if (pd->dim == 3) {   // max-weight pyramid (level 0 = a copy of cell_wmax) for the far-field traversal
        pd->wpyr.ensure(pyr_off(g.ncodes, 3, g.bits + 1));
        HIPCK(hipMemcpyAsync(pd->wpyr.p, pd->cell_wmax.p, g.ncodes * sizeof(double), hipMemcpyDeviceToDevice, s));
        for (int k = 1; k <= g.bits; ++k) {
            const int64_t nn = g.ncodes >> (3 * k);
            hipLaunchKernelGGL(k_pyr_level, dim3(nblk(nn, BS)), dim3(BS), 0, s, pd->wpyr.p + pyr_off(g.ncodes, 3, k - 1), nn,
                               8, pd->wpyr.p + pyr_off(g.ncodes, 3, k));
        }
        HIPCK(hipGetLastError());
    }
}

}  // namespace sdot_amd
