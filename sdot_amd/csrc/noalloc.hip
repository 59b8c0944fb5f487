// Allocation-free cell masses (sdot_pd_integrals_noalloc): the same grid build and
// 2D cell kernels as sdot_pd_integrals, with every intermediate array carved out
// of one fixed device array instead of the handle's growable buffers, so the call
// makes no device allocation (no hipMalloc / hipFree, no allocating library
// routine).  For callers that cannot allocate: tight serving loops, harnesses
// that forbid allocation after set-up.
//
// The array `scratch` (NA_SCRATCH_BYTES) exists once per device in each process
// and lives until the process ends.  Its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and used for every
// kernel argument; the name itself is only a host-side stand-in.  Every call on a
// device shares that device's array and keeps what the last call left there:
// everything a kernel needs at zero (counts, flags, overflow lists) is cleared by
// the call itself with fill kernels.  Calls that overlapped on one device would
// race on the array; a per-device lock held for the whole (synchronous) call
// serialises them within the process.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

using namespace sdot_amd;

namespace na {

constexpr int64_t NA_MAX_DIRACS = int64_t(1) << 18;   // 262,144 diracs
constexpr int64_t NA_MAX_CODES = int64_t(1) << 20;    // Morton codes of the grid (cells x padding to 2^bits)
constexpr int64_t NA_SCAN_TMP = 4096;                 // scan block sums for NA_MAX_CODES (1,028 needed)
constexpr int NA_MAX_DEVICES = 64;

constexpr int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Byte layout of the workspace; every slice starts on a 256-byte boundary.
struct Layout {
    int64_t bbox, flags, wglob, keys, rank, cell_start, scan_tmp, perm, inv_perm, spos, sw, cell_wmax, mass, arc, perim,
        diag, row_len, ell_col, ell_val, ovf, total;
};
constexpr Layout make_layout() {
    Layout l{};
    int64_t o = 0;
    auto take = [&o](int64_t bytes) {
        const int64_t at = o;
        o += align256(bytes);
        return at;
    };
    const int64_t N = NA_MAX_DIRACS, C = NA_MAX_CODES;
    l.bbox = take(8 * BBOX_BLOCKS * 2 * 3);
    l.flags = take(4 * 4);
    l.wglob = take(8);
    l.keys = take(4 * N);
    l.rank = take(4 * N);
    l.cell_start = take(4 * (C + 1));
    l.scan_tmp = take(4 * (C + NA_SCAN_TMP));
    l.perm = take(4 * N);
    l.inv_perm = take(4 * N);
    l.spos = take(8 * 2 * N);
    l.sw = take(8 * N);
    l.cell_wmax = take(8 * C);
    l.mass = take(8 * N);
    l.arc = take(8 * N);
    l.perim = take(8 * N);
    l.diag = take(8 * N);
    l.row_len = take(4 * N);
    l.ell_col = take(4);
    l.ell_val = take(8);
    l.ovf = take(4 * (2 * N + 2));
    l.total = o;
    return l;
}
constexpr Layout LAYOUT = make_layout();
constexpr int64_t NA_SCRATCH_BYTES = 40 << 20;   // 40 MiB
static_assert(LAYOUT.total <= NA_SCRATCH_BYTES, "workspace layout exceeds the scratch array");

}  // namespace na

static __device__ __align__(256) unsigned char scratch[na::NA_SCRATCH_BYTES];

namespace na {

std::mutex g_mu[NA_MAX_DEVICES];
unsigned char *g_ptr[NA_MAX_DEVICES] = {};

// the device address of this device's copy of `scratch` (looked up on the first call
// on the device, with that device current)
unsigned char *scratch_ptr(int dev) {
    if (!g_ptr[dev]) {
        void *p = nullptr;
        HIPCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw StatusError{SDOT_HIP_ERROR, "no device address for the fixed workspace"};
        g_ptr[dev] = static_cast<unsigned char *>(p);
    }
    return g_ptr[dev];
}

template <class T>
void carve(DevBuf<T> &b, unsigned char *base, int64_t off, int64_t count) {
    b.borrow(reinterpret_cast<T *>(base + off), (size_t)count);
}

// restores the current device when the call returns
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        HIPCK(hipGetDevice(&prev));
        if (prev != dev) HIPCK(hipSetDevice(dev));
    }
    ~DeviceScope() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

sdot_status integrals_noalloc(sdot_pd *pd, const double *d_pos, const double *d_w, int64_t n, double *d_masses) {
    if (pd->dim != 2) throw StatusError{SDOT_NOT_SUPPORTED, "integrals_noalloc: 2D only"};
    if (n < 1) throw StatusError{SDOT_INVALID_ARG, "integrals_noalloc: no diracs"};
    if (n > NA_MAX_DIRACS)
        throw StatusError{SDOT_CAPACITY_OVERFLOW, "integrals_noalloc: more than 262144 diracs (the fixed workspace's limit)"};
    if (pd->has_image || pd->shapes.size() != 1 || !pd->shapes[0].planes.empty() || !pd->htrans.empty() || pd->rf)
        throw StatusError{SDOT_NOT_SUPPORTED,
                          "integrals_noalloc: one box domain, Unit / InBallW05, no image and no translations"};
    if (pd->device < 0 || pd->device >= NA_MAX_DEVICES) throw StatusError{SDOT_NOT_SUPPORTED, "device index"};
    DeviceScope ds(pd->device);
    std::lock_guard<std::mutex> lk(g_mu[pd->device]);
    unsigned char *base = scratch_ptr(pd->device);

    // a handle-shaped view whose buffers are all slices of the workspace (or the caller's arrays)
    sdot_pd v;
    v.dim = 2;
    v.precision = pd->precision;
    v.ball = pd->ball;
    v.dpc_x100 = pd->dpc_x100;
    v.stream = pd->stream;
    v.device = pd->device;
    v.N = n;
    v.pos.borrow(const_cast<double *>(d_pos), (size_t)(2 * n));
    v.w.borrow(const_cast<double *>(d_w), (size_t)n);
    carve(v.bbox_part, base, LAYOUT.bbox, BBOX_BLOCKS * 2 * 3);
    carve(v.flags, base, LAYOUT.flags, 4);
    carve(v.wglob, base, LAYOUT.wglob, 1);
    carve(v.keys, base, LAYOUT.keys, NA_MAX_DIRACS);
    carve(v.rank, base, LAYOUT.rank, NA_MAX_DIRACS);
    carve(v.cell_start, base, LAYOUT.cell_start, NA_MAX_CODES + 1);
    carve(v.scan_tmp, base, LAYOUT.scan_tmp, NA_MAX_CODES + NA_SCAN_TMP);
    carve(v.perm, base, LAYOUT.perm, NA_MAX_DIRACS);
    carve(v.inv_perm, base, LAYOUT.inv_perm, NA_MAX_DIRACS);
    carve(v.spos, base, LAYOUT.spos, 2 * NA_MAX_DIRACS);
    carve(v.sw, base, LAYOUT.sw, NA_MAX_DIRACS);
    carve(v.cell_wmax, base, LAYOUT.cell_wmax, NA_MAX_CODES);
    carve(v.mass, base, LAYOUT.mass, NA_MAX_DIRACS);
    carve(v.arc, base, LAYOUT.arc, NA_MAX_DIRACS);
    carve(v.perim, base, LAYOUT.perim, NA_MAX_DIRACS);
    carve(v.diag, base, LAYOUT.diag, NA_MAX_DIRACS);
    carve(v.row_len, base, LAYOUT.row_len, NA_MAX_DIRACS);
    carve(v.ell_col, base, LAYOUT.ell_col, 1);
    carve(v.ell_val, base, LAYOUT.ell_val, 1);
    carve(v.ovf_list, base, LAYOUT.ovf, 2 * NA_MAX_DIRACS + 2);

    // the grid's code space must fit the workspace: decided on the host from the
    // bounding box before any code-indexed kernel runs (build_grid's ensure() of
    // cell_start / scan_tmp throws SDOT_CAPACITY_OVERFLOW otherwise)
    const hipStream_t s = v.stream;
    build_grid(&v);   // bbox, keys, counting sort, sorted positions
    fill_u32(v.flags.p, 3, 0u, s);
    refresh_weights(&v);   // sorted weights, per-cell and global max weight

    const Shape &sh = pd->shapes[0];
    CellArgs a{};
    for (int d = 0; d < 3; ++d) {
        a.dom_lo[d] = d < 2 ? sh.q0[d] : 0;
        a.dom_hi[d] = d < 2 ? sh.q1[d] : 1;
        a.img_n[d] = 1;
    }
    a.coeff = sh.coeff;
    a.img_nc = 1;
    a.N = n;
    a.spos = v.spos.p;
    a.sw = v.sw.p;
    a.cell_start = v.cell_start.p;
    a.cell_wmax = v.cell_wmax.p;
    a.wglob = v.wglob.p;
    a.g = v.g;
    a.mass = v.mass.p;
    a.arc = v.arc.p;
    a.perim = v.perim.p;
    a.diag = v.diag.p;
    a.row_len = v.row_len.p;
    a.ell_col = v.ell_col.p;
    a.ell_val = v.ell_val.p;
    a.ell_cap = 0;   // masses only: no Hessian rows
    a.ell_stride = n;
    a.flags = v.flags.p;
    run_cells2(&v, a);
    export_vec_device(&v, v.mass.p, d_masses);
    unsigned fl = 0;
    HIPCK(hipMemcpyAsync(&fl, v.flags.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (fl & FLAG_BAD_INPUT) throw StatusError{SDOT_INVALID_ARG, "non-finite weight or coincident diracs"};
    if (fl & FLAG_CELL_OVERFLOW)
        throw StatusError{SDOT_CAPACITY_OVERFLOW, "a Laguerre cell exceeds the polygon capacity (64 vertices)"};
    return (fl & FLAG_EMPTY) ? SDOT_EMPTY_CELL : SDOT_OK;
}

}  // namespace na

extern "C" sdot_status sdot_pd_integrals_noalloc(sdot_pd *pd, const double *d_positions, const double *d_weights,
                                                 int64_t nb_diracs, double *d_masses) {
    if (!pd || !d_positions || !d_weights || !d_masses) return SDOT_INVALID_ARG;
    return guard(pd, [&] { return na::integrals_noalloc(pd, d_positions, d_weights, nb_diracs, d_masses); });
}
