// C-ABI entry points (include/sdot_mi355x.h): handle life cycle, uploads, the
// cell pass driver and the result getters.  Every entry point converts HIP
// failures and invalid input into status codes.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <new>

#include <algorithm>
#include <cfloat>
#include "internal.h"
#include "pcg.h"

using namespace sdot_amd;

namespace {

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}

// The box the cells of an item are clipped to. A face of the item's box that its
// planes make redundant (their polytope reaches it at most) moves out by an eighth
// of the box's extent: the item is the same set, and no cell starts from a corner
// where a box face meets several planes (the corners of a simplex lie on its
// bounding box; a regular octahedron's can lie on the widened box), a vertex of
// four or more planes whose exact in/out tests round inconsistently in a plane cut.
// The vertices of planes + box widened by a quarter are enumerated (all pairs in
// 2D, triples in 3D); items with more than 160 planes keep their box.
void item_clip_box(Shape &sh, int dim) {
    for (int d = 0; d < 3; ++d) {
        sh.q0[d] = sh.p0[d];
        sh.q1[d] = sh.p1[d];
    }
    if (sh.planes.empty() || sh.planes.size() > 160) return;
    double ext[3] = {0, 0, 0}, L = 0;
    for (int d = 0; d < dim; ++d) {
        ext[d] = sh.p1[d] - sh.p0[d];
        L = std::max(L, ext[d]);
    }
    std::vector<std::array<double, 4>> c = sh.planes;
    for (int d = 0; d < dim; ++d) {
        std::array<double, 4> lo{0, 0, 0, -(sh.p0[d] - 0.25 * ext[d])}, hi{0, 0, 0, sh.p1[d] + 0.25 * ext[d]};
        lo[d] = -1;
        hi[d] = 1;
        c.push_back(lo);
        c.push_back(hi);
    }
    const int m = (int)c.size();
    std::vector<double> nrm(m);
    for (int i = 0; i < m; ++i) nrm[i] = std::sqrt(c[i][0] * c[i][0] + c[i][1] * c[i][1] + c[i][2] * c[i][2]);
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    bool any = false;
    auto take = [&](const double *x) {
        for (int i = 0; i < m; ++i)
            if (c[i][0] * x[0] + c[i][1] * x[1] + c[i][2] * x[2] - c[i][3] > 1e-11 * nrm[i] * L) return;
        any = true;
        for (int d = 0; d < dim; ++d) {
            mn[d] = std::min(mn[d], x[d]);
            mx[d] = std::max(mx[d], x[d]);
        }
    };
    for (int i = 0; i < m; ++i)
        for (int j = i + 1; j < m; ++j) {
            if (dim == 2) {
                const double det = c[i][0] * c[j][1] - c[i][1] * c[j][0];
                if (!(std::fabs(det) > 1e-12 * nrm[i] * nrm[j])) continue;
                const double x[3] = {(c[i][3] * c[j][1] - c[j][3] * c[i][1]) / det,
                                     (c[i][0] * c[j][3] - c[j][0] * c[i][3]) / det, 0};
                take(x);
                continue;
            }
            for (int k = j + 1; k < m; ++k) {
                const auto &a = c[i], &b = c[j], &e = c[k];
                const double bx = b[1] * e[2] - b[2] * e[1], by = b[2] * e[0] - b[0] * e[2], bz = b[0] * e[1] - b[1] * e[0];
                const double ex = e[1] * a[2] - e[2] * a[1], ey = e[2] * a[0] - e[0] * a[2], ez = e[0] * a[1] - e[1] * a[0];
                const double ax = a[1] * b[2] - a[2] * b[1], ay = a[2] * b[0] - a[0] * b[2], az = a[0] * b[1] - a[1] * b[0];
                const double det = a[0] * bx + a[1] * by + a[2] * bz;
                if (!(std::fabs(det) > 1e-12 * nrm[i] * nrm[j] * nrm[k])) continue;
                const double x[3] = {(a[3] * bx + b[3] * ex + e[3] * ax) / det, (a[3] * by + b[3] * ey + e[3] * ay) / det,
                                     (a[3] * bz + b[3] * ez + e[3] * az) / det};
                take(x);
            }
        }
    if (!any) return;   // an empty item: its cells come out empty with any box
    for (int d = 0; d < dim; ++d) {
        if (mn[d] >= sh.p0[d] - 1e-11 * L) sh.q0[d] = sh.p0[d] - 0.125 * ext[d];
        if (mx[d] <= sh.p1[d] + 1e-11 * L) sh.q1[d] = sh.p1[d] + 0.125 * ext[d];
    }
}

// the domain as seen by the kernels: one box (with density coefficient) or an image
void domain_args(sdot_pd *pd, CellArgs &a) {
    a.img = nullptr;
    a.img_nc = 1;
    a.ntrans = (int)(pd->htrans.size() / pd->dim);
    a.trans = a.ntrans > 0 ? pd->trans.p : nullptr;
    if (a.ntrans > 0 && (double)pd->N * (a.ntrans + 1) >= 2147483647.0)
        throw StatusError{SDOT_NOT_SUPPORTED, "translations: N * (translations + 1) must stay below 2^31"};
    if (pd->has_image) {
        if (!pd->shapes.empty()) throw StatusError{SDOT_NOT_SUPPORTED, "image density together with shapes"};
        for (int d = 0; d < 3; ++d) {
            a.dom_lo[d] = d < pd->dim ? pd->img_lo[d] : 0;
            a.dom_hi[d] = d < pd->dim ? pd->img_hi[d] : 1;
            a.img_n[d] = d < pd->dim ? pd->img_n[d] : 1;
        }
        a.coeff = 1;
        a.img = pd->img.p;
        a.img_nc = pd->img_nc;
        return;
    }
    if (pd->shapes.empty()) throw StatusError{SDOT_NOT_SUPPORTED, "no domain shape (add_box_shape)"};
    a.nitems = 0;
    const Shape &sh = pd->shapes[0];
    for (int d = 0; d < 3; ++d) {
        a.dom_lo[d] = d < pd->dim ? sh.q0[d] : 0;
        a.dom_hi[d] = d < pd->dim ? sh.q1[d] : 1;
        a.img_n[d] = 1;
    }
    a.coeff = sh.coeff;
    if (pd->shapes.size() == 1 && sh.planes.empty()) return;   // one box: the starting cell
    // items (ConvexPolyhedronAssembly.tcc:7-142): tables on the device
    const int K = (int)pd->shapes.size();
    std::vector<double> box(6 * K), coeff(K), planes;
    std::vector<int> off(K + 1, 0);
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int k = 0; k < K; ++k) {
        const Shape &it = pd->shapes[k];
        for (int d = 0; d < 3; ++d) {
            box[6 * k + d] = it.q0[d];
            box[6 * k + 3 + d] = it.q1[d];
            lo[d] = std::min(lo[d], it.p0[d]);
            hi[d] = std::max(hi[d], it.p1[d]);
        }
        coeff[k] = it.coeff;
        for (const auto &pl : it.planes) planes.insert(planes.end(), pl.begin(), pl.end());
        off[k + 1] = (int)(planes.size() / 4);
    }
    if (planes.empty()) planes.assign(4, 0.0);
    pd->it_box.ensure(box.size());
    pd->it_coeff.ensure(coeff.size());
    pd->it_planes.ensure(planes.size());
    pd->it_off.ensure(off.size());
    HIPCK(hipMemcpyAsync(pd->it_box.p, box.data(), box.size() * sizeof(double), hipMemcpyHostToDevice, pd->stream));
    HIPCK(hipMemcpyAsync(pd->it_coeff.p, coeff.data(), coeff.size() * sizeof(double), hipMemcpyHostToDevice, pd->stream));
    HIPCK(hipMemcpyAsync(pd->it_planes.p, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice, pd->stream));
    HIPCK(hipMemcpyAsync(pd->it_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
    a.nitems = K;
    a.it_box = pd->it_box.p;
    a.it_coeff = pd->it_coeff.p;
    a.it_planes = pd->it_planes.p;
    a.it_off = pd->it_off.p;
    if (K > 1) {   // englobing_convex_polyhedron (:60-92): a box around the items, widened by its extent
        for (int d = 0; d < 3; ++d) {
            const double e = d < pd->dim ? hi[d] - lo[d] : 0.0;
            a.dom_lo[d] = d < pd->dim ? lo[d] - e : 0;
            a.dom_hi[d] = d < pd->dim ? hi[d] + e : 1;
        }
        a.coeff = 1;
    }
}

}  // namespace

namespace sdot_amd {

std::recursive_mutex &device_alloc_mutex() {
    static std::recursive_mutex m;
    return m;
}

sdot_status compute_pass(sdot_pd *pd) {
    if (pd->N <= 0) throw StatusError{SDOT_INVALID_ARG, "no diracs"};
    if (pd->ball && pd->dim == 3)
        throw StatusError{SDOT_NOT_SUPPORTED, "ball clipping in 3D (ConvexPolyhedron3::ball_cut is TODO in the reference)"};
    const bool pieces = pd->has_image || pd->shapes.size() > 1;
    if (pd->ball && pd->has_image && pd->img_nc != 1)
        throw StatusError{SDOT_NOT_SUPPORTED, "ball clipping with a quadratic image (apply_poly has no ball in the reference)"};
    if (pd->ball && pieces && pd->want_jac)
        throw StatusError{SDOT_NOT_SUPPORTED, "position derivatives with ball clipping over several pieces"};
    if (pd->rf && pd->has_image && pd->img_nc != 1)
        throw StatusError{SDOT_NOT_SUPPORTED, "R2 / WmR2 / ExpWmR2db integration: constant pieces only"};
    if (pd->rf >= 2 && pd->dim == 3)
        throw StatusError{SDOT_NOT_SUPPORTED, "WmR2 / ExpWmR2db in 3D (ConvexPolyhedron3 integration is TODO in the reference)"};
    if (pd->rf && (pd->want_mom || pd->want_jac))
        throw StatusError{SDOT_NOT_SUPPORTED, "centroids and position derivatives: Unit / InBallW05 only"};
    if (pd->N > (int64_t)0x7FFFFFFF) throw StatusError{SDOT_INVALID_ARG, "more than 2^31-1 diracs"};
    hipStream_t s = pd->stream;
    CellArgs a{};
    domain_args(pd, a);
    for (int k = 0; k < 4; ++k)
        if (!pd->ev[k]) HIPCK(hipEventCreate(&pd->ev[k]));
    HIPCK(hipEventRecord(pd->ev[0], s));
    if (!pd->grid_valid) build_grid(pd);
    pd->flags.ensure(3);
    fill_u32(pd->flags.p, 3, 0u, s);
    refresh_weights(pd);
    HIPCK(hipEventRecord(pd->ev[1], s));

    const int64_t N = pd->N;
    const int cap0 = pd->dim == 2 ? ELL_CAP2 : ELL_CAP3, cap_wide = pd->dim == 2 ? ELL_CAP2_WIDE : ELL_CAP3_WIDE;
    if (pd->ell_cap < cap0) pd->ell_cap = cap0;
    pd->mass.ensure(N);
    pd->arc.ensure(N);
    pd->perim.ensure(N);
    pd->diag.ensure(N);
    pd->row_len.ensure(N);
    pd->ell_col.ensure((size_t)pd->ell_cap * N);
    pd->ell_val.ensure((size_t)pd->ell_cap * N);
    pd->ovf_list.ensure(2 * (size_t)N + 2);

    a.N = N;
    a.spos = pd->spos.p;
    a.sw = pd->sw.p;
    a.cell_start = pd->cell_start.p;
    a.cell_wmax = pd->cell_wmax.p;
    a.wglob = pd->wglob.p;
    a.wpyr = pd->dim == 3 ? pd->wpyr.p : nullptr;
    a.g = pd->g;
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
    a.stop_if_void = pd->stop_if_void ? 1 : 0;
    a.mom = nullptr;
    if (pd->want_mom) {
        pd->mom.ensure((size_t)pd->dim * N);
        a.mom = pd->mom.p;
    }
    pd->pass_has_mom = pd->want_mom;
    a.rf = pd->rf;
    a.rf_eps = pd->rf_eps;
    a.jblk = a.jdiag = nullptr;
    if (pd->want_jac) {
        const size_t nb = (size_t)(pd->dim + 1) * (pd->dim + 1);
        pd->jblk.ensure(nb * pd->ell_cap * N);
        pd->jdiag.ensure(nb * N);
        a.jblk = pd->jblk.p;
        a.jdiag = pd->jdiag.p;
    }
    a.ovf_list = pd->ovf_list.p;
    a.ovf_count = pd->flags.p + 1;
    if (pd->dim == 2 && pd->precision != SDOT_MIXED && !pd->want_jac && !pd->rf) {
        ensure_resume(pd, N);
        use_resume(pd, a);
    }
    pd->wctr.ensure(3 * WCTR);
    fill_u32((uint32_t *)pd->wctr.p, 6 * WCTR, 0u, s);
    a.ctr = pd->wctr.p;
    if (pd->dim == 2) run_cells2(pd, a);
    else run_cells3(pd, a);
    HIPCK(hipEventRecord(pd->ev[2], s));
    unsigned fls[3] = {0, 0, 0};
    std::vector<unsigned long long> ctr(3 * WCTR);
    HIPCK(hipMemcpyAsync(fls, pd->flags.p, 3 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(ctr.data(), pd->wctr.p, 3 * WCTR * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    const unsigned fl = fls[0];
    if (pd->dim == 2) pd->last.nb_cells_overflow = fls[1];
    double flops = 0;
    int64_t sk = 0, sa = 0;
    for (int k = 0; k < WCTR; ++k) {
        flops += (double)ctr[k];
        sk += (int64_t)ctr[WCTR + k];
        sa += (int64_t)ctr[2 * WCTR + k];
    }
    pd->last.cell_flops = flops;
    pd->last.sum_faces = sk;
    pd->last.sum_arcs = sa;
    pd->last.ms_cell_kernel = elapsed_ms(pd->ev[1], pd->ev[3]);
    pd->last.ms_grid = elapsed_ms(pd->ev[0], pd->ev[1]);
    pd->last.ms_cells = elapsed_ms(pd->ev[1], pd->ev[2]);
    pd->last.ms_rows = 0;
    pd->last.ms_total = elapsed_ms(pd->ev[0], pd->ev[2]);
    pd->last.grid_cells = (int64_t)pd->g.ncell[0] * pd->g.ncell[1] * pd->g.ncell[2];
    pd->last_flags = fl;
    // a pass that ends in an error is never reused: the next getter runs it again and
    // reports the same error instead of returning the partial results
    pd->pass_valid = !(fl & (FLAG_BAD_INPUT | FLAG_CELL_OVERFLOW | FLAG_ROW_OVERFLOW));
    if (!pd->pass_valid) pd->hint_ok = false;
    if (fl & FLAG_BAD_INPUT) throw StatusError{SDOT_INVALID_ARG, "non-finite weight or coincident diracs"};
    if (pd->stop_if_void && (fl & FLAG_EMPTY)) {   // stopped early: the per-cell results are partial
        pd->pass_valid = false;
        return SDOT_EMPTY_CELL;
    }
    if (fl & FLAG_CELL_OVERFLOW)
        throw StatusError{SDOT_CAPACITY_OVERFLOW, pd->dim == 2 ? "a Laguerre cell exceeds the polygon capacity (64 vertices)"
                                                              : "a Laguerre cell exceeds the polyhedron capacity (255 faces, 508 vertices)"};
    if (fl & FLAG_ROW_OVERFLOW) {
        if (pd->ell_cap < cap_wide) {   // rerun with wide rows (kept for the next passes)
            pd->ell_cap = cap_wide;
            return compute_pass(pd);
        }
        throw StatusError{SDOT_CAPACITY_OVERFLOW, "Hessian row longer than the row capacity"};
    }
    // a complete pass of a Newton solve in 3D: its neighbours start the next passes (a pass
    // stopped at an empty cell leaves the snapshot as it was, whatever rows it rewrote)
    if (pd->warm && pd->dim == 3) {
        snapshot_hints(pd);
        pd->hint_ok = true;
    }
    return (fl & FLAG_EMPTY) ? SDOT_EMPTY_CELL : SDOT_OK;
}
}  // namespace sdot_amd

namespace {
sdot_status ensure_pass(sdot_pd *pd) {
    if (pd->pass_valid && pd->want_mom && !pd->pass_has_mom) pd->pass_valid = false;
    if (pd->pass_valid) {
        if (pd->last_flags & (FLAG_BAD_INPUT | FLAG_ROW_OVERFLOW | FLAG_CELL_OVERFLOW)) return compute_pass(pd);
        return (pd->last_flags & FLAG_EMPTY) ? SDOT_EMPTY_CELL : SDOT_OK;
    }
    return compute_pass(pd);
}

}  // namespace

// a plain copy of n 16-byte words (grid-stride): the reference kernel for scaling
// the hardware HBM counters
__global__ void k_copy16(uint4 *dst, const uint4 *src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" {

const char *sdot_status_string(int st) {
    switch (st) {
        case SDOT_OK: return "ok";
        case SDOT_EMPTY_CELL: return "empty cell";
        case SDOT_NOT_CONVERGED: return "not converged";
        case SDOT_CAPACITY_OVERFLOW: return "capacity overflow";
        case SDOT_INVALID_ARG: return "invalid argument";
        case SDOT_HIP_ERROR: return "HIP error";
        case SDOT_OOM: return "out of memory";
        case SDOT_NOT_SUPPORTED: return "not supported";
        case SDOT_PENDING: return "pending";
        default: return "unknown status";
    }
}

const char *sdot_pd_last_error(const sdot_pd *pd) { return pd ? pd->err.c_str() : "null handle"; }

sdot_status sdot_pd_create(sdot_pd **out, int dim, int precision) {
    if (!out || (dim != 2 && dim != 3) || (precision != SDOT_FP64 && precision != SDOT_MIXED)) return SDOT_INVALID_ARG;
    *out = nullptr;
    sdot_pd *pd = new (std::nothrow) sdot_pd;
    if (!pd) return SDOT_OOM;
    pd->dim = dim;
    pd->precision = precision;
    sdot_status st = guard(pd, [&] {
        int ndev = 0;
        HIPCK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw StatusError{SDOT_HIP_ERROR, "no HIP device"};
        HIPCK(hipGetDevice(&pd->device));
        HIPCK(hipStreamCreateWithFlags(&pd->stream, hipStreamNonBlocking));
        return SDOT_OK;
    });
    if (st != SDOT_OK) {
        delete pd;
        return st;
    }
    *out = pd;
    return SDOT_OK;
}

sdot_status sdot_pd_destroy(sdot_pd *pd) {
    if (!pd) return SDOT_INVALID_ARG;
    if (pd->stream) (void)hipStreamSynchronize(pd->stream);
    for (auto &e : pd->ev)
        if (e) (void)hipEventDestroy(e);
    pd->img.release();
    pd->pos.release(); pd->w.release();
    pd->keys.release(); pd->rank.release(); pd->cell_start.release(); pd->scan_tmp.release();
    pd->perm.release(); pd->inv_perm.release();
    pd->spos.release(); pd->sw.release(); pd->cell_wmax.release(); pd->bbox_part.release(); pd->wpyr.release(); pd->wglob.release();
    pd->mass.release(); pd->arc.release(); pd->perim.release(); pd->diag.release(); pd->ell_val.release();
    pd->row_len.release(); pd->mom.release(); pd->ell_col.release(); pd->ovf_list.release(); pd->flags.release(); pd->wctr.release();
    pd->rs_hdr.release(); pd->rs_i.release(); pd->rs_d.release(); pd->hint_col.release();
    pd->csr_off.release(); pd->csr_col.release(); pd->csr_cnt.release(); pd->csr_tmp.release();
    pd->csr_val.release(); pd->out_tmp.release(); pd->scratch3.release();
    pd->csr_off2.release(); pd->csr_col2.release(); pd->csr_val2.release();
    for (auto &b : pd->sv) b.release();
    pd->red.release();
    delete pd->pcg;
    pd->pcg = nullptr;
    if (pd->stream) (void)hipStreamDestroy(pd->stream);
    delete pd;
    return SDOT_OK;
}

sdot_status sdot_pd_set_diracs(sdot_pd *pd, const double *positions, const double *weights, int64_t n) {
    if (!pd || n <= 0 || !positions || !weights) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        pd->N = n;
        pd->pos.ensure(n * pd->dim);
        pd->w.ensure(n);
        HIPCK(hipMemcpyAsync(pd->pos.p, positions, n * pd->dim * sizeof(double), hipMemcpyHostToDevice, pd->stream));
        HIPCK(hipMemcpyAsync(pd->w.p, weights, n * sizeof(double), hipMemcpyHostToDevice, pd->stream));
        HIPCK(hipStreamSynchronize(pd->stream));
        pd->grid_valid = false;
        pd->pass_valid = false;
        return SDOT_OK;
    });
}

sdot_status sdot_pd_set_diracs_device(sdot_pd *pd, const double *d_positions, const double *d_weights, int64_t n) {
    if (!pd || n <= 0 || !d_positions || !d_weights) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        pd->N = n;
        pd->pos.ensure(n * pd->dim);
        pd->w.ensure(n);
        HIPCK(hipMemcpyAsync(pd->pos.p, d_positions, n * pd->dim * sizeof(double), hipMemcpyDeviceToDevice, pd->stream));
        HIPCK(hipMemcpyAsync(pd->w.p, d_weights, n * sizeof(double), hipMemcpyDeviceToDevice, pd->stream));
        pd->grid_valid = false;
        pd->pass_valid = false;
        return SDOT_OK;
    });
}

sdot_status sdot_pd_set_weights(sdot_pd *pd, const double *weights) {
    if (!pd || !weights || pd->N <= 0) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        HIPCK(hipMemcpyAsync(pd->w.p, weights, pd->N * sizeof(double), hipMemcpyHostToDevice, pd->stream));
        HIPCK(hipStreamSynchronize(pd->stream));
        pd->pass_valid = false;
        return SDOT_OK;
    });
}

sdot_status sdot_pd_dirac_buffers(sdot_pd *pd, double **d_positions, double **d_weights) {
    if (!pd || pd->N <= 0 || !d_positions || !d_weights) return SDOT_INVALID_ARG;
    *d_positions = pd->pos.p;
    *d_weights = pd->w.p;
    return SDOT_OK;
}

sdot_status sdot_pd_mark_diracs_changed(sdot_pd *pd, int positions_changed) {
    if (!pd) return SDOT_INVALID_ARG;
    if (positions_changed) pd->grid_valid = false;
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_csr_device(sdot_pd *pd, const int64_t **d_offsets, const int64_t **d_cols, const double **d_vals) {
    if (!pd || !d_offsets || !d_cols || !d_vals) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        sdot_status st = ensure_pass(pd);
        for (int k = 0; k < 2; ++k)
            if (!pd->ev[4 + k]) HIPCK(hipEventCreate(&pd->ev[4 + k]));
        HIPCK(hipEventRecord(pd->ev[4], pd->stream));
        build_csr_device(pd);
        HIPCK(hipEventRecord(pd->ev[5], pd->stream));
        HIPCK(hipStreamSynchronize(pd->stream));
        pd->last.ms_rows = elapsed_ms(pd->ev[4], pd->ev[5]);
        *d_offsets = pd->csr_off.p;
        *d_cols = pd->csr_col.p;
        *d_vals = pd->csr_val.p;
        return st;
    });
}

sdot_status sdot_pd_add_box_shape(sdot_pd *pd, const double *p0, const double *p1, double coeff, int64_t cut_id) {
    if (!pd || !p0 || !p1) return SDOT_INVALID_ARG;
    if (cut_id != -1) {
        pd->err = "only the domain cut id -1 is supported";
        return SDOT_NOT_SUPPORTED;
    }
    Shape sh{};
    for (int d = 0; d < pd->dim; ++d) {
        if (!(p1[d] > p0[d])) return SDOT_INVALID_ARG;
        sh.p0[d] = p0[d];
        sh.p1[d] = p1[d];
    }
    sh.coeff = coeff;
    sh.cut_id = cut_id;
    item_clip_box(sh, pd->dim);
    pd->shapes.push_back(sh);
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_add_convex_polyhedron_shape(sdot_pd *pd, const double *positions, const double *normals, int64_t n,
                                                double coeff, int64_t cut_id) {
    if (!pd || !positions || !normals || n < 1) return SDOT_INVALID_ARG;
    if (cut_id != -1) {
        pd->err = "only the domain cut id -1 is supported";
        return SDOT_NOT_SUPPORTED;
    }
    const int dim = pd->dim;
    Shape sh{};
    // ConvexPolyhedronAssembly.tcc:7-33: the box of the points widened by its extent on
    // each side, cut by the planes (x - p_i) . n_i <= 0
    for (int d = 0; d < dim; ++d) {
        double lo = DBL_MAX, hi = -DBL_MAX;
        for (int64_t i = 0; i < n; ++i) {
            if (!std::isfinite(positions[dim * i + d]) || !std::isfinite(normals[dim * i + d])) return SDOT_INVALID_ARG;
            lo = std::min(lo, positions[dim * i + d]);
            hi = std::max(hi, positions[dim * i + d]);
        }
        const double e = hi - lo;
        sh.p0[d] = lo - e;
        sh.p1[d] = hi + e;
        if (!(sh.p1[d] > sh.p0[d])) { sh.p0[d] = lo - 1; sh.p1[d] = hi + 1; }
    }
    for (int64_t i = 0; i < n; ++i) {
        std::array<double, 4> pl{0, 0, 0, 0};
        double nn = 0;
        for (int d = 0; d < dim; ++d) {
            pl[d] = normals[dim * i + d];
            nn += pl[d] * pl[d];
            pl[3] += normals[dim * i + d] * positions[dim * i + d];
        }
        if (!(nn > 0)) return SDOT_INVALID_ARG;
        sh.planes.push_back(pl);
    }
    sh.coeff = coeff;
    sh.cut_id = cut_id;
    item_clip_box(sh, pd->dim);
    pd->shapes.push_back(sh);
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_add_simplex_shape(sdot_pd *pd, const double *points, double coeff, int64_t cut_id) {
    if (!pd || !points) return SDOT_INVALID_ARG;
    if (cut_id != -1) {
        pd->err = "only the domain cut id -1 is supported";
        return SDOT_NOT_SUPPORTED;
    }
    const int dim = pd->dim;
    Shape sh{};
    for (int d = 0; d < dim; ++d) {
        sh.p0[d] = DBL_MAX;
        sh.p1[d] = -DBL_MAX;
        for (int k = 0; k <= dim; ++k) {
            if (!std::isfinite(points[dim * k + d])) return SDOT_INVALID_ARG;
            sh.p0[d] = std::min(sh.p0[d], points[dim * k + d]);
            sh.p1[d] = std::max(sh.p1[d], points[dim * k + d]);
        }
        if (!(sh.p1[d] > sh.p0[d])) return SDOT_INVALID_ARG;
    }
    // one plane per facet (the points but vertex k), oriented away from vertex k
    for (int k = 0; k <= dim; ++k) {
        double f[3][3] = {}, q[3] = {};
        int m = 0;
        for (int j = 0; j <= dim; ++j) {
            if (j == k) {
                for (int d = 0; d < dim; ++d) q[d] = points[dim * j + d];
            } else {
                for (int d = 0; d < dim; ++d) f[m][d] = points[dim * j + d];
                ++m;
            }
        }
        double nv[3] = {0, 0, 0};
        if (dim == 2) {
            nv[0] = f[1][1] - f[0][1];
            nv[1] = -(f[1][0] - f[0][0]);
        } else {
            const double u[3] = {f[1][0] - f[0][0], f[1][1] - f[0][1], f[1][2] - f[0][2]};
            const double v[3] = {f[2][0] - f[0][0], f[2][1] - f[0][1], f[2][2] - f[0][2]};
            nv[0] = u[1] * v[2] - u[2] * v[1];
            nv[1] = u[2] * v[0] - u[0] * v[2];
            nv[2] = u[0] * v[1] - u[1] * v[0];
        }
        double s = 0, h = 0;
        for (int d = 0; d < dim; ++d) {
            s += (q[d] - f[0][d]) * nv[d];
            h += nv[d] * f[0][d];
        }
        if (s == 0) return SDOT_INVALID_ARG;   // degenerate simplex
        if (s > 0) {
            for (int d = 0; d < dim; ++d) nv[d] = -nv[d];
            h = -h;
        }
        sh.planes.push_back({nv[0], nv[1], nv[2], h});
    }
    sh.coeff = coeff;
    sh.cut_id = cut_id;
    item_clip_box(sh, pd->dim);
    pd->shapes.push_back(sh);
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_set_image_density(sdot_pd *pd, const double *min_pt, const double *max_pt, const double *data,
                                      const int64_t *sizes, int64_t nb_coeffs) {
    if (!pd || !min_pt || !max_pt || !data || !sizes) return SDOT_INVALID_ARG;
    // apply_poly (Polynomial.h:27-58): constant pixels, or 6 coefficients in 2D
    if (!(nb_coeffs == 1 || (nb_coeffs == 6 && pd->dim == 2))) {
        pd->err = "image coefficients per pixel: 1, or 6 in 2D";
        return SDOT_NOT_SUPPORTED;
    }
    int64_t tot = nb_coeffs;
    for (int d = 0; d < pd->dim; ++d) {
        if (sizes[d] < 1 || sizes[d] > (1 << 20) || !(max_pt[d] > min_pt[d])) return SDOT_INVALID_ARG;
        tot *= sizes[d];
    }
    return guard(pd, [&] {
        pd->img.ensure(tot);
        HIPCK(hipMemcpyAsync(pd->img.p, data, tot * sizeof(double), hipMemcpyHostToDevice, pd->stream));
        HIPCK(hipStreamSynchronize(pd->stream));
        for (int d = 0; d < 3; ++d) {
            pd->img_lo[d] = d < pd->dim ? min_pt[d] : 0;
            pd->img_hi[d] = d < pd->dim ? max_pt[d] : 1;
            pd->img_n[d] = d < pd->dim ? (int)sizes[d] : 1;
        }
        pd->img_nc = (int)nb_coeffs;
        pd->has_image = true;
        pd->pass_valid = false;
        return SDOT_OK;
    });
}

sdot_status sdot_pd_clear_shapes(sdot_pd *pd) {
    if (!pd) return SDOT_INVALID_ARG;
    pd->shapes.clear();
    pd->has_image = false;
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_set_translations(sdot_pd *pd, const double *translations, int64_t n) {
    if (!pd || n < 0 || (n > 0 && !translations)) return SDOT_INVALID_ARG;
    if (n > 64) {
        pd->err = "at most 64 translations";
        return SDOT_NOT_SUPPORTED;
    }
    const int dim = pd->dim;
    for (int64_t s = 0; s < n; ++s) {
        double nn = 0;
        for (int d = 0; d < dim; ++d) {
            if (!std::isfinite(translations[dim * s + d])) return SDOT_INVALID_ARG;
            nn += translations[dim * s + d] * translations[dim * s + d];
        }
        if (!(nn > 0)) return SDOT_INVALID_ARG;
    }
    return guard(pd, [&] {
        pd->htrans.assign(translations, translations + n * dim);
        if (n > 0) {
            pd->trans.ensure(pd->htrans.size());
            HIPCK(hipMemcpyAsync(pd->trans.p, pd->htrans.data(), pd->htrans.size() * sizeof(double),
                                 hipMemcpyHostToDevice, pd->stream));
            HIPCK(hipStreamSynchronize(pd->stream));
        }
        pd->pass_valid = false;
        return SDOT_OK;
    });
}

sdot_status sdot_pd_set_ball_clip(sdot_pd *pd, int on) {
    if (!pd) return SDOT_INVALID_ARG;
    pd->ball = on ? 1 : 0;
    pd->rf = 0;
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_set_radial_func(sdot_pd *pd, int func) {
    if (!pd) return SDOT_INVALID_ARG;
    // FunctionEnum: the ball cut (need_ball_cut) and the final integration function
    // (func_for_final_cp_integration) of each
    static const int ball[7] = {0, 1, 0, 1, 0, 1, 0}, rf[7] = {0, 0, 1, 1, 2, 2, 3};
    if (func < 0 || func > 6) return SDOT_INVALID_ARG;
    pd->ball = ball[func];
    pd->rf = rf[func];
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_set_radial_func_eps(sdot_pd *pd, int func, double eps) {
    if (!pd || !(eps > 0) || !std::isfinite(eps)) return SDOT_INVALID_ARG;
    const sdot_status st = sdot_pd_set_radial_func(pd, func);
    if (st == SDOT_OK) pd->rf_eps = eps;
    return st;
}

sdot_status sdot_pd_set_tuning(sdot_pd *pd, int dpc_x100, int block) {
    if (!pd || dpc_x100 < 0 || block < 0) return SDOT_INVALID_ARG;
    if (dpc_x100 != pd->dpc_x100) pd->grid_valid = false;
    pd->dpc_x100 = dpc_x100;
    pd->block = block;
    pd->pass_valid = false;
    return SDOT_OK;
}

sdot_status sdot_pd_set_linear_solver(sdot_pd *pd, int precond) {
    if (!pd || (precond != SDOT_PRECOND_AMG && precond != SDOT_PRECOND_JACOBI)) return SDOT_INVALID_ARG;
    pd->precond = precond;
    return SDOT_OK;
}

sdot_status sdot_pd_compute(sdot_pd *pd) {
    if (!pd) return SDOT_INVALID_ARG;
    return guard(pd, [&] { return compute_pass(pd); });
}

sdot_status sdot_pd_integrals(sdot_pd *pd, double *masses) {
    if (!pd || !masses) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        sdot_status st = ensure_pass(pd);
        export_vec(pd, pd->mass.p, masses);
        return st;
    });
}

sdot_status sdot_pd_centroids(sdot_pd *pd, double *centroids, double *masses) {
    if (!pd || !centroids) return SDOT_INVALID_ARG;
    if (pd->has_image && pd->img_nc != 1) {
        pd->err = "centroids of quadratic pixels are not offered";
        return SDOT_NOT_SUPPORTED;
    }
    return guard(pd, [&] {
        struct Want {
            sdot_pd *pd;
            explicit Want(sdot_pd *p) : pd(p) { pd->want_mom = true; }
            ~Want() { pd->want_mom = false; }
        } want(pd);
        sdot_status st = ensure_pass(pd);
        export_centroids(pd, centroids);
        if (masses) export_vec(pd, pd->mass.p, masses);
        return st;
    });
}

sdot_status sdot_pd_boundary_measures(sdot_pd *pd, double *ball_arc, double *perimeter) {
    if (!pd) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        sdot_status st = ensure_pass(pd);
        if (ball_arc) export_vec(pd, pd->arc.p, ball_arc);
        if (perimeter) export_vec(pd, pd->perim.p, perimeter);
        return st;
    });
}

sdot_status sdot_pd_der_integrals_wrt_weights(sdot_pd *pd, int64_t *offsets, int64_t *cols, double *vals,
                                              double *masses, int64_t cap, int64_t *nnz) {
    if (!pd || !nnz || (cols && (!offsets || !vals))) return SDOT_INVALID_ARG;
    if (pd->rf == 1) {
        pd->err = "R2 has no boundary items (TODO in the reference): integrals only";
        return SDOT_NOT_SUPPORTED;
    }
    return guard(pd, [&] {
        sdot_status st = ensure_pass(pd);
        export_csr(pd, offsets, cols, vals, cap, nnz);
        if (masses) export_vec(pd, pd->mass.p, masses);
        return st;
    });
}

sdot_status sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions(sdot_pd *pd, int64_t *offsets, int64_t *cols,
                                                                         double *vals, double *v_values, int64_t cap,
                                                                         int64_t *nnz) {
    if (!pd || !nnz || (cols && (!offsets || !vals))) return SDOT_INVALID_ARG;
    if (pd->has_image && pd->img_nc != 1) {
        pd->err = "position derivatives of quadratic pixels are not offered";
        return SDOT_NOT_SUPPORTED;
    }
    return guard(pd, [&] {
        // a pass that also writes the derivative blocks (and the first moments); it is a
        // full pass, kept for the other getters
        struct Restore {
            sdot_pd *pd;
            bool mom;
            ~Restore() {
                pd->want_jac = false;
                pd->want_mom = mom;
            }
        } restore{pd, pd->want_mom};
        pd->want_jac = true;
        pd->want_mom = true;
        sdot_status st = compute_pass(pd);
        export_jac(pd, offsets, cols, vals, v_values, cap, nnz);
        return st;
    });
}

sdot_status sdot_pd_last_timings(sdot_pd *pd, sdot_timings *t) {
    if (!pd || !t) return SDOT_INVALID_ARG;
    *t = pd->last;
    return SDOT_OK;
}

sdot_status sdot_ot_solve(sdot_pd *pd, const double *targets, double tol, int stop_rule, int max_iter,
                          double *weights_inout, sdot_solve_stats *stats) {
    if (!pd || !targets || !weights_inout || !(tol > 0) || max_iter < 1 || (stop_rule != 0 && stop_rule != 1))
        return SDOT_INVALID_ARG;
    if (pd->N <= 0) return SDOT_INVALID_ARG;
    if (pd->rf == 1) {
        pd->err = "R2 has no boundary items (TODO in the reference): no Newton solve";
        return SDOT_NOT_SUPPORTED;
    }
    return guard(pd, [&] { return ot_solve(pd, targets, tol, stop_rule, max_iter, weights_inout, stats); });
}

sdot_status sdot_pd_display_vtk(sdot_pd *pd, const char *filename, const char *field_name) {
    if (!pd || !filename || !filename[0]) return SDOT_INVALID_ARG;
    return guard(pd, [&] {
        const sdot_status st = ensure_pass(pd);
        if (st != SDOT_OK && st != SDOT_EMPTY_CELL) return st;
        write_vtk(pd, filename, field_name && field_name[0] ? field_name : "num_dirac");
        return st;
    });
}

sdot_status sdot_calibration_copy(void *dst, const void *src, int64_t nbytes) {
    if (!dst || !src || nbytes <= 0 || nbytes % 16) return SDOT_INVALID_ARG;
    const int64_t n = nbytes / 16;
    const unsigned nb = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, nullptr, static_cast<uint4 *>(dst),
                       static_cast<const uint4 *>(src), n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return SDOT_HIP_ERROR;
    return SDOT_OK;
}

// splitmix64 seeding xoshiro256** (Blackman & Vigna), u = (x >> 11) * 2^-53
void sdot_rng_uniform(uint64_t seed, int64_t n, double *out) {
    auto splitmix = [](uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    uint64_t s[4], sm = seed;
    for (auto &v : s) v = splitmix(sm);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        out[i] = (double)(r >> 11) * 0x1.0p-53;
    }
}

}  // extern "C"
