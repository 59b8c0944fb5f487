"""ctypes front-end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
baseline.  Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OR_OK, OR_EMPTY_CELL, OR_NOT_CONVERGED, OR_CAPACITY, OR_INVALID = 0, 1, 2, 3, 4

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


class Domain(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("p0", _d * 3), ("p1", _d * 3), ("coeff", _d),
                ("img", _p), ("img_sizes", _i64 * 3), ("img_nc", _i64),
                ("nitems", _i64), ("it_box", _p), ("it_coeff", _p), ("it_planes", _p), ("it_off", _p),
                ("ntrans", _i64), ("trans", _p), ("rf_eps", _d)]


class SolveStats(ctypes.Structure):
    _fields_ = [("newton_iters", ctypes.c_int), ("pcg_iters_total", ctypes.c_int),
                ("backtracks", ctypes.c_int), ("max_rel_mass_err", _d),
                ("max_abs_dw", _d), ("nnz", _i64)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_cp2_eval.restype = ctypes.c_int
        L.or_cp2_eval.argtypes = [_d, _d, _d, _d, ctypes.c_int, _p, ctypes.c_int, _d, _d, _d,
                                  _p, _p, _p, _p, ctypes.c_int, _p]
        L.or_cp3_eval.restype = ctypes.c_int
        L.or_cp3_eval.argtypes = [_p, _p, ctypes.c_int, _p, _p, _p, _p, _p, ctypes.c_int]
        L.or_cells.restype = ctypes.c_int
        L.or_cells.argtypes = [ctypes.c_int, _i64, _p, _p, ctypes.POINTER(Domain), ctypes.c_int, ctypes.c_int,
                               _p, _p, _p, _i64, _p, _p, _p]
        L.or_cells_range.restype = ctypes.c_int
        L.or_cells_range.argtypes = [ctypes.c_int, _i64, _p, _p, ctypes.POINTER(Domain), ctypes.c_int, ctypes.c_int,
                                     _i64, _i64, _p, _p, _p, _i64, _p, _p, _p]
        L.or_rows_to_csr.restype = _i64
        L.or_rows_to_csr.argtypes = [_i64, _i64, _p, _p, _p, _p, _p, _p]
        L.or_pcg.restype = ctypes.c_int
        L.or_pcg.argtypes = [_i64, _p, _p, _p, _p, _p, _d, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.or_ot_solve.restype = ctypes.c_int
        L.or_ot_solve.argtypes = [ctypes.c_int, _i64, _p, _p, _p, ctypes.POINTER(Domain), ctypes.c_int,
                                  ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SolveStats)]
        L.or_cp2_centroid.restype = ctypes.c_int
        L.or_cp2_centroid.argtypes = [_d, _d, _d, _d, ctypes.c_int, _p, ctypes.c_int, _d, _d, _d, _p,
                                      ctypes.POINTER(_d)]
        L.or_centroids.restype = ctypes.c_int
        L.or_centroids.argtypes = [ctypes.c_int, _i64, _p, _p, ctypes.POINTER(Domain), ctypes.c_int, ctypes.c_int,
                                   _p, _p]
        L.or_der_pos.restype = ctypes.c_int
        L.or_der_pos.argtypes = [ctypes.c_int, _i64, _p, _p, ctypes.POINTER(Domain), ctypes.c_int, ctypes.c_int, _i64,
                                 _p, _p, _p, _p]
        L.or_rng_uniform.restype = None
        L.or_rng_uniform.argtypes = [ctypes.c_uint64, _i64, _p]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def rng_uniform(seed, n):
    out = np.empty(int(n), dtype=np.float64)
    lib().or_rng_uniform(ctypes.c_uint64(seed), int(n), _ptr(out))
    return out


def cp2_eval(box, cuts, ball=None, cap=256):
    """box=(x0,y0,x1,y1); cuts=[(ox,oy,nx,ny)] with normalised normals; ball=(cx,cy,r)."""
    c = np.ascontiguousarray(np.asarray(cuts, dtype=np.float64).reshape(-1, 4))
    area = ctypes.c_double()
    nbp = ctypes.c_int()
    nit = ctypes.c_int()
    meas = np.zeros(cap)
    ids = np.zeros(cap, dtype=np.int64)
    b = ball if ball is not None else (0.0, 0.0, 0.0)
    st = lib().or_cp2_eval(box[0], box[1], box[2], box[3], len(c), _ptr(c), int(ball is not None),
                           b[0], b[1], b[2], ctypes.byref(area), ctypes.byref(nbp), _ptr(meas), _ptr(ids),
                           cap, ctypes.byref(nit))
    n = nit.value
    return dict(status=st, area=area.value, nb_points=nbp.value, measures=meas[:n].copy(), ids=ids[:n].copy())


def cp2_centroid(box, cuts, ball=None):
    c = np.ascontiguousarray(np.asarray(cuts, dtype=np.float64).reshape(-1, 4))
    out = np.zeros(2)
    area = ctypes.c_double()
    b = ball if ball is not None else (0.0, 0.0, 0.0)
    st = lib().or_cp2_centroid(box[0], box[1], box[2], box[3], len(c), _ptr(c), int(ball is not None),
                               b[0], b[1], b[2], _ptr(out), ctypes.byref(area))
    return dict(status=st, centroid=out, area=area.value)


def centroids(pos, w, dim=None, ball=False, nthreads=None, domain=None, img=None, items=None, translations=None):
    """get_centroids: (status, centroids[N, dim], masses[N])."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dim = dim or pos.shape[1]
    N = pos.shape[0]
    w = np.ascontiguousarray(w, dtype=np.float64)
    if domain is None:
        dom, keep = make_domain(dim, img=img, items=items, translations=translations)
    else:
        dom, keep = make_domain(dim, *domain, img=img, items=items, translations=translations)
    c = np.zeros(N * dim)
    m = np.zeros(N)
    st = lib().or_centroids(dim, N, _ptr(pos), _ptr(w), ctypes.byref(dom), 1 if ball else 0,
                            nthreads or os.cpu_count() or 1, _ptr(c), _ptr(m))
    del keep
    return st, c.reshape(N, dim), m


def cp3_eval(p0, p1, cuts, cap=256):
    """cuts=[(ox,oy,oz,nx,ny,nz)]; keeps (o-p).n >= 0."""
    c = np.ascontiguousarray(np.asarray(cuts, dtype=np.float64).reshape(-1, 6))
    a0 = np.asarray(p0, dtype=np.float64)
    a1 = np.asarray(p1, dtype=np.float64)
    vol = ctypes.c_double()
    nf = ctypes.c_int()
    areas = np.zeros(cap)
    ids = np.zeros(cap, dtype=np.int64)
    st = lib().or_cp3_eval(_ptr(a0), _ptr(a1), len(c), _ptr(c), ctypes.byref(vol), ctypes.byref(nf),
                           _ptr(areas), _ptr(ids), cap)
    n = nf.value
    return dict(status=st, volume=vol.value, nb_faces=n, areas=areas[:n].copy(), ids=ids[:n].copy())


def convex_polyhedron_item(positions, normals, coeff=1.0):
    """ConvexPolyhedronAssembly::add_convex_polyhedron (ConvexPolyhedronAssembly.tcc:7-33):
    the box of the points widened by its extent on each side, cut by the planes
    (x - p_k) . n_k <= 0."""
    p = np.asarray(positions, dtype=np.float64)
    n = np.asarray(normals, dtype=np.float64)
    lo, hi = p.min(axis=0), p.max(axis=0)
    e = hi - lo
    h = (p * n).sum(axis=1)
    return dict(lo=lo - e, hi=hi + e, coeff=coeff, planes=np.column_stack([n, h]))


def simplex_item(points, coeff=1.0):
    """ConvexPolyhedronAssembly::add_simplex (:36-43): dim+1 points; one plane per facet,
    oriented away from the opposite vertex."""
    q = np.asarray(points, dtype=np.float64)
    dim = q.shape[1]
    planes = []
    for k in range(dim + 1):
        f = np.delete(q, k, axis=0)
        if dim == 2:
            t = f[1] - f[0]
            nrm = np.array([t[1], -t[0]])
        else:
            nrm = np.cross(f[1] - f[0], f[2] - f[0])
        if np.dot(q[k] - f[0], nrm) > 0:
            nrm = -nrm
        planes.append(np.concatenate([nrm, [np.dot(nrm, f[0])]]))
    return dict(lo=q.min(axis=0), hi=q.max(axis=0), coeff=coeff, planes=np.array(planes))


def box_item(p0, p1, coeff=1.0):
    dim = len(p0)
    return dict(lo=np.asarray(p0, dtype=np.float64), hi=np.asarray(p1, dtype=np.float64), coeff=coeff,
                planes=np.zeros((0, dim + 1)))


def make_domain(dim, p0=None, p1=None, coeff=1.0, img=None, items=None, translations=None):
    """translations: SpZGrid::translations, a (T, dim) array of vectors."""
    d = Domain()
    d.dim = dim
    p0 = p0 if p0 is not None else [0.0] * dim
    p1 = p1 if p1 is not None else [1.0] * dim
    for k in range(dim):
        d.p0[k] = p0[k]
        d.p1[k] = p1[k]
    d.coeff = coeff
    keep = None
    if img is not None:
        keep = np.ascontiguousarray(img, dtype=np.float64)
        d.img = keep.ctypes.data
        # img indexed [z][y][x] (x fastest); a leading axis holds the coefficients
        # of a per-pixel polynomial ([k][y][x], 6 of them in 2D)
        shp = keep.shape[-dim:][::-1]
        for k in range(dim):
            d.img_sizes[k] = shp[k]
        d.img_nc = keep.shape[0] if keep.ndim == dim + 1 else 1
    else:
        d.img = None
        d.img_nc = 1
    d.nitems = 0
    if items:
        box = np.zeros((len(items), 6))
        off = np.zeros(len(items) + 1, dtype=np.int64)
        planes = []
        for k, it in enumerate(items):
            box[k, :dim] = it["lo"]
            box[k, 3:3 + dim] = it["hi"]
            for pl in np.atleast_2d(it["planes"]):
                if len(pl) == 0:
                    continue
                row = np.zeros(4)
                row[:dim] = pl[:dim]
                row[3] = pl[dim]
                planes.append(row)
            off[k + 1] = len(planes)
        coeffs = np.array([it["coeff"] for it in items], dtype=np.float64)
        pl = np.ascontiguousarray(np.array(planes).reshape(-1, 4) if planes else np.zeros((1, 4)))
        keep = (keep, box, off, coeffs, pl)
        d.nitems = len(items)
        d.it_box = box.ctypes.data
        d.it_coeff = coeffs.ctypes.data
        d.it_planes = pl.ctypes.data
        d.it_off = off.ctypes.data
    d.ntrans = 0
    if translations is not None and len(translations):
        tr = np.ascontiguousarray(np.asarray(translations, dtype=np.float64).reshape(-1, dim))
        keep = (keep, tr)
        d.ntrans = len(tr)
        d.trans = tr.ctypes.data
    return d, keep


RADIAL = {"unit": 0, "r2": 1, "wmr2": 2, "expwmr2db": 3}


def cells(pos, w, dim=None, ball=False, brute=False, nthreads=None, domain=None, img=None, cap=None, items=None,
          translations=None, radial="unit", eps=1.0):
    """radial: the final integration function -- "unit" (Unit / InBallW05), "r2" (R2 /
    R2InBallW05: masses only), "wmr2" (WmR2 / PpWmR2 with ball=True) or "expwmr2db"
    (ExpWmR2db{eps}: exp((w - |x - c|^2) / eps), no ball)."""
    """Per-cell masses, ball arc, perimeter and raw rows.  Returns dict."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dim = dim or pos.shape[1]
    N = pos.shape[0]
    w = np.ascontiguousarray(w, dtype=np.float64)
    cap = cap or (64 if dim == 2 else 160)
    if domain is None:
        dom, keep = make_domain(dim, img=img, items=items, translations=translations)
    else:
        dom, keep = make_domain(dim, *domain, img=img, items=items, translations=translations)
    dom.rf_eps = eps
    nthreads = nthreads or os.cpu_count() or 1
    masses = np.zeros(N)
    arc = np.zeros(N)
    per = np.zeros(N)
    rl = np.zeros(N, dtype=np.int64)
    rc = np.zeros(N * cap, dtype=np.int64)
    rv = np.zeros(N * cap)
    flags = (1 if ball else 0) | (2 if brute else 0) | (RADIAL[radial] << 4)
    st = lib().or_cells(dim, N, _ptr(pos), _ptr(w), ctypes.byref(dom), flags, nthreads,
                        _ptr(masses), _ptr(arc), _ptr(per), cap, _ptr(rl), _ptr(rc), _ptr(rv))
    del keep
    return dict(status=st, masses=masses, ball_arc=arc, perimeter=per, row_len=rl, rcols=rc, rvals=rv, cap=cap)


def cells_sample_seconds(pos, w, i_begin, i_end, ball=False, nthreads=None, img=None, cap=None):
    """Wall time of the oracle evaluating cells [i_begin, i_end) of the full problem."""
    import time
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dim = pos.shape[1]
    N = pos.shape[0]
    w = np.ascontiguousarray(w, dtype=np.float64)
    cap = cap or (64 if dim == 2 else 160)
    dom, keep = make_domain(dim, img=img)
    masses = np.zeros(N)
    rl = np.zeros(N, dtype=np.int64)
    rc = np.zeros(N * cap, dtype=np.int64)
    rv = np.zeros(N * cap)
    t0 = time.perf_counter()
    st = lib().or_cells_range(dim, N, _ptr(pos), _ptr(w), ctypes.byref(dom), 1 if ball else 0,
                              nthreads or os.cpu_count() or 1, int(i_begin), int(i_end), _ptr(masses), None, None,
                              cap, _ptr(rl), _ptr(rc), _ptr(rv))
    dt = time.perf_counter() - t0
    del keep
    return dt, st, masses


def der_pos(pos, w, ball=False, nthreads=None, domain=None, img=None, items=None, translations=None, cap=None):
    """get_der_centroids_and_integrals_wrt_weight_and_positions: returns (status, offsets,
    columns, values, v) in the reference's layout -- row nu n + d (centroid coordinates,
    then the mass of dirac n), column nu m + e (position coordinates, then the weight of
    dirac m), nu = dim + 1; v[nu n + d]."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    N, dim = pos.shape
    nu = dim + 1
    cap = cap or (64 if dim == 2 else 160)
    w = np.ascontiguousarray(w, dtype=np.float64)
    if domain is None:
        dom, keep = make_domain(dim, img=img, items=items, translations=translations)
    else:
        dom, keep = make_domain(dim, *domain, img=img, items=items, translations=translations)
    jlen = np.zeros(N, dtype=np.int64)
    jcols = np.zeros(N * cap, dtype=np.int64)
    jblk = np.zeros(N * cap * 16)
    jv = np.zeros(nu * N)
    st = lib().or_der_pos(dim, N, _ptr(pos), _ptr(w), ctypes.byref(dom), 1 if ball else 0,
                          nthreads or os.cpu_count() or 1, cap, _ptr(jlen), _ptr(jcols), _ptr(jblk), _ptr(jv))
    del keep
    offsets = [0]
    cols, vals = [], []
    for n in range(N):
        L = int(jlen[n])
        c = jcols[n * cap:n * cap + L]
        b = jblk[n * cap * 16:(n * cap + L) * 16].reshape(L, 16)[:, :nu * nu].reshape(L, nu, nu)
        for d in range(nu):
            for k in range(L):
                cols.extend(nu * int(c[k]) + e for e in range(nu))
                vals.extend(b[k, d, :])
            offsets.append(len(cols))
    return st, np.array(offsets, dtype=np.int64), np.array(cols, dtype=np.int64), np.array(vals), jv


def to_csr(res):
    N = len(res["masses"])
    cap = res["cap"]
    nnz_max = int(res["row_len"].sum())
    off = np.zeros(N + 1, dtype=np.int64)
    cols = np.zeros(max(nnz_max, 1), dtype=np.int64)
    vals = np.zeros(max(nnz_max, 1))
    nnz = lib().or_rows_to_csr(N, cap, _ptr(res["row_len"]), _ptr(res["rcols"]), _ptr(res["rvals"]),
                               _ptr(off), _ptr(cols), _ptr(vals))
    return off, cols[:nnz].copy(), vals[:nnz].copy()


def der_integrals_wrt_weights(pos, w, **kw):
    res = cells(pos, w, **kw)
    off, cols, vals = to_csr(res)
    return res["status"], off, cols, vals, res["masses"]


def pcg(off, cols, vals, b, rtol=1e-12, max_iter=100000):
    x = np.zeros(len(b))
    it = ctypes.c_int()
    st = lib().or_pcg(len(b), _ptr(off), _ptr(cols), _ptr(vals), _ptr(np.ascontiguousarray(b)), _ptr(x),
                      rtol, max_iter, ctypes.byref(it))
    return st, x, it.value


def ot_solve(pos, w0, targets, ball=False, tol=1e-6, stop_rule=0, max_iter=150, nthreads=None,
             domain=None, img=None, items=None, translations=None, radial="unit", eps=1.0):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dim = pos.shape[1]
    w = np.array(w0, dtype=np.float64, copy=True)
    t = np.ascontiguousarray(targets, dtype=np.float64)
    if domain is None:
        dom, keep = make_domain(dim, img=img, items=items, translations=translations)
    else:
        dom, keep = make_domain(dim, *domain, img=img, items=items, translations=translations)
    dom.rf_eps = eps
    stats = SolveStats()
    st = lib().or_ot_solve(dim, pos.shape[0], _ptr(pos), _ptr(w), _ptr(t), ctypes.byref(dom),
                           (1 if ball else 0) | (RADIAL[radial] << 4), nthreads or os.cpu_count() or 1, tol,
                           stop_rule, max_iter, ctypes.byref(stats))
    del keep
    return st, w, stats
