"""sdot_amd -- MI355X (gfx950) Laguerre-cell path behind the sdot API.

Python front-end over the C-ABI of libsdot_mi355x.so (include/sdot_mi355x.h).
Mirrors the reference's operator API for this path (paths relative to the
sdot tree):

  PowerDiagram.add_box_shape / set_image_density   Domains/ConvexPolyhedronAssembly.h:26-41,
                                                   Domains/ScaledImage.h:26
  PowerDiagram.integrals                           PowerDiagram/get_integrals.h:13
  PowerDiagram.der_integrals_wrt_weights           PowerDiagram/get_der_integrals_wrt_weights.h:16
  OptimalTransport.adjust_weights / solve          Solvers/OptimalTransportSolver.h:18-26

The HIP library is the only compute path: importing this package raises when
the built library is missing, and every call reports the library's status.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdot_mi355x.so")

SDOT_OK = 0
SDOT_EMPTY_CELL = 1
SDOT_NOT_CONVERGED = 2
SDOT_CAPACITY_OVERFLOW = 3
SDOT_INVALID_ARG = 4
SDOT_HIP_ERROR = 5
SDOT_OOM = 6
SDOT_NOT_SUPPORTED = 7
SDOT_PENDING = 8

PRECOND_AMG = 0
PRECOND_JACOBI = 1

FP64 = 0
MIXED = 1

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
_int = ctypes.c_int


class SolveStats(ctypes.Structure):
    _fields_ = [("newton_iters", _int), ("pcg_iters_total", _int), ("backtracks", _int), ("precond_levels", _int),
                ("max_rel_mass_err", _d), ("max_abs_dw", _d), ("nnz", _i64), ("seconds_cells", _d),
                ("seconds_solve", _d), ("seconds_total", _d)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Timings(ctypes.Structure):
    _fields_ = [("ms_grid", _d), ("ms_cells", _d), ("ms_rows", _d), ("ms_total", _d),
                ("nb_cells_overflow", _i64), ("nnz_offdiag", _i64), ("grid_cells", _i64),
                ("ms_cell_kernel", _d), ("cell_flops", _d), ("sum_faces", _i64), ("sum_arcs", _i64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [("newton_iters_max", _int), ("nb_converged", _int), ("pcg_iters_total", _int), ("backtracks", _int),
                ("graph_replays", _int), ("tuned_dpc_x100", _int), ("max_rel_mass_err", _d), ("seconds_cells", _d),
                ("seconds_solve", _d), ("seconds_total", _d), ("ms_pcg_iteration", _d), ("precond_levels", _int),
                ("precond_fallbacks", _int), ("tuned_cells_per_wave", _int), ("tuned_first_tier", _int),
                ("tuning_passes", _int), ("device_newton", _int), ("newton_replays", _int), ("seconds_device", _d)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("precision", _int), ("hbm_budget_bytes", _d), ("max_batch_problems", _int), ("workers_3d", _int),
                ("diracs_per_cell_x100", _int), ("tol", _d), ("stop_rule", _int), ("max_iter", _int)]


class ServerResult(ctypes.Structure):
    _fields_ = [("status", _int), ("newton_iters", _int), ("max_rel_mass_err", _d), ("t_submit", _d),
                ("t_start", _d), ("t_done", _d)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("queued_2d", _i64), ("queued_3d", _i64), ("running", _i64), ("completed", _i64),
                ("resident_bytes", _d), ("peak_resident_bytes", _d), ("budget_bytes", _d),
                ("cells_evaluated", _i64), ("ticks_2d", _i64), ("seconds", _d)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes): every symbol the C header declares
SIGNATURES = {
    "sdot_pd_create": (_int, [ctypes.POINTER(_p), _int, _int]),
    "sdot_pd_destroy": (_int, [_p]),
    "sdot_status_string": (ctypes.c_char_p, [_int]),
    "sdot_pd_last_error": (ctypes.c_char_p, [_p]),
    "sdot_pd_set_diracs": (_int, [_p, _p, _p, _i64]),
    "sdot_pd_set_weights": (_int, [_p, _p]),
    "sdot_pd_set_diracs_device": (_int, [_p, _p, _p, _i64]),
    "sdot_pd_add_box_shape": (_int, [_p, _p, _p, _d, _i64]),
    "sdot_pd_set_image_density": (_int, [_p, _p, _p, _p, _p, _i64]),
    "sdot_pd_clear_shapes": (_int, [_p]),
    "sdot_pd_set_ball_clip": (_int, [_p, _int]),
    "sdot_pd_set_translations": (_int, [_p, _p, _i64]),
    "sdot_pd_set_radial_func": (_int, [_p, _int]),
    "sdot_pd_set_radial_func_eps": (_int, [_p, _int, _d]),
    "sdot_pd_compute": (_int, [_p]),
    "sdot_pd_integrals": (_int, [_p, _p]),
    "sdot_pd_integrals_noalloc": (_int, [_p, _p, _p, _i64, _p]),
    "sdot_pd_add_convex_polyhedron_shape": (_int, [_p, _p, _p, _i64, _d, _i64]),
    "sdot_pd_add_simplex_shape": (_int, [_p, _p, _d, _i64]),
    "sdot_pd_centroids": (_int, [_p, _p, _p]),
    "sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions": (_int, [_p, _p, _p, _p, _p, _i64,
                                                                           ctypes.POINTER(_i64)]),
    "sdot_pd_boundary_measures": (_int, [_p, _p, _p]),
    "sdot_pd_der_integrals_wrt_weights": (_int, [_p, _p, _p, _p, _p, _i64, ctypes.POINTER(_i64)]),
    "sdot_ot_solve": (_int, [_p, _p, _d, _int, _int, _p, ctypes.POINTER(SolveStats)]),
    "sdot_pd_set_tuning": (_int, [_p, _int, _int]),
    "sdot_pd_set_linear_solver": (_int, [_p, _int]),
    "sdot_batch_set_linear_solver": (_int, [_p, _int]),
    "sdot_pd_last_timings": (_int, [_p, ctypes.POINTER(Timings)]),
    "sdot_rng_uniform": (None, [ctypes.c_uint64, _i64, _p]),
    "sdot_pd_dirac_buffers": (_int, [_p, ctypes.POINTER(_p), ctypes.POINTER(_p)]),
    "sdot_pd_mark_diracs_changed": (_int, [_p, _int]),
    "sdot_pd_csr_device": (_int, [_p, ctypes.POINTER(_p), ctypes.POINTER(_p), ctypes.POINTER(_p)]),
    "sdot_batch_create": (_int, [ctypes.POINTER(_p), _int, _int]),
    "sdot_batch_destroy": (_int, [_p]),
    "sdot_batch_last_error": (ctypes.c_char_p, [_p]),
    "sdot_batch_set_problems": (_int, [_p, _i64, _p, _p, _p, _p, _p, _int]),
    "sdot_batch_set_tuning": (_int, [_p, _int, _int]),
    "sdot_batch_solve": (_int, [_p, _d, _int, _int, _p, _p, _p, _p, ctypes.POINTER(BatchStats)]),
    "sdot_batch_set_schedule": (_int, [_p, _int, _int, _int, _int]),
    "sdot_calibration_copy": (_int, [_p, _p, _i64]),
    "sdot_pd_display_vtk": (_int, [_p, ctypes.c_char_p, ctypes.c_char_p]),
    "sdot_srv_create": (_int, [ctypes.POINTER(_p), ctypes.POINTER(ServerConfig)]),
    "sdot_srv_submit": (_int, [_p, _int, _i64, _p, _p, _p, _p, ctypes.POINTER(_i64)]),
    "sdot_srv_poll": (_int, [_p, _i64, ctypes.POINTER(_int)]),
    "sdot_srv_wait": (_int, [_p, _i64, _d, _p, ctypes.POINTER(ServerResult)]),
    "sdot_srv_get_stats": (_int, [_p, ctypes.POINTER(ServerStats)]),
    "sdot_srv_last_error": (ctypes.c_char_p, [_p]),
    "sdot_srv_destroy": (_int, [_p]),
}

_LIB = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C sdot_amd` "
                              "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


lib()


class SdotError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{lib().sdot_status_string(status).decode()}: {msg}")
        self.status = status


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rng_uniform(seed, n):
    """splitmix64 -> xoshiro256** uniform doubles in [0,1): the synthetic-input generator."""
    out = np.empty(int(n), dtype=np.float64)
    lib().sdot_rng_uniform(ctypes.c_uint64(seed), int(n), _ptr(out))
    return out


class PowerDiagram:
    """Laguerre (power) diagram of weighted diracs over a domain, on the GPU.

    `allow_empty=True` makes an empty cell a reported condition (`self.last_status`)
    rather than an exception, like the reference's int result."""

    def __init__(self, dim=2, precision="fp64", positions=None, weights=None, ball=False):
        self.dim = int(dim)
        self.precision = precision
        prec = FP64 if precision == "fp64" else MIXED
        h = _p()
        st = lib().sdot_pd_create(ctypes.byref(h), self.dim, prec)
        if st != SDOT_OK:
            raise SdotError(st, "sdot_pd_create")
        self._h = h
        self.last_status = SDOT_OK
        self.N = 0
        if ball:
            self.set_ball_clip(True)
        if positions is not None:
            self.set_diracs(positions, weights if weights is not None else np.zeros(len(positions)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sdot_pd_destroy(h)
            self._h = None

    def _check(self, st, what, allow_empty=False):
        self.last_status = st
        if st == SDOT_OK or (allow_empty and st == SDOT_EMPTY_CELL):
            return st
        raise SdotError(st, f"{what}: {lib().sdot_pd_last_error(self._h).decode()}")

    # -- inputs --
    def set_diracs(self, positions, weights):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, self.dim)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        if len(w) != len(pos):
            raise ValueError("positions and weights differ in length")
        self._check(lib().sdot_pd_set_diracs(self._h, _ptr(pos), _ptr(w), len(w)), "set_diracs")
        self.N = len(w)

    def set_weights(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
        if len(w) != self.N:
            raise ValueError("weights length")
        self._check(lib().sdot_pd_set_weights(self._h, _ptr(w)), "set_weights")

    def set_diracs_device(self, pos_ptr, w_ptr, n):
        """Device pointers already in HBM, allocated through the HIP runtime this library
        links (ROCm's libamdhip64.so.7; a second runtime in the process, such as the one
        torch bundles, does not share the GPU with it)."""
        self._check(lib().sdot_pd_set_diracs_device(self._h, _p(pos_ptr), _p(w_ptr), int(n)), "set_diracs_device")
        self.N = int(n)

    def add_box_shape(self, p0, p1, coeff=1.0, cut_id=-1):
        a = np.ascontiguousarray(p0, dtype=np.float64)
        b = np.ascontiguousarray(p1, dtype=np.float64)
        self._check(lib().sdot_pd_add_box_shape(self._h, _ptr(a), _ptr(b), float(coeff), int(cut_id)), "add_box_shape")

    def add_convex_polyhedron_shape(self, positions, normals, coeff=1.0, cut_id=-1):
        """ConvexPolyhedronAssembly::add_convex_polyhedron: (x - p_i) . n_i <= 0 for each i."""
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, self.dim)
        n = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, self.dim)
        self._check(lib().sdot_pd_add_convex_polyhedron_shape(self._h, _ptr(p), _ptr(n), len(p), float(coeff),
                                                              int(cut_id)), "add_convex_polyhedron_shape")

    def add_simplex_shape(self, points, coeff=1.0, cut_id=-1):
        """ConvexPolyhedronAssembly::add_simplex: dim+1 points."""
        q = np.ascontiguousarray(points, dtype=np.float64).reshape(self.dim + 1, self.dim)
        self._check(lib().sdot_pd_add_simplex_shape(self._h, _ptr(q), float(coeff), int(cut_id)),
                    "add_simplex_shape")

    def set_image_density(self, min_pt, max_pt, data):
        """data indexed [z][y][x] / [y][x] (x fastest), constant per pixel; in 2D an
        array [6][y][x] gives per-pixel quadratics c0 + c1 x + c2 y + c3 x^2 + c4 xy + c5 y^2."""
        a = np.ascontiguousarray(min_pt, dtype=np.float64)
        b = np.ascontiguousarray(max_pt, dtype=np.float64)
        img = np.ascontiguousarray(data, dtype=np.float64)
        nc = img.shape[0] if img.ndim == self.dim + 1 else 1
        sizes = np.ascontiguousarray(img.shape[-self.dim:][::-1], dtype=np.int64)
        self._keep_img = img
        self._check(lib().sdot_pd_set_image_density(self._h, _ptr(a), _ptr(b), _ptr(img), _ptr(sizes), nc),
                    "set_image_density")

    def clear_shapes(self):
        self._check(lib().sdot_pd_clear_shapes(self._h), "clear_shapes")

    def set_translations(self, translations):
        """SpZGrid::translations: every cell is also cut by the diracs translated by each
        vector (periodic images); an empty list clears them."""
        t = np.ascontiguousarray(np.asarray(translations, dtype=np.float64).reshape(-1, self.dim))
        self._check(lib().sdot_pd_set_translations(self._h, _ptr(t), len(t)), "set_translations")

    RADIAL_FUNCS = {"unit": 0, "in_ball_w05": 1, "r2": 2, "r2_in_ball_w05": 3, "wmr2": 4, "pp_wmr2": 5,
                    "exp_wmr2db": 6}

    def set_radial_func(self, name, eps=None):
        """FunctionEnum of the integrals: unit, in_ball_w05, r2, r2_in_ball_w05, wmr2, pp_wmr2,
        exp_wmr2db (ExpWmR2db{eps}: exp((w - |x - c|^2) / eps))."""
        if eps is None:
            self._check(lib().sdot_pd_set_radial_func(self._h, self.RADIAL_FUNCS[name]), "set_radial_func")
        else:
            self._check(lib().sdot_pd_set_radial_func_eps(self._h, self.RADIAL_FUNCS[name], float(eps)),
                        "set_radial_func_eps")

    def set_ball_clip(self, on=True):
        self._check(lib().sdot_pd_set_ball_clip(self._h, int(bool(on))), "set_ball_clip")

    def set_tuning(self, diracs_per_cell=0.0, block=0):
        self._check(lib().sdot_pd_set_tuning(self._h, int(round(diracs_per_cell * 100)), int(block)), "set_tuning")

    def set_linear_solver(self, precond="amg"):
        """'amg' (aggregation multigrid V-cycle, default) or 'jacobi' preconditioned CG."""
        code = PRECOND_AMG if precond == "amg" else PRECOND_JACOBI
        self._check(lib().sdot_pd_set_linear_solver(self._h, code), "set_linear_solver")

    # -- evaluation --
    def compute(self, allow_empty=True):
        return self._check(lib().sdot_pd_compute(self._h), "compute", allow_empty)

    def integrals(self, allow_empty=True):
        m = np.empty(self.N)
        self._check(lib().sdot_pd_integrals(self._h, _ptr(m)), "integrals", allow_empty)
        return m

    def integrals_noalloc(self, pos_ptr, w_ptr, n, masses_ptr, allow_empty=True):
        """Masses of n diracs given as device pointers (positions 2n interleaved, weights n)
        into the device array masses_ptr (n), with this handle's settings and no device
        allocation (2D, one box shape, n <= 262144); the handle's own pass is untouched."""
        return self._check(lib().sdot_pd_integrals_noalloc(self._h, _p(pos_ptr), _p(w_ptr), int(n), _p(masses_ptr)),
                           "integrals_noalloc", allow_empty)

    def centroids(self, allow_empty=True):
        """get_centroids (rand_ratio 0): (centroids[N, dim], masses[N]); 0 for empty cells."""
        c = np.empty((self.N, self.dim))
        m = np.empty(self.N)
        self._check(lib().sdot_pd_centroids(self._h, _ptr(c), _ptr(m)), "centroids", allow_empty)
        return c, m

    def der_centroids_and_integrals_wrt_weight_and_positions(self, allow_empty=True):
        """get_der_centroids_and_integrals_wrt_weight_and_positions: (offsets, cols, vals, v) with
        rows nu n + d (centroid coordinates, mass) and columns nu m + e (position
        coordinates, weight), nu = dim + 1."""
        nu = self.dim + 1
        nnz = _i64(0)
        f = lib().sdot_pd_der_centroids_and_integrals_wrt_weight_and_positions
        self._check(f(self._h, None, None, None, None, 0, ctypes.byref(nnz)), "der_centroids_and_integrals",
                    allow_empty)
        off = np.zeros(nu * self.N + 1, dtype=np.int64)
        cols = np.zeros(max(nnz.value, 1), dtype=np.int64)
        vals = np.zeros(max(nnz.value, 1))
        v = np.zeros(nu * self.N)
        self._check(f(self._h, _ptr(off), _ptr(cols), _ptr(vals), _ptr(v), nnz.value, ctypes.byref(nnz)),
                    "der_centroids_and_integrals", allow_empty)
        return off, cols[:nnz.value], vals[:nnz.value], v

    def boundary_measures(self, allow_empty=True):
        a = np.empty(self.N)
        p = np.empty(self.N)
        self._check(lib().sdot_pd_boundary_measures(self._h, _ptr(a), _ptr(p)), "boundary_measures", allow_empty)
        return a, p

    def der_integrals_wrt_weights(self, allow_empty=True):
        """CSR (offsets[N+1], columns, values) of d mass_i / d w_j and the masses."""
        nnz = _i64()
        self._check(lib().sdot_pd_der_integrals_wrt_weights(self._h, None, None, None, None, 0, ctypes.byref(nnz)),
                    "der_integrals_wrt_weights", allow_empty)
        cap = max(int(nnz.value), 1)
        off = np.empty(self.N + 1, dtype=np.int64)
        cols = np.empty(cap, dtype=np.int64)
        vals = np.empty(cap)
        m = np.empty(self.N)
        self._check(lib().sdot_pd_der_integrals_wrt_weights(self._h, _ptr(off), _ptr(cols), _ptr(vals), _ptr(m), cap,
                                                            ctypes.byref(nnz)), "der_integrals_wrt_weights", allow_empty)
        n = int(nnz.value)
        return off, cols[:n].copy(), vals[:n].copy(), m

    def mark_diracs_changed(self, positions_changed=True):
        self._check(lib().sdot_pd_mark_diracs_changed(self._h, int(bool(positions_changed))), "mark_diracs_changed")

    def dirac_buffers(self):
        a, b = _p(), _p()
        self._check(lib().sdot_pd_dirac_buffers(self._h, ctypes.byref(a), ctypes.byref(b)), "dirac_buffers")
        return a.value, b.value

    def display_vtk(self, filename, field_name="num_dirac", allow_empty=True):
        """display_vtk (PowerDiagram/display_vtk.h:13-25): the cells as a legacy ASCII VTK file."""
        self._check(lib().sdot_pd_display_vtk(self._h, str(filename).encode(), field_name.encode()), "display_vtk",
                    allow_empty)

    def csr_device(self, allow_empty=True):
        a, b, c = _p(), _p(), _p()
        self._check(lib().sdot_pd_csr_device(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "csr_device", allow_empty)
        return a.value, b.value, c.value

    def last_timings(self):
        t = Timings()
        self._check(lib().sdot_pd_last_timings(self._h, ctypes.byref(t)), "last_timings")
        return t.as_dict()

    # -- Newton solve --
    def ot_solve(self, targets, weights, tol=1e-6, stop_rule=0, max_iter=150):
        """OptimalTransportSolver::solve; returns (status, weights, stats dict)."""
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1)
        w = np.array(weights, dtype=np.float64, copy=True).reshape(-1)
        if len(t) != self.N or len(w) != self.N:
            raise ValueError(f"ot_solve: {len(t)} targets and {len(w)} weights for {self.N} diracs")
        stats = SolveStats()
        st = lib().sdot_ot_solve(self._h, _ptr(t), float(tol), int(stop_rule), int(max_iter), _ptr(w),
                                 ctypes.byref(stats))
        self.last_status = st
        if st not in (SDOT_OK, SDOT_NOT_CONVERGED):
            raise SdotError(st, f"ot_solve: {lib().sdot_pd_last_error(self._h).decode()}")
        return st, w, stats.as_dict()


def read_xyw(path):
    """The benchmarks' `.xyw` dirac files (benchmarks/set_up_diracs_2D.h:171-181): returns
    (positions (n, 2), weights (n,)) from whitespace-separated "x y w" triples."""
    v = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if v.size == 0:
        return np.zeros((0, 2)), np.zeros(0)
    if v.shape[1] != 3:
        raise ValueError(f"{path}: expected 3 values per line")
    return v[:, :2].copy(), v[:, 2].copy()


def write_xyw(path, positions, weights):
    p = np.asarray(positions, dtype=np.float64).reshape(-1, 2)
    w = np.asarray(weights, dtype=np.float64).reshape(-1)
    if len(p) != len(w):
        raise ValueError("one weight per position")
    np.savetxt(path, np.column_stack([p, w]), fmt="%.17g")


class OptimalTransport:
    """pysdot-style front door (`OptimalTransport(positions).adjust_weights()`)."""

    def __init__(self, positions, masses=None, domain=None, precision="fp64", ball=False, weights=None):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        self.dim = pos.shape[1]
        self.positions = pos
        n = len(pos)
        self.masses = np.full(n, 1.0 / n) if masses is None else np.asarray(masses, dtype=np.float64)
        self.weights = np.zeros(n) if weights is None else np.asarray(weights, dtype=np.float64)
        self.pd = PowerDiagram(self.dim, precision, ball=ball)
        if domain is None:
            self.pd.add_box_shape(np.zeros(self.dim), np.ones(self.dim))
        else:
            domain(self.pd)
        self.pd.set_diracs(pos, self.weights)
        self.stats = None

    def adjust_weights(self, tol=1e-6, stop_rule=0, max_iter=150):
        st, w, stats = self.pd.ot_solve(self.masses, self.weights, tol, stop_rule, max_iter)
        self.weights = w
        self.stats = stats
        self.status = st
        return w


class Batch:
    """K independent 2D problems solved together (one launch per kernel over all
    problems); per problem the rules of OptimalTransportSolver::solve
    (Solvers/OptimalTransportSolver.tcc:27-119).  A problem's result does not
    depend on the other problems of the batch."""

    def __init__(self, precision="mixed", diracs_per_cell=0.0, use_graph=True, precond="amg", device_newton=True,
                 cells_per_wave=0, first_tier=-1, cg_per_replay=16):
        h = _p()
        st = lib().sdot_batch_create(ctypes.byref(h), 2, FP64 if precision == "fp64" else MIXED)
        if st != SDOT_OK:
            raise SdotError(st, "sdot_batch_create")
        self._h = h
        self.precision = precision
        self.set_tuning(diracs_per_cell, use_graph)
        self.set_schedule(cells_per_wave, first_tier, device_newton, cg_per_replay)
        code = PRECOND_AMG if precond == "amg" else PRECOND_JACOBI
        self._check(lib().sdot_batch_set_linear_solver(self._h, code), "set_linear_solver")
        self.offsets = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sdot_batch_destroy(h)
            self._h = None

    def _check(self, st, what, ok=(SDOT_OK,)):
        if st in ok:
            return st
        raise SdotError(st, f"{what}: {lib().sdot_batch_last_error(self._h).decode()}")

    def set_tuning(self, diracs_per_cell=0.0, use_graph=True):
        self._check(lib().sdot_batch_set_tuning(self._h, int(round(diracs_per_cell * 100)), int(bool(use_graph))),
                    "set_tuning")

    def set_schedule(self, cells_per_wave=0, first_tier=-1, device_newton=True, cg_per_replay=16):
        """Pass schedule (cells per wavefront grab, capacity tier set; 0 / -1 = autotune)
        and whether Newton runs from one replayed device graph."""
        self._check(lib().sdot_batch_set_schedule(self._h, int(cells_per_wave), int(first_tier), int(bool(device_newton)),
                                                  int(cg_per_replay)), "set_schedule")

    def set_problems(self, positions, weights=None, targets=None, boxes=None, ball=False):
        """positions: list of (N_k, 2) arrays; weights/targets: lists of (N_k,) arrays
        (default 0 and 1/N_k * box area); boxes: (K, 4) array of (x0, y0, x1, y1)."""
        sizes = [len(p) for p in positions]
        off = np.zeros(len(sizes) + 1, dtype=np.int64)
        off[1:] = np.cumsum(sizes)
        pos = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 2) for p in positions]))
        if weights is None:
            w = np.zeros(off[-1])
        else:
            w = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in weights]))
        bx = None if boxes is None else np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 4)
        if targets is None:
            area = np.ones(len(sizes)) if bx is None else (bx[:, 2] - bx[:, 0]) * (bx[:, 3] - bx[:, 1])
            t = np.concatenate([np.full(n, a / n) for n, a in zip(sizes, area)])
        else:
            t = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in targets]))
        self._keep = (pos, w, t, bx, off)
        self._check(lib().sdot_batch_set_problems(self._h, len(sizes), _ptr(off), _ptr(pos), _ptr(w), _ptr(t),
                                                  None if bx is None else _ptr(bx), int(bool(ball))), "set_problems")
        self.offsets = off

    def solve(self, tol=1e-6, stop_rule=0, max_iter=150):
        """Returns (status, list of weights, list of masses, per-problem status, per-problem iterations, stats)."""
        off = self.offsets
        n, K = int(off[-1]), len(off) - 1
        w = np.empty(n)
        m = np.empty(n)
        pst = np.empty(K, dtype=np.int32)
        its = np.empty(K, dtype=np.int32)
        stats = BatchStats()
        st = lib().sdot_batch_solve(self._h, float(tol), int(stop_rule), int(max_iter), _ptr(w), _ptr(m), _ptr(pst),
                                    _ptr(its), ctypes.byref(stats))
        self._check(st, "solve", ok=(SDOT_OK, SDOT_NOT_CONVERGED))
        split = lambda a: [a[off[k]:off[k + 1]].copy() for k in range(K)]
        return st, split(w), split(m), pst, its, stats.as_dict()


class Server:
    """Serving loop: submit independent OT solve requests (2D or 3D), collect
    their converged weights.  2D requests share one continuous batch; 3D
    requests run on `workers_3d` concurrent device handles."""

    def __init__(self, precision="mixed", hbm_budget_bytes=0.0, max_batch_problems=0, workers_3d=0,
                 diracs_per_cell=0.0, tol=1e-6, stop_rule=0, max_iter=150):
        cfg = ServerConfig(FP64 if precision == "fp64" else MIXED, float(hbm_budget_bytes), int(max_batch_problems),
                           int(workers_3d), int(round(diracs_per_cell * 100)), float(tol), int(stop_rule),
                           int(max_iter))
        h = _p()
        st = lib().sdot_srv_create(ctypes.byref(h), ctypes.byref(cfg))
        if st != SDOT_OK:
            raise SdotError(st, "sdot_srv_create")
        self._h = h
        self._n = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().sdot_srv_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, positions, targets=None, weights=None, box=None):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        n, dim = pos.shape
        t = np.full(n, 1.0 / n) if targets is None else np.ascontiguousarray(targets, dtype=np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        b = None if box is None else np.ascontiguousarray(box, dtype=np.float64)
        rid = _i64()
        st = lib().sdot_srv_submit(self._h, dim, n, _ptr(pos), None if w is None else _ptr(w), _ptr(t),
                                   None if b is None else _ptr(b), ctypes.byref(rid))
        if st != SDOT_OK:
            raise SdotError(st, "sdot_srv_submit")
        self._n[rid.value] = n
        return rid.value

    def poll(self, rid):
        state = _int()
        st = lib().sdot_srv_poll(self._h, int(rid), ctypes.byref(state))
        if st != SDOT_OK:
            raise SdotError(st, "sdot_srv_poll")
        return state.value

    def wait(self, rid, timeout=-1.0):
        """Returns (weights, result dict), or None on a timeout."""
        n = self._n[rid]
        w = np.empty(n)
        res = ServerResult()
        st = lib().sdot_srv_wait(self._h, int(rid), float(timeout), _ptr(w), ctypes.byref(res))
        if st == SDOT_PENDING:
            return None
        if st != SDOT_OK:
            raise SdotError(st, "sdot_srv_wait")
        del self._n[rid]
        return w, res.as_dict()

    def stats(self):
        s = ServerStats()
        lib().sdot_srv_get_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def last_error(self):
        return lib().sdot_srv_last_error(self._h).decode()
