"""ctypes binding of libppesdg.so, the MI355X PP-ES nodal DG step for 2D MHD.

The class ``Solver`` mirrors the reference's module API (DG_solver.f90,
drive.f90, problems.f90): ``post_process``, ``semidiscretize``,
``rk_combine``, ``apply_bc``, ``solve`` (= solve_execute). Host arrays use the
reference's Fortran layouts: state (8, NX, NY, nx+2, ny+2), residual
(8, NX, NY, nx, ny), averages (8, nx+2, ny+2), NX = p + 1, NY = py + 1. ``Batch`` runs many independent
instances (own mesh, degree, BC, dt and end time) through shared launches.

There is no CPU fallback: if the HIP library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libppesdg.so")

OK, NEG_AVG, DT_COLLAPSE, PP_FAIL, BAD_ARG, HIP_ERROR, OOM, UNSUPPORTED = range(8)
BC = {"periodic": 0, "extrap": 1, "fixed": 2, "host": 3, "briowu": 4, "cloud": 5, "jet": 6}
KERNELS = {"residual": 0, "stage": 1, "oedg": 2, "limit": 3, "reduce": 4, "bc": 5}
INST_STATE = {0: "idle", 1: "active", 2: "done", 3: "failed"}
ARITH = {"exact": 0, "fast": 1}
FLUX = {"espp_hll": 0, "pp_hll": 1, "espp_lf": 2, "ec_lf": 3}
PROB = {"ot2pi": 0, "rotor": 1, "blast": 2, "otunit": 3, "blastref": 4, "otpert": 5, "alfven": 6, "khi": 7,
        "briowu": 8, "wave1d": 9, "cloud": 10, "jet": 11, "yee": 12}

_lib = None


class PPESError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: status {code}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("px", ctypes.c_int), ("py", ctypes.c_int),
                ("rk_stages", ctypes.c_int), ("left", ctypes.c_double), ("right", ctypes.c_double),
                ("bottom", ctypes.c_double), ("top", ctypes.c_double), ("gamma", ctypes.c_double),
                ("bc_kind", ctypes.c_int), ("bc_state", ctypes.c_double * 8)]


class RunStats(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt_last", ctypes.c_double), ("dt_lim", ctypes.c_double),
                ("tot_ent", ctypes.c_double), ("steps", ctypes.c_longlong), ("retries", ctypes.c_longlong),
                ("negcells", ctypes.c_longlong), ("stages", ctypes.c_longlong), ("status", ctypes.c_int),
                ("pad", ctypes.c_int)]


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C pp-es-dg-mhd_amd`")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        sig = {
            "ppes_create": [ctypes.POINTER(Config), ctypes.POINTER(vp)],
            "ppes_destroy": [vp],
            "ppes_set_gamma": [vp, ctypes.c_double],
            "ppes_geometry": [vp, dp, dp, dp, dp, dp, dp],
            "ppes_upload_state": [vp, ctypes.c_int, dp],
            "ppes_download_state": [vp, ctypes.c_int, dp],
            "ppes_upload_residual": [vp, ctypes.c_int, dp],
            "ppes_download_residual": [vp, ctypes.c_int, dp],
            "ppes_download_avg": [vp, dp],
            "ppes_init_problem": [vp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double],
            "ppes_apply_bc": [vp, ctypes.c_int],
            "ppes_error_smooth_alfven": [vp, ctypes.c_int, dp, dp],
            "ppes_set_ent_offset": [vp, ctypes.c_int],
            "ppes_set_flux": [vp, ctypes.c_int],
            "ppes_batch_set_flux": [vp, ctypes.c_int, ctypes.c_int],
            "ppes_batch_set_ent_offset": [vp, ctypes.c_int, ctypes.c_int],
            "ppes_error_smooth1d": [vp, ctypes.c_int, ctypes.c_double, dp, dp],
            "ppes_post_process": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, ip, ip, dp],
            "ppes_semidiscretize": [vp, ctypes.c_int, ctypes.c_int, dp],
            "ppes_rk_combine": [vp, ctypes.c_int, ctypes.c_double],
            "ppes_solve": [vp, ctypes.c_double, ctypes.c_longlong, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, ctypes.POINTER(RunStats)],
            "ppes_tables": [ctypes.c_int, dp, dp, dp, dp, dp],
            "ppes_ldf_nodal_operator": [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp],
            "ppes_sync": [vp],
            "ppes_solve_begin": [vp, ctypes.c_int],
            "ppes_advance": [vp, ctypes.c_double, ctypes.c_longlong, ctypes.c_double, ctypes.c_int,
                             ctypes.POINTER(RunStats)],
            "ppes_set_timing": [vp, ctypes.c_int],
            "ppes_set_arith": [vp, ctypes.c_int],
            "ppes_batch_set_arith": [vp, ctypes.c_int],
            "ppes_autotune": [vp, ctypes.c_int, ip],
            "ppes_face_tb": [vp, ip],
            "ppes_batch_face_tb": [vp, ip],
            "ppes_batch_autotune": [vp, ctypes.c_int, ip],
            "ppes_kernel_time": [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong)],
            "ppes_batch_create": [ctypes.c_int, ctypes.POINTER(vp)],
            "ppes_batch_destroy": [vp],
            "ppes_batch_add": [vp, ctypes.POINTER(Config), ip],
            "ppes_batch_remove": [vp, ctypes.c_int],
            "ppes_batch_init_problem": [vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double],
            "ppes_batch_upload_state": [vp, ctypes.c_int, ctypes.c_int, dp],
            "ppes_batch_download_state": [vp, ctypes.c_int, ctypes.c_int, dp],
            "ppes_batch_download_avg": [vp, ctypes.c_int, dp],
            "ppes_batch_start": [vp, ctypes.c_int, ctypes.c_double, ctypes.c_longlong, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int],
            "ppes_batch_step": [vp, ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ip],
            "ppes_batch_status": [vp, ctypes.c_int, ip, ctypes.POINTER(RunStats)],
            "ppes_batch_memory": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)],
            "ppes_positivity": [vp, dp, dp],
            "ppes_batch_positivity": [vp, ctypes.c_int, dp, dp],
            "ppes_batch_set_timing": [vp, ctypes.c_int],
            "ppes_batch_kernel_time": [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_longlong)],
            "ppes_set_precombine": [vp, ctypes.c_int],
            "ppes_batch_set_precombine": [vp, ctypes.c_int],
            "ppes_tiles": [vp, ctypes.c_int, ip],
            "ppes_batch_tiles": [vp, ctypes.c_int, ip],
            "ppes_set_tile": [vp, ctypes.c_int, ctypes.c_int],
            "ppes_batch_set_tile": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.ppes_semidiscretize_noalloc.argtypes = [ctypes.POINTER(Config), ctypes.c_int, ctypes.c_int, vp, vp, dp, vp]
        L.ppes_semidiscretize_noalloc.restype = ctypes.c_int
        L.ppes_instance_bytes.argtypes = [ctypes.POINTER(Config)]
        L.ppes_instance_bytes.restype = ctypes.c_longlong
        L.ppes_status_string.argtypes = [ctypes.c_int]
        L.ppes_status_string.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, what, allow=(OK,)):
    if rc not in allow:
        raise PPESError(rc, f"{what} ({lib().ppes_status_string(rc).decode()})")
    return rc


def tables(p):
    """gauss_lobatto / gauss_lobatto_matrix tables of degree p (host only)."""
    n = p + 1
    x, w = np.zeros(n), np.zeros(n)
    D, Pmn, Pnm = (np.zeros((n, n), order="F") for _ in range(3))
    _check(lib().ppes_tables(p, _dp(x), _dp(w), _dp(D), _dp(Pmn), _dp(Pnm)), "ppes_tables")
    return dict(x=x, w=w, D=D, Pmn=Pmn, Pnm=Pnm)


def ldf_nodal_operator(p, hx, hy):
    m = 2 * (p + 1) ** 2
    op = np.zeros((m, m), order="F")
    _check(lib().ppes_ldf_nodal_operator(p, hx, hy, _dp(op)), "ppes_ldf_nodal_operator")
    return op


def make_config(nx, ny, p, domain, gamma, bc="periodic", bc_state=None, rk=3, py=None):
    """Degree p in x and py in y (default p; py = 0: the reference's 1D set-up)."""
    cfg = Config()
    cfg.nx, cfg.ny, cfg.px, cfg.py, cfg.rk_stages = nx, ny, p, p if py is None else py, rk
    cfg.left, cfg.right, cfg.bottom, cfg.top = [float(v) for v in domain]
    cfg.gamma = float(gamma)
    cfg.bc_kind = BC[bc]
    if bc_state is not None:
        for v in range(8):
            cfg.bc_state[v] = float(bc_state[v])
    return cfg


def to_device_layout(a):
    """Fortran state (8, NX, NY, nx+2, ny+2) or residual (8, NX, NY, nx, ny)
    -> the device layout [v][ey][ex][k + NX l], C-contiguous."""
    return np.ascontiguousarray(np.asarray(a).transpose(0, 4, 3, 2, 1))


def from_device_layout(d):
    """Inverse of to_device_layout: (8, ey, ex, l, k) -> Fortran (8, k, l, ex, ey)."""
    return np.asfortranarray(np.asarray(d).transpose(0, 4, 3, 2, 1))


def semidiscretize_noalloc(cfg, u_dev, ut_dev, flux="espp_hll", arith="fast", stream=None):
    """ppes_semidiscretize_noalloc on device buffers the caller owns (device
    addresses as integers, arrays in the device layout) -> dt_lim."""
    dtl = ctypes.c_double()
    _check(lib().ppes_semidiscretize_noalloc(ctypes.byref(cfg), FLUX[flux], ARITH[arith], ctypes.c_void_p(u_dev),
                                              ctypes.c_void_p(ut_dev), ctypes.byref(dtl), stream),
           "ppes_semidiscretize_noalloc")
    return dtl.value


def instance_bytes(nx, ny, p, rk=3, py=None):
    """Device bytes one instance of this size holds (host-side formula)."""
    cfg = make_config(nx, ny, p, (0.0, 1.0, 0.0, 1.0), 1.4, rk=rk, py=py)
    return lib().ppes_instance_bytes(ctypes.byref(cfg))


def _stats(st):
    return {k: getattr(st, k) for k, _ in RunStats._fields_ if k != "pad"}


NSHAPES = 12  # shape ids 1-4: P1-P4 x same degree, 5-8: P1-P4 x P0, 9-10: P5-P6 x same, 11-12: P5-P6 x P0


def shape_of(sid):
    px = sid if sid <= 4 else (sid - 4 if sid <= 10 else sid - 6)
    py = px if (sid <= 4 or 9 <= sid <= 10) else 0
    return px, py


def _tuned(call, name):
    out = (ctypes.c_int * NSHAPES)()
    _check(call(out), name)
    res = {}
    for sid in range(1, NSHAPES + 1):
        if out[sid - 1]:
            px, py = shape_of(sid)
            res[f"P{px}x{py}"] = int(out[sid - 1])
    return res


class Solver:
    """One device-resident instance (init_DG_solver ... finalize_DG_solver)."""

    def __init__(self, nx, ny, p, domain, gamma, bc="periodic", bc_state=None, rk=3, arith=None, py=None):
        cfg = make_config(nx, ny, p, domain, gamma, bc, bc_state, rk, py)
        h = ctypes.c_void_p()
        _check(lib().ppes_create(ctypes.byref(cfg), ctypes.byref(h)), "ppes_create")
        self.h = h
        if arith is not None:
            self.set_arith(arith)
        self.nx, self.ny, self.p, self.py, self.rk = nx, ny, p, cfg.py, rk
        self.shape = f"P{p}x{cfg.py}"
        self.N, self.NX, self.NY = p + 1, p + 1, cfg.py + 1
        NX, NY = self.NX, self.NY
        self.shape_u = (8, NX, NY, nx + 2, ny + 2)
        self.shape_ut = (8, NX, NY, nx, ny)
        self.shape_avg = (8, nx + 2, ny + 2)

    def close(self):
        if getattr(self, "h", None):
            lib().ppes_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def geometry(self):
        hx, hy = ctypes.c_double(), ctypes.c_double()
        gx, gy = np.zeros(self.nx + 1), np.zeros(self.ny + 1)
        xx, yy = np.zeros((self.NX, self.nx), order="F"), np.zeros((self.NY, self.ny), order="F")
        _check(lib().ppes_geometry(self.h, ctypes.byref(hx), ctypes.byref(hy), _dp(gx), _dp(gy), _dp(xx), _dp(yy)),
               "ppes_geometry")
        return dict(hx=hx.value, hy=hy.value, grid_x=gx, grid_y=gy, xx=xx, yy=yy)

    def upload_state(self, slot, u):
        a = np.asfortranarray(u, dtype=np.float64)
        assert a.shape == self.shape_u
        _check(lib().ppes_upload_state(self.h, slot, _dp(a)), "ppes_upload_state")

    def download_state(self, slot=0):
        a = np.zeros(self.shape_u, order="F")
        _check(lib().ppes_download_state(self.h, slot, _dp(a)), "ppes_download_state")
        return a

    def upload_residual(self, slot, ut):
        a = np.asfortranarray(ut, dtype=np.float64)
        assert a.shape == self.shape_ut
        _check(lib().ppes_upload_residual(self.h, slot, _dp(a)), "ppes_upload_residual")

    def download_residual(self, slot):
        a = np.zeros(self.shape_ut, order="F")
        _check(lib().ppes_download_residual(self.h, slot, _dp(a)), "ppes_download_residual")
        return a

    def download_avg(self):
        a = np.zeros(self.shape_avg, order="F")
        _check(lib().ppes_download_avg(self.h, _dp(a)), "ppes_download_avg")
        return a

    def init_problem(self, prob, pert=None, amp=0.0):
        if pert is None:
            _check(lib().ppes_init_problem(self.h, PROB[prob], None, 0, 0.0), "ppes_init_problem")
        else:
            pt = np.ascontiguousarray(pert, dtype=np.float64)
            _check(lib().ppes_init_problem(self.h, PROB[prob], _dp(pt), len(pt), float(amp)), "ppes_init_problem")

    def apply_bc(self, slot):
        _check(lib().ppes_apply_bc(self.h, slot), "ppes_apply_bc")

    def error_smooth_alfven(self, slot=0):
        """compute_error_smooth_alfven (problems.f90:659-697) -> (l2, max)"""
        l2, mx = ctypes.c_double(), ctypes.c_double()
        _check(lib().ppes_error_smooth_alfven(self.h, slot, ctypes.byref(l2), ctypes.byref(mx)),
               "ppes_error_smooth_alfven")
        return l2.value, mx.value

    def error_smooth1d(self, T, slot=0):
        """compute_error_smooth1D (problems.f90:699-742) -> (l2, max)"""
        l2, mx = ctypes.c_double(), ctypes.c_double()
        _check(lib().ppes_error_smooth1d(self.h, slot, float(T), ctypes.byref(l2), ctypes.byref(mx)),
               "ppes_error_smooth1d")
        return l2.value, mx.value

    def post_process(self, slot, apply_oedg, apply_ldf, oedg_dt, comput_ent=True, ent_offset=False):
        neg, nneg, ent = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        rc = lib().ppes_post_process(self.h, slot, int(apply_oedg), int(apply_ldf), float(oedg_dt), int(ent_offset),
                                     int(comput_ent), ctypes.byref(neg), ctypes.byref(nneg), ctypes.byref(ent))
        _check(rc, "ppes_post_process", allow=(OK, NEG_AVG, PP_FAIL))
        return dict(rc=rc, neg_avg=bool(neg.value), num_negcell=nneg.value, tot_ent=ent.value)

    def semidiscretize(self, ut_slot, u_slot):
        dtl = ctypes.c_double()
        _check(lib().ppes_semidiscretize(self.h, ut_slot, u_slot, ctypes.byref(dtl)), "ppes_semidiscretize")
        return dtl.value

    def positivity(self):
        """(min density, min pressure) over every interior GLL node of every
        post-processed stage since the solve started."""
        r, p = ctypes.c_double(), ctypes.c_double()
        _check(lib().ppes_positivity(self.h, ctypes.byref(r), ctypes.byref(p)), "ppes_positivity")
        return r.value, p.value

    def rk_combine(self, stage, dt):
        _check(lib().ppes_rk_combine(self.h, stage, float(dt)), "ppes_rk_combine")

    def solve(self, t_final=-1.0, nsteps=0, cfl=0.2, oedg_mode=1, use_ldf=True, ent_offset=False):
        st = RunStats()
        rc = lib().ppes_solve(self.h, float(t_final), int(nsteps), float(cfl), int(oedg_mode), int(use_ldf),
                              int(ent_offset), ctypes.byref(st))
        _check(rc, "ppes_solve", allow=(OK, DT_COLLAPSE, PP_FAIL, NEG_AVG))
        return {k: getattr(st, k) for k, _ in RunStats._fields_ if k != "pad"}

    def solve_begin(self, use_ldf=True, ent_offset=False):
        _check(lib().ppes_set_ent_offset(self.h, int(ent_offset)), "ppes_set_ent_offset")
        _check(lib().ppes_solve_begin(self.h, int(use_ldf)), "ppes_solve_begin")

    def advance(self, t_final=-1.0, nsteps=0, cfl=0.2, oedg_mode=1):
        st = RunStats()
        rc = lib().ppes_advance(self.h, float(t_final), int(nsteps), float(cfl), int(oedg_mode), ctypes.byref(st))
        _check(rc, "ppes_advance", allow=(OK, DT_COLLAPSE, PP_FAIL))
        return {k: getattr(st, k) for k, _ in RunStats._fields_ if k != "pad"}

    def set_timing(self, on):
        _check(lib().ppes_set_timing(self.h, int(on)), "ppes_set_timing")

    def set_arith(self, mode):
        """"exact" (the reference's rounding order) or "fast" (default)"""
        _check(lib().ppes_set_arith(self.h, ARITH[mode]), "ppes_set_arith")

    def set_precombine(self, on):
        """stage combine in the residual launches of later stages (default on; same bits)"""
        _check(lib().ppes_set_precombine(self.h, int(bool(on))), "ppes_set_precombine")

    def autotune(self, reps=5):
        """Residual workgroup size by timing (after solve_begin); returns
        {"P<px>x<py>": lanes} for the tuned shape."""
        return _tuned(lambda out: lib().ppes_autotune(self.h, int(reps), out), "ppes_autotune")

    def face_tb(self):
        """OEDG face workgroup size per tuned shape, {"P<px>x<py>": lanes}."""
        return _tuned(lambda out: lib().ppes_face_tb(self.h, out), "ppes_face_tb")

    def tiles(self, kernel):
        """lanes per workgroup the step launches kernel class `kernel`
        ("residual", "stage", "oedg", "limit") at, for this solver's degree"""
        return _tuned(lambda out: lib().ppes_tiles(self.h, KERNELS[kernel], out), "ppes_tiles")[self.shape]

    def set_tile(self, kernel, lanes):
        _check(lib().ppes_set_tile(self.h, KERNELS[kernel], int(lanes)), "ppes_set_tile")

    def set_flux(self, kind):
        """face flux: espp_hll (default), pp_hll, espp_lf or ec_lf (ESPP_HLL_flux_x,
        PP_HLL_flux_x, ESPP_LF_flux_x, EC_LF_flux_x of DG_solver.f90:894-1497)"""
        _check(lib().ppes_set_flux(self.h, FLUX[kind]), "ppes_set_flux")

    def kernel_time(self, kernel):
        """(total ms, launches) of one kernel class while timing is on."""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        _check(lib().ppes_kernel_time(self.h, KERNELS[kernel], ctypes.byref(ms), ctypes.byref(n)), "ppes_kernel_time")
        return ms.value, n.value

    def sync(self):
        _check(lib().ppes_sync(self.h), "ppes_sync")


class Batch:
    """Independent instances of one RK method advanced together: one launch
    per kernel class and degree covers every active instance."""

    def __init__(self, rk=3):
        h = ctypes.c_void_p()
        _check(lib().ppes_batch_create(int(rk), ctypes.byref(h)), "ppes_batch_create")
        self.h = h
        self.rk = rk
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            lib().ppes_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add(self, nx, ny, p, domain, gamma, bc="periodic", bc_state=None, py=None):
        cfg = make_config(nx, ny, p, domain, gamma, bc, bc_state, self.rk, py)
        i = ctypes.c_int()
        _check(lib().ppes_batch_add(self.h, ctypes.byref(cfg), ctypes.byref(i)), "ppes_batch_add")
        self.shapes[i.value] = (8, p + 1, cfg.py + 1, nx + 2, ny + 2)
        return i.value

    def remove(self, i):
        _check(lib().ppes_batch_remove(self.h, int(i)), "ppes_batch_remove")
        self.shapes.pop(i, None)

    def init_problem(self, i, prob, pert=None, amp=0.0):
        if pert is None:
            rc = lib().ppes_batch_init_problem(self.h, int(i), PROB[prob], None, 0, 0.0)
        else:
            pt = np.ascontiguousarray(pert, dtype=np.float64)
            rc = lib().ppes_batch_init_problem(self.h, int(i), PROB[prob], _dp(pt), len(pt), float(amp))
        _check(rc, "ppes_batch_init_problem")

    def upload_state(self, i, slot, u):
        a = np.asfortranarray(u, dtype=np.float64)
        assert a.shape == self.shapes[i]
        _check(lib().ppes_batch_upload_state(self.h, int(i), slot, _dp(a)), "ppes_batch_upload_state")

    def download_state(self, i, slot=0):
        a = np.zeros(self.shapes[i], order="F")
        _check(lib().ppes_batch_download_state(self.h, int(i), slot, _dp(a)), "ppes_batch_download_state")
        return a

    def start(self, i, t_final=-1.0, max_steps=0, cfl=0.2, oedg_mode=1, use_ldf=True, ent_offset=False):
        _check(lib().ppes_batch_set_ent_offset(self.h, int(i), int(ent_offset)), "ppes_batch_set_ent_offset")
        rc = lib().ppes_batch_start(self.h, int(i), float(t_final), int(max_steps), float(cfl), int(oedg_mode),
                                    int(use_ldf))
        _check(rc, "ppes_batch_start")

    def step(self, nsteps=0, use_graph=True):
        """Advance every active instance; returns (batch steps run, instances still active)."""
        done, act = ctypes.c_longlong(), ctypes.c_int()
        _check(lib().ppes_batch_step(self.h, int(nsteps), int(use_graph), ctypes.byref(done), ctypes.byref(act)),
               "ppes_batch_step")
        return done.value, act.value

    def status(self, i):
        stt, st = ctypes.c_int(), RunStats()
        _check(lib().ppes_batch_status(self.h, int(i), ctypes.byref(stt), ctypes.byref(st)), "ppes_batch_status")
        d = _stats(st)
        d["state"] = INST_STATE[stt.value]
        return d

    def positivity(self, i):
        """(min density, min pressure) of instance i over every interior GLL
        node of every post-processed stage since its start."""
        r, p = ctypes.c_double(), ctypes.c_double()
        _check(lib().ppes_batch_positivity(self.h, int(i), ctypes.byref(r), ctypes.byref(p)), "ppes_batch_positivity")
        return r.value, p.value

    def memory(self, i):
        b = ctypes.c_longlong()
        _check(lib().ppes_batch_memory(self.h, int(i), ctypes.byref(b)), "ppes_batch_memory")
        return b.value

    def set_timing(self, on):
        _check(lib().ppes_batch_set_timing(self.h, int(on)), "ppes_batch_set_timing")

    def set_arith(self, mode):
        _check(lib().ppes_batch_set_arith(self.h, ARITH[mode]), "ppes_batch_set_arith")

    def set_precombine(self, on):
        _check(lib().ppes_batch_set_precombine(self.h, int(bool(on))), "ppes_batch_set_precombine")

    def autotune(self, reps=5):
        """Residual workgroup size per degree among the active instances, by
        timing; returns {"P<px>x<py>": lanes} for the tuned shapes."""
        return _tuned(lambda out: lib().ppes_batch_autotune(self.h, int(reps), out), "ppes_batch_autotune")

    def face_tb(self):
        """OEDG face workgroup size per tuned shape, {"P<px>x<py>": lanes}."""
        return _tuned(lambda out: lib().ppes_batch_face_tb(self.h, out), "ppes_batch_face_tb")

    def tiles(self, kernel):
        """{"P<px>x<py>": lanes per workgroup} at which the step launches
        kernel class `kernel` ("residual", "stage", "oedg", "limit")"""
        return _tuned(lambda out: lib().ppes_batch_tiles(self.h, KERNELS[kernel], out), "ppes_batch_tiles")

    def set_tile(self, kernel, px, py, lanes):
        _check(lib().ppes_batch_set_tile(self.h, KERNELS[kernel], int(px), int(py), int(lanes)), "ppes_batch_set_tile")

    def set_flux(self, i, kind):
        """face flux of instance i (see Solver.set_flux)"""
        _check(lib().ppes_batch_set_flux(self.h, int(i), FLUX[kind]), "ppes_batch_set_flux")

    def kernel_time(self, kernel):
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        _check(lib().ppes_batch_kernel_time(self.h, KERNELS[kernel], ctypes.byref(ms), ctypes.byref(n)),
               "ppes_batch_kernel_time")
        return ms.value, n.value
