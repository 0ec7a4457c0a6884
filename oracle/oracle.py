"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module. It loads oracle/_build/libppes_oracle.so, the plain-C restatement
of the reference solver (DG_solver.f90, drive.f90, LDF.f90, gauss.f90), and
exposes it with numpy arrays in the reference's Fortran layout.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libppes_oracle.so")

BC = {"periodic": 0, "extrap": 1, "fixed": 2, "briowu": 4, "cloud": 5, "jet": 6}
PROB = {"ot2pi": 0, "rotor": 1, "blast": 2, "otunit": 3, "blastref": 4, "otpert": 5, "alfven": 6, "khi": 7,
        "briowu": 8, "wave1d": 9, "cloud": 10, "jet": 11, "yee": 12}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 5 + [ctypes.c_int, dp]
        L.orc_create2.restype = ctypes.c_void_p
        L.orc_create2.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double] * 5 + [ctypes.c_int, dp]
        L.orc_ldf_modal_operator2.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_u", "orc_uavg"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ("orc_urk", "orc_ut"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ("orc_xx", "orc_yy"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_tables.argtypes = [ctypes.c_int] + [dp] * 5
        L.orc_ldf_modal_operator.argtypes = [ctypes.c_int, dp]
        L.orc_set_tables.argtypes = [ctypes.c_void_p] + [dp] * 5
        L.orc_init_problem.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double]
        L.orc_apply_bc.argtypes = [ctypes.c_void_p, dp]
        L.orc_post_process.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp]
        L.orc_post_process_ex.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), dp]
        L.orc_run_ex.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.POINTER(Summary)]
        L.orc_semidiscretize.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, ctypes.c_int, ctypes.POINTER(Summary)]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_error_smooth_alfven.argtypes = [ctypes.c_void_p, dp, dp, dp]
        L.orc_face_flux.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, dp]
        L.orc_set_flux.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_error_smooth1d.argtypes = [ctypes.c_void_p, dp, ctypes.c_double, dp, dp]
        L.orc_log_avg_divisor.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_log_avg_divisor.restype = ctypes.c_double
        _lib = L
    return _lib


class Summary(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt_last", ctypes.c_double), ("dt_lim", ctypes.c_double),
                ("tot_ent", ctypes.c_double), ("mass0", ctypes.c_double), ("mass", ctypes.c_double),
                ("steps", ctypes.c_int), ("retries", ctypes.c_int), ("negcells", ctypes.c_int),
                ("status", ctypes.c_int)]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tables(p):
    n = p + 1
    x, w = np.zeros(n), np.zeros(n)
    D, Pmn, Pnm = (np.zeros((n, n), order="F") for _ in range(3))
    lib().orc_tables(p, _dp(x), _dp(w), _dp(D), _dp(Pmn), _dp(Pnm))
    return dict(x=x, w=w, D=D, Pmn=Pmn, Pnm=Pnm)


def ldf_modal_operator(p):
    m = 2 * (p + 1) ** 2
    op = np.zeros((m, m), order="F")
    lib().orc_ldf_modal_operator(p, _dp(op))
    return op


def log_avg_divisor(L, R):
    return lib().orc_log_avg_divisor(float(L), float(R))


FLUX = {"espp_hll": 0, "pp_hll": 1, "espp_lf": 2, "ec_lf": 3}


def face_flux(kind, wl, wr, gamma):
    """One x-flux evaluation -> (nflux[8], nBn, sl, sr) as an array of 11."""
    out = np.zeros(11)
    a = np.ascontiguousarray(wl, dtype=np.float64)
    b = np.ascontiguousarray(wr, dtype=np.float64)
    lib().orc_face_flux(FLUX[kind], _dp(a), _dp(b), float(gamma), _dp(out))
    return out


def set_threads(n):
    lib().orc_set_threads(int(n))


class Oracle:
    """One solver instance of the restated reference: degree p in x and py in
    y (default py = p; py = 0 is the reference's 1D set-up)."""

    def __init__(self, nx, ny, p, domain, gamma, bc="periodic", bc_state=None, rk=3, py=None):
        py = p if py is None else py
        self.nx, self.ny, self.p, self.py, self.rk = nx, ny, p, py, rk
        self.N, self.NX, self.NY = p + 1, p + 1, py + 1
        st = np.zeros(8) if bc_state is None else np.asarray(bc_state, dtype=np.float64)
        self.h = lib().orc_create2(nx, ny, p, py, rk, *[float(v) for v in domain], float(gamma), BC[bc], _dp(st))
        if not self.h:
            raise ValueError("orc_create2 rejected the configuration")
        NX, NY = self.NX, self.NY
        self.shape_u = (8, NX, NY, nx + 2, ny + 2)
        self.shape_ut = (8, NX, NY, nx, ny)
        self.shape_avg = (8, nx + 2, ny + 2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _view(self, ptr, shape):
        n = int(np.prod(shape))
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape, order="F")

    @property
    def u(self):
        return self._view(lib().orc_u(self.h), self.shape_u)

    def urk(self, s):
        return self._view(lib().orc_urk(self.h, s), self.shape_u)

    def ut(self, s):
        return self._view(lib().orc_ut(self.h, s), self.shape_ut)

    @property
    def uavg(self):
        return self._view(lib().orc_uavg(self.h), self.shape_avg)

    @property
    def xx(self):
        return self._view(lib().orc_xx(self.h), (self.NX, self.nx))

    @property
    def yy(self):
        return self._view(lib().orc_yy(self.h), (self.NY, self.ny))

    def set_tables(self, t):
        f = lambda a: np.asfortranarray(a, dtype=np.float64)
        self._keep = [f(t["x"]), f(t["w"]), f(t["D"]), f(t["Pmn"]), f(t["Pnm"])]
        lib().orc_set_tables(self.h, *[_dp(a) for a in self._keep])

    def init_problem(self, prob, pert=None, amp=0.0):
        pert = np.zeros((1, 4)) if pert is None else np.ascontiguousarray(pert, dtype=np.float64)
        rc = lib().orc_init_problem(self.h, PROB[prob], _dp(pert), len(pert), float(amp))
        if rc:
            raise ValueError("unknown problem")

    def set_interior(self, u_int):
        self.u[:, :, :, 1:-1, 1:-1] = u_int

    def apply_bc(self, arr):
        assert arr.flags.f_contiguous
        lib().orc_apply_bc(self.h, _dp(arr))

    def post_process(self, arr, avg, apply_oedg, apply_ldf, oedg_dt, comput_ent=True, ent_offset=False):
        neg, nneg, ent = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        rc = lib().orc_post_process_ex(self.h, _dp(arr), _dp(avg), int(apply_oedg), int(apply_ldf),
                                       float(oedg_dt), int(ent_offset), int(comput_ent), ctypes.byref(neg),
                                       ctypes.byref(nneg), ctypes.byref(ent))
        return dict(rc=rc, neg_avg=bool(neg.value), num_negcell=nneg.value, tot_ent=ent.value)

    def semidiscretize(self, ut, arr):
        dtl = ctypes.c_double()
        lib().orc_semidiscretize(self.h, _dp(ut), ctypes.byref(dtl), _dp(arr))
        return dtl.value

    def set_flux(self, kind):
        lib().orc_set_flux(self.h, FLUX[kind])

    def error_smooth_alfven(self, arr=None):
        """compute_error_smooth_alfven (problems.f90:659-697) -> (l2, max)"""
        u = np.asfortranarray(self.u if arr is None else arr, dtype=np.float64)
        l2, mx = ctypes.c_double(), ctypes.c_double()
        lib().orc_error_smooth_alfven(self.h, _dp(u), ctypes.byref(l2), ctypes.byref(mx))
        return l2.value, mx.value

    def error_smooth1d(self, T, arr=None):
        """compute_error_smooth1D (problems.f90:699-742) -> (l2, max)"""
        u = np.asfortranarray(self.u if arr is None else arr, dtype=np.float64)
        l2, mx = ctypes.c_double(), ctypes.c_double()
        lib().orc_error_smooth1d(self.h, _dp(u), float(T), ctypes.byref(l2), ctypes.byref(mx))
        return l2.value, mx.value

    def run(self, t_final=-1.0, nsteps=0, cfl=0.2, oedg_mode=1, use_ldf=True, ent_offset=False):
        s = Summary()
        lib().orc_run_ex(self.h, float(t_final), int(nsteps), float(cfl), int(oedg_mode), int(use_ldf),
                         int(ent_offset), ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in Summary._fields_}
