"""ppes_semidiscretize_noalloc against ppes_semidiscretize.

The no-allocation entry point runs the residual kernels on device arrays the
caller owns, with its record and block map in a fixed static workspace. On
the same state it must give the residual and dt_lim of the existing entry
point bit for bit. Two calls run in a row, the second on a smaller grid, so
the workspace is reused with different contents; the output buffer starts
as NaN so that a node the kernels skip shows. Device buffers come from the
HIP runtime (hipMalloc / hipMemcpy through ctypes): the caller allocates,
the entry point does not.
"""
import ctypes

import numpy as np
import pytest

import ppesdg
from helpers import setup

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            L = ctypes.CDLL(name)
            break
        except OSError:
            L = None
    assert L is not None, "HIP runtime not found"
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DevArray:
    """A device copy of a C-contiguous float64 array."""

    def __init__(self, H, host):
        self.H, self.shape = H, host.shape
        self.nbytes = host.nbytes
        p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(p), self.nbytes) == 0
        self.ptr = p.value
        a = np.ascontiguousarray(host, dtype=np.float64)
        assert H.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.H.hipFree(self.ptr)
            self.ptr = None


def existing_entry_point(prob, nx, ny, p, py, flux, arith):
    """A non-trivial state (two RK3 steps, ghosts refilled) and its residual
    and dt_lim through ppes_semidiscretize."""
    s = setup(prob)
    g = ppesdg.Solver(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"], py=py)
    g.init_problem(prob)
    g.solve(-1.0, 2, 0.2 if prob != "blast" else 0.15, 1, True)
    g.apply_bc(0)
    g.set_arith(arith)
    lib = ppesdg.lib()
    ppesdg._check(lib.ppes_set_flux(g.h, ppesdg.FLUX[flux]), "ppes_set_flux")
    u = g.download_state(0)
    dtl = g.semidiscretize(1, 0)
    ut = g.download_residual(1)
    cfg = ppesdg.make_config(nx, ny, p, s["domain"], s["gamma"], s["bc"], s["bc_state"], py=py)
    g.close()
    return cfg, u, ut, dtl


CASES = [
    # (first call, second call on a smaller grid): (problem, nx, ny, p, py, flux, arith)
    (("rotor", 40, 36, 3, 3, "espp_hll", "fast"), ("ot2pi", 12, 10, 3, 3, "espp_hll", "fast")),
    (("blast", 30, 28, 4, 4, "espp_hll", "fast"), ("rotor", 9, 7, 1, 1, "espp_hll", "fast")),
    (("ot2pi", 33, 31, 2, 2, "espp_hll", "exact"), ("blast", 10, 11, 2, 2, "pp_hll", "fast")),
    (("briowu", 64, 2, 3, 0, "espp_hll", "fast"), ("ot2pi", 16, 16, 4, 4, "ec_lf", "exact")),
]


@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_residual_equals_existing_entry_point(first, second):
    H = hip()
    for case in (first, second):
        cfg, u, ut_ref, dtl_ref = existing_entry_point(*case)
        flux, arith = case[5], case[6]
        print(f"noalloc residual {case}", flush=True)
        du = DevArray(H, ppesdg.to_device_layout(u))
        dut = DevArray(H, np.full(ppesdg.to_device_layout(ut_ref).shape, np.nan))
        try:
            dtl = ppesdg.semidiscretize_noalloc(cfg, du.ptr, dut.ptr, flux=flux, arith=arith)
            ut = ppesdg.from_device_layout(dut.get())
        finally:
            du.free()
            dut.free()
        assert not np.isnan(ut).any(), case
        assert np.array_equal(ut, ut_ref), (case, float(np.max(np.abs(ut - ut_ref))))
        assert dtl == dtl_ref, (case, dtl, dtl_ref)


def test_noalloc_rejects_grids_beyond_the_workspace():
    """A grid needing more workgroups than the workspace's block map holds is
    refused before any device work; the limit is the header's."""
    s = setup("ot2pi")
    # P3 at the default 64-lane residual workgroup: 4 elements each
    nx = ny = 1100
    assert (nx * ny + 3) // 4 > 262144
    cfg = ppesdg.make_config(nx, ny, 3, s["domain"], s["gamma"], "periodic")
    dtl = ctypes.c_double()
    rc = ppesdg.lib().ppes_semidiscretize_noalloc(ctypes.byref(cfg), 0, 1, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                                  ctypes.byref(dtl), None)
    assert rc == ppesdg.BAD_ARG
    # and 1024^2 P3 (the single-instance benchmark grid) fits
    assert (1024 * 1024 + 3) // 4 <= 262144
