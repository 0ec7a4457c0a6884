"""The Tecplot writers of the Fortran drop-in modules (io.f90 counterpart)
against the files the reference itself wrote (fixture end_ot2pi_p1_32:
solution_0.plt from save_tecbin_finite_volume, main.f90:67, and entropy.plt
from solve_execute, drive.f90:338-341).

Given the reference's own final cell averages and grid, the solution file is
byte-identical. The entropy log is fed the values stored in the reference's
file (FP32); every byte but the FP64 min/max pair of each variable (which the
reference takes from its FP64 log) is identical."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORT = os.path.join(ROOT, "pp-es-dg-mhd_amd", "fortran")
GOLD = os.path.join(ROOT, "tests", "golden", "end_ot2pi_p1_32.npz")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

PROGRAM = """
program plt_writer
	use MHD_PDE_m, only: set_gamma
	use io
	implicit none
	integer :: u, nx, ny, n, mode
	real :: g, t
	real, allocatable :: gx(:), gy(:), avg(:,:,:), log2(:,:)
	open(newunit=u, file='in.bin', access='stream', form='unformatted', status='old')
	read(u) mode, nx, ny, n
	read(u) g, t
	if (mode == 0) then
		allocate(gx(nx+1), gy(ny+1), avg(8, 0:nx+1, 0:ny+1))
		read(u) gx, gy, avg
		call set_gamma(g)
		call save_tecbin_finite_volume(avg, nx, ny, gx, gy, 0, t)
	else
		allocate(log2(2, n))
		read(u) log2
		call save_tecbin_entlog(log2, n, 'entropy.plt', '32 32 1 1')
	end if
	close(u)
end program plt_writer
"""


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    d = tmp_path_factory.mktemp("plt")
    srcs = [os.path.join(FORT, f) for f in ("utility.f90", "MHD_PDE.f90", "io.f90")]
    with open(d / "prog.f90", "w") as f:
        f.write(PROGRAM)
    subprocess.check_call([FLANG, "-fdefault-real-8", "-cpp", "-O2", "-o", str(d / "plt_writer")] + srcs +
                          [str(d / "prog.f90")], cwd=d)
    return d


def zone_data_offset(b, nvars):
    """offset of the per-variable FP64 min/max block: after the zone marker
    299.0, nvars format ints, the passive and sharing flags and the
    connectivity-sharing -1"""
    marker = struct.pack("<f", 299.0)
    m = b.rfind(marker)
    assert m > 0
    return m + 4 + 4 * nvars + 12


def test_solution_plt_byte_identical(writer):
    z = np.load(GOLD)
    a = json.loads(str(z["args_json"]))
    nx, ny = a["nx"], a["ny"]
    with open(writer / "in.bin", "wb") as f:
        f.write(struct.pack("<4i", 0, nx, ny, 0))
        f.write(struct.pack("<2d", 5.0 / 3.0, a["tfinal"]))
        f.write(np.asfortranarray(z["grid_x"]).tobytes(order="F"))
        f.write(np.asfortranarray(z["grid_y"]).tobytes(order="F"))
        f.write(np.asfortranarray(z["avg_final"]).tobytes(order="F"))
    subprocess.check_call([str(writer / "plt_writer")], cwd=writer, stdout=subprocess.DEVNULL)
    mine = (writer / "solution_0.plt").read_bytes()
    ref = z["plt_solution_0"].tobytes()
    assert len(mine) == len(ref)
    assert mine == ref


def test_entropy_log_layout(writer):
    z = np.load(GOLD)
    ref = z["plt_entropy"].tobytes()
    off = zone_data_offset(ref, 2)
    n = (len(ref) - off - 32) // 8
    data = np.frombuffer(ref[off + 32:], dtype="<f4").reshape(2, n)
    log2 = np.asfortranarray(np.stack([data[0], data[1]]).astype(np.float64))
    with open(writer / "in.bin", "wb") as f:
        f.write(struct.pack("<4i", 1, 0, 0, n))
        f.write(struct.pack("<2d", 0.0, 0.0))
        f.write(log2.tobytes(order="F"))
    subprocess.check_call([str(writer / "plt_writer")], cwd=writer, stdout=subprocess.DEVNULL)
    mine = (writer / "entropy.plt").read_bytes()
    assert len(mine) == len(ref)
    assert mine[:off] == ref[:off]                   # header, zone header, markers
    assert mine[off + 32:] == ref[off + 32:]         # FP32 data
    mm, rr = np.frombuffer(mine[off:off + 32], "<f8"), np.frombuffer(ref[off:off + 32], "<f8")
    assert np.allclose(mm, rr, rtol=1e-6, atol=0)    # FP64 min/max of FP32 vs FP64 values
