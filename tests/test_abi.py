"""The C-ABI libraries load and export every function include/kwhip.h
declares; struct layouts of the ctypes view match the header (CPU only, no
device calls)."""

import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "kwhip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(kw_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_api():
    names = declared_functions()
    for required in ("kw_run", "kw_output_sizes", "kw_sim_create", "kw_sim_step", "kw_sim_read",
                     "kw_sim_destroy", "kw_smooth", "kw_backend", "kw_abi_version"):
        assert required in names


@pytest.mark.parametrize("backend", ["hip", "emu"])
def test_library_exports_every_symbol(backend):
    from kwhip import _abi

    path = _abi.library_path(backend)
    assert path.exists(), f"{path} not built"
    lib = C.CDLL(str(path))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_emulator_reports_its_backend():
    from kwhip import _abi

    lib = _abi.load("emu")
    assert lib.kw_abi_version() == 2
    assert lib.kw_backend().decode() == "emulator"


def test_struct_sizes_match_header():
    """Compile a probe of the header's struct sizes with the host compiler
    and compare with ctypes."""
    import subprocess
    import tempfile

    from kwhip import _abi

    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "kwhip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(kw_problem), sizeof(kw_time_source), sizeof(kw_outputs),
         sizeof(kw_sizes), sizeof(kw_kernel_stats));
  printf("%zu %zu %zu\n", offsetof(kw_problem, p_source), offsetof(kw_problem, sensor_num_points),
         offsetof(kw_problem, time_reversal_boundary_data));
  printf("%zu %zu %zu\n", sizeof(kw_engine_config), sizeof(kw_ticket_info), sizeof(kw_engine_stats));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "probe.c"
        c.write_text(src)
        exe = Path(d) / "probe"
        subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(c), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [C.sizeof(_abi.Problem), C.sizeof(_abi.TimeSource), C.sizeof(_abi.Outputs),
            C.sizeof(_abi.Sizes), C.sizeof(_abi.KernelStats),
            _abi.Problem.p_source.offset, _abi.Problem.sensor_num_points.offset,
            _abi.Problem.time_reversal_boundary_data.offset,
            C.sizeof(_abi.EngineConfig), C.sizeof(_abi.TicketInfo), C.sizeof(_abi.EngineStats)]
    assert got == want


def test_errors_are_reported_not_raised():
    """Invalid problems come back as negative codes with a message."""
    from kwhip import _abi

    lib = _abi.load("emu")
    prob = _abi.Problem()
    prob.dim = 4
    h = C.c_void_p()
    buf = C.create_string_buffer(256)
    rc = lib.kw_sim_create(C.byref(prob), C.byref(h), buf, 256)
    assert rc == _abi.KW_EINVAL
    assert b"dim" in buf.value
    prob.dim, prob.Nx, prob.Ny, prob.Nz, prob.Nt = 2, 4099, 16, 1, 3   # prime: chirp-z length 8200 > 4096
    prob.dx = prob.dy = 1e-4
    prob.dt, prob.c_ref, prob.c0_scalar, prob.rho0_scalar = 1e-8, 1500.0, 1500.0, 1000.0
    rc = lib.kw_sim_create(C.byref(prob), C.byref(h), buf, 256)
    assert rc == _abi.KW_EUNSUP
    prob.dim, prob.Nx = 1, 16     # 1-D: Ny and Nz must be 1
    rc = lib.kw_sim_create(C.byref(prob), C.byref(h), buf, 256)
    assert rc == _abi.KW_EINVAL and b"1D" in buf.value
    prob.Ny = 1
    prob.u_source[1].flag = 3      # no uy source in 1-D
    rc = lib.kw_sim_create(C.byref(prob), C.byref(h), buf, 256)
    assert rc == _abi.KW_EINVAL and b"ux only" in buf.value


def test_interpolated_sensor_points_are_checked():
    """Cartesian points interpolated linearly: simplex vertices and weights
    come together, exclude a mask index list, stay inside the grid, and are
    refused with time reversal and the split velocity."""
    import numpy as np
    from kwhip import _abi

    lib = _abi.load("emu")
    prob = _abi.Problem()
    prob.dim, prob.Nx, prob.Ny, prob.Nz, prob.Nt = 2, 16, 16, 1, 3
    prob.dx = prob.dy = 1e-4
    prob.dt, prob.c_ref, prob.c0_scalar, prob.rho0_scalar = 1e-8, 1500.0, 1500.0, 1000.0
    prob.pml_size[0] = prob.pml_size[1] = 2
    prob.use_kspace = prob.use_sg = 1
    prob.sensor_num_points = 2
    prob.record_flags = _abi.REC["p"]
    tri = np.array([1, 5, 2, 6, 18, 22], dtype=np.uint64)     # [2 x 3] column-major
    bc = np.array([0.5, 0.25, 0.25, 0.5, 0.25, 0.25])
    bad = np.array([1, 5, 2, 6, 18, 257], dtype=np.uint64)
    h = C.c_void_p()
    buf = C.create_string_buffer(256)

    def create():
        rc = lib.kw_sim_create(C.byref(prob), C.byref(h), buf, 256)
        if rc == 0:
            lib.kw_sim_destroy(h)
        return rc

    prob.sensor_tri = _abi.uptr(tri)
    assert create() == _abi.KW_EINVAL and b"together" in buf.value
    prob.sensor_bc = _abi.dptr(bc)
    assert create() == 0
    prob.sensor_tri = _abi.uptr(bad)
    assert create() == _abi.KW_EINVAL and b"range" in buf.value
    prob.sensor_tri = _abi.uptr(tri)
    ang = np.zeros(2)
    prob.directivity_angle = _abi.dptr(ang)
    assert create() == _abi.KW_EINVAL and b"binary sensor mask" in buf.value
    prob.directivity_angle = None
    prob.record_flags = _abi.REC["u_split_field"]
    assert create() == _abi.KW_EINVAL and b"binary sensor mask" in buf.value
    prob.record_flags = _abi.REC["p"]
    idx = np.array([1, 2], dtype=np.uint64)
    prob.sensor_mask_index = _abi.uptr(idx)
    assert create() == _abi.KW_EINVAL and b"NULL" in buf.value


def test_cuboid_corners_are_checked():
    """sensor_mask_corners: the point count must match the cuboids, the
    corners lie in the grid with x2 >= x1, and an index list is not given
    as well."""
    import numpy as np
    from kwhip import _abi

    lib = _abi.load("emu")
    prob = _abi.Problem()
    prob.dim, prob.Nx, prob.Ny, prob.Nz, prob.Nt = 2, 16, 16, 1, 3
    prob.dx = prob.dy = 1e-4
    prob.dt, prob.c_ref, prob.c0_scalar, prob.rho0_scalar = 1e-8, 1500.0, 1500.0, 1000.0
    prob.pml_size[0] = prob.pml_size[1] = 2
    prob.use_kspace = prob.use_sg = 1
    prob.record_flags = _abi.REC["p"]
    cor = np.array([3, 4, 5, 6, 8, 8, 9, 9], dtype=np.uint64)   # 3x3 and 2x2 points
    prob.sensor_mask_corners = _abi.uptr(cor)
    prob.sensor_num_cuboids = 2
    h = C.c_void_p()
    buf = C.create_string_buffer(256)

    def create():
        rc = lib.kw_sim_create(C.byref(prob), C.byref(h), buf, 256)
        if rc == 0:
            lib.kw_sim_destroy(h)
        return rc

    prob.sensor_num_points = 12
    assert create() == _abi.KW_EINVAL and b"number of points" in buf.value
    prob.sensor_num_points = 13
    assert create() == 0
    bad = np.array([5, 4, 3, 6, 8, 8, 9, 9], dtype=np.uint64)
    prob.sensor_mask_corners = _abi.uptr(bad)
    assert create() == _abi.KW_EINVAL and b"x2 => x1" in buf.value
    out = np.array([3, 4, 5, 17, 8, 8, 9, 9], dtype=np.uint64)
    prob.sensor_mask_corners = _abi.uptr(out)
    assert create() == _abi.KW_EINVAL and b"within the grid" in buf.value
