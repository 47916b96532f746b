"""ctypes view of the C ABI in include/kwhip.h.

The product library is ``libkwhip.so`` (HIP, gfx950).  ``libkwhip_emu.so`` is
the same host code over the CPU lane emulator.  The public API always runs
the HIP library; the test-suite alone switches to the emulator, with the
private ``_backend_override`` context.  There is no silent fallback: if the
HIP library is missing, loading fails.
"""

from __future__ import annotations

import contextlib
import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "_lib"
LIB_NAMES = {"hip": "libkwhip.so", "emu": "libkwhip_emu.so"}

KW_OK = 0
KW_EINVAL = -1
KW_EUNSUP = -2
KW_EDEVICE = -3
KW_ENOMEM = -4
KW_ESTATE = -5
KW_ENUMERIC = -6

REC = {
    "p": 1 << 0, "p_max": 1 << 1, "p_min": 1 << 2, "p_rms": 1 << 3,
    "p_max_all": 1 << 4, "p_min_all": 1 << 5, "p_final": 1 << 6,
    "u": 1 << 7, "u_max": 1 << 8, "u_min": 1 << 9, "u_rms": 1 << 10,
    "u_max_all": 1 << 11, "u_min_all": 1 << 12, "u_final": 1 << 13,
    "u_non_staggered": 1 << 14, "u_split_field": 1 << 15,
}
SRC_MODE = {"dirichlet": 0, "additive-no-correction": 1, "additive": 2}
NKINDS = 9
NOALLOC_MAX_BYTES = 128 << 20  # KW_NOALLOC_MAX_BYTES

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_u64_p = C.POINTER(C.c_uint64)


class TimeSource(C.Structure):
    _fields_ = [
        ("flag", C.c_int64),
        ("mode", C.c_int32),
        ("num_points", C.c_int64),
        ("index", c_u64_p),
        ("num_series", C.c_int64),
        ("sig_index", c_u64_p),
        ("input", c_double_p),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("Nx", C.c_int64), ("Ny", C.c_int64), ("Nz", C.c_int64),
        ("Nt", C.c_int64),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("c_ref", C.c_double),
        ("c0", c_double_p), ("c0_scalar", C.c_double),
        ("rho0", c_double_p), ("rho0_scalar", C.c_double),
        ("rho0_sg", c_double_p * 3),
        ("nonlinear_flag", C.c_int32),
        ("BonA", c_double_p), ("BonA_scalar", C.c_double),
        ("absorbing_flag", C.c_int32),
        ("alpha_coeff", c_double_p), ("alpha_coeff_scalar", C.c_double),
        ("alpha_power", C.c_double),
        ("alpha_mode", C.c_int32),
        ("alpha_sign", C.c_double * 2),
        ("alpha_sign_set", C.c_int32),
        ("pml_size", C.c_int64 * 3),
        ("pml_alpha", C.c_double * 3),
        ("use_kspace", C.c_int32),
        ("use_sg", C.c_int32),
        ("p0", c_double_p),
        ("p_source", TimeSource),
        ("u_source", TimeSource * 3),
        ("sensor_num_points", C.c_int64),
        ("sensor_mask_index", c_u64_p),
        ("record_flags", C.c_uint32),
        ("record_start_index", C.c_int64),
        ("time_reversal_boundary_data", c_double_p),
        ("sensor_tri", c_u64_p), ("sensor_bc", c_double_p),
        ("directivity_angle", c_double_p), ("directivity_pattern", C.c_int32),
        ("directivity_size", C.c_double), ("alpha_filter", c_double_p),
        ("sensor_mask_corners", c_u64_p), ("sensor_num_cuboids", C.c_int64),
        ("axisymmetric", C.c_int32),
    ]


class Sizes(C.Structure):
    _fields_ = [("num_sensor_points", C.c_int64), ("num_recorded_time_points", C.c_int64),
                ("grid_cells", C.c_int64)]


# summed split density at the sensor points (the new toolbox's density sensor)
KW_REC_RHO = 1 << 16


class Outputs(C.Structure):
    _fields_ = [
        ("p", c_float_p), ("p_max", c_double_p), ("p_min", c_double_p), ("p_rms", c_double_p),
        ("p_max_all", c_double_p), ("p_min_all", c_double_p), ("p_final", c_float_p),
        ("u", c_float_p * 3), ("u_max", c_double_p * 3), ("u_min", c_double_p * 3),
        ("u_rms", c_double_p * 3), ("u_max_all", c_double_p * 3), ("u_min_all", c_double_p * 3),
        ("u_final", c_float_p * 3), ("u_non_staggered", c_float_p * 3),
        ("u_split_p", c_float_p * 3), ("u_split_s", c_float_p * 3), ("rho", c_float_p),
    ]


class KernelStats(C.Structure):
    _fields_ = [("ms", C.c_double * NKINDS), ("launches", C.c_int64 * NKINDS),
                ("bytes", C.c_double * NKINDS)]


class EngineConfig(C.Structure):
    _fields_ = [("device_bytes", C.c_int64), ("steps_per_round", C.c_int32),
                ("admit_threads", C.c_int32)]


class TicketInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("error", C.c_int32), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("done_s", C.c_double), ("steps_done", C.c_int64),
                ("steps_total", C.c_int64), ("grid_cells", C.c_int64),
                ("device_bytes", C.c_int64)]


class EngineStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("cancelled", C.c_int64), ("queued", C.c_int64), ("running", C.c_int64),
                ("device_bytes", C.c_int64), ("device_budget", C.c_int64),
                ("cell_steps", C.c_double), ("busy_s", C.c_double), ("launches", C.c_int64)]


TICKET_STATUS = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


class KWError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kwhip error {code}: {msg}")
        self.code = code


_libs: dict[str, C.CDLL] = {}
_active = "hip"


def active_backend() -> str:
    """The library the public API runs: the HIP library, unless the
    test-suite has switched to the emulator build."""
    return _active


@contextlib.contextmanager
def _backend_override(name: str):
    """Test harness only: run the public API on another build of the library
    (``"emu"``: the CPU lane emulator) inside the ``with`` block."""
    global _active
    if name not in LIB_NAMES:
        raise ValueError(name)
    prev, _active = _active, name
    try:
        yield
    finally:
        _active = prev


def host_threads() -> int:
    """Host threads for the host-side precompute (FFT workers)."""
    return max(1, min(16, os.cpu_count() or 1))


def library_path(backend: str = "hip") -> Path:
    return LIB_DIR / LIB_NAMES[backend]


def load(backend: str = "hip") -> C.CDLL:
    """Load the kwhip library.  ``backend="hip"`` is the product; a missing
    library is an error (no fallback)."""
    if backend in _libs:
        return _libs[backend]
    path = library_path(backend)
    if not path.exists():
        raise ImportError(f"{path} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(path))
    P = C.POINTER
    lib.kw_abi_version.restype = C.c_int
    lib.kw_backend.restype = C.c_char_p
    lib.kw_kind_name.restype = C.c_char_p
    lib.kw_kind_name.argtypes = [C.c_int]
    err = [C.c_char_p, C.c_size_t]
    lib.kw_output_sizes.argtypes = [P(Problem), P(Sizes)] + err
    lib.kw_run.argtypes = [P(Problem), P(Outputs)] + err
    lib.kw_run_noalloc.argtypes = [P(Problem), P(Outputs)] + err
    lib.kw_sim_create.argtypes = [P(Problem), P(C.c_void_p)] + err
    lib.kw_tiles.argtypes = [P(Problem), P(C.c_int32)] + err
    lib.kw_sim_extend.argtypes = [C.c_void_p, C.c_int64] + err
    lib.kw_sim_change_dt.argtypes = [C.c_void_p, C.c_double, C.c_double] + err
    lib.kw_sim_init_velocity.argtypes = [C.c_void_p, c_double_p * 3] + err
    lib.kw_sim_get_fields.argtypes = [C.c_void_p, c_float_p, c_float_p * 3, c_float_p * 3] + err
    lib.kw_sim_step.argtypes = [C.c_void_p, C.c_int64] + err
    lib.kw_sim_sync.argtypes = [C.c_void_p] + err
    lib.kw_sim_read.argtypes = [C.c_void_p, P(Outputs)] + err
    lib.kw_sim_steps_done.argtypes = [C.c_void_p]
    lib.kw_sim_steps_done.restype = C.c_int64
    lib.kw_sim_device_bytes.argtypes = [C.c_void_p]
    lib.kw_sim_device_bytes.restype = C.c_int64
    lib.kw_sim_destroy.argtypes = [C.c_void_p]
    lib.kw_sim_step_timed.argtypes = [C.c_void_p, C.c_int64, P(C.c_double), P(KernelStats)] + err
    lib.kw_smooth.argtypes = [C.c_int32, P(C.c_int64), c_double_p, c_double_p, c_double_p] + err
    lib.kw_save_problem.argtypes = [P(Problem), C.c_char_p] + err
    lib.kw_run_file.argtypes = [C.c_char_p, C.c_char_p, C.c_uint32, C.c_int64] + err
    lib.kw_engine_create.argtypes = [P(EngineConfig), P(C.c_void_p)] + err
    lib.kw_submit.argtypes = [C.c_void_p, P(Problem), P(C.c_int64)] + err
    lib.kw_poll.argtypes = [C.c_void_p, C.c_int64, P(TicketInfo)] + err
    lib.kw_wait.argtypes = [C.c_void_p, C.c_int64, P(Outputs)] + err
    lib.kw_cancel.argtypes = [C.c_void_p, C.c_int64] + err
    lib.kw_engine_get_stats.argtypes = [C.c_void_p, P(EngineStats)] + err
    lib.kw_engine_destroy.argtypes = [C.c_void_p]
    if lib.kw_abi_version() != 2:
        raise ImportError("kwhip ABI version mismatch")
    _libs[backend] = lib
    return lib


def call(fn, *args):
    buf = C.create_string_buffer(1024)
    rc = fn(*args, buf, C.c_size_t(len(buf)))
    if rc != KW_OK:
        raise KWError(rc, buf.value.decode(errors="replace"))


def dptr(a):
    """double* of a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(c_double_p)


def fptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(c_float_p)


def uptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(c_u64_p)


def as_f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


_FLAT_POOL = None


def host_pool():
    """A few host threads for slab-parallel NumPy work on large grids (NumPy
    releases the GIL inside its copy and arithmetic loops)."""
    global _FLAT_POOL
    if _FLAT_POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _FLAT_POOL = ThreadPoolExecutor(max_workers=8)
    return _FLAT_POOL


def by_slabs(fn, n, k=8):
    """fn(lo, hi) over k slabs of range(n), on the host pool."""
    k = max(1, min(k, n))
    bounds = [n * i // k for i in range(k + 1)]
    list(host_pool().map(lambda i: fn(bounds[i], bounds[i + 1]), range(k)))


def fortran_flat(a):
    """Column-major (MATLAB) flattening of a grid array as float64.  Large
    row-major grids are transposed by slabs of the last axis on a few host
    threads (NumPy releases the GIL while it copies), several times faster
    than one strided copy."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim < 2 or a.size < (1 << 20) or a.flags.f_contiguous:
        return np.ascontiguousarray(a.ravel(order="F"))
    out = np.empty(a.shape, dtype=np.float64, order="F")

    def part(lo, hi):
        out[..., lo:hi] = a[..., lo:hi]

    by_slabs(part, a.shape[-1])
    return out.ravel(order="F")  # a view: out is column-major


def backend_name(lib) -> str:
    return lib.kw_backend().decode()


os.environ.setdefault("HSA_ENABLE_IPC_MODE_LEGACY", "0")
