"""Device-resident simulation handle (kw_sim_* in include/kwhip.h).

Used where the time loop is driven piecewise: benchmarking with inputs
already in HBM, batching and serving.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .solver import _parse_options, alloc_outputs, prepare, shape_outputs


def launch_tiles(kgrid, medium, source, sensor, dim, *args, **kwargs) -> tuple:
    """Pencils per block (x, y, last axis) the library launches for this
    problem's grid (kw_tiles: measured by the tile autotuner on first use of
    a grid shape, cached afterwards; y is 0 in 1-D and 2-D)."""
    opts = _parse_options(args, kwargs)
    pp = prepare(kgrid, medium, source, sensor, dim, opts)
    lib = _abi.load(pp.info.backend)
    tiles = (C.c_int32 * 3)()
    _abi.call(lib.kw_tiles, C.byref(pp.prob), tiles)
    return tuple(int(v) for v in tiles)


class Simulation:
    """Simulation(kgrid, medium, source, sensor, dim, **options): inputs are
    copied to the device and all operators built on construction."""

    def __init__(self, kgrid, medium, source, sensor, dim, *args, **kwargs):
        opts = _parse_options(args, kwargs)
        self.pp = prepare(kgrid, medium, source, sensor, dim, opts)
        self.lib = _abi.load(self.pp.info.backend)
        self.h = C.c_void_p()
        _abi.call(self.lib.kw_sim_create, C.byref(self.pp.prob), C.byref(self.h))
        self.Nt = self.pp.info.Nt
        self.steps_total = self.Nt - 1 if self.pp.info.time_rev else self.Nt

    @property
    def steps_done(self) -> int:
        return int(self.lib.kw_sim_steps_done(self.h))

    @property
    def device_bytes(self) -> int:
        return int(self.lib.kw_sim_device_bytes(self.h))

    def step(self, n: int):
        _abi.call(self.lib.kw_sim_step, self.h, C.c_int64(n))

    def sync(self):
        _abi.call(self.lib.kw_sim_sync, self.h)

    def step_timed(self, n: int, per_kernel: bool = False):
        """Advance n steps timed with device events.  Returns (ms, kernels);
        with per_kernel, kernels maps kind -> dict(ms, launches, bytes) from
        a run in which every launch is bracketed by events."""
        ms = C.c_double()
        st = _abi.KernelStats() if per_kernel else None
        _abi.call(self.lib.kw_sim_step_timed, self.h, C.c_int64(n), C.byref(ms),
                  C.byref(st) if st is not None else None)
        kern = None
        if per_kernel:
            kern = {self.lib.kw_kind_name(k).decode(): {"ms": st.ms[k], "launches": st.launches[k],
                                                        "bytes": st.bytes[k]}
                    for k in range(_abi.NKINDS)}
        return ms.value, kern

    def run(self):
        self.step(self.steps_total - self.steps_done)
        return self.read()

    def read(self):
        out, arrays = alloc_outputs(self.pp)
        _abi.call(self.lib.kw_sim_read, self.h, C.byref(out))
        return shape_outputs(self.pp, arrays)

    def close(self):
        if self.h:
            self.lib.kw_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
