"""Many simulations on one GPU: asynchronous submission, admission by device
footprint, and stepping of all admitted simulations together (one launch
per pass for every simulation with the same dimension and equation of
state).  Python view of kw_engine_* / kw_submit / kw_wait in
include/kwhip.h.

    eng = Engine()
    t = eng.submit(kgrid, medium, source, sensor, 3, PMLSize=10)
    ...                                 # submit more; they run together
    sensor_data = eng.wait(t)           # as kspaceFirstOrder3D returns it
    eng.close()

Each result equals what kspaceFirstOrder1D/2D/3D returns for the same inputs.
"""

from __future__ import annotations

import ctypes as C
import time

from . import _abi
from .solver import _parse_options, _prepare_as, alloc_outputs, prepare, shape_outputs


class Ticket:
    """A submitted simulation.  Keeps what is needed to shape its outputs."""

    def __init__(self, engine: "Engine", tid: int, pp):
        self.engine = engine
        self.id = tid
        self.pp = pp

    def poll(self) -> dict:
        return self.engine.poll(self)

    def wait(self):
        return self.engine.wait(self)

    def cancel(self):
        self.engine.cancel(self)


class Engine:
    def __init__(self, device_bytes: int = 0, steps_per_round: int = 0, admit_threads: int = 0):
        self.lib = _abi.load(_abi.active_backend())
        cfg = _abi.EngineConfig(int(device_bytes), int(steps_per_round), int(admit_threads))
        self.h = C.c_void_p()
        _abi.call(self.lib.kw_engine_create, C.byref(cfg), C.byref(self.h))

    def submit(self, kgrid, medium, source, sensor, dim: int, *args, **kwargs) -> Ticket:
        """Validate and pre-process as kspaceFirstOrder1D/2D/3D, copy the problem
        into the engine and return at once."""
        opts = _parse_options(args, kwargs)
        pp = prepare(kgrid, medium, source, sensor, dim, opts)
        tid = C.c_int64()
        _abi.call(self.lib.kw_submit, self.h, C.byref(pp.prob), C.byref(tid))
        pp.keep.clear()  # the engine holds its own copy of every array
        pp.prob = None
        return Ticket(self, tid.value, pp)

    def submit_as(self, kgrid, medium, source, sensor, *args, **kwargs) -> Ticket:
        """Validate and pre-process as kspaceFirstOrderAS (the axisymmetric
        solver), copy the problem into the engine and return at once; its
        steps run on the engine's streams beside the other simulations."""
        opts = _parse_options(args, kwargs)
        if opts.get("SaveToDisk"):
            raise ValueError("'SaveToDisk' writes a file: call kspaceFirstOrderAS")
        pp = _prepare_as(kgrid, medium, source, sensor, opts)
        tid = C.c_int64()
        _abi.call(self.lib.kw_submit, self.h, C.byref(pp.prob), C.byref(tid))
        pp.keep.clear()
        pp.prob = None
        return Ticket(self, tid.value, pp)

    def poll(self, ticket: Ticket) -> dict:
        info = _abi.TicketInfo()
        _abi.call(self.lib.kw_poll, self.h, C.c_int64(ticket.id), C.byref(info))
        return {"status": _abi.TICKET_STATUS[info.status], "error": info.error,
                "submit_s": info.submit_s, "start_s": info.start_s, "done_s": info.done_s,
                "steps_done": info.steps_done, "steps_total": info.steps_total,
                "grid_cells": info.grid_cells, "device_bytes": info.device_bytes}

    def wait(self, ticket: Ticket, with_info: bool = False):
        """Block until the simulation has finished and return its sensor data
        (raises KWError if it failed or was cancelled).  With `with_info` the
        ticket's final poll() record (engine-clock times) is returned too."""
        info = None
        if with_info:
            while True:
                info = self.poll(ticket)
                if info["status"] in ("done", "failed", "cancelled"):
                    break
                time.sleep(0.0005)
        out, arrays = alloc_outputs(ticket.pp)
        _abi.call(self.lib.kw_wait, self.h, C.c_int64(ticket.id), C.byref(out))
        res = shape_outputs(ticket.pp, arrays)
        return (res, info) if with_info else res

    def cancel(self, ticket: Ticket):
        _abi.call(self.lib.kw_cancel, self.h, C.c_int64(ticket.id))

    def stats(self) -> dict:
        st = _abi.EngineStats()
        _abi.call(self.lib.kw_engine_get_stats, self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in _abi.EngineStats._fields_}

    def close(self):
        if self.h:
            self.lib.kw_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
