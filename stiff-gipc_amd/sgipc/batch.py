"""Batched Newton steps: independent scenes assembled into one triplet store
at disjoint vertex ranges, converted together, one MAS build over the union
and one batched PCG launch sequence in which every scene runs its own PCG
(sgipc_solve_batched).  Also the baseline it is measured against: the same
scenes solved one after another, each with its own plain FP64 Newton step.
"""
import ctypes as C

import numpy as np

from . import capi
from .scene import union_scene
from .sim import GroundSim
from .system import _p


class BatchedNewton:
    def __init__(self, scenes, device="cuda:0", precond=capi.PRECOND_MAS_FP32):
        self.scenes = scenes
        self.sc, self.row_start = union_scene(scenes)
        self.S = len(scenes)
        self.sim = GroundSim(self.sc, device)
        self.sim.sys.precond = precond
        self.stats = (capi.SolveStats * self.S)()

    @property
    def sys(self):
        return self.sim.sys

    def assemble(self):
        self.sim._detect()
        self.sim._assemble()

    def convert(self):
        return self.sys.convert()

    def build_precond(self):
        self.sys.build_precond()

    def autotune(self, precision=capi.PRECISION_MIXED):
        g, pg = C.c_int(0), C.c_int(0)
        self.sys.ctx.call("sgipc_autotune", _p(self.sys.b), self.S, self.row_start.ctypes.data, precision,
                          C.byref(g), C.byref(pg))
        return g.value, pg.value

    def set_mode(self, persistent, threads=1024, max_workgroups=0):
        """Batched form: the graph-captured launch sequence (False) or one
        persistent launch, a workgroup of `threads` per scene (True)."""
        self.sys.ctx.call("sgipc_set_batch_mode", int(bool(persistent)), threads, max_workgroups)

    def mode(self):
        p, t = C.c_int(0), C.c_int(0)
        self.sys.ctx.call("sgipc_get_batch_mode", C.byref(p), C.byref(t))
        return bool(p.value), t.value

    def solve(self, precision=capi.PRECISION_MIXED, tol=1e-6, stop_rule=capi.STOP_TRUE_REL, max_iter=0):
        o = capi.SolveOpts(precision, stop_rule, tol, 0.3, max_iter, 8)
        self.sys.ctx.call("sgipc_solve_batched", self.S, self.row_start.ctypes.data, _p(self.sys.b),
                          _p(self.sys.dx), C.byref(o), self.stats)
        return self.stats

    def newton_step(self, **kw):
        self.assemble()
        self.convert()
        self.build_precond()
        return self.solve(**kw)

    def scene_slice(self, s):
        a, b = int(self.row_start[s]), int(self.row_start[s + 1])
        return slice(3 * a, 3 * b)


class SequentialNewton:
    """The scenes one after another, each its own context and plain FP64
    Newton step (MAS FP64, true relative residual)."""

    def __init__(self, scenes, device="cuda:0"):
        self.sims = [GroundSim(s, device) for s in scenes]

    def newton_step(self, tol=1e-6):
        its = []
        for sim in self.sims:
            sim._detect()
            sim._assemble()
            s = sim.sys
            s.convert()
            s.build_precond()
            st = s.solve(tol=tol, stop_rule=capi.STOP_TRUE_REL, precision=capi.PRECISION_FP64)
            its.append((st.iterations, st.rel_residual, st.converged))
        return its
