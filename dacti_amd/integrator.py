"""Integrator facade over the C ABI (include/dacti_gpu.h).

Mirrors dacti::integrator::Integrator<Model, Limiter, ErrorEstimator>
(src/lib/integrator/integrator.hpp:14-225): construction from a case and its
initial tree runs the constructor's initialisation on the GPU, integrate()
is one global step (impl/integrate.hpp:11-87) and the read-side calls
(n_active_cells, get_nodes, for_each_active_cell, global_dt,
max_acti_level) match the reference's names and meaning.  Errors raised by
the C ABI become RuntimeError; a NaN state is reported through
force_break, as the reference does (impl/cell_update.hpp:52-57).
"""
import ctypes as C

import numpy as np

from . import _lib
from .scene import build_tree

METADATA_NAMES = ("acti_level", "error", "max_dt", "cfl")  # integrator.hpp:30
MAX_LEVEL = 16


class Integrator:
    NumMeta = 4
    MAX_LEVEL = MAX_LEVEL
    METADATA_NAMES = METADATA_NAMES

    def __init__(self, case, tree=None):
        self.case = case
        self.lib = _lib.load()
        self.tree = tree if tree is not None else build_tree(case)
        self._cs = _lib.case_struct(case)
        self._ts = _lib.tree_struct(self.tree)
        h = C.c_void_p()
        rc = self.lib.dacti_create(C.byref(self._cs), C.byref(self._ts), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"dacti_create failed ({rc}): {self.lib.dacti_last_error(None).decode()}")
        self.h = h
        self.NumDims = case.dim

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.dacti_last_error(self.h).decode()}")

    def integrate(self, substep_hook=None, signal=True):
        """One global step; returns (global_dt, force_break).

        substep_hook(integrator, level, acti_time, n_substeps, dt) runs after
        every LTS substep, as in impl/integrate.hpp:64 (through
        dacti_set_substep_hook); an exception it raises is re-raised here
        after the step."""
        dt = C.c_double()
        fb = C.c_int32()
        if substep_hook is None:
            self._check(self.lib.dacti_integrate(self.h, int(bool(signal)), C.byref(dt), C.byref(fb)), "integrate")
            return dt.value, bool(fb.value)
        errors = []

        def tramp(user, h, level, acti_time, n_substeps, dtl):
            if errors:
                return
            try:
                substep_hook(self, level, acti_time, n_substeps, dtl)
            except BaseException as e:  # noqa: BLE001 - re-raised below
                errors.append(e)

        cb = _lib.SUBSTEP_HOOK(tramp)
        self._check(self.lib.dacti_set_substep_hook(self.h, cb, None), "set_substep_hook")
        try:
            rc = self.lib.dacti_integrate(self.h, int(bool(signal)), C.byref(dt), C.byref(fb))
        finally:
            self.lib.dacti_set_substep_hook(self.h, _lib.SUBSTEP_HOOK(), None)
        self._check(rc, "integrate")
        if errors:
            raise errors[0]
        return dt.value, bool(fb.value)

    def run(self, steps, signal=True):
        done = C.c_int64()
        self._check(self.lib.dacti_run(self.h, steps, int(bool(signal)), C.byref(done)), "run")
        return done.value

    def n_active_cells(self):
        n = C.c_int64()
        self._check(self.lib.dacti_n_active(self.h, C.byref(n)), "n_active")
        return n.value

    def n_nodes(self):
        n = C.c_int64()
        self._check(self.lib.dacti_n_nodes(self.h, C.byref(n)), "n_nodes")
        return n.value

    def get_nodes(self):
        """Node pool as host arrays (node-major), Integrator::get_nodes()."""
        n = self.n_nodes()
        arrs = _lib.alloc_node_arrays(n, self.case.model, self.case.dim)
        v = _lib.node_view(arrs)
        self._check(self.lib.dacti_export_nodes(self.h, C.byref(v)), "export_nodes")
        return arrs

    def for_each_active_cell(self, action):
        """action(i, node_index, nodes) over the leaves of the LTS clusters and
        buffers, in level order (integrator.hpp:172-191)."""
        nodes = self.get_nodes()
        cl = nodes["cluster"]
        i = 0
        for key in list(range(MAX_LEVEL)) + [MAX_LEVEL + l for l in range(MAX_LEVEL)]:
            for idx in np.nonzero(cl == key)[0]:
                action(i, int(idx), nodes)
                i += 1

    def stats(self):
        s = _lib.StatsStruct()
        self._check(self.lib.dacti_get_stats(self.h, C.byref(s)), "get_stats")
        return s

    def schedule(self):
        """(acti_time, acti_counter) per LTS level (integrator.hpp:48-49)."""
        t = (C.c_double * MAX_LEVEL)()
        c = (C.c_int32 * MAX_LEVEL)()
        self._check(self.lib.dacti_get_schedule(self.h, t, c), "get_schedule")
        return list(t), list(c)

    def set_launch(self, block=256, grid=0, graphs=False):
        """Launch configuration of this integrator's stepping (results do not depend on it)."""
        self._check(self.lib.dacti_set_launch(self.h, block, grid, int(bool(graphs))), "set_launch")

    def set_dense(self, mode):
        """Flux op for dense clusters: 0 never, 1 when a cluster fills half
        the pool (default), 2 always (results do not depend on it)."""
        self._check(self.lib.dacti_set_dense(self.h, int(mode)), "set_dense")

    def set_remesh(self, on):
        """Pause (False) or resume (True) refinement marks; an extension of
        the reference, which remeshes in every substep."""
        self._check(self.lib.dacti_set_remesh(self.h, int(bool(on))), "set_remesh")

    def mark_split(self, ids):
        """Flag node ids to be split at the next remesh (active leaves that may split)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._check(self.lib.dacti_mark_split(self.h, ids.ctypes.data_as(C.POINTER(C.c_int32)), len(ids)),
                    "mark_split")

    def set_timing(self, on):
        self._check(self.lib.dacti_set_timing(self.h, int(bool(on))), "set_timing")

    @property
    def global_dt(self):
        return self.stats().global_dt

    @property
    def max_acti_level(self):
        return self.stats().max_acti_level

    def close(self):
        if getattr(self, "h", None):
            self.lib.dacti_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Independent integrators of one model stepped together (dacti_batch_*).

    Every op of an LTS substep is one launch over the cluster lists of all
    instances; each instance takes part exactly when its own integrate()
    would run that substep, so it evolves as it would alone.  While the batch
    exists, step the integrators only through it; their read side
    (get_nodes, stats, n_active_cells) stays valid.
    """

    def __init__(self, integrators):
        self.members = list(integrators)
        self.lib = _lib.load()
        arr = (C.c_void_p * len(self.members))(*[g.h for g in self.members])
        b = C.c_void_p()
        rc = self.lib.dacti_batch_create(arr, len(self.members), C.byref(b))
        if rc != 0:
            raise RuntimeError(f"dacti_batch_create failed ({rc}): {self.lib.dacti_batch_last_error(None).decode()}")
        self.b = b

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.lib.dacti_batch_last_error(self.b).decode()}")

    def integrate(self, signal=True):
        """One global step of every instance; returns force_break."""
        fb = C.c_int32()
        self._check(self.lib.dacti_batch_integrate(self.b, int(bool(signal)), C.byref(fb)), "batch integrate")
        return bool(fb.value)

    def run(self, steps, signal=True):
        done = C.c_int64()
        self._check(self.lib.dacti_batch_run(self.b, steps, int(bool(signal)), C.byref(done)), "batch run")
        return done.value

    def set_launch(self, block=256, grid=0, graphs=False):
        self._check(self.lib.dacti_batch_set_launch(self.b, block, grid, int(bool(graphs))), "set_launch")

    def launch(self):
        b, g, gr = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.dacti_batch_get_launch(self.b, C.byref(b), C.byref(g), C.byref(gr)), "get_launch")
        return b.value, g.value, bool(gr.value)

    def autotune(self):
        """Time one candidate launch configuration per following global step
        and keep the fastest (the results do not depend on it)."""
        self._check(self.lib.dacti_batch_autotune(self.b), "autotune")

    def tune_results(self):
        n = C.c_int32(16)
        bl, gr, gp = (C.c_int32 * 16)(), (C.c_int32 * 16)(), (C.c_int32 * 16)()
        ms = (C.c_double * 16)()
        self._check(self.lib.dacti_batch_tune_results(self.b, C.byref(n), bl, gr, gp, ms), "tune_results")
        return [(bl[k], gr[k], bool(gp[k]), ms[k]) for k in range(min(n.value, 16))]

    def stats(self):
        s = _lib.StatsStruct()
        self._check(self.lib.dacti_batch_get_stats(self.b, C.byref(s)), "batch stats")
        return s

    def set_timing(self, on):
        self._check(self.lib.dacti_batch_set_timing(self.b, int(bool(on))), "batch timing")

    def close(self):
        if getattr(self, "b", None):
            self.lib.dacti_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
