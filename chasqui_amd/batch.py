"""Batches: many independent meshes (`Mesh` objects of one scheme) advanced by shared launches.

Each RK stage is one launch per kernel class over a device list of (run, tile) items (classes
that no run needs, such as the z faces of 2D runs, are left out), captured once in a hipGraph
and replayed; every run keeps its own dt, t, tlim, CFL, boundaries and size,
and its result is bitwise the one `Mesh.step` gives it alone. `tune()` is the autotuner of the
per-class grid sizes.
"""
import ctypes as C

from ._capi import check, lib

GRID_CLASSES = ["flux_x1", "flux_x2", "flux_x3", "corner", "update", "bc_x1", "bc_x2", "bc_x3", "c2p",
                "repair", "fofc_rerun"]


class Batch:
    def __init__(self, capacity):
        L = lib()
        h = C.c_void_p()
        check(L.chq_batch_create(int(capacity), C.byref(h)), "chq_batch_create")
        self.h, self.L = h, L
        self.meshes = []          # keeps the contexts alive while they are batched

    def close(self):
        if getattr(self, "h", None):
            self.L.chq_batch_destroy(self.h)
            self.h = None
            self.meshes = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, mesh):
        check(self.L.chq_batch_add(self.h, mesh.h), "chq_batch_add")
        self.meshes.append(mesh)

    def remove(self, mesh):
        check(self.L.chq_batch_remove(self.h, mesh.h), "chq_batch_remove")
        self.meshes.remove(mesh)

    def __len__(self):
        return self.L.chq_batch_size(self.h)

    def step(self, n=1):
        check(self.L.chq_batch_step(self.h, int(n)), "chq_batch_step")

    def step_timed(self, n):
        ms = C.c_double()
        check(self.L.chq_batch_step_timed(self.h, int(n), C.byref(ms)), "chq_batch_step_timed")
        return ms.value

    def sync(self):
        check(self.L.chq_batch_sync(self.h), "chq_batch_sync")

    def unfinished(self):
        """(count, per-run flags in table order) of runs that have not reached tlim / nlim."""
        n = len(self)
        flags = (C.c_int * max(n, 1))()
        cnt = C.c_int()
        check(self.L.chq_batch_unfinished(self.h, C.byref(cnt), flags), "chq_batch_unfinished")
        return cnt.value, [flags[i] for i in range(n)]

    def table_index(self, mesh):
        for i in range(len(self)):
            if self.L.chq_batch_run(self.h, i) == mesh.h.value:
                return i
        raise KeyError("mesh is not in this batch")

    def use_graph(self, on=True):
        check(self.L.chq_batch_use_graph(self.h, int(bool(on))), "chq_batch_use_graph")

    def tune(self, reps=3):
        best = C.c_double()
        check(self.L.chq_batch_tune(self.h, int(reps), C.byref(best)), "chq_batch_tune")
        return best.value

    def grids(self):
        out = {}
        for q in range(len(GRID_CLASSES)):
            g, name = C.c_int(), C.c_char_p()
            check(self.L.chq_batch_grid(self.h, q, C.byref(g), C.byref(name)), "chq_batch_grid")
            out[name.value.decode()] = g.value
        return out

    def set_grid(self, cls, grid):
        q = GRID_CLASSES.index(cls) if isinstance(cls, str) else int(cls)
        check(self.L.chq_batch_set_grid(self.h, q, int(grid)), "chq_batch_set_grid")
