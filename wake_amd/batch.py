"""Batched stepping of many domains (libwake_hip.so, wake_batch_*).

    b = Batch(lattice=9, collision="MRT", chunk=100)
    ids = [b.add(g, steps=20000, forces=(cx, cy, 0.0, 1.5 * D)) for g in grids]
    b.tune()           # optional: pick the launch shape on real steps
    ms = b.run(20000)  # every member takes its own 20000 steps
    F = b.forces(ids[0])  # (steps, 3): Fx, Fy, torque (2D) / F (3D)

Members share lattice, storage, collision and LES coupling; each keeps its
own size, geometry, boundaries, tau and body force, and gives the same
populations as when stepped alone.  A member must not be stepped on its own
while it is in a batch.
"""
import ctypes

import numpy as np

from . import _lib
from .grid import Storage, _coll


class Batch:
    def __init__(self, lattice=19, storage=Storage.F32, collision="MRT", use_les=False, chunk=100, thermal=False):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.wake_batch_create(int(lattice), int(storage), _coll(collision), int(bool(use_les)),
                                       int(bool(thermal)), int(chunk), ctypes.byref(h)), "wake_batch_create")
        self._L = L
        self._h = h
        self._members = {}  # slot -> grid (keeps the domain alive while it is a member)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.wake_batch_destroy(h)
            self._h = None

    def add(self, grid, steps, g=(0.0, 0.0, 0.0), smag=0.14, forces=None):
        """forces: None or (cx, cy, cz, r_max) of the momentum-exchange links."""
        slot = ctypes.c_int()
        fc = forces if forces is not None else (0.0, 0.0, 0.0, 0.0)
        _lib.check(self._L.wake_batch_add(self._h, grid._h, int(steps), float(g[0]), float(g[1]), float(g[2]),
                                          float(smag), int(forces is not None), float(fc[0]), float(fc[1]),
                                          float(fc[2]), float(fc[3]), ctypes.byref(slot)), "wake_batch_add")
        self._members[slot.value] = grid
        return slot.value

    def remove(self, slot):
        _lib.check(self._L.wake_batch_remove(self._h, int(slot)), "wake_batch_remove")
        self._members.pop(slot, None)

    def run(self, steps, use_graph=True):
        """Run `steps` launches; returns the device time in ms."""
        ms = ctypes.c_float()
        _lib.check(self._L.wake_batch_run(self._h, int(steps), int(bool(use_graph)), ctypes.byref(ms)),
                   "wake_batch_run")
        return ms.value

    def launch(self, steps, use_graph=True):
        """Enqueue one chunk (min(steps, chunk) launches) and return at once."""
        _lib.check(self._L.wake_batch_launch(self._h, int(steps), int(bool(use_graph))), "wake_batch_launch")

    def wait(self):
        """Complete the chunk in flight; returns its device time in ms."""
        ms = ctypes.c_float()
        _lib.check(self._L.wake_batch_wait(self._h, ctypes.byref(ms)), "wake_batch_wait")
        return ms.value

    def remaining(self, slot):
        r = ctypes.c_longlong()
        _lib.check(self._L.wake_batch_remaining(self._h, int(slot), ctypes.byref(r)), "wake_batch_remaining")
        return r.value

    def forces(self, slot):
        n = ctypes.c_longlong()
        _lib.check(self._L.wake_batch_forces(self._h, int(slot), None, 0, ctypes.byref(n)), "wake_batch_forces")
        out = np.zeros((n.value, 3), np.float64)
        if n.value:
            _lib.check(self._L.wake_batch_forces(self._h, int(slot), _lib.ptr(out), n.value, ctypes.byref(n)),
                       "wake_batch_forces")
        return out

    def tune(self, probe_steps=4):
        """Autotune the launch shape on real steps; returns (block, grid_factor)."""
        b, gf = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._L.wake_batch_tune(self._h, int(probe_steps), ctypes.byref(b), ctypes.byref(gf)),
                   "wake_batch_tune")
        return b.value, gf.value

    def set_launch(self, block, grid_factor=0):
        _lib.check(self._L.wake_batch_set_launch(self._h, int(block), int(grid_factor)), "wake_batch_set_launch")

    def info(self):
        out = (ctypes.c_longlong * 4)()
        _lib.check(self._L.wake_batch_info(self._h, out), "wake_batch_info")
        return {"members": out[0], "nodes": out[1], "block": out[2], "grid_factor": out[3]}
