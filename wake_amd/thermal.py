"""The reference's D2Q9 temperature solver (src/thermal/thermal_solver.cpp,
class ThermalSolver2D) on the GPU, through the wake_thermal2d C ABI.

One step collides the temperature populations with the fluid grid's current
velocity, applies the isothermal south / north walls, streams with the fluid
grid's cell types (wall velocity zero) and computes the per-node Boussinesq
acceleration beta (T - t_ref) g_gravity.
"""
import ctypes

import numpy as np

from . import _lib


class ThermalSolver2D:
    def __init__(self, nx, ny, tau_thermal=0.8, t_ref=0.5, beta=0.05):
        self._L = _lib.lib()
        self._nx, self._ny = nx, ny
        h = ctypes.c_void_p()
        _lib.check(self._L.wake_thermal2d_create(nx, ny, tau_thermal, t_ref, beta, ctypes.byref(h)),
                   "ThermalSolver2D")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.wake_thermal2d_destroy(self._h)
            self._h = None

    def initialize(self, t_init=0.5):
        _lib.check(self._L.wake_thermal2d_initialize(self._h, t_init), "initialize")

    def set_isothermal_wall_south(self, t_hot):
        _lib.check(self._L.wake_thermal2d_set_walls(self._h, 1, t_hot, 0, 0.0), "set_isothermal_wall_south")

    def set_isothermal_wall_north(self, t_cold):
        _lib.check(self._L.wake_thermal2d_set_walls(self._h, 0, 0.0, 1, t_cold), "set_isothermal_wall_north")

    def step(self, fluid_grid, g_gravity=-0.001):
        _lib.check(self._L.wake_thermal2d_step(self._h, fluid_grid._h, g_gravity), "step")

    def _read(self, which):
        out = np.empty(self._nx * self._ny, np.float32)
        _lib.check(self._L.wake_thermal2d_read(self._h, which, _lib.ptr(out)), "read")
        return out

    def temperature_data(self):
        return self._read(0)

    def buoyancy_data(self):
        return self._read(1)

    def temperature(self, x, y):
        return float(self.temperature_data()[y * self._nx + x])

    def _stats(self):
        out = (ctypes.c_double * 2)()
        _lib.check(self._L.wake_thermal2d_stats(self._h, out), "stats")
        return out[0], out[1]

    def compute_mean_temperature(self):
        return self._stats()[0]

    def compute_nusselt_number(self):
        return self._stats()[1]
