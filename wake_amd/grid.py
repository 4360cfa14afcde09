"""LatticeGrid2D / LatticeGrid3D over the device-resident HIP path.

Method names, arguments and defaults follow the reference's C++ classes
(include/wake/grid/lattice_grid_2d.hpp:14-113 and lattice_grid_3d.hpp:21-123);
the *_data() accessors return numpy copies of the host mirrors instead of
raw pointers.  Extensions the reference lacks (fully periodic 2D stepping,
per-axis periodicity, 3D open boundaries, D3Q27, FP16/BF16 storage,
LES-coupled collision, device momentum exchange) are extra methods.
"""
import ctypes
from enum import IntEnum

import numpy as np

from . import _lib


class CellType(IntEnum):
    """include/wake/core/types.hpp:69-77"""
    Fluid = 0
    Solid = 1
    Inflow = 2
    Outflow = 3
    Periodic = 4
    Symmetry = 5
    MovingWall = 6


class CollisionType(IntEnum):
    """include/wake/core/types.hpp:80-85 (Entropic has no kernel)."""
    BGK = 0
    TRT = 1
    MRT = 2


class Storage(IntEnum):
    F32 = 0
    F16 = 1  # FP16 of f_q - w_q, FP32 arithmetic
    BF16 = 2


DEFAULT_SMAGORINSKY_CONSTANT = 0.14  # include/wake/core/constants.hpp
RHO_0 = 1.0
CS2 = 1.0 / 3.0


def _coll(c):
    if isinstance(c, str):
        return int(CollisionType[c])
    return int(c)


class _Domain:
    def __init__(self, dim, lattice, nx, ny, nz, tau, storage=Storage.F32, plane_pad=-1):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.wake_domain_create(dim, lattice, nx, ny, nz, tau, int(storage), plane_pad, ctypes.byref(h)),
                   "wake_domain_create")
        self._L = L
        self._h = h
        info = self.info()
        self._nx, self._ny, self._nz, self.Q = info[0], info[1], info[2], info[3]
        self.dim = dim

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.wake_domain_destroy(h)
            self._h = None

    def _c(self, rc, what):
        _lib.check(rc, what)

    # ------------------------------------------------------------- info
    def info(self):
        out = (ctypes.c_longlong * 8)()
        self._c(self._L.wake_domain_info(self._h, out), "info")
        return list(out)

    def nx(self):
        return self._nx

    def ny(self):
        return self._ny

    def total_nodes(self):
        return self._nx * self._ny * self._nz

    def current_step(self):
        return self.info()[5]

    def export_vti(self, path, binary=True, with_analytics=False):
        """vtk_exporter.cpp:8-191: the current fields as a .vti file (raw
        appended binary by default, or the reference's ASCII)."""
        self._c(self._L.wake_domain_export_vti(self._h, str(path).encode(), 1 if binary else 0,
                                               1 if with_analytics else 0), "export_vti")

    def set_buffer_addressing(self, enable=True):
        """16-bit storage: buffer-descriptor addressing (default, buffers
        under 4 GiB) or 64-bit pointers (as larger buffers use)."""
        self._c(self._L.wake_domain_set_buffer_addressing(self._h, 1 if enable else 0), "set_buffer_addressing")

    def is_periodic(self):
        out = ctypes.c_int(0)
        self._c(self._L.wake_domain_is_periodic(self._h, ctypes.byref(out)), "is_periodic")
        return bool(out.value)

    def plane_stride(self):
        return self.info()[4]

    def device_bytes(self):
        return self.info()[7]

    # ------------------------------------------------------- configuration
    def set_relaxation_time(self, tau):
        self._tau = tau
        self._c(self._L.wake_domain_set_relaxation_time(self._h, tau), "set_relaxation_time")

    def relaxation_time(self):
        return self._tau

    def kinematic_viscosity(self):
        return CS2 * (self._tau - 0.5)

    def set_periodic(self, periodic):
        self._c(self._L.wake_domain_set_periodic(self._h, int(bool(periodic))), "set_periodic")

    def set_periodic_axes(self, per_y, per_z=False):
        self._c(self._L.wake_domain_set_periodic_axes(self._h, int(per_y), int(per_z)), "set_periodic_axes")

    def set_les_coupled(self, coupled=True):
        self._c(self._L.wake_domain_set_les_coupled(self._h, int(bool(coupled))), "set_les_coupled")

    def set_inflow_west(self, u_in_x, u_in_y=0.0):
        self._c(self._L.wake_domain_set_inflow_west(self._h, u_in_x, u_in_y), "set_inflow_west")

    def set_inflow_parabolic(self, u_max):
        self._c(self._L.wake_domain_set_inflow_parabolic(self._h, u_max), "set_inflow_parabolic")

    def set_inflow_turbulence(self, amplitude, seed):
        self._c(self._L.wake_domain_set_inflow_turbulence(self._h, amplitude, seed), "set_inflow_turbulence")

    def set_outflow_east(self, rho_out=RHO_0):
        self._c(self._L.wake_domain_set_outflow_east(self._h, rho_out), "set_outflow_east")

    def set_convective_outflow_east(self, u_conv):
        self._c(self._L.wake_domain_set_convective_outflow_east(self._h, u_conv), "set_convective_outflow_east")

    def add_noise(self, amplitude, seed):
        self._c(self._L.wake_domain_add_noise(self._h, amplitude, seed), "add_noise")

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(-1)
        if f.size != self.Q * self.total_nodes():
            raise ValueError("populations must be Q planes of nx*ny*nz")
        self._c(self._L.wake_domain_set_populations(self._h, _lib.ptr(f)), "set_populations")

    def set_block_x(self, bx):
        self._c(self._L.wake_domain_set_block_x(self._h, bx), "set_block_x")

    def set_multistep(self, on=True):
        """Let run() take 8 steps per launch where the domain allows it
        (fully periodic 2D, all cells Fluid, no open boundary, force probe
        or temperature) and single launches are bound by the launch rate
        (up to 2^21 nodes in FP32, 2^17 with 16-bit storage); the result is
        the same as single steps."""
        self._c(self._L.wake_domain_set_multistep(self._h, int(bool(on))), "set_multistep")

    # --------------------------------------------------------------- steps
    def run(self, steps, collision=CollisionType.MRT, gx=0.0, gy=0.0, gz=0.0, use_les=False,
            smag=DEFAULT_SMAGORINSKY_CONSTANT, timed=False):
        ms = ctypes.c_float(0.0)
        self._c(self._L.wake_domain_run(self._h, steps, _coll(collision), gx, gy, gz, int(use_les), smag,
                                        ctypes.byref(ms) if timed else None), "run")
        return ms.value

    def kernel_time(self):
        """Average device ms of the fused step kernel per launch in the last
        timed run() (HIP events around every 16th launch of that run)."""
        t = ctypes.c_float()
        self._c(self._L.wake_domain_kernel_time(self._h, ctypes.byref(t)), "kernel_time")
        return t.value

    def synchronize(self):
        self._c(self._L.wake_domain_synchronize(self._h), "synchronize")

    # ----------------------------------------------------------- read-back
    def _read(self, field, count):
        out = np.empty(count, np.float32)
        self._c(self._L.wake_domain_read(self._h, field, _lib.ptr(out)), "read")
        return out

    def rho_data(self):
        return self._read(0, self.total_nodes())

    def ux_data(self):
        return self._read(1, self.total_nodes())

    def uy_data(self):
        return self._read(2, self.total_nodes())

    def speed_data(self):
        return self._read(4, self.total_nodes())

    def f_src_data(self):
        return self._read(5, self.Q * self.total_nodes()).reshape(self.Q, -1)

    def vorticity_data(self):
        """z vorticity (2D) on the device (analytics_kernels.ispc:4-29)."""
        return self._read(10, self.total_nodes())

    def vorticity_q_data(self):
        """3D: (omega_x, omega_y, omega_z, Q) (analytics_kernels.ispc:33-110)."""
        return tuple(self._read(k, self.total_nodes()) for k in (8, 9, 10, 11))

    def temperature_data(self):
        """Temperature of a thermal (D2Q9 + D2Q5) domain."""
        return self._read(7, self.total_nodes())

    def cell_type_data(self):
        out = np.empty(self.total_nodes(), np.uint8)
        self._c(self._L.wake_domain_read_cell_types(self._h, _lib.ptr(out)), "read_cell_types")
        return out

    def update_macroscopic(self):
        self._c(self._L.wake_domain_update_macroscopic(self._h), "update_macroscopic")

    def _diag(self):
        out = np.zeros(3, np.float64)
        self._c(self._L.wake_domain_diagnostics(self._h, _lib.ptr(out)), "diagnostics")
        return out

    def compute_total_mass(self):
        return float(self._diag()[0])

    def compute_total_kinetic_energy(self):
        return float(self._diag()[1])

    def compute_max_speed(self):
        return float(self._diag()[2])

    def momentum_exchange(self, cx, cy, cz, r_max):
        out = np.zeros(3, np.float64)
        self._c(self._L.wake_domain_momentum_exchange(self._h, cx, cy, cz, r_max, _lib.ptr(out)),
                "momentum_exchange")
        return out

    def nonfinite_seen(self):
        """True if a step since the state was last set produced a non-finite
        density (the step kernels' divergence flag)."""
        v = ctypes.c_int(0)
        self._c(self._L.wake_domain_nonfinite_seen(self._h, ctypes.byref(v)), "nonfinite_seen")
        return bool(v.value)

    def count_nonfinite(self):
        """(count, first node) of nodes whose density or speed is not finite."""
        n, first = ctypes.c_longlong(0), ctypes.c_longlong(-1)
        self._c(self._L.wake_domain_count_nonfinite(self._h, ctypes.byref(n), ctypes.byref(first)),
                "count_nonfinite")
        return n.value, first.value

    def check_finite(self):
        """Raise FloatingPointError if a density or velocity is not finite
        (the reference throws NumericalDivergenceException)."""
        n, first = self.count_nonfinite()
        if n:
            raise FloatingPointError(f"non-finite density or velocity at {n} node(s), first at node {first}")

    def set_force_probe(self, cx, cy, cz, r_max):
        """Keep the momentum exchange of every step from now on (computed in
        the step launch itself); read it with force_history()."""
        self._c(self._L.wake_domain_set_force_probe(self._h, cx, cy, cz, r_max), "set_force_probe")

    def clear_force_probe(self):
        self._c(self._L.wake_domain_clear_force_probe(self._h), "clear_force_probe")

    def force_history(self):
        """(first_step, F): F[i] = forces of step first_step + i, shape (K, 3)
        float64, up to the current step."""
        first = ctypes.c_longlong(0)
        count = ctypes.c_longlong(0)
        self._c(self._L.wake_domain_force_history(self._h, None, 0, ctypes.byref(first), ctypes.byref(count)),
                "force_history")
        out = np.zeros((count.value, 3), np.float64)
        self._c(self._L.wake_domain_force_history(self._h, _lib.ptr(out), count.value, ctypes.byref(first),
                                                  ctypes.byref(count)), "force_history")
        return first.value, out


class LatticeGrid2D(_Domain):
    """D2Q9 grid (lattice_grid_2d.hpp:14-113).  step() streams with walls in
    y like the reference; set_periodic(True) selects fully periodic streaming."""

    def __init__(self, nx, ny, relaxation_time=0.6, storage=Storage.F32, plane_pad=-1):
        super().__init__(2, 9, nx, ny, 1, relaxation_time, storage, plane_pad)
        self._tau = relaxation_time

    def initialize(self, rho_init=RHO_0, ux_init=0.0, uy_init=0.0):
        self._c(self._L.wake_domain_initialize(self._h, rho_init, ux_init, uy_init, 0.0), "initialize")

    def initialize_fields(self, rho, ux, uy):
        r, a, b = (np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in (rho, ux, uy))
        self._c(self._L.wake_domain_initialize_fields(self._h, _lib.ptr(r), _lib.ptr(a), _lib.ptr(b), None),
                "initialize_fields")

    def initialize_custom(self, init_fn):
        """init_fn(x, y) -> (rho, ux, uy), evaluated on the host per node."""
        n = self.total_nodes()
        rho, ux, uy = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.float32)
        for y in range(self._ny):
            for x in range(self._nx):
                i = y * self._nx + x
                rho[i], ux[i], uy[i] = init_fn(x, y)
        self.initialize_fields(rho, ux, uy)

    def step(self, collision=CollisionType.MRT, gx=0.0, gy=0.0, use_les=False, smag_const=DEFAULT_SMAGORINSKY_CONSTANT):
        self._c(self._L.wake_domain_step(self._h, _coll(collision), gx, gy, 0.0, int(use_les), smag_const), "step")

    def set_cell_type(self, x, y, cell_type):
        self._c(self._L.wake_domain_set_cell_type(self._h, x, y, 0, int(cell_type)), "set_cell_type")

    def get_cell_type(self, x, y):
        return CellType(self._L.wake_domain_get_cell_type(self._h, x, y, 0))

    def set_solid_box(self, x0, y0, x1, y1):
        self._c(self._L.wake_domain_set_solid_box(self._h, x0, y0, 0, x1, y1, 0), "set_solid_box")

    def enable_thermal(self, tau_g, gbeta, t_ref=0.5, t_init=0.5):
        """D2Q5 temperature coupled by a_y = gbeta (T - t_ref); alpha = (tau_g - 1/2) / 3."""
        self._c(self._L.wake_domain_enable_thermal(self._h, tau_g, gbeta, t_ref, t_init), "enable_thermal")

    def set_temperature_field(self, t):
        a = np.ascontiguousarray(t, np.float32)
        self._c(self._L.wake_domain_set_temperature_field(self._h, _lib.ptr(a)), "set_temperature_field")

    def set_wall_temperature(self, x0, y0, x1, y1, t_wall):
        """Anti-bounce-back temperature of the solid nodes in the box; NaN = adiabatic."""
        self._c(self._L.wake_domain_set_wall_temperature(self._h, x0, y0, x1, y1, t_wall), "set_wall_temperature")

    def set_solid_naca(self, code, chord, aoa_deg, lead_x, lead_y):
        """naca_airfoil.cpp:152-166 rasterize of NACA `code` (e.g. "0012")."""
        self._c(self._L.wake_domain_set_solid_naca(self._h, int(code), chord, aoa_deg, lead_x, lead_y, 0, 0),
                "set_solid_naca")

    def set_solid_circle(self, cx, cy, radius):
        self._c(self._L.wake_domain_set_solid_circle(self._h, cx, cy, radius), "set_solid_circle")

    def set_moving_wall_north(self, u_wall_x):
        self._c(self._L.wake_domain_set_moving_wall_top(self._h, u_wall_x, 0.0), "set_moving_wall_north")

    def f_dst_data(self):
        return self._read(6, self.Q * self.total_nodes()).reshape(self.Q, -1)

    def rho(self, x, y):
        return float(self.rho_data()[y * self._nx + x])

    def ux(self, x, y):
        return float(self.ux_data()[y * self._nx + x])

    def uy(self, x, y):
        return float(self.uy_data()[y * self._nx + x])

    def momentum_exchange_2d(self, cx, cy, r_max):
        """(Fx, Fy, torque_z), karman_vortex.cpp:107-114 call on device."""
        return self.momentum_exchange(cx, cy, 0.0, r_max)


class LatticeGrid3D(_Domain):
    """D3Q19 (or D3Q27) grid (lattice_grid_3d.hpp:21-123)."""

    def __init__(self, nx, ny, nz, relaxation_time=0.6, lattice=19, storage=Storage.F32, plane_pad=-1):
        super().__init__(3, lattice, nx, ny, nz, relaxation_time, storage, plane_pad)
        self._tau = relaxation_time

    def nz(self):
        return self._nz

    def initialize(self, rho_init=RHO_0, ux_init=0.0, uy_init=0.0, uz_init=0.0):
        self._c(self._L.wake_domain_initialize(self._h, rho_init, ux_init, uy_init, uz_init), "initialize")

    def initialize_fields(self, rho, ux, uy, uz):
        r, a, b, c = (np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in (rho, ux, uy, uz))
        self._c(self._L.wake_domain_initialize_fields(self._h, _lib.ptr(r), _lib.ptr(a), _lib.ptr(b), _lib.ptr(c)),
                "initialize_fields")

    def initialize_custom(self, init_fn):
        """init_fn(x, y, z) -> (rho, ux, uy, uz), evaluated on the host per node."""
        n = self.total_nodes()
        arr = [np.empty(n, np.float32) for _ in range(4)]
        for z in range(self._nz):
            for y in range(self._ny):
                for x in range(self._nx):
                    i = (z * self._ny + y) * self._nx + x
                    arr[0][i], arr[1][i], arr[2][i], arr[3][i] = init_fn(x, y, z)
        self.initialize_fields(*arr)

    def step(self, collision=CollisionType.MRT, gx=0.0, gy=0.0, gz=0.0, use_les=False,
             smag_const=DEFAULT_SMAGORINSKY_CONSTANT):
        self._c(self._L.wake_domain_step(self._h, _coll(collision), gx, gy, gz, int(use_les), smag_const), "step")

    def set_cell_type(self, x, y, z, cell_type):
        self._c(self._L.wake_domain_set_cell_type(self._h, x, y, z, int(cell_type)), "set_cell_type")

    def get_cell_type(self, x, y, z):
        return CellType(self._L.wake_domain_get_cell_type(self._h, x, y, z))

    def set_solid_box(self, x0, y0, z0, x1, y1, z1):
        self._c(self._L.wake_domain_set_solid_box(self._h, x0, y0, z0, x1, y1, z1), "set_solid_box")

    def set_solid_sphere(self, cx, cy, cz, radius):
        self._c(self._L.wake_domain_set_solid_sphere(self._h, cx, cy, cz, radius), "set_solid_sphere")

    def set_solid_naca(self, code, chord, aoa_deg, lead_x, lead_y, z_start=0, z_end=None):
        """naca_airfoil.cpp:152-190 voxelize_extruded of NACA `code` (e.g. "0012")."""
        z1 = self.nz() - 1 if z_end is None else z_end
        self._c(self._L.wake_domain_set_solid_naca(self._h, int(code), chord, aoa_deg, lead_x, lead_y, z_start, z1),
                "set_solid_naca")

    def set_solid_ellipsoid(self, cx, cy, cz, rx, ry, rz, axes=None):
        """primitives.cpp:73-100; axes: optional 3x3 whose rows are the body
        axes in world coordinates (rotated ellipsoid)."""
        a = None if axes is None else np.ascontiguousarray(np.asarray(axes, np.float32).reshape(9))
        self._c(self._L.wake_domain_set_solid_ellipsoid(self._h, cx, cy, cz, rx, ry, rz, _lib.ptr(a)),
                "set_solid_ellipsoid")

    def voxelize_mesh(self, triangles, fill_interior=True):
        """obj_loader.cpp:208-400 voxelize_into_grid on the device.
        triangles: (ntri, 3, 3) vertex coordinates in grid units."""
        t = np.ascontiguousarray(np.asarray(triangles, np.float32).reshape(-1, 9))
        self._c(self._L.wake_domain_voxelize_mesh(self._h, _lib.ptr(t), t.shape[0], 1 if fill_interior else 0),
                "voxelize_mesh")

    def set_solid_cylinder_z(self, cx, cy, radius):
        self._c(self._L.wake_domain_set_solid_cylinder_z(self._h, cx, cy, radius), "set_solid_cylinder_z")

    def set_moving_wall_top(self, u_wall_x, u_wall_z=0.0):
        self._c(self._L.wake_domain_set_moving_wall_top(self._h, u_wall_x, u_wall_z), "set_moving_wall_top")

    def build_curved_sphere_links(self, cx, cy, cz, radius):
        n = ctypes.c_int(0)
        self._c(self._L.wake_domain_build_curved_sphere_links(self._h, cx, cy, cz, radius, ctypes.byref(n)),
                "build_curved_sphere_links")
        self._nlinks = n.value
        return n.value

    def curved_links(self):
        n = getattr(self, "_nlinks", 0)
        idx, q, d = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
        self._c(self._L.wake_domain_get_links(self._h, _lib.ptr(idx), _lib.ptr(q), _lib.ptr(d)), "get_links")
        return idx, q, d

    def uz_data(self):
        return self._read(3, self.total_nodes())

    def rho(self, x, y, z):
        return float(self.rho_data()[(z * self._ny + y) * self._nx + x])
