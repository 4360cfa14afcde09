"""ctypes binding of the CPU oracle (oracle/liboracle{32,64}.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  It never touches the GPU.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_libs = {}


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(bits=32, path=None):
    """Load liboracle32.so (float), liboracle64.so (double) or, for bits =
    "32fma", the float build with fused multiply-adds; `path` loads another
    float build (the host-native one the CPU baseline times) under the key
    `bits`."""
    if bits in _libs:
        return _libs[bits]
    if path is None:
        path = os.path.join(ORACLE_DIR, f"liboracle{bits}.so")
        if not os.path.exists(path):
            build_oracle()
    L = ctypes.CDLL(path)
    real = ctypes.c_double if bits == 64 else ctypes.c_float
    P = ctypes.c_void_p
    I = ctypes.c_int
    L.oracle_grid_create.restype = P
    L.oracle_grid_create.argtypes = [I, I, I, I, I, real]
    L.oracle_grid_destroy.argtypes = [P]
    L.oracle_grid_initialize.argtypes = [P, real, real, real, real]
    L.oracle_grid_initialize_fields.argtypes = [P, P, P, P, P]
    L.oracle_grid_add_noise.argtypes = [P, real, ctypes.c_uint64]
    L.oracle_grid_set_cell_type.argtypes = [P, I, I, I, I]
    L.oracle_grid_set_periodic.argtypes = [P, I]
    L.oracle_grid_set_periodic_axes.argtypes = [P, I, I]
    L.oracle_grid_set_solid_box.argtypes = [P, I, I, I, I, I, I]
    L.oracle_grid_set_solid_sphere.argtypes = [P, real, real, real, real]
    L.oracle_grid_set_solid_ellipsoid.argtypes = [P, real, real, real, real, real, real, P]
    L.oracle_grid_set_solid_naca.argtypes = [P, I, real, real, real, real, I, I]
    L.oracle_grid_enable_thermal.argtypes = [P, real, real, real, real]
    L.oracle_grid_set_temperature_field.argtypes = [P, P]
    L.oracle_grid_set_wall_temperature.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, real]
    L.oracle_grid_temperature.argtypes = [P]
    L.oracle_grid_temperature.restype = P
    L.oracle_grid_set_solid_circle.argtypes = [P, real, real, real]
    L.oracle_grid_set_solid_cylinder_z.argtypes = [P, real, real, real]
    L.oracle_grid_set_moving_wall_top.argtypes = [P, real, real]
    L.oracle_grid_set_inflow_west.argtypes = [P, real, real]
    L.oracle_grid_set_inflow_parabolic.argtypes = [P, real]
    L.oracle_grid_set_inflow_turbulence.argtypes = [P, real, ctypes.c_uint64]
    L.oracle_grid_set_outflow_east.argtypes = [P, real]
    L.oracle_grid_set_convective_outflow_east.argtypes = [P, real]
    L.oracle_grid_set_les_coupled.argtypes = [P, I]
    L.oracle_grid_build_curved_sphere_links.argtypes = [P, real, real, real, real]
    L.oracle_grid_build_curved_sphere_links.restype = I
    L.oracle_grid_num_links.argtypes = [P]
    L.oracle_grid_num_links.restype = I
    L.oracle_grid_get_links.argtypes = [P, P, P, P]
    L.oracle_grid_step.argtypes = [P, I, real, real, real, I, real]
    L.oracle_grid_run.argtypes = [P, I, I, real, real, real, I, real]
    for name in ("f_src", "f_dst", "rho", "ux", "uy", "uz", "speed", "cell_type"):
        fn = getattr(L, "oracle_grid_" + name)
        fn.restype = P
        fn.argtypes = [P]
    for name in ("total_mass", "total_kinetic_energy", "max_speed"):
        fn = getattr(L, "oracle_grid_" + name)
        fn.restype = ctypes.c_double
        fn.argtypes = [P]
    L.oracle_grid_momentum_exchange.argtypes = [P, real, real, real, real, P]
    L.oracle_set_threads.argtypes = [I]
    L.oracle_rng_splitmix64.restype = ctypes.c_uint64
    L.oracle_rng_splitmix64.argtypes = [ctypes.c_uint64]
    L.oracle_rng_u01.restype = ctypes.c_double
    L.oracle_rng_u01.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.oracle_tri_box_overlap.argtypes = [P, P, P]
    L.oracle_tri_box_overlap.restype = I
    L.oracle_voxelize_mesh.argtypes = [P, I, I, I, P, ctypes.c_longlong, I]
    L.oracle_voxelize_mesh.restype = ctypes.c_longlong
    L.real = real
    L.np_real = np.float64 if bits == 64 else np.float32
    _libs[bits] = L
    return L


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def call(bits, name, *args):
    """Call oracle_<name> with numpy arrays passed by pointer and Python
    floats converted to the build's real type."""
    L = lib(bits)
    fn = getattr(L, "oracle_" + name)
    cargs = []
    for a in args:
        if isinstance(a, np.ndarray):
            cargs.append(ptr(a))
        elif isinstance(a, float):
            cargs.append(L.real(a))
        elif isinstance(a, int):
            cargs.append(ctypes.c_int(a))
        else:
            cargs.append(a)
    return fn(*cargs)


def tri_box_overlap(tri, center, half, bits=32):
    """oracle_tri_box_overlap: triangle (3x3 vertices) against a box."""
    L = lib(bits)
    t = np.ascontiguousarray(np.asarray(tri, L.np_real).reshape(9))
    c = np.ascontiguousarray(np.asarray(center, L.np_real).reshape(3))
    h = np.ascontiguousarray(np.asarray(half, L.np_real).reshape(3))
    return bool(L.oracle_tri_box_overlap(ptr(t), ptr(c), ptr(h)))


def voxelize_mesh(flags, tris, fill_interior=True, bits=32):
    """oracle_voxelize_mesh into `flags` ((nz, ny, nx) uint8, in place);
    returns the number of Solid writes."""
    L = lib(bits)
    assert flags.dtype == np.uint8 and flags.ndim == 3 and flags.flags.c_contiguous
    t = np.ascontiguousarray(np.asarray(tris, L.np_real).reshape(-1, 9))
    nz, ny, nx = flags.shape
    return int(L.oracle_voxelize_mesh(ptr(flags), nx, ny, nz, ptr(t), t.shape[0], 1 if fill_interior else 0))


COLL = {"BGK": 0, "TRT": 1, "MRT": 2}


class OracleGrid:
    """Python handle over oracle_grid (LatticeGrid2D/3D step semantics)."""

    def __init__(self, dim, nx, ny, nz=1, tau=0.6, lattice=None, bits=32):
        self.L = lib(bits)
        self.bits = bits
        self.dim = dim
        self.lattice = lattice or (9 if dim == 2 else 19)
        self.nx, self.ny, self.nz = nx, ny, (1 if dim == 2 else nz)
        self.N = nx * ny * self.nz
        self.Q = self.lattice
        self.h = self.L.oracle_grid_create(dim, self.lattice, nx, ny, self.nz, tau)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_grid_destroy(self.h)
            self.h = None

    def _arr(self, name, count, dtype=None):
        p = getattr(self.L, "oracle_grid_" + name)(self.h)
        dt = dtype or self.L.np_real
        buf = (ctypes.c_char * (count * np.dtype(dt).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dt).copy()

    def f_src(self):
        return self._arr("f_src", self.Q * self.N).reshape(self.Q, self.N)

    def f_dst(self):
        return self._arr("f_dst", self.Q * self.N).reshape(self.Q, self.N)

    def rho(self):
        return self._arr("rho", self.N)

    def ux(self):
        return self._arr("ux", self.N)

    def uy(self):
        return self._arr("uy", self.N)

    def uz(self):
        return self._arr("uz", self.N)

    def speed(self):
        return self._arr("speed", self.N)

    def cell_type(self):
        return self._arr("cell_type", self.N, np.uint8)

    def initialize(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        self.L.oracle_grid_initialize(self.h, rho, ux, uy, uz)

    def initialize_fields(self, rho, ux, uy, uz=None):
        r = self.L.np_real
        rho, ux, uy = (np.ascontiguousarray(a, dtype=r) for a in (rho, ux, uy))
        uz = None if uz is None else np.ascontiguousarray(uz, dtype=r)
        self.L.oracle_grid_initialize_fields(self.h, ptr(rho), ptr(ux), ptr(uy), ptr(uz))

    def add_noise(self, amp, seed):
        self.L.oracle_grid_add_noise(self.h, amp, seed)

    def set_cell_type(self, x, y, z, t):
        self.L.oracle_grid_set_cell_type(self.h, x, y, z, t)

    def set_periodic(self, p):
        self.L.oracle_grid_set_periodic(self.h, int(p))

    def set_periodic_axes(self, py, pz):
        self.L.oracle_grid_set_periodic_axes(self.h, int(py), int(pz))

    def set_solid_box(self, *a):
        if self.dim == 2 and len(a) == 4:  # LatticeGrid2D::set_solid_box(x0, y0, x1, y1)
            a = (a[0], a[1], 0, a[2], a[3], 0)
        self.L.oracle_grid_set_solid_box(self.h, *a)

    def set_solid_sphere(self, cx, cy, cz, r):
        self.L.oracle_grid_set_solid_sphere(self.h, cx, cy, cz, r)

    def enable_thermal(self, tau_g, gbeta, t_ref, t_init):
        self.L.oracle_grid_enable_thermal(self.h, tau_g, gbeta, t_ref, t_init)
        self._thermal = True

    def set_temperature_field(self, t):
        a = np.ascontiguousarray(t, np.float64 if self.bits == 64 else np.float32)
        self.L.oracle_grid_set_temperature_field(self.h, a.ctypes.data_as(ctypes.c_void_p))

    def set_wall_temperature(self, x0, y0, x1, y1, t):
        self.L.oracle_grid_set_wall_temperature(self.h, x0, y0, x1, y1, t)

    def temperature(self):
        return self._arr("temperature", self.N)

    def set_solid_naca(self, code, chord, aoa_deg, lead_x, lead_y, z_start=0, z_end=None):
        self.L.oracle_grid_set_solid_naca(self.h, int(code), chord, aoa_deg, lead_x, lead_y, z_start,
                                          self.nz - 1 if z_end is None else z_end)

    def set_solid_ellipsoid(self, cx, cy, cz, rx, ry, rz, axes=None):
        a = None
        if axes is not None:
            a = np.ascontiguousarray(np.asarray(axes).reshape(9), np.float64 if self.bits == 64 else np.float32)
            self._keep = a
        self.L.oracle_grid_set_solid_ellipsoid(self.h, cx, cy, cz, rx, ry, rz,
                                               None if a is None else a.ctypes.data_as(ctypes.c_void_p))

    def set_solid_circle(self, cx, cy, r):
        self.L.oracle_grid_set_solid_circle(self.h, cx, cy, r)

    def set_solid_cylinder_z(self, cx, cy, r):
        self.L.oracle_grid_set_solid_cylinder_z(self.h, cx, cy, r)

    def set_moving_wall_top(self, uwx, uwz=0.0):
        self.L.oracle_grid_set_moving_wall_top(self.h, uwx, uwz)

    def set_inflow_west(self, ux, uy=0.0):
        self.L.oracle_grid_set_inflow_west(self.h, ux, uy)

    def set_inflow_parabolic(self, umax):
        self.L.oracle_grid_set_inflow_parabolic(self.h, umax)

    def set_inflow_turbulence(self, amp, seed):
        self.L.oracle_grid_set_inflow_turbulence(self.h, amp, seed)

    def set_outflow_east(self, rho_out=1.0):
        self.L.oracle_grid_set_outflow_east(self.h, rho_out)

    def set_convective_outflow_east(self, u_conv):
        self.L.oracle_grid_set_convective_outflow_east(self.h, u_conv)

    def set_les_coupled(self, c):
        self.L.oracle_grid_set_les_coupled(self.h, int(c))

    def build_curved_sphere_links(self, cx, cy, cz, r):
        return self.L.oracle_grid_build_curved_sphere_links(self.h, cx, cy, cz, r)

    def links(self):
        n = self.L.oracle_grid_num_links(self.h)
        idx = np.zeros(n, np.int32)
        q = np.zeros(n, np.int32)
        d = np.zeros(n, self.L.np_real)
        self.L.oracle_grid_get_links(self.h, ptr(idx), ptr(q), ptr(d))
        return idx, q, d

    def step(self, collision="MRT", gx=0.0, gy=0.0, gz=0.0, use_les=False, smag=0.14):
        self.L.oracle_grid_step(self.h, COLL[collision], gx, gy, gz, int(use_les), smag)

    def run(self, steps, collision="MRT", gx=0.0, gy=0.0, gz=0.0, use_les=False, smag=0.14):
        self.L.oracle_grid_run(self.h, steps, COLL[collision], gx, gy, gz, int(use_les), smag)

    def total_mass(self):
        return self.L.oracle_grid_total_mass(self.h)

    def total_kinetic_energy(self):
        return self.L.oracle_grid_total_kinetic_energy(self.h)

    def max_speed(self):
        return self.L.oracle_grid_max_speed(self.h)

    def momentum_exchange(self, cx, cy, cz, r_max):
        out = np.zeros(3, np.float64)
        self.L.oracle_grid_momentum_exchange(self.h, cx, cy, cz, r_max, ptr(out))
        return out


# Lattice tables as the reference lists them (lattice_models.hpp:14-110);
# test-side copies used to build closed-form expectations.
D2Q9_C = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
D2Q9_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
D3Q19_C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                    [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def feq_np(c, w, rho, u, dtype=np.float32):
    """Second-order equilibrium in the reference's float expression order."""
    dt = dtype
    rho = dt(rho)
    u = [dt(v) for v in u]
    usq = dt(0)
    for v in u:
        usq = dt(usq + v * v)
    out = []
    for q in range(len(w)):
        cu = dt(0)
        for d in range(len(u)):
            cu = dt(cu + dt(c[q][d]) * u[d])
        wq = dt(w[q])
        out.append(dt(wq * rho * (dt(1) + dt(3) * cu + dt(4.5) * cu * cu - dt(1.5) * usq)))
    return np.array(out, dtype=dt)
