"""GPU parity of the reference's D2Q9 temperature kernels, its IBM kernels
(the ISPC-named exports) and the device ThermalSolver2D with the oracle.

Kernel exports: FP32 against the FP32 oracle within a few ulps of the
field's scale (the device cosine and FMA contraction round differently).  Solver: 300 coupled steps
of a Rayleigh-Benard cell (fluid TRT, uniform buoyancy from the mean
temperature, as thermal_convection.cpp:63-71) against the oracle's
composition of the same kernels, to the FP32 bar.
"""
import numpy as np
import pytest

import oracle_bind as ob
from parity import check_fp32_parity

pytestmark = pytest.mark.gpu

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def close(a, b, ulps=8):
    """max |a - b| within `ulps` FP32 ulps of max |b| (over-relaxation
    cancels, so a per-element relative bound is not meaningful)."""
    err = float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))
    return err <= ulps * 2.0 ** -23 * float(np.max(np.abs(b)))


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def test_thermal_exports_match_oracle(wa):
    K = wa.kernels
    rng = np.random.default_rng(5)
    nx, ny = 48, 20
    n = nx * ny
    g = (W[:, None] * rng.uniform(0.2, 1.0, (9, n))).astype(np.float32)
    ux, uy = rng.uniform(-0.05, 0.05, n).astype(np.float32), rng.uniform(-0.05, 0.05, n).astype(np.float32)
    ct = np.zeros(n, np.uint8)
    ct[rng.random(n) < 0.1] = 1
    outs = [np.zeros_like(g) for _ in range(2)]
    Ts = [np.full(n, 7.0, np.float32) for _ in range(2)]
    K.kernel("thermal_collide_d2q9", g, outs[0], ct, ux, uy, Ts[0], 0.77, n)
    ob.call(32, "thermal_collide_d2q9", g, outs[1], ct, ux, uy, Ts[1], 0.77, n)
    assert close(outs[0], outs[1])
    assert close(Ts[0], Ts[1]) and np.all(Ts[0][ct == 1] == 7.0)
    for yw, tw in ((0, 1.0), (ny - 1, 0.0), (7, 0.4)):
        a, b = g.copy(), g.copy()
        K.kernel("thermal_apply_isothermal_wall", a, nx, ny, yw, tw)
        ob.call(32, "thermal_apply_isothermal_wall", b, nx, ny, yw, tw)
        assert close(a, b)
    with pytest.raises(RuntimeError):
        K.kernel("thermal_apply_isothermal_wall", g.copy(), nx, ny, ny, 1.0)
    gy = [np.full(n, 0.1, np.float32) for _ in range(2)]
    K.kernel("thermal_compute_buoyancy_force_2d", Ts[0], gy[0], 0.5, 0.05, -0.001, n)
    ob.call(32, "thermal_compute_buoyancy_force_2d", Ts[1], gy[1], 0.5, 0.05, -0.001, n)
    assert close(gy[0] - 0.1, gy[1] - 0.1, ulps=16)


def test_ibm_exports_match_oracle(wa):
    K = wa.kernels
    rng = np.random.default_rng(6)
    nx, ny, m = 64, 40, 300
    ux = rng.normal(0, 0.05, nx * ny).astype(np.float32)
    uy = rng.normal(0, 0.05, nx * ny).astype(np.float32)
    # markers everywhere, including across the x seam and next to the y edges
    mx = rng.uniform(-3, nx + 3, m).astype(np.float32)
    my = rng.uniform(-1, ny + 1, m).astype(np.float32)
    res = [[np.zeros(m, np.float32) for _ in range(2)] for _ in range(2)]
    K.kernel("ibm_interpolate_velocity_2d", ux, uy, mx, my, res[0][0], res[0][1], nx, ny, m)
    ob.call(32, "ibm_interpolate_velocity_2d", ux, uy, mx, my, res[1][0], res[1][1], nx, ny, m)
    for k in range(2):
        assert np.allclose(res[0][k], res[1][k], rtol=1e-5, atol=2e-7)
    fx = rng.normal(0, 0.01, m).astype(np.float32)
    fy = rng.normal(0, 0.01, m).astype(np.float32)
    ds = rng.uniform(0.5, 1.5, m).astype(np.float32)
    gx = [np.full(nx * ny, 1e-3, np.float32) for _ in range(2)]
    gy = [np.zeros(nx * ny, np.float32) for _ in range(2)]
    K.kernel("ibm_spread_force_2d", mx, my, fx, fy, ds, gx[0], gy[0], nx, ny, m)
    ob.call(32, "ibm_spread_force_2d", mx, my, fx, fy, ds, gx[1], gy[1], nx, ny, m)
    assert np.allclose(gx[0], gx[1], rtol=1e-5, atol=2e-8)
    assert np.allclose(gy[0], gy[1], rtol=1e-5, atol=2e-8)
    # deterministic: a second call adds the same amounts bit for bit
    again = [np.full(nx * ny, 1e-3, np.float32), np.zeros(nx * ny, np.float32)]
    K.kernel("ibm_spread_force_2d", mx, my, fx, fy, ds, again[0], again[1], nx, ny, m)
    assert np.array_equal(again[0], gx[0]) and np.array_equal(again[1], gy[0])
    with pytest.raises(RuntimeError):
        K.kernel("ibm_interpolate_velocity_2d", ux, uy, np.array([np.nan], np.float32), my[:1].copy(),
                 res[0][0], res[0][1], nx, ny, 1)


class OracleThermal:
    """thermal_solver.cpp:83-123 composed from the oracle's kernels (the
    streaming with a zero wall velocity)."""

    def __init__(self, nx, ny, tau, t_ref, beta, bits):
        self.nx, self.ny, self.n, self.tau, self.t_ref, self.beta, self.bits = nx, ny, nx * ny, tau, t_ref, beta, bits
        dt = np.float64 if bits == 64 else np.float32
        self.g_src = np.repeat((W * t_ref)[:, None], self.n, 1).astype(dt)
        self.g_dst = self.g_src.copy()
        self.T = np.full(self.n, t_ref, dt)
        self.gy = np.zeros(self.n, dt)

    def step(self, fluid, g_gravity, hot, cold):
        b = self.bits
        ct = np.ascontiguousarray(fluid.cell_type())
        ux, uy = np.ascontiguousarray(fluid.ux()), np.ascontiguousarray(fluid.uy())
        ob.call(b, "thermal_collide_d2q9", self.g_src, self.g_dst, ct, ux, uy, self.T, self.tau, self.n)
        ob.call(b, "thermal_apply_isothermal_wall", self.g_dst, self.nx, self.ny, 0, hot)
        ob.call(b, "thermal_apply_isothermal_wall", self.g_dst, self.nx, self.ny, self.ny - 1, cold)
        ob.call(b, "d2q9_stream_with_boundaries", self.g_dst, self.g_src, ct, self.nx, self.ny, 0.0, 0.0)
        self.gy[:] = 0
        ob.call(b, "thermal_compute_buoyancy_force_2d", self.T, self.gy, self.t_ref, self.beta, g_gravity, self.n)


def test_thermal_solver_rayleigh_benard_matches_oracle(wa):
    nx, ny, steps = 96, 34, 300
    tau_f, tau_t, beta, gg = 0.56, 0.53, 0.05, -0.001
    fluid = wa.LatticeGrid2D(nx, ny, tau_f)
    fluid.initialize(1.0, 0.0, 0.0)
    th = wa.ThermalSolver2D(nx, ny, tau_t, 0.5, beta)
    th.initialize(0.5)
    th.set_isothermal_wall_south(1.0)
    th.set_isothermal_wall_north(0.0)
    ors = []
    for bits in (32, 64):
        of = ob.OracleGrid(2, nx, ny, tau=tau_f, lattice=9, bits=bits)
        of.initialize(1.0, 0.0, 0.0)
        ors.append((of, OracleThermal(nx, ny, tau_t, 0.5, beta, bits)))
    for s in range(steps):
        th.step(fluid, gg)
        gy_fluid = -0.0005 * 0.05 * (th.compute_mean_temperature() - 0.5)
        fluid.step("TRT", 0.0, gy_fluid)
        for of, ot in ors:
            ot.step(of, gg, 1.0, 0.0)
            of.step("TRT", 0.0, -0.0005 * 0.05 * (float(np.mean(ot.T, dtype=np.float64)) - 0.5))
    T = th.temperature_data()
    (o32f, o32), (o64f, o64) = ors
    check_fp32_parity(T, o32.T, o64.T, "thermal T")
    check_fp32_parity(th.buoyancy_data(), o32.gy, o64.gy, "buoyancy")
    check_fp32_parity(fluid.ux_data(), o32f.ux(), o64f.ux(), "fluid ux")
    # heat enters at the hot wall: the first row above it is warmer than the mean
    t2 = T.reshape(ny, nx)
    assert t2[1].mean() > 0.5 > t2[ny - 2].mean()
    assert th.compute_nusselt_number() > 0
