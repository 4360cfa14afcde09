"""CPU pins of the oracle's restatement of the reference's D2Q9 temperature
kernels (thermal_kernels.ispc:4-81) and immersed-boundary kernels
(ibm_kernels.ispc:4-96), each by a property fixed by the mathematics:

  * collision conserves the temperature sum_q g (sum_q w_q (1 + 3 c.u) = 1)
    and leaves its equilibrium w_q T (1 + 3 c.u) in place; solid nodes copy;
  * the isothermal wall, applied in place for q = 0..8 in turn, leaves
    g_1, g_2, g_5, g_6 as they were (to rounding) and sets g_0 and the
    opposite directions to -g + 2 w T_wall; at T_wall equal to the wall's
    own equilibrium temperature with g at rest equilibrium it is a no-op;
  * interpolation reproduces a constant field (the 4-point cosine delta is a
    partition of unity) and is periodic in x;
  * spreading adds F ds (1 + (1 - cos(pi f / 2)) / 4) in total, f the
    fractional part of y: the row floor(y) - 2 carries a weight because the
    reference's row weight has no |r| < 2 cutoff (ibm_kernels.ispc:80).
"""
import math

import numpy as np
import pytest

import oracle_bind as ob

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]


@pytest.mark.parametrize("bits", [32, 64])
def test_thermal_collide_conserves_and_keeps_equilibrium(bits):
    rng = np.random.default_rng(1)
    n = 500
    dt = np.float64 if bits == 64 else np.float32
    g = (W[:, None] * rng.uniform(0.2, 1.2, n)[None, :] * rng.uniform(0.8, 1.2, (9, n))).astype(dt)
    ux, uy = rng.uniform(-0.05, 0.05, n).astype(dt), rng.uniform(-0.05, 0.05, n).astype(dt)
    ct = np.zeros(n, np.uint8)
    ct[::7] = 1
    out = np.zeros_like(g)
    T = np.full(n, -1.0, dt)
    ob.call(bits, "thermal_collide_d2q9", np.ascontiguousarray(g), out, ct, ux, uy, T, 0.83, n)
    fl = ct == 0
    tol = 1e-12 if bits == 64 else 2e-6
    assert np.allclose(T[fl], g.sum(0)[fl], rtol=tol)
    assert np.allclose(out.sum(0)[fl], g.sum(0)[fl], rtol=tol)
    assert np.array_equal(out[:, ~fl], g[:, ~fl]) and np.all(T[~fl] == -1.0)
    # equilibrium in, equilibrium out
    t0 = rng.uniform(0.3, 1.0, n)
    geq = (W[:, None] * t0[None, :] * (1 + 3 * (CX[:, None] * ux[None, :] + CY[:, None] * uy[None, :]))).astype(dt)
    out2 = np.zeros_like(geq)
    ob.call(bits, "thermal_collide_d2q9", np.ascontiguousarray(geq), out2, np.zeros(n, np.uint8), ux, uy,
            np.zeros(n, dt), 0.6, n)
    assert np.allclose(out2, geq, rtol=1e-11 if bits == 64 else 5e-6)


def test_isothermal_wall_in_place_sequence():
    nx, ny, yw, tw = 12, 5, 4, 0.25
    rng = np.random.default_rng(2)
    g = (W[:, None] * rng.uniform(0.3, 1.0, (9, nx * ny))).astype(np.float64)
    before = g.copy()
    ob.call(64, "thermal_apply_isothermal_wall", g, nx, ny, yw, tw)
    row = slice(yw * nx, (yw + 1) * nx)
    other = np.ones(nx * ny, bool)
    other[row] = False
    assert np.array_equal(g[:, other], before[:, other])
    b = before[:, row]
    assert np.allclose(g[0, row], -b[0] + 2 * W[0] * tw, rtol=1e-14)
    for q in (1, 2, 5, 6):
        assert np.allclose(g[OPP[q], row], -b[q] + 2 * W[q] * tw, rtol=1e-13)
        assert np.allclose(g[q, row], b[q], rtol=1e-13)
    # at rest equilibrium of T_wall the wall changes nothing
    h = np.repeat((W * tw)[:, None], nx * ny, 1)
    h0 = h.copy()
    ob.call(64, "thermal_apply_isothermal_wall", h, nx, ny, 0, tw)
    assert np.allclose(h, h0, rtol=1e-14)


def test_buoyancy_adds_linear_term():
    n = 64
    T = np.linspace(0.0, 1.0, n).astype(np.float32)
    gy = np.full(n, 0.5, np.float32)
    ob.call(32, "thermal_compute_buoyancy_force_2d", T, gy, 0.5, 0.05, -0.001, n)
    assert np.allclose(gy, 0.5 + 0.05 * (T - 0.5) * -0.001, rtol=1e-6)


def test_ibm_interpolation_partition_of_unity_and_periodic_x():
    nx, ny = 40, 30
    ux = np.full(nx * ny, 0.037, np.float64)
    uy = np.full(nx * ny, -0.021, np.float64)
    rng = np.random.default_rng(3)
    mx = rng.uniform(0, nx, 50)
    my = rng.uniform(3, ny - 3, 50)  # whole stencil inside the y range
    ou, ov = np.zeros(50), np.zeros(50)
    ob.call(64, "ibm_interpolate_velocity_2d", ux, uy, mx, my, ou, ov, nx, ny, 50)
    assert np.allclose(ou, 0.037, rtol=1e-13) and np.allclose(ov, -0.021, rtol=1e-13)
    # a periodic field sampled at x and x + nx
    xs = np.arange(nx)
    f = np.tile(np.sin(2 * np.pi * xs / nx), ny)
    a, b = np.zeros(50), np.zeros(50)
    ob.call(64, "ibm_interpolate_velocity_2d", f, f, mx, my, a, np.zeros(50), nx, ny, 50)
    ob.call(64, "ibm_interpolate_velocity_2d", f, f, mx + nx, my, b, np.zeros(50), nx, ny, 50)
    assert np.allclose(a, b, atol=1e-12)
    # near the y walls the missing rows lower the weight sum
    c = np.zeros(1)
    ob.call(64, "ibm_interpolate_velocity_2d", np.ones(nx * ny), np.ones(nx * ny), np.array([10.3]),
            np.array([0.4]), c, np.zeros(1), nx, ny, 1)
    assert 0.5 < c[0] < 1.0


def test_ibm_spread_total_force_closed_form():
    nx, ny = 32, 28
    for fy_frac in (0.0, 0.25, 0.5, 0.9):
        mx, my = np.array([11.37]), np.array([13.0 + fy_frac])
        fx, fy, ds = np.array([0.02]), np.array([-0.01]), np.array([0.7])
        gx, gy = np.zeros(nx * ny), np.zeros(nx * ny)
        ob.call(64, "ibm_spread_force_2d", mx, my, fx, fy, ds, gx, gy, nx, ny, 1)
        extra = 0.25 * (1 - math.cos(0.5 * math.pi * fy_frac))
        assert gx.sum() == pytest.approx(0.02 * 0.7 * (1 + extra), rel=1e-12)
        assert gy.sum() == pytest.approx(-0.01 * 0.7 * (1 + extra), rel=1e-12)
        # the quirk row: floor(y) - 2 gets a weight unless y is an integer
        assert (np.abs(gx.reshape(ny, nx)[11]).sum() > 0) == (fy_frac > 0)
    # two markers add
    mx2, my2 = np.array([5.5, 5.5]), np.array([9.0, 9.0])
    g2 = np.zeros(nx * ny)
    ob.call(64, "ibm_spread_force_2d", mx2, my2, np.array([1.0, 2.0]), np.zeros(2), np.ones(2), g2, np.zeros(nx * ny),
            nx, ny, 2)
    assert g2.sum() == pytest.approx(3.0, rel=1e-12)
