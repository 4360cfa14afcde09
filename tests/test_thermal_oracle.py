"""CPU pins of the oracle's thermal extension (D2Q5 temperature coupled to
the D2Q9 flow by a Boussinesq force, anti-bounce-back wall temperatures).

No reference test covers it (the reference's ThermalSolver2D is a D2Q9
passive scalar, src/thermal/thermal_solver.cpp:83-123); the pins are closed
forms and a published benchmark:
* a linear temperature between two anti-bounce-back walls at rest is an
  exact fixed point of BGK D2Q5 with halfway walls at x = 1/2, nx - 3/2;
* adiabatic (bounce-back) walls conserve the total temperature;
* the differentially heated square cavity at Ra = 1e4, Pr = 0.71 has mean
  Nusselt number 2.243 (de Vahl Davis 1983, benchmark solution).
"""
import math

import numpy as np
import pytest

import oracle_bind as ob


def cavity(bits, n, ra=1e4, pr=0.71, U=0.1, coll="MRT", gbeta_on=True):
    """Square cavity of n x n nodes, solid frame; hot wall x = 0 (T = 1),
    cold wall x = n-1 (T = 0), adiabatic top and bottom; H = n - 2."""
    H = n - 2
    nu = U * H * math.sqrt(pr / ra)
    alpha = nu / pr
    gbeta = U * U / H if gbeta_on else 0.0
    o = ob.OracleGrid(2, n, n, 1, tau=3 * nu + 0.5, lattice=9, bits=bits)
    o.initialize(1.0, 0.0, 0.0)
    for k in range(n):
        for (x, y) in ((0, k), (n - 1, k), (k, 0), (k, n - 1)):
            o.set_cell_type(x, y, 0, 1)
    o.enable_thermal(3 * alpha + 0.5, gbeta, 0.5, 0.5)
    o.set_wall_temperature(0, 0, 0, n - 1, 1.0)
    o.set_wall_temperature(n - 1, 0, n - 1, n - 1, 0.0)
    return o, alpha, H


def linear_profile(n):
    x = np.arange(n, dtype=np.float64)
    t = 1.0 - (x - 0.5) / (n - 2)
    return np.tile(t, n)


def test_thermal_conduction_linear_profile_is_fixed_point():
    n = 20
    o, alpha, H = cavity(64, n, gbeta_on=False)
    o.set_temperature_field(linear_profile(n))
    o.run(300, "MRT")
    fl = o.cell_type() == 0
    T = o.temperature()
    assert np.max(np.abs(T[fl] - linear_profile(n)[fl])) < 1e-12
    assert np.max(np.abs(o.ux())) < 1e-14 and np.max(np.abs(o.uy())) < 1e-14
    # FP32: the same fixed point to rounding
    o32, _, _ = cavity(32, n, gbeta_on=False)
    o32.set_temperature_field(linear_profile(n))
    o32.run(300, "MRT")
    assert np.max(np.abs(o32.temperature()[fl] - linear_profile(n)[fl])) < 2e-6


def test_thermal_adiabatic_box_conserves_temperature():
    n = 24
    o = ob.OracleGrid(2, n, n, 1, tau=0.7, lattice=9, bits=64)
    o.initialize(1.0, 0.0, 0.0)
    for k in range(n):
        for (x, y) in ((0, k), (n - 1, k), (k, 0), (k, n - 1)):
            o.set_cell_type(x, y, 0, 1)
    o.enable_thermal(0.8, 1e-3, 0.5, 0.5)
    rng = np.random.default_rng(3)
    o.set_temperature_field(rng.uniform(0.0, 1.0, n * n))
    fl = o.cell_type() == 0
    s0 = o.temperature()[fl].sum()
    o.run(500, "MRT")
    assert abs(o.temperature()[fl].sum() / s0 - 1) < 1e-12
    assert np.max(np.abs(o.uy())) > 1e-6  # the buoyancy did drive a flow


def nusselt(o, alpha, H, n):
    """Mean horizontal heat flux through the interior vertical mid-lines,
    u_x T - alpha dT/dx at x + 1/2, scaled by alpha dT / H."""
    T = o.temperature().reshape(n, n)
    ux = o.ux().reshape(n, n)
    rows = slice(1, n - 1)
    q = []
    for x in range(1, n - 2):
        flux = 0.5 * (ux[rows, x] * T[rows, x] + ux[rows, x + 1] * T[rows, x + 1]) - alpha * (T[rows, x + 1] - T[rows, x])
        q.append(flux.sum() / H)
    return float(np.mean(q)) * H / alpha


@pytest.mark.slow
def test_thermal_cavity_ra1e4_nusselt_de_vahl_davis():
    n = 42
    o, alpha, H = cavity(64, n)
    steps = int(4 * H * H / alpha)
    o.run(steps, "MRT")
    nu = nusselt(o, alpha, H, n)
    assert abs(nu / 2.243 - 1) < 0.03, nu
