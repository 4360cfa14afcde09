"""GPU tests of the coupled D2Q9 flow + D2Q5 temperature step (thermal
cavities) against the oracle's restatement and the published cavity
benchmark, alone and in a batch."""
import math

import numpy as np
import pytest

import oracle_bind as ob
from parity import FIELD_TOL, check_fp32_parity, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def cavity(target, n, ra=1e4, pr=0.71, U=0.1, oracle_bits=None, gbeta_on=True, storage=0):
    """Differentially heated square cavity, n x n nodes with a solid frame:
    hot wall x = 0 (T = 1), cold wall x = n - 1 (T = 0), adiabatic top and
    bottom, H = n - 2, free-fall velocity U = sqrt(gbeta dT H)."""
    H = n - 2
    nu = U * H * math.sqrt(pr / ra)
    alpha = nu / pr
    gbeta = U * U / H if gbeta_on else 0.0
    if oracle_bits:
        g = ob.OracleGrid(2, n, n, 1, tau=3 * nu + 0.5, lattice=9, bits=oracle_bits)
        cell = lambda x, y: g.set_cell_type(x, y, 0, 1)
    else:
        g = target.LatticeGrid2D(n, n, 3 * nu + 0.5, storage=storage)
        cell = lambda x, y: g.set_cell_type(x, y, 1)
    g.initialize(1.0, 0.0, 0.0)
    for k in range(n):
        for (x, y) in ((0, k), (n - 1, k), (k, 0), (k, n - 1)):
            cell(x, y)
    g.enable_thermal(3 * alpha + 0.5, gbeta, 0.5, 0.5)
    g.set_wall_temperature(0, 0, 0, n - 1, 1.0)
    g.set_wall_temperature(n - 1, 0, n - 1, n - 1, 0.0)
    return g, alpha, H


def nusselt(T, ux, alpha, H, n):
    T = T.reshape(n, n)
    ux = ux.reshape(n, n)
    rows = slice(1, n - 1)
    q = []
    for x in range(1, n - 2):
        flux = 0.5 * (ux[rows, x] * T[rows, x] + ux[rows, x + 1] * T[rows, x + 1]) - alpha * (T[rows, x + 1] - T[rows, x])
        q.append(flux.sum() / H)
    return float(np.mean(q)) * H / alpha


def test_thermal_cavity_parity_with_oracle(wa):
    n, steps = 42, 2000
    g, alpha, H = cavity(wa, n)
    o, _, _ = cavity(None, n, oracle_bits=32)
    o64, _, _ = cavity(None, n, oracle_bits=64)
    ofma, _, _ = cavity(None, n, oracle_bits="32fma")
    for _ in range(steps):
        g.step("MRT")
    for x in (o, o64, ofma):
        x.run(steps, "MRT")
    fl = o.cell_type() == 0
    assert rel_l2(g.temperature_data()[fl], o.temperature()[fl]) <= FIELD_TOL
    assert rel_l2(g.rho_data()[fl], o.rho()[fl]) <= FIELD_TOL
    vel = lambda a, b: np.concatenate([a[fl], b[fl]])
    # the convection onset amplifies rounding: the two FP32 oracle builds
    # (separate and fused multiply-adds) end 1.9e-5 apart in u at 2000 steps
    check_fp32_parity(vel(g.ux_data(), g.uy_data()), vel(o.ux(), o.uy()), vel(o64.ux(), o64.uy()), "cavity u",
                      o32fma=vel(ofma.ux(), ofma.uy()))


def test_thermal_conduction_fixed_point_on_gpu(wa):
    n = 20
    g, alpha, H = cavity(wa, n, gbeta_on=False)
    x = np.arange(n, dtype=np.float64)
    lin = np.tile(1.0 - (x - 0.5) / (n - 2), n)
    g.set_temperature_field(lin.astype(np.float32))
    g.run(300, "MRT")
    fl = g.cell_type_data() == 0
    assert np.max(np.abs(g.temperature_data()[fl] - lin[fl])) < 2e-6
    assert np.max(np.abs(g.ux_data())) == 0.0


def test_thermal_cavity_nusselt_ra1e4(wa):
    n = 42
    g, alpha, H = cavity(wa, n)
    g.run(int(4 * H * H / alpha), "MRT")
    nu = nusselt(g.temperature_data().astype(np.float64), g.ux_data().astype(np.float64), alpha, H, n)
    print(f"Nu(Ra=1e4) = {nu:.4f} (de Vahl Davis 2.243)")
    assert abs(nu / 2.243 - 1) < 0.03


def test_thermal_batch_matches_alone(wa):
    sizes, steps = (34, 42, 26), 300
    batch = wa.Batch(lattice=9, collision="MRT", thermal=True, chunk=64)
    gs = []
    for n in sizes:
        g, _, _ = cavity(wa, n)
        batch.add(g, steps)
        gs.append(g)
    batch.run(steps)
    for n, g in zip(sizes, gs):
        solo = wa.Batch(lattice=9, collision="MRT", thermal=True, chunk=50)
        g1, _, _ = cavity(wa, n)
        solo.add(g1, steps)
        solo.run(steps)
        assert np.array_equal(g.temperature_data(), g1.temperature_data())
        assert np.array_equal(g.f_src_data(), g1.f_src_data())
        p, _, _ = cavity(wa, n)
        p.run(steps, "MRT")
        assert rel_l2(g.temperature_data(), p.temperature_data()) <= 1e-6
        assert rel_l2(g.f_src_data(), p.f_src_data()) <= 1e-6
