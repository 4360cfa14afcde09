"""GPU checks of the benchmark configuration at its full size, 256^3, where
the oracle cannot follow: properties that hold at any size.

* Translation invariance.  With periodic y and z, moving the sphere by
  whole cells along those axes moves every field by the same amount, bit
  for bit: the per-node arithmetic is the same, only the indexing differs
  (padded plane stride, wrap rows, BFL records, inlet and outlet planes).
  The inlet noise is keyed on the node index, so it is off here.
* Conservation.  BGK in a fully periodic box without force conserves mass
  and momentum up to FP32 rounding: 30 steps from a seeded perturbation of
  every population.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 256


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def sphere_run(wa, cy, cz, storage, steps=30):
    g = wa.LatticeGrid3D(N, N, N, 0.505, storage=storage)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.05, 0.0, 0.0)
    g.set_les_coupled(True)
    g.set_inflow_west(0.05, 0.0)
    g.set_convective_outflow_east(0.05)
    g.set_solid_sphere(64.0, cy, cz, 24.0)
    nl = g.build_curved_sphere_links(64.0, cy, cz, 24.0)
    for _ in range(steps):
        g.step("MRT", 0.0, 0.0, 0.0, True, 0.17)
    shape = (N, N, N)
    out = (nl, g.ux_data().reshape(shape).copy(), g.uy_data().reshape(shape).copy(),
           g.rho_data().reshape(shape).copy())
    del g
    return out


@pytest.mark.parametrize("storage", [0, 1])
def test_sphere_256_translation_invariance(wa, storage):
    dy, dz = 37, -61
    nl_a, ux_a, uy_a, rho_a = sphere_run(wa, 128.0, 128.0, storage)
    nl_b, ux_b, uy_b, rho_b = sphere_run(wa, 128.0 + dy, 128.0 + dz, storage)
    assert nl_a == nl_b == 41322
    assert np.isfinite(ux_a).all()
    for a, b, name in ((ux_a, ux_b, "ux"), (uy_a, uy_b, "uy"), (rho_a, rho_b, "rho")):
        shifted = np.roll(a, (dz, dy), axis=(0, 1))  # arrays are [z, y, x]
        diff = int(np.count_nonzero(shifted.view(np.uint32) != b.view(np.uint32)))
        assert diff == 0, f"{name}: {diff} nodes differ after the shift"
    # the flow is not trivial: the sphere has slowed the stream behind it
    assert ux_a[128, 128, 100] < 0.9 * 0.05


def test_bgk_periodic_256_conserves_mass_and_momentum(wa):
    g = wa.LatticeGrid3D(N, N, N, 0.7)
    g.initialize(1.0, 0.02, -0.01, 0.005)
    g.add_noise(1e-3, 11)
    g.update_macroscopic()
    rho0 = g.rho_data().astype(np.float64)
    m0 = rho0.sum()
    j0 = [(rho0 * u.astype(np.float64)).sum() for u in (g.ux_data(), g.uy_data(), g.uz_data())]
    for _ in range(30):
        g.step("BGK", 0.0, 0.0, 0.0, False, 0.14)
    rho = g.rho_data().astype(np.float64)
    m = rho.sum()
    j = [(rho * u.astype(np.float64)).sum() for u in (g.ux_data(), g.uy_data(), g.uz_data())]
    n = float(N ** 3)
    print(f"mass drift {abs(m - m0) / m0:.2e}, momentum drift "
          f"{max(abs(a - b) for a, b in zip(j, j0)) / n:.2e} per node")
    # FP32: the rounded weights do not sum to exactly 1 (2.4e-7 off, summed
    # in order), and relaxation feeds that into the mass every step (6.4e-7
    # over these 30 steps); an indexing or streaming error moves far more
    assert abs(m - m0) / m0 < 2e-6
    assert max(abs(a - b) for a, b in zip(j, j0)) / n < 1e-7
    assert abs(g.compute_total_mass() - m) / m < 1e-6
