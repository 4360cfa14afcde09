"""Fully periodic D2Q9 domains run 8 steps per launch (k_step_tb2d): each
work-group pull-collides its 16 x 16 nodes eight times in LDS from a
32 x 32 tile of the post-collision state.  Every node update is the same
pull, collision and storage rounding as one single-step launch, so a run
must give the same bits as the same number of single steps (the reference
grid steps one launch at a time, lattice_grid_2d.cpp step()).  Sizes that
are not multiples of the 16-node tile, domains smaller than the tile's halo
(the periodic wrap then repeats inside one tile), every collision with and
without the coupled LES, a body force and the three storages are covered,
as is the Taylor-Green decay against its closed form and the oracle.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from parity import FIELD_TOL, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def _tg_fields(nx, ny, u0=0.04):
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    kx, ky = 2 * math.pi / nx, 2 * math.pi / ny
    ux = (u0 * np.sin(kx * x) * np.cos(ky * y)).astype(np.float32).ravel()
    uy = (-u0 * np.cos(kx * x) * np.sin(ky * y)).astype(np.float32).ravel()
    return np.ones(nx * ny, np.float32), ux, uy


def _grid(wa, nx, ny, tau, storage, multistep):
    g = wa.LatticeGrid2D(nx, ny, tau, storage=storage)
    g.set_periodic(True)
    g.set_les_coupled(True)
    g.initialize_fields(*_tg_fields(nx, ny))
    g.add_noise(1e-4, 7)
    g.set_multistep(multistep)
    return g


@pytest.mark.parametrize("storage", [0, 1, 2])
@pytest.mark.parametrize("nx,ny", [(53, 37), (64, 48), (5, 7), (2, 9), (17, 1)])
def test_multistep_run_equals_single_steps(wa, storage, nx, ny):
    for coll, les, g in (("BGK", False, (0.0, 0.0)), ("TRT", False, (1e-6, -2e-6)), ("MRT", False, (0.0, 0.0)),
                         ("MRT", True, (2e-6, 0.0))):
        steps = 8 * 3 + 5  # three multi-step launches and five single steps
        runs = []
        for multistep in (True, False):
            h = _grid(wa, nx, ny, 0.56, storage, multistep)
            h.run(steps, coll, gx=g[0], gy=g[1], use_les=les, smag=0.17)
            runs.append((h.f_src_data().copy(), h.ux_data().copy(), h.current_step()))
        a, b = runs
        assert a[2] == b[2] == steps
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)), f"{coll} les={les} {nx}x{ny} st{storage}"
        assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_multistep_taylor_green_decay_and_oracle(wa):
    n, tau, steps = 128, 0.8, 400
    g = _grid(wa, n, n, tau, 0, True)
    g.set_les_coupled(False)
    g.initialize_fields(*_tg_fields(n, n))
    e0 = g.compute_total_kinetic_energy()
    g.run(steps, "BGK")
    nu = (tau - 0.5) / 3.0
    k = 2 * math.pi / n
    ratio = g.compute_total_kinetic_energy() / e0
    # the lattice decay rate differs from nu k^2 by O(k^2): 1.9e-3 at 256^2 and 1010 steps
    assert abs(ratio - math.exp(-4 * nu * k * k * steps)) < 2e-2 * ratio
    o = ob.OracleGrid(2, n, n, tau=tau)
    o.set_periodic(True)
    o.initialize_fields(*_tg_fields(n, n))
    o.run(steps, "BGK")
    err = rel_l2(np.concatenate([g.ux_data(), g.uy_data()]), np.concatenate([o.ux(), o.uy()]))
    assert err <= FIELD_TOL, err


def test_domains_with_solids_or_boundaries_keep_single_steps(wa):
    """A solid node, or a wall, makes the domain ineligible: run() is then
    the same as stepping, whatever the switch says."""
    for setup in ("solid", "walls"):
        runs = []
        for multistep in (True, False):
            h = wa.LatticeGrid2D(40, 24, 0.6)
            h.set_periodic(setup == "solid")
            h.initialize_fields(*_tg_fields(40, 24))
            if setup == "solid":
                h.set_cell_type(10, 10, wa.CellType.Solid)
            h.set_multistep(multistep)
            h.run(19, "MRT")
            runs.append(h.f_src_data().copy())
        assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))
