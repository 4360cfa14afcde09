"""GPU tests of the batched multi-domain step (wake_batch_* through ctypes).

A member of a batch must give what the same domain gives stepped alone: the
same kernel code runs per node, so a batch of one and a batch of many are
bit-identical, the plain single-domain kernel agrees to FP32 rounding, and
the oracle's restatement of the reference step agrees to the 1e-5 bar (or,
on a Zou-He channel that amplifies rounding past it, to the FP32-floor
criterion of parity.check_fp32_parity).
"""
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


def cylinder_channel(target, nx, dy_off, umax=0.05, re=100.0, storage=0, oracle=False):
    """Batched-wake geometry: NY = NX/4, cylinder D = NY/8 at (NX/5, NY/2 +
    offset), solid walls at y = 0 and NY - 1, parabolic Zou-He inlet, Zou-He
    pressure outlet, MRT.  Returns (grid, (cx, cy, D))."""
    ny = nx // 4
    D = ny / 8.0
    tau = 3.0 * (umax * D / re) + 0.5
    if oracle:
        g = ob.OracleGrid(2, nx, ny, 1, tau=tau, lattice=9, bits=target)
        cell = lambda x, y, t: g.set_cell_type(x, y, 0, t)
    else:
        g = target.LatticeGrid2D(nx, ny, tau, storage=storage)
        cell = lambda x, y, t: g.set_cell_type(x, y, t)
    g.initialize(1.0, umax, 0.0)
    g.set_inflow_parabolic(umax)
    g.set_outflow_east(1.0)
    for x in range(nx):
        cell(x, 0, 1)
        cell(x, ny - 1, 1)
    cx, cy = nx / 5.0, ny / 2.0 + dy_off
    g.set_solid_circle(cx, cy, D / 2)
    return g, (cx, cy, D)


SIZES = [(128, 1), (160, -2), (96, 0)]


def test_batch_members_match_batch_of_one_and_plain_step(wa):
    steps = [35, 20, 35]
    batch = wa.Batch(lattice=9, collision="MRT", chunk=10)
    grids = []
    for (nx, off), k in zip(SIZES, steps):
        g, _ = cylinder_channel(wa, nx, off)
        batch.add(g, k)
        grids.append(g)
    assert batch.run(40) > 0.0
    for (nx, off), k, g in zip(SIZES, steps, grids):
        assert g.current_step() == k
        # batch of one
        solo = wa.Batch(lattice=9, collision="MRT", chunk=7)
        g1, _ = cylinder_channel(wa, nx, off)
        solo.add(g1, k)
        solo.run(k)
        assert np.array_equal(g.f_src_data(), g1.f_src_data()), f"nx={nx}: batch of one differs"
        # plain single-domain path
        g2, _ = cylinder_channel(wa, nx, off)
        for _ in range(k):
            g2.step("MRT")
        assert rel_l2(g.f_src_data(), g2.f_src_data()) <= 1e-6
        assert rel_l2(g.ux_data(), g2.ux_data()) <= 1e-5


def test_batch_member_vs_oracle_and_forces(wa):
    nx, off = SIZES[0]
    steps = 300
    g, (cx, cy, D) = cylinder_channel(wa, nx, off)
    o, _ = cylinder_channel(32, nx, off, oracle=True)
    o64, _ = cylinder_channel(64, nx, off, oracle=True)
    batch = wa.Batch(lattice=9, collision="MRT", chunk=50)
    slot = batch.add(g, steps, forces=(cx, cy, 0.0, 1.5 * D))
    batch.run(steps)
    F = batch.forces(slot)
    assert F.shape == (steps, 3)
    o.run(steps, "MRT")
    o64.run(steps, "MRT")
    fl = o.cell_type() == 0
    assert rel_l2(g.rho_data()[fl], o.rho()[fl]) <= FIELD_TOL
    vel = lambda a: np.concatenate([a.ux()[fl], a.uy()[fl]])
    # a Zou-He channel: the FP32 oracle ends 3.4e-5 from FP64 at 300 steps
    check_fp32_parity(np.concatenate([g.ux_data()[fl], g.uy_data()[fl]]), vel(o), vel(o64), "batched channel u")
    fo = o.momentum_exchange(cx, cy, 0.0, 1.5 * D)
    assert abs(F[-1, 0] - fo[0]) <= 1e-4 * abs(fo[0])
    assert abs(F[-1, 1] - fo[1]) <= 1e-4 * abs(fo[0])
    # the recorded history is the force after each step
    g2, _ = cylinder_channel(wa, nx, off)
    for s in range(5):
        g2.step("MRT")
        f2 = g2.momentum_exchange_2d(cx, cy, 1.5 * D)
        assert abs(F[s, 0] - f2[0]) <= 1e-5 * abs(f2[0])


def test_batch_tune_and_graph_do_not_change_results(wa):
    a = wa.Batch(lattice=9, collision="MRT", chunk=8)
    b = wa.Batch(lattice=9, collision="MRT", chunk=8)
    ga, gb = [], []
    for nx, off in SIZES:
        g, _ = cylinder_channel(wa, nx, off)
        a.add(g, 60)
        ga.append(g)
        g, _ = cylinder_channel(wa, nx, off)
        b.add(g, 60)
        gb.append(g)
    shape = a.tune(probe_steps=2)
    assert shape[0] in (64, 128, 256)
    a.run(60)
    b.set_launch(64, 8)
    b.run(60, use_graph=False)
    for x, y in zip(ga, gb):
        assert x.current_step() == 60 and y.current_step() == 60
        assert np.array_equal(x.f_src_data(), y.f_src_data())


def test_batch_membership_changes_and_mixed_progress(wa):
    batch = wa.Batch(lattice=9, collision="MRT", chunk=10)
    g0, _ = cylinder_channel(wa, 128, 0)
    s0 = batch.add(g0, 25)
    batch.run(10)
    g1, _ = cylinder_channel(wa, 96, 1)
    s1 = batch.add(g1, 30)  # joins while g0 is running
    batch.run(20)
    assert batch.remaining(s0) == 0 and batch.remaining(s1) == 10
    batch.remove(s0)
    batch.run(10)
    assert g0.current_step() == 25 and g1.current_step() == 30
    ref0, _ = cylinder_channel(wa, 128, 0)
    ref1, _ = cylinder_channel(wa, 96, 1)
    r = wa.Batch(lattice=9, collision="MRT", chunk=10)
    r.add(ref0, 25)
    r.add(ref1, 30)
    r.run(30)
    assert np.array_equal(g0.f_src_data(), ref0.f_src_data())
    assert np.array_equal(g1.f_src_data(), ref1.f_src_data())


def _sphere3d(wa, n, storage=0):
    g = wa.LatticeGrid3D(2 * n, n, n, 0.505, storage=storage)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.05, 0.0, 0.0)
    g.set_les_coupled(True)
    g.set_inflow_west(0.05, 0.0)
    g.set_inflow_turbulence(0.01, 2)
    g.set_convective_outflow_east(0.05)
    c = (n / 2.0, n / 2.0, n / 2.0)
    g.set_solid_sphere(*c, n / 8.0)
    g.build_curved_sphere_links(*c, n / 8.0)
    return g


@pytest.mark.parametrize("shape", [None, (64, 4)], ids=["tiles", "persistent"])
@pytest.mark.parametrize("storage", [0, 1, 2])
def test_batch_3d_les_bfl_matches_plain(wa, storage, shape):
    """3D members (BFL sphere, turbulent inlet, convective outlet, coupled
    LES) stepped in a batch match the plain step, with one work-group per row
    segment and with a persistent grid that walks the segments of both
    members (4 work-groups per CU of 64 lanes)."""
    steps = 40
    batch = wa.Batch(lattice=19, storage=storage, collision="MRT", use_les=True, chunk=16)
    gs = [_sphere3d(wa, n, storage) for n in (24, 32)]
    for g in gs:
        batch.add(g, steps, smag=0.17)
    if shape is not None:
        batch.set_launch(*shape)
        assert batch.info()["block"] == shape[0] and batch.info()["grid_factor"] == shape[1]
    batch.run(steps)
    for g, n in zip(gs, (24, 32)):
        p = _sphere3d(wa, n, storage)
        for _ in range(steps):
            p.step("MRT", 0.0, 0.0, 0.0, True, 0.17)
        # 16-bit storage rounds each population to the format, so a last-bit
        # difference in FP32 can move a stored value by one FP16 / BF16 step
        assert rel_l2(g.f_src_data(), p.f_src_data()) <= {0: 1e-6, 1: 1e-5, 2: 1e-4}[storage]


def _wing(wa, nx, storage):
    from bench import wing_setup

    ny = nz = nx // 2
    return wing_setup(wa.LatticeGrid3D(nx, ny, nz, 0.52, lattice=27, storage=storage), nx, ny, nx / 4.0)


@pytest.mark.parametrize("shape", [None, (64, 4)], ids=["tiles", "persistent"])
@pytest.mark.parametrize("storage", [0, 1, 2])
def test_batch_d3q27_wing_matches_plain(wa, storage, shape):
    """D3Q27 TRT + coupled LES wing channels of two sizes (walls, a NACA
    section, velocity inlet, convective outlet, rows of 48 and 80 nodes)
    stepped in a batch match the plain step, under both launch shapes."""
    steps = 30
    sizes = (48, 80)
    batch = wa.Batch(lattice=27, storage=storage, collision="TRT", use_les=True, chunk=16)
    gs = [_wing(wa, nx, storage) for nx in sizes]
    for g in gs:
        batch.add(g, steps, smag=0.17)
    if shape is not None:
        batch.set_launch(*shape)
    batch.run(steps)
    for g, nx in zip(gs, sizes):
        p = _wing(wa, nx, storage)
        for _ in range(steps):
            p.step("TRT", 0.0, 0.0, 0.0, True, 0.17)
        assert np.all(np.isfinite(p.f_src_data()))
        assert rel_l2(g.f_src_data(), p.f_src_data()) <= {0: 1e-6, 1: 1e-5, 2: 1e-4}[storage]


def test_forces_member_without_links_records_zeros(wa):
    """A member with forces but no wall link (no body, or r_max too small)
    records zero forces, also after an earlier member left the batch and the
    slots moved."""
    batch = wa.Batch(lattice=9, collision="MRT", chunk=10)
    g0, (cx, cy, D) = cylinder_channel(wa, 128, 1)
    s0 = batch.add(g0, 40, forces=(cx, cy, 0.0, 1.5 * D))
    empty = wa.LatticeGrid2D(96, 24, 0.7)
    empty.initialize(1.0, 0.03, 0.0)
    s1 = batch.add(empty, 40, forces=(48.0, 12.0, 0.0, 5.0))
    g2, (cx2, cy2, D2) = cylinder_channel(wa, 160, -2)
    s2 = batch.add(g2, 40, forces=(cx2, cy2, 0.0, 0.1))  # r_max below every link midpoint
    # the default width is the one wasting the fewest lanes over the members'
    # rows (128x32, 96x24, 160x40: 64), reported before the first launch
    assert batch.info()["block"] == 64
    batch.run(15)
    assert np.all(batch.forces(s1) == 0.0) and np.all(batch.forces(s2) == 0.0)
    assert batch.forces(s0).shape == (15, 3) and batch.forces(s0)[-1, 0] != 0.0
    batch.remove(s0)
    batch.run(25)
    F1, F2 = batch.forces(s1), batch.forces(s2)
    assert F1.shape == (40, 3) and F2.shape == (40, 3)
    assert np.all(F1 == 0.0) and np.all(F2 == 0.0)
