"""Per-step momentum exchange on the GPU.

The reference's scenarios take the momentum-exchange force after every step
on the post-stream populations (sphere_3d_wake.cpp:94-114,
karman_vortex.cpp:104-120; the kernel is forces_momentum_exchange.ispc:6-138).
Here that force is computed inside the step launch that forms each
post-stream state (the force probe) and kept as a per-step history.  Checked:

* the history is bit-identical to calling momentum_exchange() after every
  step, across ring folds, mid-run reads, a configuration change that
  materialises the state, reduced-precision storage and a geometry change;
* the history matches the oracle's momentum exchange after every step
  (C_D within 1e-4 relative, C_L within 1e-4 of the drag scale);
* the reference-shaped sphere scenario driven by SimulationRunner through
  the C++ API exports the same per-step coefficients as the oracle.
"""
import csv
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-4


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def sphere_channel(wa, storage, nx=64, ny=48, nz=48, R=6.0):
    """The 256^3 sphere workload's boundary set at reduced size: MRT +
    coupled LES, Zou-He inlet with 1% seeded turbulence, convective outlet,
    periodic y/z, BFL sphere."""
    g = wa.LatticeGrid3D(nx, ny, nz, 0.505, storage=storage)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.05, 0.0, 0.0)
    g.set_les_coupled(True)
    g.set_inflow_west(0.05, 0.0)
    g.set_inflow_turbulence(0.01, 2)
    g.set_convective_outflow_east(0.05)
    c = (16.0, 24.0, 24.0)
    g.set_solid_sphere(*c, R)
    g.build_curved_sphere_links(*c, R)
    return g, c, R


@pytest.mark.parametrize("storage", [0, 1], ids=["fp32", "fp16"])
def test_force_history_equals_per_step_calls(wa, storage):
    """Probe history vs momentum_exchange() after every step: the same bits,
    over 150 steps (two ring folds), with a history read at step 70 and a
    configuration change at step 100 (the state is materialised)."""
    a, c, R = sphere_channel(wa, storage)
    b, _, _ = sphere_channel(wa, storage)
    a.set_force_probe(*c, 2 * R)
    calls = [b.momentum_exchange(*c, 2 * R)]
    for s in range(1, 151):
        a.step("MRT", 0.0, 0.0, 0.0, True, 0.17)
        b.step("MRT", 0.0, 0.0, 0.0, True, 0.17)
        calls.append(b.momentum_exchange(*c, 2 * R))
        if s == 70:
            first, mid = a.force_history()
            assert first == 0 and mid.shape == (71, 3)
            assert np.array_equal(mid, np.array(calls))
        if s == 100:
            for g in (a, b):
                g.set_inflow_turbulence(0.02, 3)
    first, hist = a.force_history()
    assert first == 0 and hist.shape == (151, 3)
    calls = np.array(calls)
    assert np.all(np.isfinite(hist)) and calls[-1, 0] > 0.0
    bad = np.flatnonzero(np.any(hist != calls, axis=1))
    assert bad.size == 0, f"steps {bad[:10]} differ: {hist[bad[:3]]} vs {calls[bad[:3]]}"


def test_force_history_geometry_change_and_no_links(wa):
    """A probe set on a domain without a body records zeros; a sphere placed
    mid-run (new links, new row layout) is picked up from that step on."""
    a = wa.LatticeGrid3D(40, 24, 24, 0.6)
    b = wa.LatticeGrid3D(40, 24, 24, 0.6)
    c = (12.0, 12.0, 12.0)
    for g in (a, b):
        g.initialize(1.0, 0.04, 0.0, 0.0)
    a.set_force_probe(*c, 10.0)
    calls = []
    for s in range(1, 101):
        if s == 30:
            for g in (a, b):
                g.set_solid_sphere(*c, 4.0)
                g.build_curved_sphere_links(*c, 4.0)
        for g in (a, b):
            g.step("BGK")
        calls.append(b.momentum_exchange(*c, 10.0))
    first, hist = a.force_history()
    assert first == 0 and hist.shape == (101, 3)
    assert np.all(hist[:30] == 0.0)
    assert np.array_equal(hist[1:], np.array(calls))
    assert hist[-1, 0] > 0.0


def test_force_history_2d_with_torque_vs_oracle(wa):
    """2D cylinder channel (TRT, Zou-He inlet and outlet, walls): the history
    (Fx, Fy, torque) against the oracle's momentum exchange after every
    step, and bit-identical to per-step calls."""
    nx, ny, D = 160, 64, 12.0
    cx, cy = 40.0, 31.0
    u = 0.04
    tau = 3.0 * u * D / 150.0 + 0.5
    g = wa.LatticeGrid2D(nx, ny, tau)
    b = wa.LatticeGrid2D(nx, ny, tau)
    o = ob.OracleGrid(2, nx, ny, tau=tau)
    for h in (g, b, o):
        h.initialize(1.0, u, 0.0)
        h.set_inflow_west(u, 0.0)
        h.set_outflow_east(1.0)
        h.set_solid_circle(cx, cy, D / 2)
    g.set_force_probe(cx, cy, 0.0, 1.5 * D)
    fo, fb = [o.momentum_exchange(cx, cy, 0.0, 1.5 * D)], [b.momentum_exchange_2d(cx, cy, 1.5 * D)]
    for _ in range(400):
        g.step("TRT", 0.0, 0.0)
        b.step("TRT", 0.0, 0.0)
        o.step("TRT")
        fo.append(o.momentum_exchange(cx, cy, 0.0, 1.5 * D))
        fb.append(b.momentum_exchange_2d(cx, cy, 1.5 * D))
    first, hist = g.force_history()
    fo, fb = np.array(fo), np.array(fb)
    assert first == 0 and hist.shape == fo.shape
    assert np.array_equal(hist, fb)
    scale = np.max(np.abs(fo[50:, 0]))
    dx = np.abs(hist[:, 0] - fo[:, 0]) / np.maximum(np.abs(fo[:, 0]), scale)
    dy = np.abs(hist[:, 1] - fo[:, 1]) / scale
    dt = np.abs(hist[:, 2] - fo[:, 2]) / (scale * D)
    print(f"2D history vs oracle: Fx {dx.max():.2e}, Fy {dy.max():.2e}, torque {dt.max():.2e}")
    assert dx.max() <= FORCE_TOL and dy.max() <= FORCE_TOL and dt.max() <= FORCE_TOL


def _oracle_sphere_scenario(nx, ny, nz, re, u, r, steps):
    """sphere_3d_wake.cpp:24-114 on the oracle: per-step C_D, C_L."""
    f32 = np.float32
    nu = f32(f32(u) * f32(f32(2.0) * f32(r))) / f32(re)
    tau = float(f32(f32(3.0) * nu) + f32(0.5))
    o = ob.OracleGrid(3, nx, ny, nz, tau=tau)
    cx, cy, cz = nx * 0.25, ny * 0.5, nz * 0.5
    o.set_periodic(True)
    o.initialize(1.0, u, 0.0, 0.0)
    o.set_solid_sphere(cx, cy, cz, r)
    o.build_curved_sphere_links(cx, cy, cz, r)
    q = 0.5 * u * u * math.pi * r * r
    out = [o.momentum_exchange(cx, cy, cz, 2 * r) / q]
    for _ in range(steps):
        o.step("MRT", use_les=re >= 500.0, smag=0.12)
        out.append(o.momentum_exchange(cx, cy, cz, 2 * r) / q)
    return np.array(out)


@pytest.mark.parametrize("re", [200.0, 600.0])
def test_sphere_scenario_through_simulation_runner(wa, tmp_path, re):
    """The reference's sphere-wake scenario on the C++ API, run by
    SimulationRunner with the force of every step: the exported per-step
    C_D / C_L against the oracle running the same scenario (Re 600 takes the
    reference's LES switch)."""
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "grid_api_check")
    steps = 300
    r = subprocess.run([exe, "sphere_runner", str(tmp_path), str(steps), str(re)], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(tmp_path / "sphere_aero.csv") as fh:
        rows = list(csv.DictReader(fh))
    assert [int(x["step"]) for x in rows] == list(range(steps + 1))
    cd = np.array([float(x["cd"]) for x in rows])
    cl = np.array([float(x["cl"]) for x in rows])
    ref = _oracle_sphere_scenario(64, 32, 32, re, 0.05, 6.0, steps)
    ecd = np.abs(cd - ref[:, 0]) / np.abs(ref[:, 0])
    scale = np.max(np.abs(ref[:, 0]))
    ecl = np.abs(cl - ref[:, 1]) / scale
    print(f"Re {re}: C_D {cd[-1]:.5f} (oracle {ref[-1, 0]:.5f}); max rel C_D error {ecd.max():.2e}, "
          f"C_L error / drag scale {ecl.max():.2e}")
    assert ecd.max() <= FORCE_TOL and ecl.max() <= FORCE_TOL
    with open(tmp_path / "sphere_perf.csv") as fh:
        perf = list(csv.DictReader(fh))
    assert [int(x["step"]) for x in perf] == list(range(10, steps + 1, 10))
