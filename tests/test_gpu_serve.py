"""GPU tests of the continuous-batching serving loop (wake_amd.serve).

A request's result must not depend on what else is on the GPU: the same
request served alone and in a mix gives the same numbers bit for bit (its
domain's arithmetic is the same in any batch and any chunking).  Admission
must respect the HBM budget, and a request whose error budget admits FP16
storage stays within 1e-3 of the FP64 oracle.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from parity import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    import wake_amd
    from wake_amd import serve

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return serve


def mix(sv):
    rng = np.random.default_rng(11)
    R = sv.SimRequest
    return [
        R(0, "bluff3d", 48, 160, re=200.0, semi_axes=(9.0, 6.0, 5.0), axes=sv.random_rotation(rng), tol=1e-3),
        R(1, "cavity2d", 40, 300, ra=2e4, tol=1e-5),
        R(2, "bluff3d", 40, 230, re=500.0, semi_axes=(6.0, 6.0, 4.0), axes=sv.random_rotation(rng), tol=1e-5),
        R(3, "cavity2d", 56, 120, ra=1e5, tol=1e-3),
    ]


def test_serve_results_same_alone_and_in_mix(sv):
    reqs = mix(sv)
    s = sv.Server(budget_bytes=8e9, chunk=32)
    s.serve(reqs, realtime=False)
    assert len(s.done) == len(reqs)
    together = {r.rid: r.result for r in s.done}
    for r in mix(sv):
        one = sv.Server(budget_bytes=8e9, chunk=50)
        one.serve([r], realtime=False)
        alone = one.done[0].result
        assert alone == together[r.rid], f"request {r.rid}: {alone} alone vs {together[r.rid]} in the mix"
    assert all(math.isfinite(v) for res in together.values() for v in res.values())


def test_serve_admission_respects_budget(sv):
    reqs = mix(sv)
    need = [sv.bytes_needed(r) for r in reqs]
    budget = max(need) + min(need)  # at most the largest plus the smallest at once
    s = sv.Server(budget_bytes=budget, chunk=32)
    s.serve(reqs, realtime=False)
    assert len(s.done) == len(reqs)
    assert s.peak_used <= budget
    # somebody had to wait for memory to free up
    assert max(r.admitted for r in s.done) >= min(r.finished for r in s.done)


def test_serve_fp16_cavity_within_1e3_of_fp64(sv):
    r = sv.SimRequest(0, "cavity2d", 42, 1500, ra=1e4, tol=1e-3)
    s = sv.Server(budget_bytes=2e9, chunk=100)
    s.serve([r], realtime=False)
    g = sv.make_domain(sv.SimRequest(0, "cavity2d", 42, 1500, ra=1e4, tol=1e-3))
    g.run(1500, "MRT")
    n, H = 42, 40
    nu = sv.U_FREEFALL * H * math.sqrt(sv.PR / 1e4)
    alpha = nu / sv.PR
    o = ob.OracleGrid(2, n, n, 1, tau=3 * nu + 0.5, lattice=9, bits=64)
    o.initialize(1.0, 0.0, 0.0)
    for k in range(n):
        for (x, y) in ((0, k), (n - 1, k), (k, 0), (k, n - 1)):
            o.set_cell_type(x, y, 0, 1)
    o.enable_thermal(3 * alpha + 0.5, sv.U_FREEFALL ** 2 / H, 0.5, 0.5)
    o.set_wall_temperature(0, 0, 0, n - 1, 1.0)
    o.set_wall_temperature(n - 1, 0, n - 1, n - 1, 0.0)
    o.run(1500, "MRT")
    fl = o.cell_type() == 0
    eu = rel_l2(np.concatenate([g.ux_data()[fl], g.uy_data()[fl]]), np.concatenate([o.ux()[fl], o.uy()[fl]]))
    et = rel_l2(g.temperature_data()[fl], o.temperature()[fl])
    print(f"FP16 cavity vs FP64 oracle: u {eu:.2e}, T {et:.2e}")
    assert eu <= 1e-3 and et <= 1e-3
    # the served request's Nusselt number agrees with the plain run's
    assert abs(s.done[0].result["Nu"] / sv.result_of(r, g)["Nu"] - 1) < 1e-4


def test_serve_least_work_admission(sv):
    """Shortest-job-first admission: with room for one request at a time the
    requests finish in order of node-steps, and each result equals the one
    the arrival-order policy gives (a request's result does not depend on
    when or beside what it runs)."""
    reqs = mix(sv)
    s = sv.Server(budget_bytes=8e9, chunk=32, policy="least_work", fill_nodes=1)
    s.serve(reqs, realtime=False)
    work = {r.rid: sv.node_count(r) * r.steps for r in reqs}
    finished = [r.rid for r in sorted(s.done, key=lambda r: r.finished)]
    assert finished == sorted(work, key=work.get)
    a = sv.Server(budget_bytes=8e9, chunk=32, policy="arrival")
    a.serve(mix(sv), realtime=False)
    assert {r.rid: r.result for r in s.done} == {r.rid: r.result for r in a.done}
