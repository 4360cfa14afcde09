"""Divergence detection on the device (the reference's scenarios throw
NumericalDivergenceException, include/wake/core/error.hpp:14-18, after a
host scan of the densities).

The step kernels raise a per-domain flag when a density comes out of the
collision non-finite; check_finite() counts non-finite densities and speeds
on the device.  A NaN planted in one population (from the host, so every
kernel runs normally) must trip both, alone, in a batch and in the serving
loop, where the request is retired with an error and its slot freed.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def _plant_nan(g, node, q=1):
    f = g.f_src_data().copy()
    f[q, node] = np.nan
    g.set_populations(f)


@pytest.mark.parametrize("storage", [0, 1])
def test_flag_and_count_single_domain(wa, storage):
    g = wa.LatticeGrid3D(32, 16, 16, 0.6, storage=storage)
    g.initialize(1.0, 0.03, 0.0, 0.0)
    g.run(5, "MRT")
    assert not g.nonfinite_seen()
    assert g.count_nonfinite() == (0, -1)
    g.check_finite()
    node = (8 * 16 + 5) * 32 + 7
    _plant_nan(g, node)
    assert g.count_nonfinite()[0] >= 1  # the planted state itself
    g.step("MRT")
    assert g.nonfinite_seen()
    n, first = g.count_nonfinite()
    assert n > 1 and 0 <= first <= node
    with pytest.raises(FloatingPointError):
        g.check_finite()
    g.initialize(1.0, 0.03, 0.0, 0.0)
    assert not g.nonfinite_seen() and g.count_nonfinite() == (0, -1)


def test_flag_in_a_batch_marks_only_the_diverged_member(wa):
    a = wa.LatticeGrid2D(64, 32, 0.7)
    b = wa.LatticeGrid2D(96, 32, 0.7)
    for g in (a, b):
        g.initialize(1.0, 0.04, 0.0)
    _plant_nan(b, 10 * 96 + 20, q=3)
    batch = wa.Batch(lattice=9, collision="BGK", chunk=10)
    batch.add(a, 30)
    batch.add(b, 30)
    batch.run(30)
    assert not a.nonfinite_seen() and a.count_nonfinite()[0] == 0
    assert b.nonfinite_seen() and b.count_nonfinite()[0] > 0


def test_serving_loop_retires_a_diverged_request(wa, monkeypatch):
    from wake_amd import serve

    reqs = serve.request_stream(4, rate=1000.0, seed=3, step_scale=0.02)
    bad = reqs[1].rid
    made = serve.make_domain

    def make(req, storage=None):
        g = made(req, storage)
        if req.rid == bad:
            _plant_nan(g, g.total_nodes() // 2)
        return g

    monkeypatch.setattr(serve, "make_domain", make)
    s = serve.Server(budget_bytes=50e9, chunk=20)
    s.serve(reqs, realtime=False)
    assert s.diverged == [bad]
    res = {r.rid: r.result for r in s.done}
    assert set(res) == {r.rid for r in reqs}
    assert "error" in res[bad] and "diverged" in res[bad]["error"]
    for r in reqs:
        if r.rid != bad:
            assert "error" not in res[r.rid]
            assert all(math.isfinite(v) for v in res[r.rid].values())
    assert s.used == 0 and not s.active
