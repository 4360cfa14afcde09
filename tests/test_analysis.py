"""CPU tests of the host-side force analysis (wake_amd.analysis) against the
reference's known answers (tests/unit/test_forces_momentum_exchange.cpp:70-100)
and an independent restatement of its radix-2 transform
(src/forces/strouhal_analyzer.cpp:8-37)."""
import cmath
import math

import numpy as np
import pytest

from wake_amd import analysis as an


def _radix2_fft_c64(x):
    """strouhal_analyzer.cpp:8-37 in complex64: bit reversal, then butterflies
    whose twiddle is advanced by repeated multiplication."""
    a = [np.complex64(v) for v in x]
    n = len(a)
    j = 0
    for i in range(1, n):
        bit = n >> 1
        while j & bit:
            j ^= bit
            bit >>= 1
        j ^= bit
        if i < j:
            a[i], a[j] = a[j], a[i]
    length = 2
    while length <= n:
        ang = np.float32(-2 * math.pi) / np.float32(length)
        wlen = np.complex64(complex(math.cos(ang), math.sin(ang)))
        for i in range(0, n, length):
            w = np.complex64(1)
            for k in range(length // 2):
                u = a[i + k]
                v = np.complex64(a[i + k + length // 2] * w)
                a[i + k] = np.complex64(u + v)
                a[i + k + length // 2] = np.complex64(u - v)
                w = np.complex64(w * wlen)
        length <<= 1
    return a


def _reference_peak_bin(sig):
    sig = np.asarray(sig, np.float32)
    mean = np.float32(sig.sum(dtype=np.float32) / np.float32(sig.size))
    n = 1
    while n < sig.size:
        n <<= 1
    n = min(n, 4096)
    m = min(n, sig.size)
    x = np.zeros(n, np.float32)
    for i in range(m):
        w = np.float32(0.5) * (np.float32(1) - np.float32(math.cos(np.float32(2 * math.pi) * np.float32(i) / np.float32(m - 1))))
        x[i] = (sig[i] - mean) * w
    X = _radix2_fft_c64(x)
    best, pk = np.float32(0), 1
    for k in range(1, n // 2):
        p = np.float32(X[k].real * X[k].real + X[k].imag * X[k].imag)
        if p > best:
            best, pk = p, k
    return pk, n


def test_reference_known_answer_sinusoid():
    f_true, dt = 0.025, 1.0
    t = np.arange(512, dtype=np.float32)
    sig = (np.float32(0.35) * np.sin(np.float32(2 * np.pi * f_true) * t)
           + np.float32(0.02) * np.cos(np.float32(2 * np.pi * 0.1) * t)).astype(np.float32)
    f = an.analyze_frequency(sig, dt)
    assert abs(f - f_true) <= 0.003  # the reference's bound
    assert f == pytest.approx(13 / 512, rel=1e-6)  # the bin nearest 0.025 * 512 = 12.8
    st = an.strouhal_number(f, 20.0, 0.1)
    assert abs(st - 5.0) <= 0.6


@pytest.mark.parametrize("seed", range(6))
def test_peak_bin_matches_radix2_restatement(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(40, 200))
    t = np.arange(n)
    sig = (np.sin(2 * np.pi * rng.uniform(0.02, 0.3) * t) + 0.3 * rng.standard_normal(n)).astype(np.float32)
    k, nfft = _reference_peak_bin(sig)
    assert an.analyze_frequency(sig, 0.5) == pytest.approx(k / (nfft * 0.5), rel=1e-6)


def test_only_first_4096_samples_are_analysed():
    t = np.arange(6000, dtype=np.float64)
    sig = np.where(t < 4096, np.sin(2 * np.pi * 0.05 * t), np.sin(2 * np.pi * 0.2 * t) * 10).astype(np.float32)
    assert an.analyze_frequency(sig, 1.0) == pytest.approx(round(0.05 * 4096) / 4096, rel=1e-6)


def test_short_or_bad_input_gives_zero():
    assert an.analyze_frequency(np.ones(15, np.float32), 1.0) == 0.0
    assert an.analyze_frequency(np.ones(64, np.float32), 0.0) == 0.0
    assert an.strouhal_number(0.1, 10.0, 0.0) == 0.0


def test_coefficients_closed_form():
    c = an.aerodynamic_coefficients(0.02, -0.01, 0.0, 0.5, 1.0, 0.05, 8.0, 2.0)
    q = 0.5 * 0.05 * 0.05
    assert float(c["cd"]) == pytest.approx(0.02 / (q * 8.0), rel=1e-6)
    assert float(c["cl"]) == pytest.approx(-0.01 / (q * 8.0), rel=1e-6)
    assert float(c["cm"]) == pytest.approx(0.5 / (q * 8.0 * 2.0), rel=1e-6)
    z = an.aerodynamic_coefficients(1.0, 1.0, 0.0, 0.0, 1.0, 1e-5, 1.0)
    assert float(z["cd"]) == 0.0 and float(z["cl"]) == 0.0


def test_serve_results_jsonl(tmp_path):
    """Per-request result records of the serving loop (CPU: no domain)."""
    import json

    from wake_amd import serve

    a = serve.SimRequest(3, "cavity2d", 64, 100, ra=1e4, tol=1e-3, arrival=0.25)
    a.admitted, a.finished, a.result = 0.5, 1.75, {"Nu": 2.1}
    b = serve.SimRequest(1, "bluff3d", 64, 200, re=300.0, semi_axes=(9.0, 6.0, 5.0), tol=1e-5, arrival=0.0)
    b.admitted, b.finished, b.result = 0.0, 2.0, {"C_D": 0.9, "C_L": 0.01}
    p = tmp_path / "results.jsonl"
    serve.write_results_jsonl([a, b], p)
    rows = [json.loads(line) for line in p.read_text().splitlines()]
    assert [r["rid"] for r in rows] == [1, 3]
    assert rows[0]["storage"] == "f32" and rows[1]["storage"] == "f16"
    assert rows[1]["time_to_result_s"] == 1.5 and rows[1]["result"] == {"Nu": 2.1}
    assert rows[0]["semi_axes"] == [9.0, 6.0, 5.0] and rows[1]["ra"] == 1e4


def test_serve_policy_names():
    """The serving loop's admission policies (CPU: nothing is admitted)."""
    from wake_amd import serve

    assert serve.Server(policy="arrival").policy == "arrival"
    assert serve.Server().policy == "least_work"
    with pytest.raises(ValueError):
        serve.Server(policy="random")
    r3 = serve.SimRequest(0, "bluff3d", 64, 100)
    r2 = serve.SimRequest(1, "cavity2d", 64, 100)
    assert serve.node_count(r3) == 64 ** 3 and serve.node_count(r2) == 64 ** 2


def test_serve_rejects_requests_larger_than_the_budget():
    """A request that could never fit the HBM budget is rejected at submit
    (CPU: sizes only, nothing is allocated), so the loop cannot wait on it
    forever."""
    from wake_amd import serve

    s = serve.Server(budget_bytes=1e6)
    big = serve.SimRequest(0, "bluff3d", 128, 100)
    s.submit(big)
    assert not s.queue and s.rejected == [big] and "budget" in big.result["error"]
    assert s.serve([], realtime=False) >= 0.0
