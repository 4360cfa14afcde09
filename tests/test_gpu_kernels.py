"""GPU parity of the reference's kernel exports (include/wake_hip.h part 1)
against the oracle on the same seeded inputs.

Streaming is a permutation with copies, so it must be bit-exact; the
arithmetic kernels agree to FP32 rounding (the GPU may fuse multiply-adds),
checked with a relative tolerance of 2e-6 on each output array.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from oracle_bind import D2Q9_C, D2Q9_W, D3Q19_C, D3Q19_W, call

pytestmark = pytest.mark.gpu

W27 = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
RTOL = 2e-6


@pytest.fixture(scope="module")
def K():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd.kernels


def close(a, b, rtol=RTOL):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.max(np.abs(b)), 1e-30)
    err = np.max(np.abs(a - b)) / scale
    assert err <= rtol, f"max rel err {err:.3e} > {rtol}"


def pops(W, count, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    return (W[:, None] * (1 + amp * rng.standard_normal((len(W), count)))).astype(np.float32)


def cells(count, seed, frac=0.1, moving=0.0):
    rng = np.random.default_rng(seed)
    ct = np.zeros(count, np.uint8)
    r = rng.random(count)
    ct[r < frac] = 1
    ct[(r >= frac) & (r < frac + moving)] = 6
    return ct


# ------------------------------------------------------------------ moments
@pytest.mark.parametrize("lat", [9, 19, 27])
def test_moments(K, lat):
    W = {9: D2Q9_W, 19: D3Q19_W, 27: W27}[lat]
    n = 1000
    f = pops(W, n, 1).ravel()
    outs = [[np.zeros(n, np.float32) for _ in range(4)] for _ in range(2)]
    if lat == 9:
        K.kernel("d2q9_compute_macroscopic_moments", f, *outs[0][:3], n)
        call(32, "d2q9_compute_macroscopic_moments", f, *outs[1][:3], n)
        K.kernel("d2q9_compute_macroscopic_moments_forced", f, *outs[0][:3], 1e-4, -2e-4, n)
        call(32, "d2q9_compute_macroscopic_moments_forced", f, *outs[1][:3], 1e-4, -2e-4, n)
        sp = [np.zeros(n, np.float32) for _ in range(2)]
        K.kernel("d2q9_compute_velocity_magnitude", outs[0][1], outs[0][2], sp[0], n)
        call(32, "d2q9_compute_velocity_magnitude", outs[1][1], outs[1][2], sp[1], n)
        close(sp[0], sp[1])
    elif lat == 19:
        K.kernel("d3q19_compute_macroscopic_moments_forced", f, *outs[0], 1e-4, -2e-4, 3e-4, n)
        call(32, "d3q19_compute_macroscopic_moments_forced", f, *outs[1], 1e-4, -2e-4, 3e-4, n)
        sp = [np.zeros(n, np.float32) for _ in range(2)]
        K.kernel("d3q19_compute_velocity_magnitude", *outs[0][1:], sp[0], n)
        call(32, "d3q19_compute_velocity_magnitude", *outs[1][1:], sp[1], n)
        close(sp[0], sp[1])
    else:
        K.kernel("d3q27_compute_macroscopic_moments", f, *outs[0], n)
        call(32, "d3q27_compute_macroscopic_moments", f, *outs[1], n)
    for a, b in zip(outs[0], outs[1]):
        close(a, b)


def test_moments_reference_known_answers(K):
    # tests/unit/test_d2q9_moments.cpp:28-61 and test_d3q19_moments.cpp:27-64
    n = 256
    fe = ob.feq_np(D2Q9_C, D2Q9_W, 1.05, (0.08, -0.04))
    f = np.repeat(fe[:, None], n, axis=1).astype(np.float32).ravel()
    rho, ux, uy, sp = (np.zeros(n, np.float32) for _ in range(4))
    K.kernel("d2q9_compute_macroscopic_moments", f, rho, ux, uy, n)
    K.kernel("d2q9_compute_velocity_magnitude", ux, uy, sp, n)
    assert np.all(np.abs(rho - 1.05) <= 1e-5) and np.all(np.abs(ux - 0.08) <= 1e-5)
    assert np.all(np.abs(uy + 0.04) <= 1e-5) and np.all(np.abs(sp - math.hypot(0.08, 0.04)) <= 1e-5)
    n = 512
    fe = ob.feq_np(D3Q19_C, D3Q19_W, 1.02, (0.05, -0.03, 0.04))
    f = np.repeat(fe[:, None], n, axis=1).astype(np.float32).ravel()
    rho, ux, uy, uz = (np.zeros(n, np.float32) for _ in range(4))
    K.kernel("d3q19_compute_macroscopic_moments", f, rho, ux, uy, uz, n)
    for a, v in ((rho, 1.02), (ux, 0.05), (uy, -0.03), (uz, 0.04)):
        assert np.all(np.abs(a - v) <= 1e-5)


# ---------------------------------------------------------------- collision
COLLIDE = [
    ("d2q9_collide_bgk", 9, (0.7, 1e-4, -2e-4)),
    ("d2q9_collide_trt", 9, (0.7, 1.2, 1e-4, -2e-4)),
    ("d2q9_collide_mrt", 9, (0.7, 1.63, 1e-4, -2e-4)),
    ("d3q19_collide_bgk", 19, (0.75, 1e-4, -2e-4, 3e-4)),
    ("d3q19_collide_trt", 19, (0.75, 1.3, 1e-4, -2e-4, 3e-4)),
    ("d3q19_collide_mrt", 19, (0.75, 1.19, 1e-4, -2e-4, 3e-4)),
    ("d3q27_collide", 27, (0.8, 1.1, 1e-4, -2e-4, 3e-4)),
]


@pytest.mark.parametrize("name,lat,params", COLLIDE, ids=[c[0] for c in COLLIDE])
def test_collide(K, name, lat, params):
    W = {9: D2Q9_W, 19: D3Q19_W, 27: W27}[lat]
    n = 2048
    f = pops(W, n, 2).ravel()
    ct = cells(n, 3)
    nm = 3 if lat == 9 else 4
    res = []
    for side in (0, 1):
        out = np.zeros_like(f)
        m = [np.full(n, -7.0, np.float32) for _ in range(nm)]
        if side == 0:
            K.kernel(name, f, out, ct, *m, *params, n)
        else:
            call(32, name, f, out, ct, *m, *params, n)
        res.append((out, m))
    close(res[0][0], res[1][0])
    for a, b in zip(res[0][1], res[1][1]):
        close(a, b)
    solid = ct == 1
    assert np.array_equal(res[0][0].reshape(lat, n)[:, solid], f.reshape(lat, n)[:, solid])


def test_collision_reference_conservation(K):
    # tests/unit/test_d2q9_collision.cpp and test_d3q19_collision.cpp, on the GPU
    count = 32
    f = np.zeros((9, count), np.float32)
    for q in range(9):
        for i in range(count):
            f[q, i] = np.float32(D2Q9_W[q]) + np.float32(0.01) * np.float32(math.sin(q + i))
    ct = np.zeros(count, np.uint8)
    ct[-1] = 1
    fl = f.ravel()
    r0 = f[:, 0].sum()
    for name, args in (("d2q9_collide_bgk", (0.7, 0.0, 0.0)), ("d2q9_collide_trt", (0.7, 1.2, 0.0, 0.0)),
                       ("d2q9_collide_mrt", (0.7, 1.6, 0.0, 0.0))):
        out = np.zeros_like(fl)
        m = [np.zeros(count, np.float32) for _ in range(3)]
        K.kernel(name, fl, out, ct, *m, *args, count)
        o = out.reshape(9, count)
        assert abs(o[:, 0].sum() - r0) <= 1e-5
        assert abs(D2Q9_C[:, 0] @ o[:, 0] - D2Q9_C[:, 0] @ f[:, 0]) <= 1e-5
        assert np.all(np.abs(o[:, -1] - f[:, -1]) <= 1e-6)


# ---------------------------------------------------------------- streaming
@pytest.mark.parametrize("lat", [9, 19, 27])
def test_stream_periodic_bit_exact(K, lat):
    nx, ny, nz = (33, 17, 1) if lat == 9 else (20, 9, 7)
    N = nx * ny * nz
    f = np.random.default_rng(4).random(lat * N).astype(np.float32)
    a, b = np.zeros_like(f), np.zeros_like(f)
    if lat == 9:
        K.kernel("d2q9_stream_periodic", f, a, nx, ny)
        call(32, "d2q9_stream_periodic", f, b, nx, ny)
    else:
        K.kernel(f"d3q{lat}_stream_periodic", f, a, nx, ny, nz)
        call(32, f"d3q{lat}_stream_periodic", f, b, nx, ny, nz)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("lat", [9, 19])
def test_stream_with_boundaries(K, lat):
    nx, ny, nz = (40, 23, 1) if lat == 9 else (24, 13, 11)
    N = nx * ny * nz
    f = np.random.default_rng(5).random(lat * N).astype(np.float32)
    ct = cells(N, 6, 0.1, 0.05)
    a, b = np.zeros_like(f), np.zeros_like(f)
    if lat == 9:
        K.kernel("d2q9_stream_with_boundaries", f, a, ct, nx, ny, 0.05, 0.0)
        call(32, "d2q9_stream_with_boundaries", f, b, ct, nx, ny, 0.05, 0.0)
    else:
        K.kernel("d3q19_stream_with_boundaries", f, a, ct, nx, ny, nz, 0.05, 0.0, 0.02)
        call(32, "d3q19_stream_with_boundaries", f, b, ct, nx, ny, nz, 0.05, 0.0, 0.02)
    assert np.max(np.abs(a - b)) <= 1e-7


# --------------------------------------------------------------- boundaries
def test_bounce_back_and_zou_he(K):
    nx, ny = 30, 21
    N = nx * ny
    f = pops(D2Q9_W, N, 7).ravel()
    ct = cells(N, 8, 0.2, 0.1)
    a, b = f.copy(), f.copy()
    K.kernel("d2q9_apply_bounce_back", a, ct, 0.03, -0.01, N)
    call(32, "d2q9_apply_bounce_back", b, ct, 0.03, -0.01, N)
    close(a, b)
    for name, args in (("d2q9_apply_zou_he_inflow_west", (0.05, 0.01)), ("d2q9_apply_zou_he_outflow_east", (1.01,))):
        a, b = f.copy(), f.copy()
        m = [[np.zeros(N, np.float32) for _ in range(3)] for _ in range(2)]
        K.kernel(name, a, *m[0], *args, nx, ny)
        call(32, name, b, *m[1], *args, nx, ny)
        close(a, b)
        for x, y in zip(m[0], m[1]):
            close(x, y)
    prev = pops(D2Q9_W, N, 9).ravel()
    a, b = f.copy(), f.copy()
    K.kernel("d2q9_apply_convective_outflow_east", a, prev, 0.07, nx, ny)
    call(32, "d2q9_apply_convective_outflow_east", b, prev, 0.07, nx, ny)
    close(a, b)


@pytest.mark.parametrize("lat", [9, 19])
def test_smagorinsky_tau(K, lat):
    W = D2Q9_W if lat == 9 else D3Q19_W
    n = 1500
    f = pops(W, n, 10, 0.1).ravel()
    rng = np.random.default_rng(11)
    rho = (1 + 0.01 * rng.standard_normal(n)).astype(np.float32)
    u = [(0.03 * rng.standard_normal(n)).astype(np.float32) for _ in range(3)]
    a, b = np.zeros(n, np.float32), np.zeros(n, np.float32)
    if lat == 9:
        K.kernel("d2q9_compute_smagorinsky_tau", f, rho, u[0], u[1], a, 0.505, 0.17, n)
        call(32, "d2q9_compute_smagorinsky_tau", f, rho, u[0], u[1], b, 0.505, 0.17, n)
    else:
        K.kernel("d3q19_compute_smagorinsky_tau", f, rho, *u, a, 0.505, 0.17, n)
        call(32, "d3q19_compute_smagorinsky_tau", f, rho, *u, b, 0.505, 0.17, n)
    close(a, b)


def test_curved_bfl(K):
    nx, ny, nz = 24, 20, 18
    g = ob.OracleGrid(3, nx, ny, nz, tau=0.8)
    g.set_solid_sphere(11.3, 9.7, 8.9, 4.4)
    n = g.build_curved_sphere_links(11.3, 9.7, 8.9, 4.4)
    idx, q, d = g.links()
    assert n > 100
    N = nx * ny * nz
    A = pops(D3Q19_W, N, 12).ravel()
    S = pops(D3Q19_W, N, 13).ravel()
    rng = np.random.default_rng(14)
    uw = [(0.01 * rng.standard_normal(n)).astype(np.float32) for _ in range(3)]
    a, b = S.copy(), S.copy()
    K.kernel("d3q19_apply_curved_bfl", a, A, idx, q, d, *uw, nx, ny, nz, n)
    call(32, "d3q19_apply_curved_bfl", b, A, idx, q, d, *uw, nx, ny, nz, n)
    close(a, b)


def test_momentum_exchange_kernels(K):
    import ctypes

    # reference known answers (tests/unit/test_forces_momentum_exchange.cpp:27-68)
    NX = NY = 16
    N = NX * NY
    f = np.tile(D2Q9_W.astype(np.float32)[:, None], (1, N)).ravel().copy()
    ct = np.zeros(N, np.uint8)
    for y in (7, 8):
        for x in (7, 8):
            ct[y * NX + x] = 1
    o = [np.zeros(1, np.float32) for _ in range(3)]
    K.kernel("d2q9_compute_momentum_exchange_forces", f, ct, NX, NY, 7.5, 7.5, 0.0, *o)
    assert abs(o[0][0]) <= 1e-5 and abs(o[1][0]) <= 1e-5
    f[1 * N + 7 * NX + 6] += np.float32(0.05)
    K.kernel("d2q9_compute_momentum_exchange_forces", f, ct, NX, NY, 7.5, 7.5, 0.0, *o)
    assert abs(o[0][0] - 0.1) <= 1e-5
    # random 3D field vs oracle (FP64 sums on both sides)
    nx, ny, nz = 30, 22, 20
    N = nx * ny * nz
    g = ob.OracleGrid(3, nx, ny, nz, tau=0.8)
    g.set_solid_sphere(14.0, 10.5, 9.5, 5.0)
    ct = g.cell_type()
    f = pops(D3Q19_W, N, 15).ravel()
    a = [np.zeros(1, np.float32) for _ in range(3)]
    b = [ctypes.c_float() for _ in range(3)]
    K.kernel("d3q19_compute_momentum_exchange_forces", f, ct, nx, ny, nz, 14.0, 10.5, 9.5, 10.0, *a)
    ob.lib(32).oracle_d3q19_compute_momentum_exchange_forces(
        ob.ptr(f), ob.ptr(ct), nx, ny, nz, ctypes.c_float(14.0), ctypes.c_float(10.5), ctypes.c_float(9.5),
        ctypes.c_float(10.0), *[ctypes.byref(v) for v in b])
    for x, y in zip(a, b):
        assert abs(float(x[0]) - y.value) <= 1e-5 * max(1.0, abs(y.value))


# ---------------------------------------------------------------- analytics
def test_vorticity_and_q_criterion_exports(K):
    """analytics_kernels.ispc:4-110 on random fields vs the oracle."""
    rng = np.random.default_rng(21)
    nx, ny, nz = 37, 21, 13
    u2 = [(0.05 * rng.standard_normal(nx * ny)).astype(np.float32) for _ in range(2)]
    w, wo = np.zeros(nx * ny, np.float32), np.zeros(nx * ny, np.float32)
    K.kernel("compute_vorticity_2d", *u2, w, nx, ny)
    call(32, "compute_vorticity_2d", *u2, wo, nx, ny)
    close(w, wo)
    n = nx * ny * nz
    u3 = [(0.05 * rng.standard_normal(n)).astype(np.float32) for _ in range(3)]
    g = [np.zeros(n, np.float32) for _ in range(4)]
    o = [np.zeros(n, np.float32) for _ in range(4)]
    K.kernel("compute_vorticity_and_q_criterion_3d", *u3, *g, nx, ny, nz)
    call(32, "compute_vorticity_and_q_criterion_3d", *u3, *o, nx, ny, nz)
    for a, b in zip(g, o):
        close(a, b, 4e-6)


def test_domain_vorticity_on_device():
    """The grid's device-side vorticity / Q of its own velocity field equal
    the oracle's analytics applied to the same (read-back) field."""
    import wake_amd

    g = wake_amd.LatticeGrid3D(40, 24, 20, 0.6)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.04, 0.0, 0.0)
    g.set_solid_sphere(12.0, 12.0, 10.0, 4.0)
    g.run(30, "MRT")
    ox, oy, oz, q = g.vorticity_q_data()
    u = [np.ascontiguousarray(a) for a in (g.ux_data(), g.uy_data(), g.uz_data())]
    o = [np.zeros_like(u[0]) for _ in range(4)]
    call(32, "compute_vorticity_and_q_criterion_3d", *u, *o, 40, 24, 20)
    for a, b in zip((ox, oy, oz, q), o):
        close(a, b, 4e-6)
    g2 = wake_amd.LatticeGrid2D(64, 32, 0.6)
    g2.initialize(1.0, 0.05, 0.0)
    g2.set_solid_circle(16.0, 16.0, 4.0)
    g2.run(40, "MRT")
    w = g2.vorticity_data()
    wo = np.zeros_like(w)
    call(32, "compute_vorticity_2d", np.ascontiguousarray(g2.ux_data()), np.ascontiguousarray(g2.uy_data()), wo, 64, 32)
    close(w, wo)
