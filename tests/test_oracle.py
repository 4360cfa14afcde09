"""Pins for the CPU oracle (oracle/wake_oracle.c).

The reference ships no stored vectors; its own tests are inline known-answer
and invariant checks (tests/unit/*.cpp, tests/validation/*.cpp).  Each test
here reproduces one of those, or checks the oracle against a closed form or
a brute-force computation that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) would fail.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle_bind as ob
from oracle_bind import D2Q9_C, D2Q9_W, D3Q19_C, D3Q19_W, OracleGrid, call, feq_np

REF = "/root/reference"


# ----------------------------------------------------------------- tables
def test_oracle_rng_splitmix64_known_values(oracle32):
    # splitmix64 reference outputs for state 0 (Vigna's published generator:
    # the first three outputs of a splitmix64 stream seeded with 0).
    L = oracle32
    state = 0
    expected = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for e in expected:
        assert L.oracle_rng_splitmix64(state) == e
        state = (state + 0x9E3779B97F4A7C15) % (1 << 64)
    u = [L.oracle_rng_u01(7, i) for i in range(20000)]
    assert 0.0 <= min(u) and max(u) < 1.0
    assert abs(np.mean(u) - 0.5) < 0.01


def _d3q19_m_closed_form():
    """d'Humieres D3Q19 basis from its polynomials in c (the reference's sign
    convention for m16..m18 is c_x(c_y^2-c_z^2), c_y(c_z^2-c_x^2),
    c_z(c_y^2-c_x^2))."""
    rows = []
    for cx, cy, cz in D3Q19_C:
        c2 = cx * cx + cy * cy + cz * cz
        rows.append([
            1, 19 * c2 - 30, (21 * c2 * c2 - 53 * c2 + 24) / 2,
            cx, (5 * c2 - 9) * cx, cy, (5 * c2 - 9) * cy, cz, (5 * c2 - 9) * cz,
            3 * cx * cx - c2, (3 * c2 - 5) * (3 * cx * cx - c2),
            cy * cy - cz * cz, (3 * c2 - 5) * (cy * cy - cz * cz),
            cx * cy, cy * cz, cx * cz,
            cx * (cy * cy - cz * cz), cy * (cz * cz - cx * cx), cz * (cy * cy - cx * cx),
        ])
    return np.array(rows, dtype=np.float64).T  # M[k][q]


def _oracle_velocities(lattice):
    """The oracle's own c table, read back through its moments kernel: node q
    holds a unit population q and nothing else, so its u is c_q exactly."""
    Q = {9: 9, 19: 19, 27: 27}[lattice]
    f = np.eye(Q, dtype=np.float64).reshape(Q, Q)  # f[q, node] = (q == node)
    rho, ux, uy, uz = (np.zeros(Q, np.float64) for _ in range(4))
    if lattice == 9:
        call(64, "d2q9_compute_macroscopic_moments", f.ravel(), rho, ux, uy, Q)
        return np.stack([ux, uy], 1)
    call(64, f"d3q{lattice}_compute_macroscopic_moments", f.ravel(), rho, ux, uy, uz, Q)
    return np.stack([ux, uy, uz], 1)


@pytest.mark.parametrize("lattice", [9, 19, 27])
def test_oracle_tables(lattice):
    """lattice_models.hpp:14-110: the velocity set (distinct, rest first;
    D3Q19 = the 19 of |c|^2 <= 2), the weights (read back as the rest
    equilibrium) with sum w = 1, sum w c = 0, sum w c c = I/3, and the
    opposite table as an involution with c_opp = -c, read back through
    halfway bounce-back (a populated node whose every non-periodic pull source
    is out of range receives its own opposite populations)."""
    C = _oracle_velocities(lattice)
    Q, D = C.shape
    assert np.array_equal(C, np.round(C)) and np.all(np.abs(C) <= 1)
    assert np.array_equal(C[0], np.zeros(D))
    assert len({tuple(c) for c in C}) == Q
    full = {tuple(v - 1 for v in c) for c in np.ndindex(*([3] * D))}
    want = {c for c in full if sum(v * v for v in c) <= 2} if lattice == 19 else full
    assert {tuple(int(v) for v in c) for c in C} == want
    g = OracleGrid(2 if lattice == 9 else 3, 2, 2, 1 if lattice == 9 else 2, tau=0.6,
                   lattice=None if lattice in (9, 19) else 27, bits=64)
    g.initialize(1.0, 0.0, 0.0, 0.0)
    w = g.f_src()[:, 0]
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.all(np.abs(w @ C) < 1e-15)
    assert np.allclose(C.T @ (w[:, None] * C), np.eye(D) / 3.0, atol=1e-15)
    # opposites: distinct values on one node of a 1 x 1 (x 1) domain with
    # y (and z) walls; x is always periodic, so only the +-x pair streams in
    A = np.arange(1, Q + 1, dtype=np.float64)
    S = np.zeros(Q, np.float64)
    ct = np.zeros(1, np.uint8)
    if lattice == 9:
        call(64, "d2q9_stream_general", A, S, ct, 1, 1, 0, 0.0, 0.0)
    else:
        call(64, "stream3d_general", lattice, A, S, ct, 1, 1, 1, 0, 0, 0.0, 0.0, 0.0)
    for q in range(Q):
        if np.any(C[q, 1:] != 0):
            o = int(S[q]) - 1
            assert np.array_equal(C[o], -C[q]), (q, o)
            assert int(S[o]) - 1 == q
        else:
            assert S[q] == A[q]


def test_oracle_mrt_forward_matches_dhumieres_basis(oracle64):
    m = np.zeros(361, np.float64)
    oracle64.oracle_d3q19_m(ob.ptr(m))
    M = m.reshape(19, 19)
    Mc = _d3q19_m_closed_form()
    assert np.array_equal(M, Mc)
    G = Mc @ Mc.T  # the basis is orthogonal
    assert np.allclose(G - np.diag(np.diag(G)), 0)


def test_oracle_mrt_inverse_is_rounded_inverse(oracle32):
    out = np.zeros(361, np.float32)
    oracle32.oracle_d3q19_minv(ob.ptr(out))
    Minv = out.reshape(19, 19)
    exact = np.linalg.inv(_d3q19_m_closed_form())
    expect = (np.round(exact * 1e7) / 1e7).astype(np.float32)
    expect[expect == 0] = 0.0
    assert np.array_equal(Minv.view(np.uint32), expect.view(np.uint32))
    assert np.count_nonzero(Minv) == 213


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not mounted")
def test_oracle_mrt_inverse_equals_reference_literals(oracle32):
    # The reference's table is data in src/ispc/d3q19_collision_mrt.ispc:4-24;
    # compare the regenerated table with it bit for bit (CPU only).
    txt = open(os.path.join(REF, "src/ispc/d3q19_collision_mrt.ispc")).read()
    body = txt[txt.index("d3q19_Minv"):txt.index("};")]
    lit = np.array([float(v) for v in re.findall(r"(-?\d+\.\d+)f", body)], dtype=np.float32)
    out = np.zeros(361, np.float32)
    oracle32.oracle_d3q19_minv(ob.ptr(out))
    assert np.array_equal(lit.view(np.uint32), out.view(np.uint32))


def test_oracle_mrt_mass_drift_signature():
    # At rest (rho = 1, u = 0) one MRT collision changes the node mass by
    # sum_k colsum(Minv)_k * meq_k - 1 with meq = (1, -11, 3, 0, ...) because
    # the table is rounded: 19*r(1/19) - 11*colsum_1 + 3*colsum_2 - 1 = -1.1e-6.
    exact = np.linalg.inv(_d3q19_m_closed_form())
    R7 = np.round(exact * 1e7) / 1e7
    drift = R7[:, 0].sum() * 1 + R7[:, 1].sum() * -11 + R7[:, 2].sum() * 3 - 1.0
    assert abs(drift - (-1.1e-6)) < 0.05e-6
    g = OracleGrid(3, 4, 4, 4, tau=0.8, bits=64)
    g.set_periodic(True)
    m0 = g.f_src().sum()
    g.step("MRT")
    m1 = g.f_src().sum()
    assert abs((m1 - m0) / g.N - drift) < 1e-12


# ---------------------------------------------------------------- moments
def test_oracle_moments_d2q9_analytic_feq(oracle32):
    # tests/unit/test_d2q9_moments.cpp:28-61
    n = 16 * 16
    fe = feq_np(D2Q9_C, D2Q9_W, 1.05, (0.08, -0.04))
    f = np.repeat(fe[:, None], n, axis=1).astype(np.float32).ravel()
    rho, ux, uy, sp = (np.zeros(n, np.float32) for _ in range(4))
    call(32, "d2q9_compute_macroscopic_moments", f, rho, ux, uy, n)
    call(32, "d2q9_compute_velocity_magnitude", ux, uy, sp, n)
    assert np.all(np.abs(rho - 1.05) <= 1e-5)
    assert np.all(np.abs(ux - 0.08) <= 1e-5)
    assert np.all(np.abs(uy + 0.04) <= 1e-5)
    assert np.all(np.abs(sp - math.hypot(0.08, 0.04)) <= 1e-5)
    # forced: u = (j + rho g / 2) / rho
    call(32, "d2q9_compute_macroscopic_moments_forced", f, rho, ux, uy, 1e-3, -2e-3, n)
    assert np.all(np.abs(ux - (0.08 + 0.5e-3)) <= 1e-6)
    assert np.all(np.abs(uy - (-0.04 - 1e-3)) <= 1e-6)


def test_oracle_moments_d3q19_analytic_feq(oracle32):
    # tests/unit/test_d3q19_moments.cpp:27-64
    n = 8 * 8 * 8
    fe = feq_np(D3Q19_C, D3Q19_W, 1.02, (0.05, -0.03, 0.04))
    f = np.repeat(fe[:, None], n, axis=1).astype(np.float32).ravel()
    rho, ux, uy, uz, sp = (np.zeros(n, np.float32) for _ in range(5))
    call(32, "d3q19_compute_macroscopic_moments", f, rho, ux, uy, uz, n)
    call(32, "d3q19_compute_velocity_magnitude", ux, uy, uz, sp, n)
    for a, v in ((rho, 1.02), (ux, 0.05), (uy, -0.03), (uz, 0.04), (sp, math.sqrt(0.05**2 + 0.03**2 + 0.04**2))):
        assert np.all(np.abs(a - v) <= 1e-5)
    # integer-valued populations: the sums are exact
    fi = np.arange(19 * n, dtype=np.float32) % 7
    call(32, "d3q19_compute_macroscopic_moments", fi, rho, ux, uy, uz, n)
    F = fi.reshape(19, n).astype(np.float64)
    assert np.array_equal(rho, F.sum(0).astype(np.float32))
    assert np.allclose(ux, (D3Q19_C[:, 0] @ F) / F.sum(0), rtol=1e-6)


# -------------------------------------------------------------- collision
def _perturbed(Q, W, count, amp, mul):
    f = np.zeros((Q, count), np.float32)
    for q in range(Q):
        for i in range(count):
            f[q, i] = np.float32(W[q]) * np.float32(1.0) + np.float32(amp) * np.float32(math.sin(q * mul + i))
    return f


def _mom(f, C):
    return f.sum(0), (C.T.astype(np.float64) @ f.astype(np.float64))


def test_oracle_collision_d2q9_conservation_and_solid():
    # tests/unit/test_d2q9_collision.cpp:23-134
    count = 32
    f = _perturbed(9, D2Q9_W, count, 0.01, 1)
    ct = np.zeros(count, np.uint8)
    ct[-1] = 1
    out = np.zeros_like(f)
    rho, ux, uy = (np.zeros(count, np.float32) for _ in range(3))
    r0, j0 = _mom(f, D2Q9_C)
    for name, args in (("d2q9_collide_bgk", (0.7, 0.0, 0.0)), ("d2q9_collide_trt", (0.7, 1.2, 0.0, 0.0)),
                       ("d2q9_collide_mrt", (0.7, 1.6, 0.0, 0.0))):
        call(32, name, f, out, ct, rho, ux, uy, *args, count)
        r1, j1 = _mom(out, D2Q9_C)
        assert abs(r1[0] - r0[0]) <= 1e-5 and np.all(np.abs(j1[:, 0] - j0[:, 0]) <= 1e-5), name
        assert np.all(np.abs(out[:, -1] - f[:, -1]) <= 1e-6), name
        assert np.all(np.abs(r1[:-1] - r0[:-1]) <= 1e-5)


def test_oracle_collision_d3q19_conservation():
    # tests/unit/test_d3q19_collision.cpp:25-116
    count = 16
    f = _perturbed(19, D3Q19_W, count, 0.005, 3)
    ct = np.zeros(count, np.uint8)
    ct[-1] = 1
    out = np.zeros_like(f)
    rho, ux, uy, uz = (np.zeros(count, np.float32) for _ in range(4))
    r0, j0 = _mom(f, D3Q19_C)
    call(32, "d3q19_collide_bgk", f, out, ct, rho, ux, uy, uz, 0.75, 0.0, 0.0, 0.0, count)
    r1, j1 = _mom(out, D3Q19_C)
    assert abs(r1[0] - r0[0]) <= 1e-5 and np.all(np.abs(j1[:, 0] - j0[:, 0]) <= 1e-5)
    call(32, "d3q19_collide_trt", f, out, ct, rho, ux, uy, uz, 0.75, 1.3, 0.0, 0.0, 0.0, count)
    r1, j1 = _mom(out, D3Q19_C)
    assert abs(r1[0] - r0[0]) <= 1e-5 and abs(j1[0, 0] - j0[0, 0]) <= 1e-5
    call(32, "d3q19_collide_mrt", f, out, ct, rho, ux, uy, uz, 0.75, 1.2, 0.0, 0.0, 0.0, count)
    r1, j1 = _mom(out, D3Q19_C)
    assert abs(r1[0] - r0[0]) <= 1e-5 and abs(j1[0, 0] - j0[0, 0]) <= 1e-5
    assert np.array_equal(out[:, -1], f[:, -1])


def test_oracle_bgk_contracts_nonequilibrium_exactly():
    # with g = 0, f' - feq = (1 - omega)(f - feq) node by node
    count = 64
    rng = np.random.default_rng(3)
    f = (D3Q19_W[:, None] * (1 + 0.05 * rng.standard_normal((19, count)))).astype(np.float64)
    out = np.zeros_like(f)
    ct = np.zeros(count, np.uint8)
    rho, ux, uy, uz = (np.zeros(count, np.float64) for _ in range(4))
    tau = 0.83
    call(64, "d3q19_collide_bgk", f, out, ct, rho, ux, uy, uz, tau, 0.0, 0.0, 0.0, count)
    r = f.sum(0)
    u = (D3Q19_C.T @ f) / r
    cu = D3Q19_C @ u
    usq = (u * u).sum(0)
    feq = D3Q19_W[:, None] * r * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    assert np.allclose(out - feq, (1 - 1 / tau) * (f - feq), atol=1e-14)


def test_oracle_trt_equal_rates_is_bgk():
    count = 40
    rng = np.random.default_rng(5)
    for Q, W, C, pre in ((19, D3Q19_W, D3Q19_C, "d3q19"),):
        f = (W[:, None] * (1 + 0.05 * rng.standard_normal((Q, count))))
        out_b, out_t = np.zeros_like(f), np.zeros_like(f)
        ct = np.zeros(count, np.uint8)
        bufs = [np.zeros(count) for _ in range(4)]
        call(64, pre + "_collide_bgk", f, out_b, ct, *bufs, 0.9, 1e-4, -2e-4, 3e-4, count)
        call(64, pre + "_collide_trt", f, out_t, ct, *bufs, 0.9, 0.9, 1e-4, -2e-4, 3e-4, count)
        assert np.allclose(out_b, out_t, atol=1e-15)


def test_oracle_d3q27_reference_pairs_vs_geometric():
    # The reference's D3Q27 opposite table is not c_opp = -c for the corners;
    # its d3q27_collide export therefore does not conserve z momentum while
    # the geometric pairing used by the grid does.
    count = 8
    rng = np.random.default_rng(11)
    w27 = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
    c27 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                    [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1],
                    [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1], [1, -1, -1],
                    [-1, -1, -1]])
    f = w27[:, None] * (1 + 0.05 * rng.standard_normal((27, count)))
    out = np.zeros_like(f)
    ct = np.zeros(count, np.uint8)
    bufs = [np.zeros(count) for _ in range(4)]
    call(64, "d3q27_collide", f, out, ct, *bufs, 0.8, 1.1, 0.0, 0.0, 0.0, count)
    assert np.allclose(out.sum(0), f.sum(0), atol=1e-14)  # mass conserved
    jz0, jz1 = c27[:, 2] @ f, c27[:, 2] @ out
    assert np.max(np.abs(jz1 - jz0)) > 1e-6  # the reference pairing leaks z momentum
    # geometric pairing (grid path, per-node tau variant): conserves all momentum
    tau = np.full(count, 0.8)
    call(64, "collide_tau_node", 27, 1, f, out, ct, tau, 0.0, 0.0, 0.0, count)
    assert np.allclose(c27.T @ out, c27.T @ f, atol=1e-14)


# -------------------------------------------------------------- streaming
def test_oracle_streaming_periodic_is_a_shift():
    nx, ny, nz = 5, 4, 3
    N = nx * ny * nz
    rng = np.random.default_rng(1)
    src = rng.random((19, N))
    dst = np.zeros_like(src)
    call(64, "d3q19_stream_periodic", src, dst, nx, ny, nz)
    S = src.reshape(19, nz, ny, nx)
    D = dst.reshape(19, nz, ny, nx)
    for q, (cx, cy, cz) in enumerate(D3Q19_C):
        assert np.array_equal(D[q], np.roll(S[q], shift=(cz, cy, cx), axis=(0, 1, 2)))
    # all-fluid general streaming with every axis periodic equals the periodic kernel
    ct = np.zeros(N, np.uint8)
    dst2 = np.zeros_like(src)
    call(64, "stream3d_general", 19, src, dst2, ct, nx, ny, nz, 1, 1, 0.0, 0.0, 0.0)
    assert np.array_equal(dst, dst2)
    # D2Q9
    src2 = rng.random((9, nx * ny))
    d2 = np.zeros_like(src2)
    call(64, "d2q9_stream_periodic", src2, d2, nx, ny)
    for q, (cx, cy) in enumerate(D2Q9_C):
        assert np.array_equal(d2.reshape(9, ny, nx)[q], np.roll(src2.reshape(9, ny, nx)[q], (cy, cx), (0, 1)))


def test_oracle_streaming_walls_conserve_mass_in_closed_box():
    # y/z out of range and solid neighbours reflect: with x periodic the box
    # is closed, so every plane-summed population total is conserved.
    nx, ny, nz = 6, 5, 4
    N = nx * ny * nz
    rng = np.random.default_rng(2)
    ct = np.zeros(N, np.uint8)
    ct[rng.random(N) < 0.15] = 1
    src = rng.random((19, N))
    src[:, ct == 1] = 0.0  # solids carry nothing into the sum
    dst = np.zeros_like(src)
    call(64, "d3q19_stream_with_boundaries", src, dst, ct, nx, ny, nz, 0.0, 0.0, 0.0)
    fl = ct == 0
    assert abs(dst[:, fl].sum() - src[:, fl].sum()) < 1e-9


def test_oracle_moving_wall_adds_momentum():
    nx, ny = 4, 4
    N = nx * ny
    ct = np.zeros(N, np.uint8)
    ct[(ny - 1) * nx:] = 6  # moving wall row
    src = np.tile(D2Q9_W[:, None], (1, N))
    dst = np.zeros_like(src)
    uw = 0.1
    call(64, "d2q9_stream_with_boundaries", src, dst, ct, nx, ny, uw, 0.0)
    # node below the lid, direction q pulled from the lid: f_opp + 2 w_q 3 c_q.u_w
    y = ny - 2
    for q in (4, 7, 8):  # c_y = -1 populations come from y + 1 (the lid)
        for x in range(nx):
            i = y * nx + x
            o = [0, 3, 4, 1, 2, 7, 8, 5, 6][q]
            assert abs(dst[q, i] - (src[o, i] + 2 * D2Q9_W[q] * 3 * D2Q9_C[q][0] * uw)) < 1e-15


# ----------------------------------------------------------- validations
def test_oracle_poiseuille_reference_validation():
    # tests/validation/test_poiseuille_convergence.cpp:12-73
    NX, NY, tau, gx = 8, 33, 0.8, 1e-4
    nu = (1.0 / 3.0) * (tau - 0.5)
    g = OracleGrid(2, NX, NY, tau=tau)
    for x in range(NX):
        g.set_cell_type(x, 0, 0, 1)
        g.set_cell_type(x, NY - 1, 0, 1)
    g.run(7000, "TRT", gx, 0.0)
    ux = g.ux().reshape(NY, NX).astype(np.float64)
    H = NY - 2
    num = den = 0.0
    for y in range(1, NY - 1):
        us = ux[y].mean()
        s = y - 0.5
        ua = gx / (2 * nu) * s * (H - s)
        num += (us - ua) ** 2
        den += ua**2
    assert math.sqrt(num / den) <= 0.015


def test_oracle_taylor_green_3d_reference_validation():
    # tests/validation/test_taylor_green_decay.cpp:12-63
    L_, tau, u0 = 32, 0.8, 0.04
    nu = (tau - 0.5) / 3
    k = 2 * math.pi / L_
    g = OracleGrid(3, L_, L_, L_, tau=tau)
    g.set_periodic(True)
    z, y, x = np.meshgrid(np.arange(L_), np.arange(L_), np.arange(L_), indexing="ij")
    f32 = np.float32
    kx, ky, kz = (f32(k) * a.astype(f32) for a in (x, y, z))
    ux = f32(u0) * np.sin(kx) * np.cos(ky) * np.cos(kz)
    uy = -f32(u0) * np.cos(kx) * np.sin(ky) * np.cos(kz)
    uz = np.zeros_like(ux)
    p = (f32(u0 * u0 / 16)) * (np.cos(2 * kx) + np.cos(2 * ky)) * (np.cos(2 * kz) + 2)
    rho = 1 + p * 3
    g.initialize_fields(rho.ravel(), ux.ravel(), uy.ravel(), uz.ravel())
    E0 = g.total_kinetic_energy()
    for s in range(1, 101):
        g.step("MRT")
        if s % 20 == 0:
            assert abs(g.total_kinetic_energy() / E0 - math.exp(-6 * k * k * nu * s)) <= 0.025


def test_oracle_taylor_green_2d_periodic_decay():
    # 2D Taylor-Green through the periodic extension: KE ~ exp(-4 nu k^2 t)
    n, tau, u0 = 64, 0.8, 0.04
    nu = (tau - 0.5) / 3
    k = 2 * math.pi / n
    g = OracleGrid(2, n, n, tau=tau, bits=64)
    g.set_periodic(True)
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ux = u0 * np.sin(k * x) * np.cos(k * y)
    uy = -u0 * np.cos(k * x) * np.sin(k * y)
    g.initialize_fields(np.ones(n * n), ux.ravel(), uy.ravel())
    E0 = g.total_kinetic_energy()
    g.run(400, "BGK")
    ratio = g.total_kinetic_energy() / E0
    assert abs(ratio / math.exp(-4 * nu * k * k * 400) - 1) < 0.01


# ------------------------------------------------------------------- LES
def test_oracle_les_equilibrium_and_closed_form():
    count = 50
    rng = np.random.default_rng(4)
    rho = 1 + 0.01 * rng.standard_normal(count)
    u = 0.03 * rng.standard_normal((3, count))
    cu = D3Q19_C @ u
    usq = (u * u).sum(0)
    feq = D3Q19_W[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    tau = np.zeros(count)
    call(64, "d3q19_compute_smagorinsky_tau", feq, rho, u[0].copy(), u[1].copy(), u[2].copy(), tau, 0.505, 0.17, count)
    assert np.allclose(tau, 0.505, atol=1e-15)
    # perturbed: Pi = sum c c (f - feq), Q = sqrt(2 Pi:Pi), tau = (t0 + sqrt(t0^2 + 18 sqrt2 Cs^2 Q / rho)) / 2
    f = feq + 1e-3 * rng.standard_normal(feq.shape)
    call(64, "d3q19_compute_smagorinsky_tau", f, rho, u[0].copy(), u[1].copy(), u[2].copy(), tau, 0.505, 0.17, count)
    neq = f - feq
    Pi = np.einsum("qa,qb,qn->abn", D3Q19_C, D3Q19_C, neq)
    Qs = np.sqrt(2 * (Pi * Pi).sum((0, 1)))
    expect = 0.5 * (0.505 + np.sqrt(0.505**2 + 18 * math.sqrt(2) * 0.17**2 * Qs / rho))
    assert np.allclose(tau, expect, rtol=1e-12)
    # the per-node collision with tau_node == tau reproduces the scalar kernel
    ct = np.zeros(count, np.uint8)
    a, b = np.zeros_like(f), np.zeros_like(f)
    bufs = [np.zeros(count) for _ in range(4)]
    call(64, "d3q19_collide_mrt", f, a, ct, *bufs, 0.7, 1.19, 0.0, 0.0, 0.0, count)
    call(64, "collide_tau_node", 19, 2, f, b, ct, np.full(count, 0.7), 0.0, 0.0, 0.0, count)
    assert np.array_equal(a, b)


# ------------------------------------------------------------------ Zou-He
def test_oracle_zou_he_2d_identities():
    nx, ny = 6, 7
    N = nx * ny
    rng = np.random.default_rng(6)
    f = D2Q9_W[:, None] * (1 + 0.02 * rng.standard_normal((9, N)))
    rho, ux, uy = (np.zeros(N) for _ in range(3))
    uin = 0.05
    call(64, "d2q9_apply_zou_he_inflow_west", f, rho, ux, uy, uin, 0.01, nx, ny)
    for y in range(1, ny - 1):
        i = y * nx
        r = f[:, i].sum()
        assert abs(r - rho[i]) < 1e-14
        assert abs((D2Q9_C[:, 0] @ f[:, i]) - r * uin) < 1e-14
        assert abs((D2Q9_C[:, 1] @ f[:, i]) - r * 0.01) < 1e-14
    call(64, "d2q9_apply_zou_he_outflow_east", f, rho, ux, uy, 1.0, nx, ny)
    for y in range(1, ny - 1):
        i = y * nx + nx - 1
        assert abs(f[:, i].sum() - 1.0) < 1e-14
        assert abs(D2Q9_C[:, 1] @ f[:, i]) < 1e-14


def test_oracle_zou_he_3d_identities():
    # EXTENSION: the 3D velocity inlet gives sum f = rho, sum f c = rho u.
    for lat in (19, 27):
        g = OracleGrid(3, 6, 5, 5, tau=0.7, lattice=lat, bits=64)
        g.set_periodic_axes(True, True)
        g.set_inflow_west(0.05, 0.0)
        g.set_convective_outflow_east(0.05)
        g.add_noise(1e-3, 9)
        g.step("TRT")
        f = g.f_src()
        C = D3Q19_C if lat == 19 else None
        if C is None:
            C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                          [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                          [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1],
                          [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1],
                          [1, -1, -1], [-1, -1, -1]])
        for z in range(5):
            for y in range(5):
                i = (z * 5 + y) * 6
                r = f[:, i].sum()
                j = C.T @ f[:, i]
                assert abs(j[0] - r * 0.05) < 1e-13 and abs(j[1]) < 1e-13 and abs(j[2]) < 1e-13


def test_oracle_regularized_inlet():
    # EXTENSION: the 3D inlet node is the projection feq + 4.5 w (cc - I/3):Pi;
    # projecting again changes nothing and the third-order moments vanish.
    g = OracleGrid(3, 8, 6, 6, tau=0.6, bits=64)
    g.set_periodic_axes(True, True)
    g.set_inflow_west(0.05, 0.0)
    g.set_inflow_turbulence(0.01, 4)
    g.add_noise(1e-3, 5)
    g.step("MRT")
    f = g.f_src()
    C = D3Q19_C.astype(np.float64)
    W = D3Q19_W
    for z in range(6):
        for y in range(6):
            fi = f[:, (z * 6 + y) * 8]
            rho = fi.sum()
            u = C.T @ fi / rho
            cu = C @ u
            feq = W * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u @ u))
            Pi = np.einsum("qa,qb,q->ab", C, C, fi - feq)
            Qt = np.einsum("qa,qb->qab", C, C) - np.eye(3) / 3
            reg = feq + 4.5 * W * np.einsum("qab,ab->q", Qt, Pi)
            assert np.allclose(fi, reg, atol=1e-15)
            m16 = (C[:, 0] * (C[:, 1] ** 2 - C[:, 2] ** 2)) @ fi
            m17 = (C[:, 1] * (C[:, 2] ** 2 - C[:, 0] ** 2)) @ fi
            m18 = (C[:, 2] * (C[:, 1] ** 2 - C[:, 0] ** 2)) @ fi
            assert max(abs(m16), abs(m17), abs(m18)) < 1e-15


def test_oracle_convective_fixed_point():
    nx, ny = 5, 4
    N = nx * ny
    f = np.tile(D2Q9_W[:, None], (1, N))
    prev = f.copy()
    call(64, "d2q9_apply_convective_outflow_east", f, prev, 0.07, nx, ny)
    assert np.allclose(f, prev, atol=1e-16)


# --------------------------------------------------------------------- MEM
def test_oracle_mem_reference_known_answers(oracle32):
    # tests/unit/test_forces_momentum_exchange.cpp:27-68
    NX = NY = 16
    N = NX * NY
    f = np.tile(D2Q9_W.astype(np.float32)[:, None], (1, N)).ravel().copy()
    ct = np.zeros(N, np.uint8)
    for y in (7, 8):
        for x in (7, 8):
            ct[y * NX + x] = 1
    out = [ctypes.c_float(), ctypes.c_float(), ctypes.c_float()]
    fn = oracle32.oracle_d2q9_compute_momentum_exchange_forces
    fn(ob.ptr(f), ob.ptr(ct), NX, NY, ctypes.c_float(7.5), ctypes.c_float(7.5), ctypes.c_float(0.0),
       *[ctypes.byref(o) for o in out])
    assert abs(out[0].value) <= 1e-5 and abs(out[1].value) <= 1e-5
    f[1 * N + (7 * NX + 6)] += np.float32(0.05)
    fn(ob.ptr(f), ob.ptr(ct), NX, NY, ctypes.c_float(7.5), ctypes.c_float(7.5), ctypes.c_float(0.0),
       *[ctypes.byref(o) for o in out])
    assert abs(out[0].value - 0.1) <= 1e-5


def test_oracle_mem_3d_symmetric_body_at_rest():
    g = OracleGrid(3, 20, 20, 20, tau=0.8, bits=64)
    g.set_solid_sphere(9.5, 9.5, 9.5, 4.0)
    out = g.momentum_exchange(9.5, 9.5, 9.5, 8.0)
    assert np.all(np.abs(out) < 1e-12)
    g.step("MRT")  # still at rest (up to the MRT drift, which is isotropic)
    out = g.momentum_exchange(9.5, 9.5, 9.5, 8.0)
    assert np.all(np.abs(out) < 1e-10)


# ------------------------------------------------------------ flags, links
def test_oracle_sphere_flags_integer_count():
    # integer centre and radius: voxels are the lattice points of the ball
    g = OracleGrid(3, 24, 24, 24, tau=0.8)
    g.set_solid_sphere(12.0, 11.0, 10.0, 6.0)
    ct = g.cell_type().reshape(24, 24, 24)
    z, y, x = np.meshgrid(np.arange(24), np.arange(24), np.arange(24), indexing="ij")
    expect = ((x - 12) ** 2 + (y - 11) ** 2 + (z - 10) ** 2 <= 36).astype(np.uint8)
    assert np.array_equal(ct, expect)


# Lattice points with x^2 + y^2 <= r^2 (Gauss's circle problem, OEIS A000328)
# and with x^2 + y^2 + z^2 <= r^2 (OEIS A000605), r = 0 .. 8.
GAUSS_CIRCLE = [1, 5, 13, 29, 49, 81, 113, 149, 197]
BALL_POINTS = [1, 7, 33, 123, 257, 515, 925, 1419, 2109]


def test_oracle_shape_setters_published_counts():
    """lattice_grid_2d.cpp:143-159 (circle), lattice_grid_3d.cpp:184-203
    (z cylinder) and :160-182 (sphere): with an integer radius on a lattice
    point the Solid cells are the lattice points of the disc / ball, whose
    counts are published (an independent check of the inclusive FP32
    distance rule and of the centre offsets)."""
    for r in range(1, 9):
        g2 = OracleGrid(2, 32, 32, 1, tau=0.8)
        g2.set_solid_circle(16.0, 15.0, float(r))
        assert int((g2.cell_type() == 1).sum()) == GAUSS_CIRCLE[r], r
        g3 = OracleGrid(3, 24, 24, 20, tau=0.8)
        g3.set_solid_cylinder_z(12.0, 11.0, float(r))
        assert int((g3.cell_type() == 1).sum()) == 20 * GAUSS_CIRCLE[r], r
        gs = OracleGrid(3, 24, 24, 24, tau=0.8)
        gs.set_solid_sphere(12.0, 11.0, 10.0, float(r))
        assert int((gs.cell_type() == 1).sum()) == BALL_POINTS[r], r


def test_oracle_grid_setters_and_diagnostics_closed_forms():
    """Which cells the grid setters mark (lattice_grid_3d.cpp:205-215,
    lattice_grid_2d.cpp:161-184), the Fluid-only diagnostics
    (lattice_grid_3d.cpp:382-410) on a uniform state, and the parabolic inlet
    extension against the closed form of its row mean."""
    g = OracleGrid(3, 12, 10, 8, tau=0.8, bits=64)
    g.initialize(1.1, 0.03, 0.04, 0.0)
    g.set_solid_box(2, 2, 2, 4, 4, 4)  # 27 cells leave the sums
    nf = 12 * 10 * 8 - 27
    assert abs(g.total_mass() - 1.1 * nf) < 1e-9 * nf
    assert abs(g.total_kinetic_energy() - 0.5 * 1.1 * 0.05 ** 2 * nf) < 1e-12 * nf
    assert abs(g.max_speed() - 0.05) < 1e-12
    g.set_moving_wall_top(0.1)  # the whole y = ny - 1 plane
    ct = g.cell_type().reshape(8, 10, 12)
    assert (ct[:, 9, :] == 6).all() and not (ct[:, :9, :] == 6).any()
    g2 = OracleGrid(2, 9, 7, 1, tau=0.8)
    g2.set_inflow_west(0.05, 0.0)  # every row of x = 0
    g2.set_outflow_east(1.0)  # every row of x = nx - 1
    c2 = g2.cell_type().reshape(7, 9)
    assert (c2[:, 0] == 2).all() and (c2[:, 8] == 3).all() and (c2[:, 1:8] == 0).all()
    g3 = OracleGrid(2, 9, 7, 1, tau=0.8)
    g3.set_moving_wall_top(0.1)  # the north row
    c3 = g3.cell_type().reshape(7, 9)
    assert (c3[6] == 6).all() and not (c3[:6] == 6).any()
    # parabolic inlet: u(s) = 4 u s (H - s) / H^2 at s = y - 1/2 on rows
    # 1 .. ny-2 (H = ny - 2 rows between halfway walls) is symmetric, and its
    # row mean is u (2/3 + 1 / (3 H^2)) -- an off-by-half row offset fails both
    nx, ny, umax = 12, 18, 0.06
    H = ny - 2
    o = OracleGrid(2, nx, ny, 1, tau=0.8, bits=64)
    o.initialize(1.0, 0.0, 0.0)
    for x in range(nx):
        o.set_cell_type(x, 0, 0, 1)
        o.set_cell_type(x, ny - 1, 0, 1)
    o.set_inflow_parabolic(umax)
    o.step("BGK")
    u_in = o.ux().reshape(ny, nx)[1:ny - 1, 0]
    assert np.allclose(u_in, u_in[::-1], rtol=0, atol=1e-14)
    assert abs(u_in.mean() - umax * (2 / 3 + 1 / (3 * H * H))) < 1e-12
    # the two rows next to the centre, s = H/2 -+ 1/2: 4 u (H^2/4 - 1/4) / H^2
    assert abs(u_in.max() - umax * (1 - 1 / (H * H))) < 1e-12


def test_oracle_links_brute_force():
    nx = ny = nz = 20
    cx, cy, cz, R = 9.3, 10.1, 9.7, 4.6
    g = OracleGrid(3, nx, ny, nz, tau=0.8)
    g.set_solid_sphere(cx, cy, cz, R)
    g.build_curved_sphere_links(cx, cy, cz, R)
    idx, q, d = g.links()
    ct = g.cell_type()
    # the link set is exactly {(fluid node, q) : neighbour in range and solid}
    expect = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = (z * ny + y) * nx + x
                if ct[i] != 0:
                    continue
                for qq in range(1, 19):
                    c = D3Q19_C[qq]
                    X, Y, Z = (x + c[0]) % nx, y + c[1], z + c[2]
                    if 0 <= Y < ny and 0 <= Z < nz and ct[(Z * ny + Y) * nx + X] == 1:
                        expect.append((i, qq))
    assert [(int(a), int(b)) for a, b in zip(idx, q)] == expect
    # delta is where the segment x_f + t c meets the sphere, clamped to [0.01, 0.99]
    for i, qq, dd in zip(idx, q, d):
        z, rem = divmod(int(i), nx * ny)
        y, x = divmod(rem, nx)
        c = D3Q19_C[qq].astype(np.float64)
        p = np.array([x - cx, y - cy, z - cz])
        a, b, c0 = c @ c, 2 * p @ c, p @ p - R * R
        disc = b * b - 4 * a * c0
        t = 0.5
        if disc >= 0:
            t1 = (-b - math.sqrt(disc)) / (2 * a)
            t2 = (-b + math.sqrt(disc)) / (2 * a)
            t = t1 if 0 <= t1 <= 1 else (t2 if 0 <= t2 <= 1 else 0.5)
        t = min(max(t, 0.01), 0.99)
        assert abs(dd - t) < 2e-5


def test_oracle_bfl_half_delta_is_halfway_bounce_back():
    nx, ny, nz = 8, 6, 6
    N = nx * ny * nz
    rng = np.random.default_rng(8)
    A = rng.random((19, N))
    ct = np.zeros(N, np.uint8)
    ct[(3 * ny + 3) * nx + 4] = 1
    S = np.zeros_like(A)
    call(64, "d3q19_stream_with_boundaries", A, S, ct, nx, ny, nz, 0.0, 0.0, 0.0)
    links = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = (z * ny + y) * nx + x
                if ct[i]:
                    continue
                for q in range(1, 19):
                    c = D3Q19_C[q]
                    X, Y, Z = (x + c[0]) % nx, y + c[1], z + c[2]
                    if 0 <= Y < ny and 0 <= Z < nz and ct[(Z * ny + Y) * nx + X] == 1:
                        links.append((i, q))
    L = len(links)
    li = np.array([a for a, _ in links], np.int32)
    lq = np.array([b for _, b in links], np.int32)
    zeros = np.zeros(L)
    for delta in (0.5, 0.5 - 1e-12):
        S2 = S.copy()
        call(64, "d3q19_apply_curved_bfl", S2, A, li, lq, np.full(L, delta), zeros, zeros, zeros, nx, ny, nz, L)
        assert np.allclose(S2, S, atol=1e-11)


# ----------------------------------------------------------- FP32 vs FP64
def test_oracle_fp32_fp64_agree_on_sphere_flow():
    out = []
    for bits in (32, 64):
        g = OracleGrid(3, 24, 16, 16, tau=0.6, bits=bits)
        g.initialize(1.0, 0.04, 0.0, 0.0)
        g.set_solid_sphere(8.0, 8.0, 8.0, 3.0)
        g.build_curved_sphere_links(8.0, 8.0, 8.0, 3.0)
        g.run(100, "MRT", 1e-6, 0.0, 0.0)
        out.append(np.stack([g.rho(), g.ux(), g.uy(), g.uz()]).astype(np.float64))
    err = np.linalg.norm(out[0] - out[1]) / np.linalg.norm(out[1])
    assert err < 1e-5


def test_ellipsoid_voxelizer_axes_conventions(oracle32):
    """primitives.cpp:73-100 restated; the rotated form (an extension) must
    reduce to it: identity axes change nothing, swapping two body axes equals
    swapping the semi-axes, and the voxel count follows 4/3 pi rx ry rz."""
    def flags(rx, ry, rz, axes=None):
        o = ob.OracleGrid(3, 48, 40, 36, bits=32)
        o.set_solid_ellipsoid(24.3, 20.1, 17.7, rx, ry, rz, axes)
        return o.cell_type().reshape(36, 40, 48)

    base = flags(11.0, 7.0, 5.0)
    assert np.array_equal(flags(11.0, 7.0, 5.0, np.eye(3)), base)
    swap = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1]], float)
    assert np.array_equal(flags(7.0, 11.0, 5.0, swap), base)
    vol = (base == 1).sum()
    assert abs(vol / (4.0 / 3.0 * np.pi * 11 * 7 * 5) - 1) < 0.03
    # a 45 degree turn about z keeps the z extent and widens the y extent
    c, s = np.cos(np.pi / 4), np.sin(np.pi / 4)
    rot = flags(11.0, 7.0, 5.0, np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]]))
    assert abs((rot == 1).sum() / vol - 1) < 0.03
    ys = lambda f: np.nonzero((f == 1).any(axis=(0, 2)))[0]
    zs = lambda f: np.nonzero((f == 1).any(axis=(1, 2)))[0]
    assert np.array_equal(zs(rot), zs(base)) and len(ys(rot)) > len(ys(base))


def test_naca_voxelizer_shape(oracle32):
    """naca_airfoil.cpp:113-190 restated: NACA 0012 at 0 deg is symmetric
    about its chord line, its section area is 0.685 t c^2 (the 4-digit
    thickness polynomial integrates to 0.6851 t), its thickest section spans
    0.12 c, and a positive angle of attack puts the trailing edge below the
    leading edge."""
    c, lx, ly = 96.0, 40.0, 64.0
    o = ob.OracleGrid(2, 192, 128, bits=32)
    o.set_solid_naca(12, c, 0.0, lx, ly)
    f = o.cell_type().reshape(128, 192)
    assert np.array_equal(f[64 + 1:64 + 20], f[64 - 1:64 - 20:-1])
    area = (f == 1).sum()
    assert abs(area / (0.6851 * 0.12 * c * c) - 1) < 0.03
    col = int(lx + 0.3 * c)
    assert abs((f[:, col] == 1).sum() - 0.12 * c) <= 2
    o8 = ob.OracleGrid(2, 192, 128, bits=32)
    o8.set_solid_naca(12, c, 8.0, lx, ly)
    g = o8.cell_type().reshape(128, 192)
    assert abs((g == 1).sum() / area - 1) < 0.03
    ys = lambda x: np.nonzero(g[:, x] == 1)[0].mean()
    assert ys(int(lx + 0.9 * c)) < ys(int(lx + 0.1 * c)) - 0.5 * 0.8 * c * np.sin(np.radians(8))
    # extruded: every z slab equals the 2D section
    o3 = ob.OracleGrid(3, 192, 128, 6, bits=32)
    o3.set_solid_naca(12, c, 8.0, lx, ly)
    h = o3.cell_type().reshape(6, 128, 192)
    assert all(np.array_equal(h[z], g) for z in range(6))


def test_vorticity_and_q_criterion_closed_forms():
    """analytics_kernels.ispc:4-110 restated: on periodic sine fields the
    central differences have closed forms (sin(k(y+1)) - sin(k(y-1)) =
    2 sin k cos ky), so omega and Q are known exactly; pure shear has Q = 0."""
    nx, ny, nz = 16, 12, 10
    kx, ky = 2 * np.pi / nx, 2 * np.pi / ny
    y2, x2 = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ux = np.ascontiguousarray((-np.sin(ky * y2)).ravel())
    uy = np.ascontiguousarray(np.sin(kx * x2).ravel())
    w = np.zeros(nx * ny)
    ob.call(64, "compute_vorticity_2d", ux, uy, w, nx, ny)
    expect = np.sin(kx) * np.cos(kx * x2) + np.sin(ky) * np.cos(ky * y2)
    assert np.max(np.abs(w - expect.ravel())) < 1e-12
    z3, y3, x3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = [np.ascontiguousarray(a.ravel()) for a in (-np.sin(ky * y3), np.sin(kx * x3), np.zeros_like(x3, float))]
    out = [np.zeros(nx * ny * nz) for _ in range(4)]
    ob.call(64, "compute_vorticity_and_q_criterion_3d", *u, *out, nx, ny, nz)
    dux_dy = -np.sin(ky) * np.cos(ky * y3)
    duy_dx = np.sin(kx) * np.cos(kx * x3)
    assert np.max(np.abs(out[2] - (duy_dx - dux_dy).ravel())) < 1e-12
    assert np.max(np.abs(out[0])) == 0.0 and np.max(np.abs(out[1])) == 0.0
    assert np.max(np.abs(out[3] - (-dux_dy * duy_dx).ravel())) < 1e-12
    shear = [np.ascontiguousarray(np.sin(ky * y3).ravel()), np.zeros(nx * ny * nz), np.zeros(nx * ny * nz)]
    ob.call(64, "compute_vorticity_and_q_criterion_3d", *shear, *out, nx, ny, nz)
    assert np.max(np.abs(out[3])) == 0.0


def test_d2q9_apply_bounce_back_closed_form():
    """d2q9_boundaries.ispc:5-39 restated: on Solid and MovingWall nodes the
    populations are reflected (f_q <- f_opp(q) + 2 w_q rho0 3 c_q.u_w), so the
    mass is kept, the momentum becomes -j + 2 u_w (sum w c c = I/3), applying
    it twice with u_w = 0 is the identity, and Fluid nodes are untouched."""
    n = 64
    rng = np.random.default_rng(5)
    f0 = (D2Q9_W[:, None] * (1 + 0.1 * rng.standard_normal((9, n)))).ravel()
    ct = np.zeros(n, np.uint8)
    ct[::3] = 1
    ct[1::7] = 6
    f = f0.copy()
    ob.call(64, "d2q9_apply_bounce_back", f, ct, 0.0, 0.0, n)
    ob.call(64, "d2q9_apply_bounce_back", f, ct, 0.0, 0.0, n)
    assert np.array_equal(f, f0)
    uw = (0.03, -0.02)
    f = f0.copy()
    ob.call(64, "d2q9_apply_bounce_back", f, ct, *uw, n)
    a, b = f0.reshape(9, n), f.reshape(9, n)
    wall = ct != 0
    assert np.array_equal(a[:, ~wall], b[:, ~wall])
    assert np.allclose(b[:, wall].sum(0), a[:, wall].sum(0), atol=1e-15)
    cx, cy = D2Q9_C[:, 0][:, None], D2Q9_C[:, 1][:, None]
    for c, u in ((cx, uw[0]), (cy, uw[1])):
        assert np.allclose((c * b[:, wall]).sum(0), -(c * a[:, wall]).sum(0) + 2 * u, atol=1e-15)
