"""Pins for the oracle's curved-wall (BFL), 3D momentum-exchange, convective
outlet and D2Q9 Smagorinsky functions (oracle/wake_oracle.c).

None of these restates the function under test.  Each checks a property the
mathematics fixes, chosen so that a plausible restatement mistake fails it:

* BFL (d3q19_curved.ispc:7-71).  Bouzidi's rule is a linear interpolation
  along the link.  For delta < 1/2 the unknown population is the outgoing
  population evaluated at the point the particle leaves from,
  x_f - (1 - 2 delta) c_q; for delta >= 1/2 it interpolates between the
  reflected particle, which lands at x_f + (2 delta - 1) c_q, and the
  opposite population one cell back.  So a field linear along the link is
  reproduced at x_f exactly.  The moving-wall terms are fixed by Galilean
  consistency: a uniform state at the equilibrium of the wall velocity must
  be reproduced exactly, which holds only with the coefficient 2 (delta < 1/2)
  and 1/delta (delta >= 1/2) and the sign the reference uses.  A wrong
  neighbour, swapped populations or a wrong wall coefficient fails one of the
  two for some delta in (0.01, 0.99).
* 3D momentum exchange (forces_momentum_exchange.ispc:71-138).  A single
  extra delta on population q at a node linked to the wall adds exactly
  2 delta c_q to F; the same extra delta on a node not linked along q, on a
  link whose midpoint lies outside r_max, or across a non-periodic domain
  edge, adds nothing.
* Convective outlet (d2q9_boundaries.ispc:125-146).  It is the implicit
  upwind discretisation of df/dt + U df/dx = 0, so the new outlet value makes
  the residual (f_new - f_prev) + U (f_new - f_upstream) vanish; its limits
  are alpha = U / (1 + U) and beta = 1 / (1 + U).
* D2Q9 Smagorinsky tau (d2q9_les.ispc:6-57): equilibrium gives tau_0
  exactly; otherwise the closed form on a brute-force Pi_neq computed with
  numpy's einsum.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_bind as ob
from oracle_bind import D2Q9_C, D2Q9_W, D3Q19_C, D3Q19_W, call

OPP19 = np.array([int(np.flatnonzero((D3Q19_C == -c).all(1))[0]) for c in D3Q19_C])


def _feq(C, W, rho, u):
    cu = C @ u
    return W * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * float(u @ u))


def _bfl_one(A, node, q, delta, uw, dims):
    """Apply the oracle's BFL to a single link and return S[opp q][node]."""
    nx, ny, nz = dims
    N = nx * ny * nz
    S = np.zeros((19, N))
    call(64, "d3q19_apply_curved_bfl", S, np.ascontiguousarray(A), np.array([node], np.int32),
         np.array([q], np.int32), np.array([delta]), np.array([uw[0]]), np.array([uw[1]]), np.array([uw[2]]),
         nx, ny, nz, 1)
    written = np.flatnonzero(S.ravel())
    # exactly one value is written, at plane opp(q), node x_f
    assert set(written.tolist()) <= {OPP19[q] * N + node}
    return S[OPP19[q], node]


DELTAS = (0.01, 0.1, 0.25, 0.4, 0.4999, 0.5, 0.6, 0.75, 0.9, 0.99)
# a face, an edge (xy) and an edge (yz) direction
LINK_QS = (1, 7, 15)


@pytest.mark.parametrize("q", LINK_QS)
@pytest.mark.parametrize("delta", DELTAS)
def test_bfl_reproduces_uniform_equilibrium_of_the_wall_velocity(q, delta):
    dims = (7, 7, 7)
    N = 7 ** 3
    x, y, z = 3, 3, 3
    node = (z * 7 + y) * 7 + x
    uw = np.array([0.031, -0.017, 0.023])
    fe = _feq(D3Q19_C, D3Q19_W, 1.0, uw)
    A = np.tile(fe[:, None], (1, N))
    got = _bfl_one(A, node, q, delta, uw, dims)
    assert abs(got - fe[OPP19[q]]) < 1e-15
    # the same state at rest with the wall moving is NOT a fixed point: the
    # wall term is really applied (guards against a dropped term)
    A0 = np.tile(_feq(D3Q19_C, D3Q19_W, 1.0, np.zeros(3))[:, None], (1, N))
    assert abs(_bfl_one(A0, node, q, delta, uw, dims) - D3Q19_W[q]) > 1e-4


@pytest.mark.parametrize("q", LINK_QS)
@pytest.mark.parametrize("delta", DELTAS)
def test_bfl_interpolates_a_linear_field_at_the_departure_point(q, delta):
    """Populations linear in the position s along c_q: g(s) = a + b s.  The
    unknown S_opp(x_f) is g at x_f's own coordinate carried along the ray:
    delta < 1/2: the outgoing particle left s_f - (1 - 2 delta) and arrives
    after reflection at s_f, so S = g_q(s_f - (1 - 2 delta));
    delta >= 1/2: interpolation at s_f between the reflected particle at
    s_f + (2 delta - 1) (value g_q(s_f)) and the opposite population at s_f
    (carried to s_f - 1), weighted by the distances to s_f."""
    nx = ny = nz = 9
    N = nx * ny * nz
    c = D3Q19_C[q]
    a_q, b_q = 0.07, 0.013  # outgoing population along the ray
    a_o, b_o = 0.05, -0.021  # opposite population
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    # one link along c_q advances s by exactly one
    s = (xx * c[0] + yy * c[1] + zz * c[2]).ravel().astype(np.float64) / float(c @ c)
    rng = np.random.default_rng(11)
    A = rng.random((19, N))  # unrelated values everywhere else
    A[q] = a_q + b_q * s
    A[OPP19[q]] = a_o + b_o * s
    x, y, z = 4, 4, 4
    node = (z * ny + y) * nx + x
    sf = s[node]
    got = _bfl_one(A, node, q, delta, (0.0, 0.0, 0.0), (nx, ny, nz))
    if delta < 0.5:
        expect = a_q + b_q * (sf - (1.0 - 2.0 * delta))
    else:
        # reflected particle at distance 2 delta - 1 from x_f carries g_q(s_f);
        # the opposite population's value at x_f is carried one cell away.
        # Linear interpolation in distance d from x_f over [-(2 delta - 1), 1]:
        d0, d1 = 2.0 * delta - 1.0, 1.0
        v0 = a_q + b_q * sf
        v1 = a_o + b_o * sf
        expect = (d1 * v0 + d0 * v1) / (d0 + d1)
    assert abs(got - expect) < 1e-14, (got, expect)


def test_bfl_limits_and_upstream_neighbour():
    """delta -> 0: the value of A_q one cell upstream; delta -> 1: the mean of
    A_q and A_opp at x_f; delta < 1/2 at the domain's y edge falls back to
    A_q(x_f) (d3q19_curved.ispc:55-59)."""
    nx, ny, nz = 6, 5, 5
    N = nx * ny * nz
    rng = np.random.default_rng(12)
    A = rng.random((19, N))
    q = 3  # +y
    x, y, z = 2, 2, 2
    node = (z * ny + y) * nx + x
    up = (z * ny + (y - 1)) * nx + x
    assert abs(_bfl_one(A, node, q, 1e-9, (0, 0, 0), (nx, ny, nz)) - A[q, up]) < 1e-8
    assert abs(_bfl_one(A, node, q, 1.0, (0, 0, 0), (nx, ny, nz)) - 0.5 * (A[q, node] + A[OPP19[q], node])) < 1e-15
    # the upstream cell x_f - c_q is out of range at y = 0
    edge = (z * ny + 0) * nx + x
    assert abs(_bfl_one(A, edge, q, 0.3, (0, 0, 0), (nx, ny, nz)) - A[q, edge]) < 1e-15
    # x wraps periodically: x = 0 along +x reads x = nx - 1
    x0 = (z * ny + y) * nx + 0
    xw = (z * ny + y) * nx + nx - 1
    d = 0.2
    got = _bfl_one(A, x0, 1, d, (0, 0, 0), (nx, ny, nz))
    assert abs(got - (2 * d * A[1, x0] + (1 - 2 * d) * A[1, xw])) < 1e-15


# ------------------------------------------------------------------ 3D MEM
def _mem3d(f, ct, dims, centre, r_max):
    nx, ny, nz = dims
    out = [ctypes.c_double() for _ in range(3)]
    call(64, "d3q19_compute_momentum_exchange_forces", np.ascontiguousarray(f), ct, nx, ny, nz, float(centre[0]),
         float(centre[1]), float(centre[2]), float(r_max), *[ctypes.byref(o) for o in out])
    return np.array([o.value for o in out])


def test_mem3d_single_population_impulses():
    nx, ny, nz = 16, 14, 12
    N = nx * ny * nz
    g = ob.OracleGrid(3, nx, ny, nz, tau=0.8, bits=64)
    cx, cy, cz, R = 7.3, 6.8, 5.9, 3.2
    g.set_solid_sphere(cx, cy, cz, R)
    ct = g.cell_type().copy()
    rng = np.random.default_rng(5)
    f = np.tile(D3Q19_W[:, None], (1, N)) + 1e-3 * rng.standard_normal((19, N))
    base = _mem3d(f, ct, (nx, ny, nz), (cx, cy, cz), 2 * R)

    def fluid(i):
        return ct[i] == 0

    def solid_at(x, y, z):
        return 0 <= y < ny and 0 <= z < nz and ct[(z * ny + y) * nx + (x % nx)] == 1

    checked = 0
    for q in (1, 6, 9, 15, 18):
        c = D3Q19_C[q]
        # a fluid node whose neighbour along c_q is solid
        cand = [(x, y, z) for z in range(nz) for y in range(ny) for x in range(nx)
                if fluid((z * ny + y) * nx + x) and solid_at(x + c[0], y + c[1], z + c[2])]
        assert cand, q
        x, y, z = cand[len(cand) // 2]
        i = (z * ny + y) * nx + x
        d = 0.037
        f2 = f.copy()
        f2[q, i] += d
        got = _mem3d(f2, ct, (nx, ny, nz), (cx, cy, cz), 2 * R) - base
        assert np.allclose(got, 2 * d * c, atol=1e-12), (q, got)
        # the opposite population at the same node points away from the wall
        f3 = f.copy()
        f3[OPP19[q], i] += d
        if not solid_at(x - c[0], y - c[1], z - c[2]):
            assert np.allclose(_mem3d(f3, ct, (nx, ny, nz), (cx, cy, cz), 2 * R), base, atol=1e-12)
        # a radius that leaves out this link's midpoint
        mid = np.array([x, y, z]) + 0.5 * c - np.array([cx, cy, cz])
        r_small = 0.999 * float(np.sqrt(mid @ mid))
        got_small = _mem3d(f2, ct, (nx, ny, nz), (cx, cy, cz), r_small) - \
            _mem3d(f, ct, (nx, ny, nz), (cx, cy, cz), r_small)
        assert np.allclose(got_small, 0.0, atol=1e-12)
        checked += 1
    assert checked == 5
    # z is not periodic: a solid cell at z = nz - 1 links the fluid node
    # below it along +z (q = 5), but not the node at z = 0 across the edge
    ct2 = np.zeros(N, np.uint8)
    ct2[((nz - 1) * ny + 3) * nx + 4] = 1
    rest = np.tile(D3Q19_W[:, None], (1, N))
    base2 = _mem3d(rest, ct2, (nx, ny, nz), (0, 0, 0), 0.0)
    below = ((nz - 2) * ny + 3) * nx + 4
    f4 = rest.copy()
    f4[5, below] += 0.05
    assert np.allclose(_mem3d(f4, ct2, (nx, ny, nz), (0, 0, 0), 0.0) - base2, [0, 0, 0.1], atol=1e-12)
    across = (0 * ny + 3) * nx + 4
    f5 = rest.copy()
    f5[6, across] += 0.05
    assert np.allclose(_mem3d(f5, ct2, (nx, ny, nz), (0, 0, 0), 0.0) - base2, 0.0, atol=1e-12)


# --------------------------------------------------------- convective outlet
def test_convective_outlet_is_implicit_upwind_advection():
    nx, ny = 7, 6
    N = nx * ny
    rng = np.random.default_rng(9)
    U = 0.083
    f = rng.random((9, N))
    prev = rng.random((9, N))
    out = f.copy()
    call(64, "d2q9_apply_convective_outflow_east", out, prev, U, nx, ny)
    F, P, O = (a.reshape(9, ny, nx) for a in (f, prev, out))
    # outlet rows 1..ny-2: (f_new - f_prev) + U (f_new - f_new(x - 1)) = 0
    res = (O[:, 1:-1, -1] - P[:, 1:-1, -1]) + U * (O[:, 1:-1, -1] - F[:, 1:-1, -2])
    assert np.abs(res).max() < 1e-15
    # everything else untouched: the two corner rows and all inner columns
    assert np.array_equal(O[:, :, :-1], F[:, :, :-1])
    assert np.array_equal(O[:, 0, -1], F[:, 0, -1]) and np.array_equal(O[:, -1, -1], F[:, -1, -1])
    # limits: prev = 0, f = 1 gives alpha; prev = 1, f = 0 gives beta
    one, zero = np.ones((9, N)), np.zeros((9, N))
    a = one.copy()
    call(64, "d2q9_apply_convective_outflow_east", a, zero, U, nx, ny)
    assert np.allclose(a.reshape(9, ny, nx)[:, 1:-1, -1], U / (1 + U), atol=1e-16)
    b = zero.copy()
    call(64, "d2q9_apply_convective_outflow_east", b, one, U, nx, ny)
    assert np.allclose(b.reshape(9, ny, nx)[:, 1:-1, -1], 1 / (1 + U), atol=1e-16)
    # a ramp in x: the upstream value read is the x = nx - 2 column
    ramp = np.tile(np.arange(nx, dtype=np.float64), (9, ny, 1)).reshape(9, N)
    r = ramp.copy()
    call(64, "d2q9_apply_convective_outflow_east", r, zero, U, nx, ny)
    assert np.allclose(r.reshape(9, ny, nx)[:, 1:-1, -1], U / (1 + U) * (nx - 2), atol=1e-14)


# ------------------------------------------------------------ D2Q9 LES tau
def test_d2q9_smagorinsky_tau_equilibrium_and_brute_force():
    count = 64
    rng = np.random.default_rng(21)
    rho = 1 + 0.02 * rng.standard_normal(count)
    u = 0.04 * rng.standard_normal((2, count))
    cu = D2Q9_C @ u
    usq = (u * u).sum(0)
    feq = D2Q9_W[:, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    tau = np.zeros(count)
    t0, cs = 0.52, 0.17
    call(64, "d2q9_compute_smagorinsky_tau", np.ascontiguousarray(feq), rho, u[0].copy(), u[1].copy(), tau, t0, cs,
         count)
    assert np.allclose(tau, t0, atol=1e-15)
    f = feq + 2e-3 * rng.standard_normal(feq.shape)
    call(64, "d2q9_compute_smagorinsky_tau", np.ascontiguousarray(f), rho, u[0].copy(), u[1].copy(), tau, t0, cs,
         count)
    Pi = np.einsum("qa,qb,qn->abn", D2Q9_C, D2Q9_C, f - feq)
    Qs = np.sqrt(2 * (Pi * Pi).sum((0, 1)))
    expect = 0.5 * (t0 + np.sqrt(t0 ** 2 + 18 * math.sqrt(2) * cs ** 2 * Qs / rho))
    assert np.allclose(tau, expect, rtol=1e-12)
    # larger stress, larger tau (the eddy viscosity is monotone in |Pi|)
    f2 = feq + 2.0 * (f - feq)
    tau2 = np.zeros(count)
    call(64, "d2q9_compute_smagorinsky_tau", np.ascontiguousarray(f2), rho, u[0].copy(), u[1].copy(), tau2, t0, cs,
         count)
    assert np.all(tau2 > tau)
