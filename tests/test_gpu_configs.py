"""The benchmark configurations at reduced size against the oracle.

* Config 3: D3Q27 TRT (Lambda = 1/4) + coupled Smagorinsky LES channel with a
  NACA 0012 wing extruded over the span at 8 degrees (naca_airfoil.cpp:111-190;
  the collision is d3q27_collision_mrt.ispc:5-91), Zou-He inlet u = 0.04,
  convective outlet, walls in y, periodic span, Re 2000 on the chord --
  384x192x192 with chord 96 scaled by 1/4 to 96x48x48 with chord 24.  FP32
  storage against the FP32 oracle; FP16 / BF16 deviation storage against the
  FP64 oracle, with C_D and C_L.
* Config 2's boundary set (MRT + coupled LES, Zou-He inlet with seeded 1%
  turbulence, convective outlet, periodic y/z, BFL sphere) for 1000 steps.
* Config 4's collision (D2Q9 MRT + coupled LES, Cs 0.17) on a batch member
  at the batched workload's hardest point (NX 128, D 4, Re 300).

Fields: 1e-5 relative L2, or where the flow amplifies rounding past it the
FP32-floor criterion of parity.check_fp32_parity (stated there).  Forces:
C_D within 1e-4 relative and C_L within 1e-4 of the drag, or within 1.25x
the spread of the two FP32 oracle builds on a flow whose rounding spread
exceeds that.
"""
import numpy as np
import pytest

import oracle_bind as ob
from bench import wing_setup
from parity import check_fp32_parity, rel_l2

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-4
WING = dict(nx=96, ny=48, nz=48, chord=24.0, u=0.04, re=2000.0)


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def _wing_tau():
    return 3.0 * WING["u"] * WING["chord"] / WING["re"] + 0.5


def _wing_oracle(bits, steps, every):
    o = wing_setup(ob.OracleGrid(3, WING["nx"], WING["ny"], WING["nz"], tau=_wing_tau(), lattice=27, bits=bits),
                   WING["nx"], WING["ny"], WING["chord"])
    hist = []
    for _ in range(steps // every):
        o.run(every, "TRT", use_les=True, smag=0.17)
        hist.append(o.momentum_exchange(0.0, 0.0, 0.0, 0.0))
    return o, np.array(hist)


def _wing_gpu(wa, storage, steps):
    g = wing_setup(wa.LatticeGrid3D(WING["nx"], WING["ny"], WING["nz"], _wing_tau(), lattice=27, storage=storage),
                   WING["nx"], WING["ny"], WING["chord"])
    g.set_force_probe(0.0, 0.0, 0.0, 0.0)
    g.run(steps, "TRT", use_les=True, smag=0.17)
    first, hist = g.force_history()
    assert first == 0 and hist.shape == (steps + 1, 3)
    return g, hist


def _coeffs(F):
    q = 0.5 * WING["u"] ** 2 * WING["chord"] * WING["nz"]
    return F[:, 0] / q, F[:, 1] / q


def _vel(h, fl, gpu):
    if gpu:
        return np.concatenate([h.ux_data()[fl], h.uy_data()[fl], h.uz_data()[fl]])
    return np.concatenate([h.ux()[fl], h.uy()[fl], h.uz()[fl]])


STEPS3, EVERY3 = 200, 10


@pytest.fixture(scope="module")
def wing_oracles():
    return {b: _wing_oracle(b, STEPS3, EVERY3) for b in (32, "32fma", 64)}


def test_config3_wing_fp32_vs_oracle(wa, wing_oracles):
    g, hist = _wing_gpu(wa, 0, STEPS3)
    (o32, h32), (ofm, hfm), (o64, h64) = (wing_oracles[b] for b in (32, "32fma", 64))
    assert np.array_equal(g.cell_type_data(), o32.cell_type())
    fl = o32.cell_type() == 0
    check_fp32_parity(_vel(g, fl, True), _vel(o32, fl, False), _vel(o64, fl, False), "wing u",
                      o32fma=_vel(ofm, fl, False))
    cd_g, cl_g = _coeffs(hist[EVERY3::EVERY3])
    cd_o, cl_o = _coeffs(h32)
    cd_f, cl_f = _coeffs(hfm)
    scale = np.abs(cd_o)
    e_cd = np.abs(cd_g - cd_o) / scale
    e_cl = np.abs(cl_g - cl_o) / scale
    s_cd = np.abs(cd_f - cd_o) / scale
    s_cl = np.abs(cl_f - cl_o) / scale
    print(f"wing FP32: C_D {cd_g[-1]:.5f} (oracle {cd_o[-1]:.5f}), C_L {cl_g[-1]:.5f} (oracle {cl_o[-1]:.5f}); "
          f"max error C_D {e_cd.max():.2e}, C_L {e_cl.max():.2e}; FP32-build spread C_D {s_cd.max():.2e}, "
          f"C_L {s_cl.max():.2e}")
    assert np.all(e_cd <= np.maximum(FORCE_TOL, 1.25 * s_cd))
    assert np.all(e_cl <= np.maximum(FORCE_TOL, 1.25 * s_cl))


# Bounds of the 16-bit storages against FP64 after 200 steps (velocity
# relative L2 and force coefficients relative to the drag).  FP16 keeps 11
# significant bits of f - w, BF16 8; at this Reynolds number tau is 0.50144,
# next to the stability limit, where the flow amplifies the storage rounding
# (measured on MI355X: FP16 3.0e-3 in u and 7e-4 in C_D; FP32 storage 2.5e-5).
PRECISION_BOUNDS = {1: {"u": 5e-3, "force": 1e-2}, 2: {"u": 5e-2, "force": 5e-2}}


@pytest.mark.parametrize("storage,name", [(1, "fp16"), (2, "bf16")])
def test_config3_wing_reduced_precision_vs_fp64(wa, wing_oracles, storage, name):
    o64, h64 = wing_oracles[64]
    fl = o64.cell_type() == 0
    ref = _vel(o64, fl, False)
    cd_r, cl_r = _coeffs(h64)
    out = {}
    for st in (0, storage):
        g, hist = _wing_gpu(wa, st, STEPS3)
        cd, cl = _coeffs(hist[EVERY3::EVERY3])
        out[st] = (rel_l2(_vel(g, fl, True), ref), np.max(np.abs(cd - cd_r) / np.abs(cd_r)),
                   np.max(np.abs(cl - cl_r) / np.abs(cd_r)))
    print(f"wing {name} vs FP64 after {STEPS3} steps: u rel L2 {out[storage][0]:.2e}, C_D {out[storage][1]:.2e}, "
          f"C_L {out[storage][2]:.2e} (FP32 storage: {out[0][0]:.2e}, {out[0][1]:.2e}, {out[0][2]:.2e})")
    b = PRECISION_BOUNDS[storage]
    assert out[storage][0] <= b["u"]
    assert out[storage][1] <= b["force"] and out[storage][2] <= b["force"]


def _config2_set(h, R, c):
    h.set_periodic_axes(True, True)
    h.initialize(1.0, 0.05, 0.0, 0.0)
    h.set_les_coupled(True)
    h.set_inflow_west(0.05, 0.0)
    h.set_inflow_turbulence(0.01, 2)
    h.set_convective_outflow_east(0.05)
    h.set_solid_sphere(*c, R)
    return h.build_curved_sphere_links(*c, R)


def test_config2_boundary_set_1000_steps(wa):
    nx, ny, nz, R = 64, 48, 48, 6.0
    c = (16.0, 24.0, 24.0)
    g = wa.LatticeGrid3D(nx, ny, nz, 0.505)
    os_ = {b: ob.OracleGrid(3, nx, ny, nz, tau=0.505, bits=b) for b in (32, "32fma", 64)}
    nl = {_config2_set(h, R, c) for h in (g, *os_.values())}
    assert len(nl) == 1
    g.set_force_probe(*c, 2 * R)
    forces = {b: [] for b in os_}
    for _ in range(10):
        g.run(100, "MRT", use_les=True, smag=0.17)
        for b, o in os_.items():
            o.run(100, "MRT", use_les=True, smag=0.17)
            forces[b].append(o.momentum_exchange(*c, 2 * R))
    o32, o64, ofm = os_[32], os_[64], os_["32fma"]
    fl = o32.cell_type() == 0
    check_fp32_parity(_vel(g, fl, True), _vel(o32, fl, False), _vel(o64, fl, False), "config-2 set u, 1000 steps",
                      o32fma=_vel(ofm, fl, False))
    _, hist = g.force_history()
    fg = hist[100::100, 0]
    f32, ffm = (np.array(forces[b])[:, 0] for b in (32, "32fma"))
    e, s = np.abs(fg - f32) / np.abs(f32), np.abs(ffm - f32) / np.abs(f32)
    print(f"config-2 set drag every 100 steps: max error {e.max():.2e}, FP32-build spread {s.max():.2e}")
    assert np.all(e <= np.maximum(FORCE_TOL, 1.25 * s))


def _les_member(target, nx, re, oracle_bits=None, storage=0):
    ny = nx // 4
    D = ny / 8.0
    umax = 0.05
    tau = 3.0 * (2.0 / 3.0 * umax) * D / re + 0.5
    if oracle_bits is not None:
        g = ob.OracleGrid(2, nx, ny, 1, tau=tau, lattice=9, bits=oracle_bits)
        cell = lambda x, y, t: g.set_cell_type(x, y, 0, t)
    else:
        g = target.LatticeGrid2D(nx, ny, tau, storage=storage)
        cell = lambda x, y, t: g.set_cell_type(x, y, t)
    H = ny - 2
    s = np.arange(ny, dtype=np.float64) - 0.5
    prof = np.clip(4.0 * umax * s * (H - s) / (H * H), 0.0, None)
    g.initialize_fields(np.ones(nx * ny, np.float32), np.repeat(prof, nx).astype(np.float32),
                        np.zeros(nx * ny, np.float32))
    g.set_les_coupled(True)
    g.set_inflow_parabolic(umax)
    g.set_outflow_east(1.0)
    for x in range(nx):
        cell(x, 0, 1)
        cell(x, ny - 1, 1)
    cx, cy = nx / 5.0, ny / 2.0 + 1.0
    g.set_solid_circle(cx, cy, D / 2)
    return g, (cx, cy, D)


def test_config4_les_coupled_mrt_batch_member_vs_oracle(wa):
    """The batched-wake member's collision (D2Q9 MRT + coupled LES) against
    the oracle, at NX 128 (D 4) and Re 300, where plain MRT diverges."""
    nx, re, steps = 128, 300.0, 400
    g, (cx, cy, D) = _les_member(wa, nx, re)
    b = wa.Batch(lattice=9, collision="MRT", use_les=True, chunk=50)
    slot = b.add(g, steps, smag=0.17, forces=(cx, cy, 0.0, 1.5 * D))
    b.run(steps)
    F = b.forces(slot)
    os_ = {bits: _les_member(wa, nx, re, oracle_bits=bits)[0] for bits in (32, "32fma", 64)}
    fo = {bits: [] for bits in os_}
    for k in range(steps // 50):
        for bits, o in os_.items():
            o.run(50, "MRT", use_les=True, smag=0.17)
            fo[bits].append(o.momentum_exchange(cx, cy, 0.0, 1.5 * D))
    o32, o64, ofm = os_[32], os_[64], os_["32fma"]
    fl = o32.cell_type() == 0
    vel = lambda h, gpu: (np.concatenate([h.ux_data()[fl], h.uy_data()[fl]]) if gpu
                          else np.concatenate([h.ux()[fl], h.uy()[fl]]))
    assert rel_l2(g.rho_data()[fl], o32.rho()[fl]) <= 1e-5
    check_fp32_parity(vel(g, True), vel(o32, False), vel(o64, False), "LES-coupled D2Q9 member u",
                      o32fma=vel(ofm, False))
    fg = F[49::50, 0]
    f32, ffm = (np.array(fo[bits])[:, 0] for bits in (32, "32fma"))
    e, s = np.abs(fg - f32) / np.abs(f32), np.abs(ffm - f32) / np.abs(f32)
    print(f"LES-coupled member drag every 50 steps: max error {e.max():.2e}, FP32-build spread {s.max():.2e}")
    assert np.all(e <= np.maximum(FORCE_TOL, 1.25 * s))


@pytest.mark.parametrize("storage,name", [(1, "fp16"), (2, "bf16")])
def test_config4_batch_member_reduced_precision_vs_fp64(wa, storage, name):
    """The batched-wake member with 16-bit deviation storage in a batch,
    against the FP64 oracle (NX 128, D 4, Re 300, 400 steps).  tau is 0.50133
    here, next to the stability limit as in the wing channel, and the flow
    amplifies the storage rounding the same way (measured on MI355X: FP16
    3.0e-3 in u and 8e-4 in the drag, BF16 2.9e-2 and 9e-3), so the wing's
    bounds apply.  FP16 storage
    misses the 1e-3 field bound on this flow, which is why the batched-wake
    workload keeps FP32 storage."""
    nx, re, steps = 128, 300.0, 400
    g, (cx, cy, D) = _les_member(wa, nx, re, storage=storage)
    b = wa.Batch(lattice=9, storage=storage, collision="MRT", use_les=True, chunk=50)
    slot = b.add(g, steps, smag=0.17, forces=(cx, cy, 0.0, 1.5 * D))
    b.run(steps)
    F = b.forces(slot)
    o64, _ = _les_member(wa, nx, re, oracle_bits=64)
    fo = []
    for _ in range(steps // 50):
        o64.run(50, "MRT", use_les=True, smag=0.17)
        fo.append(o64.momentum_exchange(cx, cy, 0.0, 1.5 * D))
    fl = o64.cell_type() == 0
    eu = rel_l2(np.concatenate([g.ux_data()[fl], g.uy_data()[fl]]), np.concatenate([o64.ux()[fl], o64.uy()[fl]]))
    f64 = np.array(fo)[:, 0]
    ef = float(np.max(np.abs(F[49::50, 0] - f64) / np.abs(f64)))
    print(f"LES-coupled D2Q9 member, {name} storage vs FP64 after {steps} steps: u rel L2 {eu:.2e}, "
          f"drag max rel error {ef:.2e}")
    bounds = PRECISION_BOUNDS[storage]
    assert eu <= bounds["u"]
    assert ef <= bounds["force"]
