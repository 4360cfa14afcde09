"""GPU parity of the device mesh voxelizer (obj_loader.cpp:208-400) with the
oracle: the cell types after LatticeGrid3D.voxelize_mesh equal the oracle's
bit for bit, for closed and open meshes, with and without the interior
fill, over existing cell types, at 256^3, and the grid then steps like the
oracle grid with the same cell types.  The device BFL link builder
(lattice_grid_3d.cpp:217-290) gives the oracle's links bit for bit.
"""
import time

import numpy as np
import pytest

import meshes
import oracle_bind as ob
from parity import check_fp32_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def rotation(a, b):
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])
    return rz @ rx


def device_vs_oracle(wa, shape, tris, fill, prepare=None):
    nx, ny, nz = shape
    g = wa.LatticeGrid3D(nx, ny, nz, 0.6)
    if prepare:
        prepare(g)
    before = g.cell_type_data().reshape(nz, ny, nx).copy()
    t0 = time.perf_counter()
    g.voxelize_mesh(tris, fill)
    dt = time.perf_counter() - t0
    ref = before.copy()
    writes = ob.voxelize_mesh(ref, tris, fill)
    got = g.cell_type_data().reshape(nz, ny, nx)
    diff = int(np.count_nonzero(got != ref))
    assert diff == 0, f"{diff} cells differ from the oracle"
    return g, got, writes, dt


def test_reference_cube(wa):
    _, tris = meshes.parse_obj(meshes.CUBE_OBJ)
    tris = meshes.fit_into_box(tris, 10.0, 22.0)
    g, got, writes, _ = device_vs_oracle(wa, (32, 32, 32), tris, True)
    assert got[16, 16, 16] == 1 and got[2, 2, 2] == 0 and got[30, 30, 30] == 0
    assert got.sum() == 13 ** 3 and writes > 0
    assert not g.is_periodic()


@pytest.mark.parametrize("fill", [True, False])
def test_rotated_sphere_and_torus(wa, fill):
    sph = meshes.icosphere(3, 13.0, (25.3, 19.8, 22.1))
    sph = ((sph - (25.3, 19.8, 22.1)) @ rotation(0.4, 1.1).T + (25.3, 19.8, 22.1)).astype(np.float32)
    device_vs_oracle(wa, (52, 40, 44), sph, fill)
    tor = meshes.torus(11.0, 4.0, 64, 32, (30.5, 27.2, 20.9), rotation(0.9, 0.5))
    device_vs_oracle(wa, (60, 56, 42), tor, fill)


def test_open_shell_and_triangle_soup(wa):
    opened = meshes.box_mesh(6.5, 5.0, 7.25, 30.0, 22.5, 19.0, open_face=3)
    _, got, _, _ = device_vs_oracle(wa, (40, 30, 28), opened, True)
    assert got[13, 14, 18] == 0  # the interior leaks out through the missing face
    rng = np.random.default_rng(3)
    soup = np.concatenate([rng.uniform(-8, 48, (150, 3, 3)),                  # large, partly outside
                           rng.uniform(5, 35, (400, 1, 3)) + rng.normal(0, 1.5, (400, 3, 3))]).astype(np.float32)
    for fill in (False, True):
        device_vs_oracle(wa, (41, 37, 39), soup, fill)


def test_over_existing_cell_types(wa):
    """Solid blocks the exterior walk; a moving wall and inflow cells do not."""
    sph = meshes.icosphere(2, 9.0, (20.0, 16.0, 14.0))

    def prepare(g):
        g.set_solid_box(0, 0, 0, 39, 31, 1)
        g.set_moving_wall_top(0.05, 0.0)
        g.set_inflow_west(0.04, 0.0)

    device_vs_oracle(wa, (40, 32, 28), sph, True, prepare)


def test_mesh_outside_domain_keeps_periodic_mode(wa):
    far = meshes.box_mesh(100.0, 100.0, 100.0, 120.0, 120.0, 120.0)
    g, got, writes, _ = device_vs_oracle(wa, (24, 20, 16), far, True)
    assert writes == 0 and got.sum() == 0 and g.is_periodic()
    g.voxelize_mesh(np.zeros((0, 3, 3), np.float32), True)
    assert g.is_periodic()


def test_sphere_mesh_256_cubed(wa):
    tris = meshes.icosphere(5, 80.0, (127.3, 128.6, 126.9))  # 20480 triangles
    g, got, _, dt = device_vs_oracle(wa, (256, 256, 256), tris, True)
    print(f"256^3, {len(tris)} triangles: voxelize_mesh {dt * 1e3:.1f} ms (upload, kernels, cell-type download)")
    assert got[128, 128, 128] == 1 and got[0, 0, 0] == 0


def test_steps_like_oracle_after_voxelize(wa):
    """A voxelized mesh obstacle in a channel: 200 TRT steps match the oracle
    grid given the same cell types."""
    nx, ny, nz = 48, 28, 26
    tris = meshes.icosphere(2, 6.5, (14.2, 13.7, 12.9))
    g = wa.LatticeGrid3D(nx, ny, nz, 0.56)
    o = ob.OracleGrid(3, nx, ny, nz, tau=0.56, bits=32)
    o64 = ob.OracleGrid(3, nx, ny, nz, tau=0.56, bits=64)
    for t in (g, o, o64):
        t.set_periodic_axes(False, True)
        t.initialize(1.0, 0.04, 0.0, 0.0)
        t.set_inflow_west(0.04, 0.0)
        t.set_convective_outflow_east(0.04)
    g.voxelize_mesh(tris, True)
    ct = g.cell_type_data().reshape(nz, ny, nx)
    ref = o.cell_type().reshape(nz, ny, nx).copy()
    ob.voxelize_mesh(ref, tris, True)
    assert np.array_equal(ct, ref)
    for z, y, x in zip(*np.nonzero(ct == 1)):
        o.set_cell_type(int(x), int(y), int(z), 1)
        o64.set_cell_type(int(x), int(y), int(z), 1)
    assert np.array_equal(o.cell_type(), g.cell_type_data())
    for _ in range(200):
        g.step("TRT", 0.0, 0.0, 0.0, False, 0.14)
    o.run(200, "TRT")
    o64.run(200, "TRT")
    fl = o.cell_type() == 0
    vel = lambda a, b, c: np.concatenate([a[fl], b[fl], c[fl]])
    check_fp32_parity(vel(g.ux_data(), g.uy_data(), g.uz_data()), vel(o.ux(), o.uy(), o.uz()),
                      vel(o64.ux(), o64.uy(), o64.uz()), "mesh channel u")


@pytest.mark.parametrize("lattice,shape,per", [(19, (72, 40, 36), (True, True)), (27, (48, 36, 30), (False, True)),
                                               (19, (40, 44, 38), (False, False))])
def test_device_link_builder_bit_exact(wa, lattice, shape, per):
    """lattice_grid_3d.cpp:217-290 on the device: the same links, order and
    wall fractions (bit for bit) as the oracle, with periodic axes and a
    sphere cut by the domain edge."""
    nx, ny, nz = shape
    c, R = (20.3, ny - 6.6, nz / 2 - 0.4), 9.7
    g = wa.LatticeGrid3D(nx, ny, nz, 0.6, lattice=lattice)
    o = ob.OracleGrid(3, nx, ny, nz, tau=0.6, lattice=lattice, bits=32)
    for t in (g, o):
        t.set_periodic_axes(*per)
        t.set_solid_sphere(*c, R)
    assert g.build_curved_sphere_links(*c, R) == o.build_curved_sphere_links(*c, R) > 0
    gi, gq, gd = g.curved_links()
    oi, oq, od = o.links()
    assert np.array_equal(gi, oi) and np.array_equal(gq, oq)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_device_link_builder_sphere_256(wa):
    """The 256^3 sphere workload's links (R = 24 at (64, 128, 128), periodic
    y/z): equal to the oracle's, built in milliseconds."""
    n, c, R = 256, (64.0, 128.0, 128.0), 24.0
    g = wa.LatticeGrid3D(n, n, n, 0.505)
    g.set_periodic_axes(True, True)
    g.set_solid_sphere(*c, R)
    t0 = time.perf_counter()
    ng = g.build_curved_sphere_links(*c, R)
    dt = time.perf_counter() - t0
    flags = g.cell_type_data().reshape(n, n, n)
    o = ob.OracleGrid(3, n, n, n, tau=0.505, bits=32)
    o.set_periodic_axes(True, True)
    o.set_solid_sphere(*c, R)
    assert np.array_equal(o.cell_type().reshape(n, n, n), flags)
    t0 = time.perf_counter()
    no = o.build_curved_sphere_links(*c, R)
    dto = time.perf_counter() - t0
    print(f"256^3 sphere links: {ng}; device {dt * 1e3:.1f} ms, oracle host scan {dto * 1e3:.0f} ms")
    assert ng == no
    gi, gq, gd = g.curved_links()
    oi, oq, od = o.links()
    assert np.array_equal(gi, oi) and np.array_equal(gq, oq)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))
