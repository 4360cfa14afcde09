"""GPU tests of the on-disk formats (vtk_exporter.cpp:8-223): the .vti of a
stepped domain holds its fields exactly (binary) or to the reference's text
precision (ASCII), in the reference's array order and XML layout."""
import numpy as np
import pytest

import vti

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wa():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() <= 0:
        pytest.fail("no GPU visible to the HIP runtime")
    return wake_amd


def stepped_3d(wa):
    g = wa.LatticeGrid3D(40, 24, 20, 0.6)
    g.set_periodic_axes(True, True)
    g.initialize(1.0, 0.04, 0.0, 0.0)
    g.set_inflow_west(0.04, 0.0)
    g.set_convective_outflow_east(0.04)
    g.set_solid_sphere(12.0, 12.0, 10.0, 4.5)
    g.run(60, "MRT")
    return g


def test_vti_3d_binary_exact_with_analytics(wa, tmp_path):
    g = stepped_3d(wa)
    p = tmp_path / "sphere.vti"
    g.export_vti(p, binary=True, with_analytics=True)
    arrays, ext, head = vti.read_vti(p)
    assert ext == [0, 39, 0, 23, 0, 19]
    assert list(arrays) == ["Velocity", "Speed", "Density", "Q_Criterion", "Vorticity", "CellType"]
    vel = arrays["Velocity"]
    assert np.array_equal(vel[:, 0], g.ux_data()) and np.array_equal(vel[:, 1], g.uy_data())
    assert np.array_equal(vel[:, 2], g.uz_data())
    assert np.array_equal(arrays["Speed"][:, 0], g.speed_data())
    assert np.array_equal(arrays["Density"][:, 0], g.rho_data())
    assert np.array_equal(arrays["CellType"][:, 0], g.cell_type_data())
    ox, oy, oz, q = g.vorticity_q_data()
    assert np.array_equal(arrays["Q_Criterion"][:, 0], q)
    assert np.array_equal(arrays["Vorticity"], np.stack([ox, oy, oz], 1))
    assert '<PointData Vectors="Velocity" Scalars="Speed">' in head


def test_vti_ascii_reference_layout(wa, tmp_path):
    g = stepped_3d(wa)
    p = tmp_path / "sphere_ascii.vti"
    g.export_vti(p, binary=False)
    text = p.read_text()
    lines = text.splitlines()
    assert lines[0] == '<?xml version="1.0"?>'
    assert lines[1] == '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">'
    assert lines[2] == '  <ImageData WholeExtent="0 39 0 23 0 19" Origin="0 0 0" Spacing="1 1 1">'
    assert lines[-1] == "</VTKFile>"
    # one text line per x row
    start = lines.index('        <DataArray type="Float32" Name="Speed" NumberOfComponents="1" format="ascii">')
    assert lines[start + 24 * 20 + 1] == "        </DataArray>"
    arrays, _, _ = vti.read_vti(p)
    assert list(arrays) == ["Velocity", "Speed", "Density", "CellType"]
    # default stream formatting: 6 significant digits
    assert np.allclose(arrays["Speed"][:, 0], g.speed_data(), rtol=6e-6, atol=1e-12)
    assert np.allclose(arrays["Density"][:, 0], g.rho_data(), rtol=6e-6)
    assert np.array_equal(arrays["CellType"][:, 0], g.cell_type_data())


def test_vti_2d_with_vorticity(wa, tmp_path):
    g = wa.LatticeGrid2D(64, 32, 0.6)
    g.initialize(1.0, 0.05, 0.0)
    g.set_inflow_west(0.05, 0.0)
    g.set_outflow_east(1.0)
    g.set_solid_circle(16.0, 16.0, 4.0)
    g.run(100, "TRT")
    p = tmp_path / "karman.vti"
    g.export_vti(p, binary=True, with_analytics=True)
    arrays, ext, _ = vti.read_vti(p)
    assert ext == [0, 63, 0, 31, 0, 0]
    assert list(arrays) == ["Velocity", "Speed", "Density", "Vorticity", "CellType"]
    assert np.array_equal(arrays["Velocity"][:, 2], np.zeros(64 * 32, np.float32))
    assert np.array_equal(arrays["Vorticity"][:, 0], g.vorticity_data())
    with pytest.raises(RuntimeError):
        g.export_vti(tmp_path / "missing_dir" / "x.vti")
