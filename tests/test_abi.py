"""C-ABI checks that need no GPU: the library loads, exports every function
include/wake_hip.h declares, compiles in the same D3Q19 MRT table as the
oracle, and the device path fails loudly when no GPU is usable."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wake_hip.h")
LIB = os.path.join(ROOT, "wake_amd", "libwake_hip.so")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([a-z_0-9]+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return set(line.split()[-1] for line in out.splitlines() if " T " in line)


def test_library_exports_every_declared_function():
    decl = declared_functions()
    assert len(decl) >= 60
    missing = [d for d in decl if d not in exported_symbols()]
    assert not missing, missing


def test_kernel_exports_match_reference_names():
    # every export of src/ispc: the 28 LBM exports the hot path replaces, the
    # two analytics exports that read its velocity field, the three thermal
    # D2Q9 exports and the two immersed-boundary exports
    names = ["d2q9_compute_macroscopic_moments", "d2q9_compute_macroscopic_moments_forced",
             "d2q9_compute_velocity_magnitude", "d2q9_collide_bgk", "d2q9_collide_trt", "d2q9_collide_mrt",
             "d2q9_stream_periodic", "d2q9_stream_with_boundaries", "d2q9_apply_bounce_back",
             "d2q9_apply_zou_he_inflow_west", "d2q9_apply_zou_he_outflow_east",
             "d2q9_apply_convective_outflow_east", "d2q9_compute_smagorinsky_tau",
             "d2q9_compute_momentum_exchange_forces", "d3q19_compute_macroscopic_moments",
             "d3q19_compute_macroscopic_moments_forced", "d3q19_compute_velocity_magnitude", "d3q19_collide_bgk",
             "d3q19_collide_trt", "d3q19_collide_mrt", "d3q19_stream_periodic", "d3q19_stream_with_boundaries",
             "d3q19_compute_smagorinsky_tau", "d3q19_apply_curved_bfl", "d3q19_compute_momentum_exchange_forces",
             "d3q27_collide", "d3q27_stream_periodic", "d3q27_compute_macroscopic_moments",
             # analytics_kernels.ispc (vorticity and Q-criterion)
             "compute_vorticity_2d", "compute_vorticity_and_q_criterion_3d",
             # thermal_kernels.ispc and ibm_kernels.ispc
             "thermal_collide_d2q9", "thermal_compute_buoyancy_force_2d", "thermal_apply_isothermal_wall",
             "ibm_interpolate_velocity_2d", "ibm_spread_force_2d"]
    assert len(names) == 35
    ex = exported_symbols()
    assert all(n in ex for n in names)
    import wake_amd._lib as wl
    assert set(names) == set(wl.KERNEL_SIGS)


REF_ISPC = "/root/reference/src/ispc"


def _reference_exports():
    """name -> signature letters (p pointer, i int, f float) of every
    `export void` in the reference's ISPC sources: arrays and `&` outputs are
    pointers at the ABI level, `uniform int` / `uniform float` scalars."""
    import glob

    out = {}
    for path in sorted(glob.glob(os.path.join(REF_ISPC, "*.ispc"))):
        txt = re.sub(r"//[^\n]*", "", open(path).read())
        for name, params in re.findall(r"export\s+void\s+(\w+)\s*\(([^)]*)\)", txt):
            sig = ""
            for prm in (x.strip() for x in params.split(",") if x.strip()):
                if "[" in prm or "&" in prm or "*" in prm:
                    sig += "p"
                elif re.search(r"\bfloat\b", prm):
                    sig += "f"
                elif re.search(r"\b(int|int32)\b", prm):
                    sig += "i"
                else:
                    raise AssertionError(f"unparsed parameter {prm!r} of {name}")
            out[name] = sig
    return out


@pytest.mark.skipif(not os.path.isdir(REF_ISPC), reason="the reference tree is not mounted here")
def test_kernel_exports_match_reference_prototypes():
    """Every ISPC export's parameter list (kinds and order) equals the one
    this library binds and include/wake_hip.h declares."""
    import wake_amd._lib as wl

    ref = _reference_exports()
    assert set(ref) == set(wl.KERNEL_SIGS), set(ref) ^ set(wl.KERNEL_SIGS)
    bad = {n: (ref[n], wl.KERNEL_SIGS[n]) for n in ref if ref[n] != wl.KERNEL_SIGS[n]}
    assert not bad, bad
    # the header's parameter lists give the same kinds
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name, sig in ref.items():
        m = re.search(r"\bvoid\s+" + name + r"\s*\(([^)]*)\)", txt)
        assert m, name
        kinds = ""
        for prm in (x.strip() for x in m.group(1).split(",")):
            kinds += "p" if ("[" in prm or "*" in prm) else ("f" if prm.startswith("float") else "i")
        assert kinds == sig, (name, kinds, sig)


def test_compiled_mrt_table_equals_oracle():
    L = ctypes.CDLL(LIB)
    dev = np.zeros(361, np.float32)
    L.wake_hip_d3q19_minv(dev.ctypes.data_as(ctypes.c_void_p))
    ora = np.zeros(361, np.float32)
    ob.lib(32).oracle_d3q19_minv(ob.ptr(ora))
    assert np.array_equal(dev.view(np.uint32), ora.view(np.uint32))


def test_package_imports_and_fails_loudly_without_gpu():
    import wake_amd

    if wake_amd._lib.lib().wake_hip_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        wake_amd.LatticeGrid3D(8, 8, 8, 0.8)
    with pytest.raises(RuntimeError):
        f = np.zeros(9 * 4, np.float32)
        r = np.zeros(4, np.float32)
        wake_amd.kernels.kernel("d2q9_compute_macroscopic_moments", f, r, r.copy(), r.copy(), 4)


def test_hbm_probe_rejects_invalid_arguments():
    """Checked on the host before any GPU call, so it runs here too."""
    import wake_amd

    L = wake_amd._lib.lib()
    ms = ctypes.c_float(-1.0)
    for args in ((3, 64, 64, 64, 5), (-1, 64, 64, 64, 5), (0, 0, 64, 64, 5), (1, 2048, 64, 64, 5), (1, 64, 64, 70000, 5),
                 (0, 64, 64, 64, 0)):
        assert L.wake_hip_hbm_probe(*args, ctypes.byref(ms)) != 0, args
        assert L.wake_hip_last_error()
    assert L.wake_hip_hbm_probe(0, 64, 64, 64, 5, None) != 0
    assert ms.value == -1.0
    L.wake_hip_clear_error()


def test_missing_library_raises(tmp_path, monkeypatch):
    import importlib

    import wake_amd._lib as wl

    monkeypatch.setattr(wl, "_lib", None)
    monkeypatch.setattr(wl, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        wl.lib()
    monkeypatch.undo()
    importlib.reload(wl)


def test_kernel_exports_on_empty_ranges_do_nothing():
    """An ISPC `foreach` over an empty range does nothing, so the exports
    return at once for count <= 0 or a zero dimension: no GPU call (none is
    usable here, so any would record an error), outputs untouched, and the
    momentum-exchange sums come back as zero."""
    import wake_amd

    K = wake_amd.kernels
    f = np.full(27 * 4, 7.0, np.float32)
    out = np.full(27 * 4, 3.0, np.float32)
    r, u, v, w = (np.full(4, 5.0, np.float32) for _ in range(4))
    ct = np.zeros(4, np.uint8)
    for count in (0, -3):
        K.kernel("d2q9_compute_macroscopic_moments", f, r, u, v, count)
        K.kernel("d3q19_collide_mrt", f, out, ct, r, u, v, w, 0.6, 1.19, 0.0, 0.0, 0.0, count)
        K.kernel("d3q27_collide", f, out, ct, r, u, v, w, 0.6, 0.7, 0.0, 0.0, 0.0, count)
        K.kernel("d3q19_compute_smagorinsky_tau", f, r, u, v, w, out, 0.6, 0.17, count)
        K.kernel("d2q9_compute_velocity_magnitude", u, v, out, count)
        K.kernel("d2q9_apply_bounce_back", f, ct, 0.0, 0.0, count)
        K.kernel("thermal_collide_d2q9", f, out, ct, u, v, r, 0.8, count)
        K.kernel("thermal_compute_buoyancy_force_2d", r, out, 0.5, 0.05, -0.001, count)
    for dims in ((0, 4, 4), (4, 0, 4), (4, 4, 0)):
        K.kernel("d3q19_stream_periodic", f, out, *dims)
        K.kernel("d3q27_stream_periodic", f, out, *dims)
        K.kernel("d3q19_stream_with_boundaries", f, out, ct, *dims, 0.0, 0.0, 0.0)
        K.kernel("compute_vorticity_and_q_criterion_3d", u, v, w, r, r, r, r, *dims)
    K.kernel("d2q9_apply_zou_he_inflow_west", f, r, u, v, 0.05, 0.0, 4, 2)
    K.kernel("d2q9_apply_zou_he_outflow_east", f, r, u, v, 1.0, 0, 4)
    K.kernel("d2q9_apply_convective_outflow_east", f, f, 0.05, 4, 1)
    assert (out == 3.0).all() and (r == 5.0).all() and (u == 5.0).all() and (f == 7.0).all()
    F = [np.full(1, 9.0, np.float32) for _ in range(3)]
    K.kernel("d3q19_compute_momentum_exchange_forces", f, ct, 0, 4, 4, 1.0, 1.0, 1.0, 2.0, *F)
    assert all(x[0] == 0.0 for x in F)
    K.kernel("d3q19_apply_curved_bfl", f, f, np.zeros(1, np.int32), np.ones(1, np.int32),
             np.full(1, 0.5, np.float32), r, r, r, 4, 4, 4, 0)


def test_curved_bfl_export_rejects_links_outside_the_grid():
    """The BFL kernel addresses f by the caller's link node and direction;
    a link outside the grid is an error before any GPU call."""
    import wake_amd

    K = wake_amd.kernels
    n = 4 * 4 * 4
    f = np.zeros(19 * n, np.float32)
    d = np.full(2, 0.5, np.float32)
    z = np.zeros(2, np.float32)
    for nodes, qs in (([0, n], [1, 2]), ([0, -1], [1, 2]), ([0, 5], [1, 19]), ([0, 5], [-2, 3])):
        with pytest.raises(RuntimeError, match="outside"):
            K.kernel("d3q19_apply_curved_bfl", f, f, np.array(nodes, np.int32), np.array(qs, np.int32), d, z, z, z,
                     4, 4, 4, 2)
