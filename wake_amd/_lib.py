"""ctypes binding of libwake_hip.so (include/wake_hip.h).

The library is loaded from this package directory only; if it is missing
the import fails with instructions to build it -- there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WAKE_HIP_LIB selects another build of the library (a code-generation
# variant of the same sources, for A/B timing); the default is the in-tree one.
LIB_PATH = os.environ.get("WAKE_HIP_LIB") or os.path.join(HERE, "libwake_hip.so")

_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double
LL = ctypes.c_longlong
U64 = ctypes.c_uint64

# Part 1: kernel exports -- signature letters: p pointer, i int, f float
KERNEL_SIGS = {
    "d2q9_compute_macroscopic_moments": "ppppi",
    "d2q9_compute_macroscopic_moments_forced": "ppppffi",
    "d2q9_compute_velocity_magnitude": "pppi",
    "d2q9_collide_bgk": "ppppppfffi",
    "d2q9_collide_trt": "ppppppffffi",
    "d2q9_collide_mrt": "ppppppffffi",
    "d2q9_stream_periodic": "ppii",
    "d2q9_stream_with_boundaries": "pppiiff",
    "d2q9_apply_bounce_back": "ppffi",
    "d2q9_apply_zou_he_inflow_west": "ppppffii",
    "d2q9_apply_zou_he_outflow_east": "ppppfii",
    "d2q9_apply_convective_outflow_east": "ppfii",
    "d2q9_compute_smagorinsky_tau": "pppppffi",
    "d2q9_compute_momentum_exchange_forces": "ppiifffppp",
    "d3q19_compute_macroscopic_moments": "pppppi",
    "d3q19_compute_macroscopic_moments_forced": "pppppfffi",
    "d3q19_compute_velocity_magnitude": "ppppi",
    "d3q19_collide_bgk": "pppppppffffi",
    "d3q19_collide_trt": "pppppppfffffi",
    "d3q19_collide_mrt": "pppppppfffffi",
    "d3q19_stream_periodic": "ppiii",
    "d3q19_stream_with_boundaries": "pppiiifff",
    "d3q19_compute_smagorinsky_tau": "ppppppffi",
    "d3q19_apply_curved_bfl": "ppppppppiiii",
    "d3q19_compute_momentum_exchange_forces": "ppiiiffffppp",
    "d3q27_collide": "pppppppfffffi",
    "d3q27_stream_periodic": "ppiii",
    "d3q27_compute_macroscopic_moments": "pppppi",
    "compute_vorticity_2d": "pppii",
    "compute_vorticity_and_q_criterion_3d": "pppppppiii",
    "thermal_collide_d2q9": "ppppppfi",
    "thermal_compute_buoyancy_force_2d": "ppfffi",
    "thermal_apply_isothermal_wall": "piiif",
    "ibm_interpolate_velocity_2d": "ppppppiii",
    "ibm_spread_force_2d": "pppppppiii",
}
_CT = {"p": P, "i": I, "f": F}

# Part 2: device-resident domains
DOMAIN_SIGS = {
    "wake_domain_bytes_needed": ([I, I, I, I, I, I, LL, I], LL),
    "wake_domain_create": ([I, I, I, I, I, F, I, LL, P], I),
    "wake_domain_destroy": ([P], None),
    "wake_domain_initialize": ([P, F, F, F, F], I),
    "wake_domain_initialize_fields": ([P, P, P, P, P], I),
    "wake_domain_add_noise": ([P, F, U64], I),
    "wake_domain_set_populations": ([P, P], I),
    "wake_domain_set_cell_type": ([P, I, I, I, I], I),
    "wake_domain_get_cell_type": ([P, I, I, I], I),
    "wake_domain_set_periodic": ([P, I], I),
    "wake_domain_set_periodic_axes": ([P, I, I], I),
    "wake_domain_set_solid_box": ([P, I, I, I, I, I, I], I),
    "wake_domain_set_solid_sphere": ([P, F, F, F, F], I),
    "wake_domain_set_solid_ellipsoid": ([P, F, F, F, F, F, F, P], I),
    "wake_domain_set_solid_naca": ([P, I, F, F, F, F, I, I], I),
    "wake_domain_voxelize_mesh": ([P, P, LL, I], I),
    "wake_domain_is_periodic": ([P, P], I),
    "wake_domain_set_buffer_addressing": ([P, I], I),
    "wake_domain_export_vti": ([P, ctypes.c_char_p, I, I], I),
    "wake_domain_set_solid_circle": ([P, F, F, F], I),
    "wake_domain_set_solid_cylinder_z": ([P, F, F, F], I),
    "wake_domain_set_moving_wall_top": ([P, F, F], I),
    "wake_domain_set_inflow_west": ([P, F, F], I),
    "wake_domain_set_inflow_parabolic": ([P, F], I),
    "wake_domain_set_inflow_turbulence": ([P, F, U64], I),
    "wake_domain_set_outflow_east": ([P, F], I),
    "wake_domain_set_convective_outflow_east": ([P, F], I),
    "wake_domain_set_les_coupled": ([P, I], I),
    "wake_domain_set_relaxation_time": ([P, F], I),
    "wake_domain_enable_thermal": ([P, F, F, F, F], I),
    "wake_domain_set_temperature_field": ([P, P], I),
    "wake_domain_set_wall_temperature": ([P, I, I, I, I, F], I),
    "wake_domain_build_curved_sphere_links": ([P, F, F, F, F, P], I),
    "wake_domain_get_links": ([P, P, P, P], I),
    "wake_domain_step": ([P, I, F, F, F, I, F], I),
    "wake_domain_run": ([P, I, I, F, F, F, I, F, P], I),
    "wake_domain_synchronize": ([P], I),
    "wake_domain_kernel_time": ([P, P], I),
    "wake_domain_read": ([P, I, P], I),
    "wake_domain_read_cell_types": ([P, P], I),
    "wake_domain_update_macroscopic": ([P], I),
    "wake_domain_diagnostics": ([P, P], I),
    "wake_domain_momentum_exchange": ([P, F, F, F, F, P], I),
    "wake_domain_set_force_probe": ([P, F, F, F, F], I),
    "wake_domain_clear_force_probe": ([P], I),
    "wake_domain_force_history": ([P, P, LL, P, P], I),
    "wake_domain_nonfinite_seen": ([P, P], I),
    "wake_domain_count_nonfinite": ([P, P, P], I),
    "wake_domain_info": ([P, P], I),
    "wake_domain_set_block_x": ([P, I], I),
    "wake_domain_set_multistep": ([P, I], I),
    "wake_hip_last_error": ([], ctypes.c_char_p),
    "wake_hip_clear_error": ([], I),
    "wake_hip_device_count": ([], I),
    "wake_hip_d3q19_minv": ([P], None),
    "wake_hip_hbm_probe": ([I, I, I, I, I, P], I),
}

# Part 2: the reference's D2Q9 temperature solver
THERMAL_SIGS = {
    "wake_thermal2d_create": ([I, I, F, F, F, P], I),
    "wake_thermal2d_destroy": ([P], I),
    "wake_thermal2d_initialize": ([P, F], I),
    "wake_thermal2d_set_walls": ([P, I, F, I, F], I),
    "wake_thermal2d_step": ([P, P, F], I),
    "wake_thermal2d_read": ([P, I, P], I),
    "wake_thermal2d_stats": ([P, P], I),
}

# Part 3: batches of domains
BATCH_SIGS = {
    "wake_batch_create": ([I, I, I, I, I, I, P], I),
    "wake_batch_destroy": ([P], None),
    "wake_batch_add": ([P, P, LL, F, F, F, F, I, F, F, F, F, P], I),
    "wake_batch_remove": ([P, I], I),
    "wake_batch_run": ([P, LL, I, P], I),
    "wake_batch_launch": ([P, LL, I], I),
    "wake_batch_wait": ([P, P], I),
    "wake_batch_remaining": ([P, I, P], I),
    "wake_batch_forces": ([P, I, P, LL, P], I),
    "wake_batch_tune": ([P, I, P, P], I),
    "wake_batch_set_launch": ([P, I, I], I),
    "wake_batch_info": ([P, P], I),
}

EXTRA_SIGS = {}


def lib():
    """Load libwake_hip.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python wake_amd/build.py` (HIP gfx950 build of the LBM kernels)")
    L = ctypes.CDLL(LIB_PATH)
    for name, sig in KERNEL_SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = [_CT[c] for c in sig]
        fn.restype = None
    for name, (args, res) in {**DOMAIN_SIGS, **THERMAL_SIGS, **BATCH_SIGS, **EXTRA_SIGS}.items():
        if not hasattr(L, name):
            continue
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error():
    return lib().wake_hip_last_error().decode()


def check(rc, what=""):
    if rc != 0:
        msg = last_error()
        lib().wake_hip_clear_error()
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("arrays passed to the HIP library must be C-contiguous")
        return a.ctypes.data_as(ctypes.c_void_p)
    return a


def kernel(name, *args):
    """Call a Part 1 kernel export with numpy arrays / scalars; raises if the
    GPU runtime reported a failure."""
    L = lib()
    fn = getattr(L, name)
    L.wake_hip_clear_error()
    fn(*[ptr(a) if isinstance(a, np.ndarray) or a is None else a for a in args])
    msg = L.wake_hip_last_error().decode()
    if L.wake_hip_clear_error() != 0:
        raise RuntimeError(f"{name}: {msg}")
