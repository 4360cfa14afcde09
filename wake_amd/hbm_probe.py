"""The practical HBM bandwidth ceilings of the step's access pattern on the
current GPU (wake_hip_hbm_probe), reported beside the step kernel.

The contiguous-copy probe (shape 0) also calibrates the HBM byte counters:
it reads and writes a known number of bytes per launch, so its FETCH_SIZE /
WRITE_SIZE in the same profiled run relate the counters to bytes moved.
"""
import ctypes

BYTES_PER_NODE = 152  # 19 floats read and written per node (FP32)


def ceiling_d2q9(nx=1024, ny=256, nz=128, reps=20):
    """The D2Q9 pull over 9 SoA planes of floats (72 bytes per node) on the
    current GPU, in GB/s, over nz independent nx x ny layers."""
    from . import _lib

    L = _lib.lib()
    ms = ctypes.c_float()
    if L.wake_hip_hbm_probe(5, nx, ny, nz, reps, ctypes.byref(ms)) != 0:
        raise RuntimeError(L.wake_hip_last_error().decode())
    return 72 * nx * ny * nz / (ms.value * 1e-3) / 1e9


def ceilings_d3q27(nx=384, ny=192, nz=192, reps=20):
    """The D3Q27 pull over 27 SoA planes on the current GPU, in GB/s:
    floats (216 bytes per node) and 16-bit values (108 bytes per node)."""
    from . import _lib

    L = _lib.lib()
    out = {}
    for shape, key, bpn in ((3, "pull_d3q27_gbps", 216), (4, "pull_d3q27_16bit_gbps", 108)):
        ms = ctypes.c_float()
        if L.wake_hip_hbm_probe(shape, nx, ny, nz, reps, ctypes.byref(ms)) != 0:
            raise RuntimeError(L.wake_hip_last_error().decode())
        out[key] = bpn * nx * ny * nz / (ms.value * 1e-3) / 1e9
    return out


def ceilings(nx=256, ny=256, nz=256, reps=20):
    """The practical HBM ceilings on the current GPU, in GB/s: a contiguous
    copy in the fastest copy shape and the D3Q19 pull over 19 SoA planes of
    floats, both at 152 bytes per node, and the same pull with 16-bit values
    at 76 bytes per node."""
    from . import _lib

    L = _lib.lib()
    out = {}
    for shape, key, bpn in ((0, "copy_contiguous_gbps", BYTES_PER_NODE), (1, "pull_d3q19_gbps", BYTES_PER_NODE),
                            (2, "pull_d3q19_16bit_gbps", BYTES_PER_NODE // 2)):
        ms = ctypes.c_float()
        if L.wake_hip_hbm_probe(shape, nx, ny, nz, reps, ctypes.byref(ms)) != 0:
            raise RuntimeError(L.wake_hip_last_error().decode())
        out[key] = bpn * nx * ny * nz / (ms.value * 1e-3) / 1e9
    return out
