"""HBM byte-counter calibration (one plain device copy of a known size),
and the practical bandwidth ceilings of the step's access pattern.

    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT -o copy -- python -m wake_amd.hbm_probe --mib 1024

The copy reads and writes exactly `--mib` MiB once (torch's elementwise copy
kernel, dispatched `--reps` times); the FETCH_SIZE / WRITE_SIZE it reports
per dispatch, divided by that size, is the factor by which the counters
relate to bytes moved on this GPU.
"""
import argparse
import ctypes

BYTES_PER_NODE = 152  # 19 floats read and written per node (FP32)


def ceilings(nx=256, ny=256, nz=256, reps=20):
    """The practical HBM ceilings on the current GPU, in GB/s
    (wake_hip_hbm_probe): a contiguous copy in the fastest copy shape and
    the D3Q19 pull over 19 SoA planes of floats, both at 152 bytes per node,
    and the same pull with 16-bit values at 76 bytes per node."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    import torch

    n = a.mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB x {a.reps}", flush=True)


if __name__ == "__main__":
    main()
