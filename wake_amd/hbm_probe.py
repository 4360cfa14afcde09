"""HBM byte-counter calibration: one plain device copy of a known size.

    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT -o copy -- python -m wake_amd.hbm_probe --mib 1024

The copy reads and writes exactly `--mib` MiB once (torch's elementwise copy
kernel, dispatched `--reps` times); the FETCH_SIZE / WRITE_SIZE it reports
per dispatch, divided by that size, is the factor by which the counters
relate to bytes moved on this GPU.
"""
import argparse


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
