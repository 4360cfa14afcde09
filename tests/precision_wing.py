"""Precision study of the D3Q27 wing channel: FP32, FP16 and BF16 population
storage on the GPU against the FP64 oracle, on the same flow at reduced size.

    python tests/precision_wing.py [--scale 4] [--steps 5000]

Flow: NACA 0012 at 8 deg, D3Q27 TRT (Lambda = 1/4) + coupled LES, Re 2000 on
the chord, inlet u = 0.04, convective outlet, walls at y = 0 and ny - 1,
periodic span -- the 384x192x192 wing workload scaled down by --scale so the
FP64 oracle finishes on the host.  Prints one JSON line: the FP64 reference
C_D, C_L and Strouhal number over the second half of the run, and per
storage the velocity relative L2 error at checkpoints (1/8, 1/4, 1/2 and the
whole run) and the relative errors of the force statistics.

A script, not a test (pytest collects test_*.py only): at the default 5000
steps the oracle takes minutes.  The GPU runs use the batched path with the
same step kernels as a plain grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


def log(msg):
    print(f"[precision_wing] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=4)
    ap.add_argument("--steps", type=int, default=5000)
    args = ap.parse_args()

    import oracle_bind as ob

    import wake_amd
    from bench import wing_setup
    from wake_amd import analysis

    u, re = 0.04, 2000.0
    names = {0: "fp32", 1: "fp16", 2: "bf16"}
    k = args.scale
    sx, sy, sz, sc = 384 // k, 192 // k, 192 // k, 96.0 / k
    stau = 3.0 * u * sc / re + 0.5
    steps, every = args.steps, 5
    area = 0.5 * u * u * sc * sz

    def coeffs(F):
        return F[:, 0] / area, F[:, 1] / area

    # field errors while the runs are still correlated, force statistics
    # over the whole history
    checks = sorted({c for c in (steps // 8, steps // 4, steps // 2, steps) if c > 0})
    res, fields = {}, {}
    fl = None
    for st in (0, 1, 2):
        g = wing_setup(wake_amd.LatticeGrid3D(sx, sy, sz, stau, lattice=27, storage=st), sx, sy, sc)
        b = wake_amd.Batch(lattice=27, storage=st, collision="TRT", use_les=True, chunk=100)
        b.add(g, steps, smag=0.17, forces=(0.0, 0.0, 0.0, 0.0))
        done = 0
        for c in checks:
            b.run(c - done)
            done = c
            fl = g.cell_type_data() == 0
            fields[(names[st], c)] = np.concatenate([g.ux_data()[fl], g.uy_data()[fl], g.uz_data()[fl]])
        F = b.forces(0)[every - 1::every]
        res[names[st]] = coeffs(F)
        del b, g
    log(f"{sx}x{sy}x{sz}: GPU runs done; FP64 oracle {steps} steps on the host")
    t0 = time.perf_counter()
    o = wing_setup(ob.OracleGrid(3, sx, sy, sz, tau=stau, lattice=27, bits=64), sx, sy, sc)
    hist, ref_fields = [], {}
    for i in range(steps // every):
        o.run(every, "TRT", use_les=True, smag=0.17)
        hist.append(o.momentum_exchange(0.0, 0.0, 0.0, 0.0))
        if (i + 1) * every in checks:
            ref_fields[(i + 1) * every] = np.concatenate([o.ux()[fl], o.uy()[fl], o.uz()[fl]])
    log(f"FP64 oracle done in {time.perf_counter() - t0:.0f} s")
    ref_cd, ref_cl = coeffs(np.array(hist))
    half = len(ref_cd) // 2

    def st_of(cl):
        return analysis.strouhal_number(analysis.analyze_frequency(cl[half:].astype(np.float32), float(every)), sc, u)

    ref = {"C_D": float(np.mean(ref_cd[half:])), "C_L": float(np.mean(ref_cl[half:])), "St": st_of(ref_cl)}
    errors = {}
    for nm, (cd, cl) in res.items():
        v = {"C_D": float(np.mean(cd[half:])), "C_L": float(np.mean(cl[half:])), "St": st_of(cl)}
        l2 = {str(c): float(np.linalg.norm(fields[(nm, c)] - ref_fields[c]) / np.linalg.norm(ref_fields[c]))
              for c in checks}
        errors[nm] = {"velocity_rel_l2_at_step": l2,
                      **{f"{q}_rel": (abs(v[q] / ref[q] - 1) if ref[q] else None) for q in v}, **v}
        log(f"{nm} vs FP64: {errors[nm]}")
    print(json.dumps({"size": [sx, sy, sz], "chord": sc, "steps": steps, "fp64_reference": ref, "errors": errors}),
          flush=True)


if __name__ == "__main__":
    main()
