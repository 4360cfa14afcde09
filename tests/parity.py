"""Shared parity criteria of the GPU tests.

FIELD_TOL is the 1e-5 relative-L2 bar on density and velocity.  Some flows
amplify rounding (Zou-He boundaries feed the inlet error back every step): the
FP32 oracle itself then ends further than the bar from the FP64 oracle.  On
such a flow the GPU is held to half that FP32 floor against the FP32 oracle
and to 1.25 floors against the FP64 oracle: two independent FP32 roundings
at most half a floor apart put the GPU within sqrt(1 + 1/4) = 1.12 floors of
FP64.
"""
import numpy as np

FIELD_TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def check_fp32_parity(gpu, o32, o64, what="u", o32fma=None):
    """Assert the bar, or the FP32-floor criterion where the floor exceeds it.

    With o32fma (the FP32 oracle built with fused multiply-adds, as the
    GPU compiler and the reference's fast-math ISPC build contract them) the
    floor is the larger distance of the two FP32 builds from FP64, and the
    GPU may sit as far from the FP32 oracle as 1.25x the spread between the
    two FP32 builds: both are faithful FP32 evaluations of the same step."""
    floor = rel_l2(o32, o64)
    spread = 0.0
    if o32fma is not None:
        floor = max(floor, rel_l2(o32fma, o64))
        spread = rel_l2(o32fma, o32)
    e32, e64 = rel_l2(gpu, o32), rel_l2(gpu, o64)
    print(f"{what}: GPU vs FP32 oracle {e32:.2e}, GPU vs FP64 {e64:.2e}, FP32 floor {floor:.2e}, "
          f"FMA spread {spread:.2e}")
    assert e32 <= FIELD_TOL or (e32 <= max(0.5 * floor, 1.25 * spread) and e64 <= 1.25 * floor), \
        f"{what}: GPU vs FP32 {e32:.2e}, vs FP64 {e64:.2e}, floor {floor:.2e}, spread {spread:.2e}"
    return e32, e64, floor
