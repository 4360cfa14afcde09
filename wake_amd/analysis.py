"""Host-side force analysis: aerodynamic coefficients and Strouhal number.

Both run on the CPU over the force histories the GPU records, as in the
reference:

* aerodynamic_coefficients -- include/wake/forces/momentum_exchange.hpp:20-37:
  C = F / (1/2 rho u^2 A), C_m = T / (1/2 rho u^2 A c), zero when the dynamic
  pressure or the area is at most 1e-8 (FP32).
* analyze_frequency / strouhal_number -- src/forces/strouhal_analyzer.cpp:39-89:
  fewer than 16 samples or dt <= 0 gives 0; mean removed; length padded to
  the next power of two, capped at 4096, and only the first min(n, len)
  samples are used; Hann window over those samples; the peak of |X_k|^2 over
  bins 1 .. n/2 - 1 (first maximum); f = k / (n dt); St = f D / U (0 when
  U <= 1e-8).  The transform is numpy's FFT instead of the reference's
  radix-2 loop, the same DFT up to rounding, so the chosen bin can differ
  only where two bins hold equal power to rounding.
"""
import numpy as np


def aerodynamic_coefficients(fx, fy, fz, torque_z, rho_inf, u_inf, area, chord=1.0):
    f32 = np.float32
    q = f32(0.5) * f32(rho_inf) * f32(u_inf) * f32(u_inf)
    out = {"fx": f32(fx), "fy": f32(fy), "fz": f32(fz), "torque_z": f32(torque_z),
           "cd": f32(0), "cl": f32(0), "cm": f32(0)}
    if q > f32(1e-8) and f32(area) > f32(1e-8):
        out["cd"] = f32(fx) / (q * f32(area))
        out["cl"] = f32(fy) / (q * f32(area))
        out["cm"] = f32(torque_z) / (q * f32(area) * f32(chord))
    return out


def analyze_frequency(signal, dt):
    """Dominant frequency of `signal` sampled every dt (first maximum of the
    one-sided power spectrum, DC excluded)."""
    sig = np.asarray(signal, np.float32)
    if sig.size < 16 or dt <= 0:
        return 0.0
    mean = np.float32(sig.sum(dtype=np.float32) / np.float32(sig.size))
    n = 1
    while n < sig.size:
        n <<= 1
    n = min(n, 4096)
    m = min(n, sig.size)
    i = np.arange(m, dtype=np.float32)
    window = np.float32(0.5) * (np.float32(1) - np.cos(np.float32(2 * np.pi) * i / np.float32(m - 1)))
    a = np.zeros(n, np.float64)
    a[:m] = ((sig[:m] - mean) * window).astype(np.float64)
    power = np.abs(np.fft.fft(a)[: n // 2]) ** 2
    k = 1 + int(np.argmax(power[1:])) if n // 2 > 1 else 1
    df = np.float32(1.0) / (np.float32(n) * np.float32(dt))
    return float(np.float32(k) * df)


def strouhal_number(f_shed, diameter, u_inf):
    if u_inf <= 1e-8:
        return 0.0
    return float(np.float32(f_shed) * np.float32(diameter) / np.float32(u_inf))
