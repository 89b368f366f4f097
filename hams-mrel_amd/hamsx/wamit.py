"""WAMIT-style scaling of field-point values (the .6p pressure/elevation
files), as WamitNondimensMulti (src/PressureElevationMulti.f90:352-402)
applies it to the values CalPressureMulti / CalElevationMulti return.

Host-side formatting only: the values themselves come from the device
(``FrequencySolver.field_values``)."""
from __future__ import annotations

import numpy as np

G = 9.80665


def nondim_field(values, z, nbody: int, w1: float, amp: float = 1.0, rho: float = 1000.0, refl: float = 1.0,
                 limit: bool = False):
    """values: P x (6 nbody + nbeta) complex, of the output drivers' kind per
    point (pressure where |z| > 1e-6, elevation on z = 0, the AUTO kind);
    z: the P point depths.  Returns the scaled values the .6p files print:
    radiation columns conjugated after (-i w) amp / REFL^MEXP (MEXP = 0 for
    modes 1-3, 1 for 4-6), diffraction columns as (-Im, -Re); magnitudes below
    1e-15 become 0 first."""
    v = np.array(values, dtype=np.complex128)
    z = np.asarray(z, dtype=np.float64)
    pressure = np.abs(z) > np.float32(1.0e-6)
    nfac = np.where(pressure, rho * amp if limit else rho * G * amp, amp)[:, None]
    nd = 6 * nbody
    out = np.empty_like(v)
    for col in range(v.shape[1]):
        x = v[:, col] / nfac[:, 0]
        if col < nd:
            mexp = 0 if col % 6 < 3 else 1
            x = x * amp / refl ** (mexp + 1) if limit else x * (-1j * w1) * amp / refl ** mexp
        x = np.where(np.abs(x) < 1.0e-15, 0.0, x)
        out[:, col] = np.conj(x) if col < nd else (-x.imag) + 1j * (-x.real)
    return out
