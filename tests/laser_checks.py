"""The laser envelope's vacuum-diffraction validation (validation/run_validation.sh:194-207, checked by
validation/laser_check.py:18-29), restated for the tests: the deck validation/laser_vacuum.in, the laser.txt rows
(Simulation::write_laser_diag, src/Simulation.cpp:1301-1314, printed with 10 significant digits) and the printed
line the reference's validation run holds (validation/reference_quick.txt:59-60 and 69-70)."""
import io

import numpy as np

# validation/laser_vacuum.in: uniform grid 0.05 to r = 14, xi from -10 to 0 by 0.05 (201 slices; internal xi from 0),
# dt 2 and 40 steps (41 diagnostic rows, steps 0..40), k0 20, a0 0.1, w0 2, L0 1.5, xi0 -5 (internal 5), focus 40,
# plasma density 1e-9 (its susceptibility, uniform here)
VAC = dict(R=14.0, dr=0.05, K=201, xi_min=0.0, dxi=0.05, dt=2.0, rows=41, k0=20.0, a0=0.1, w0=2.0, L0=1.5, xi0=5.0,
           focus=40.0, chi=1e-9)

# validation/reference_quick.txt:60 (BDF2, the default) and :70 (trapezoid)
PRINTED = {
    "bdf2": "spot size 7.1e-04, peak amplitude 7.0e-04, int |a|^2 conserved to 1.9e-06; centroid slip 0.0502 "
            "(group velocity 1 - 1/(k0 w0)^2: 0.0500)",
    "trapezoid": "spot size 7.0e-04, peak amplitude 7.0e-04, int |a|^2 conserved to 0.0e+00; centroid slip 0.0501 "
                 "(group velocity 1 - 1/(k0 w0)^2: 0.0500)",
}


def diag_row(n, t, S):
    """One laser.txt row of a one-rank run from Laser::sums S (src/Simulation.cpp:1301-1308)."""
    e = S[0] if S[0] > 0 else 1.0
    xc = S[1] / e
    L = np.sqrt(max(0.0, S[4] / e - xc * xc))
    w = np.sqrt(2 * S[2] / e)
    return f"{n} " + " ".join("%.10g" % v for v in (t, S[3], w, xc, L, S[0]))


def vacuum_line(rows, k0=VAC["k0"], w0=VAC["w0"], zf=VAC["focus"], a0=VAC["a0"]):
    """validation/laser_check.py:18-29 on the laser.txt rows."""
    a = np.loadtxt(io.StringIO("\n".join(rows)))
    t = a[:, 1]
    zR = k0 * w0 ** 2 / 2
    s = (t - zf) / zR
    w_th = w0 * np.sqrt(1 + s ** 2)
    a_th = a0 / np.sqrt(1 + s ** 2)
    slip = a[0, 4] - a[-1, 4]
    return (f"spot size {np.abs(a[:, 3] / w_th - 1).max():.1e}, peak amplitude {np.abs(a[:, 2] / a_th - 1).max():.1e}, "
            f"int |a|^2 conserved to {np.abs(a[:, 6] / a[0, 6] - 1).max():.1e}; centroid slip {slip:.4f} "
            f"(group velocity 1 - 1/(k0 w0)^2: {t[-1] / (k0 * w0) ** 2:.4f})")
