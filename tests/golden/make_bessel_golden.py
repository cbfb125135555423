"""Write tests/golden/besselkx.npz: e^x K_nu(x) for nu = 0..5 at 30 significant
digits (mpmath), on log-spaced x in [1e-3, 1e5] plus the branch points 1 and 1e4.
These pin the oracle's and the kernels' scaled Bessel K, which stand in for
SpecialFunctions.besselkx (AMOS, not available here) at the reference's call
sites src/eos.jl:29 and src2d/transport2d.jl:23-105.
"""
import os

import mpmath
import numpy as np

mpmath.mp.dps = 30
xs = np.unique(np.concatenate([np.logspace(-3, 5, 241), [1.0, np.nextafter(1.0, 2), 1e4,
                                                         np.nextafter(1e4, 0)]]))
vals = np.empty((6, xs.size))
for nu in range(6):
    for j, x in enumerate(xs):
        xm = mpmath.mpf(float(x))
        vals[nu, j] = float(mpmath.besselk(nu, xm) * mpmath.exp(xm))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "besselkx.npz")
np.savez(out, x=xs, kx=vals)
print(out, xs.size)
