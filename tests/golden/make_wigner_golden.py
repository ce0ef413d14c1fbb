"""Write tests/golden/wigner_mpmath.json: high-precision Wigner d values.

The values come from the explicit finite sum for d^j_{m'm}(beta) (Wigner's
formula, e.g. Edmonds 1957 eq. 4.1.15 / the Wikipedia "Wigner D-matrix"
article), evaluated in mpmath at enough decimal digits to make the
alternating sum exact to double precision.  GSHTrans's d^l_{mn}(theta)
(GSHTrans/src/Wigner.h:52-95, first lower index = order m, second = upper
index n) is that d^j_{m'm} with j = l, m' = m, m = n; its Ortho
normalisation multiplies by sqrt(2l+1)/(2 sqrt(pi)) (Wigner.h:475-485).

The file is independent of both the reference code and the oracle: it pins
the oracle's recurrence (tests/test_oracle_wigner.py) at sampled points up
to l = 1024, including Gauss-Legendre colatitudes of the L = 1024 grid.

Run:  python tests/golden/make_wigner_golden.py
"""
import json
import os
import random

import mpmath as mp
import numpy as np


def wigner_d(l, m, n, theta):
    """d^l_{m n}(theta) by Wigner's explicit sum (mpmath, high precision)."""
    mp.mp.dps = int(40 + 0.65 * l)
    th = mp.mpf(theta)
    c = mp.cos(th / 2)
    s = mp.sin(th / 2)
    # d^j_{m'm}: j=l, m'=m, m=n
    j, mp_, mm = l, m, n
    pref = mp.sqrt(mp.factorial(j + mp_) * mp.factorial(j - mp_) *
                   mp.factorial(j + mm) * mp.factorial(j - mm))
    smin = max(0, mm - mp_)
    smax = min(j + mm, j - mp_)
    tot = mp.mpf(0)
    for k in range(smin, smax + 1):
        num = (-1) ** (mp_ - mm + k) * c ** (2 * j + mm - mp_ - 2 * k) * s ** (mp_ - mm + 2 * k)
        den = (mp.factorial(j + mm - k) * mp.factorial(k) *
               mp.factorial(mp_ - mm + k) * mp.factorial(j - mp_ - k))
        tot += num / den
    return pref * tot


def ortho(l, d):
    return d * mp.sqrt(2 * l + 1) / (2 * mp.sqrt(mp.pi))


def main():
    rng = random.Random(20261202)
    x, _ = np.polynomial.legendre.leggauss(1025)
    gl1024 = np.arccos(-x)
    samples = []
    # Structured points at L=1024 Gauss-Legendre colatitudes.
    for it in (0, 1, 7, 300, 512, 1024):
        th = float(gl1024[it])
        for (l, m, n) in [(1024, 0, 0), (1024, 3, 2), (1024, -5, 2), (1024, 1000, 0),
                          (1024, -1024, 2), (1024, 1023, -2), (1000, 999, 1),
                          (700, -2, -2), (513, 200, 1), (2, 1, 2), (2, -2, -2)]:
            samples.append((l, m, n, th))
    # Random points.
    for _ in range(60):
        l = rng.choice([rng.randint(0, 20), rng.randint(20, 300), rng.randint(300, 1024)])
        n = rng.randint(-min(4, l), min(4, l))
        m = rng.randint(-l, l)
        th = rng.uniform(0.0, np.pi)
        samples.append((l, m, n, th))
    out = []
    for (l, m, n, th) in samples:
        d = wigner_d(l, m, n, th)
        out.append({"l": l, "m": m, "n": n, "theta": th,
                    "d": float(d), "D_ortho": float(ortho(l, d))})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wigner_mpmath.json")
    with open(path, "w") as fh:
        json.dump({"source": "mpmath Wigner explicit sum", "samples": out}, fh, indent=0)
    print("wrote", path, len(out), "samples")


if __name__ == "__main__":
    main()
