"""Writes boys_golden.json: F_m(T) for m = 0..40 at the T grid of the
reference's Boys test (tests/unit/test-core-ints.cc:28-47: T = 0 and
T = 10^(i/2)/1000, i = 0..11) plus points that exercise every regime of the
interpolation (interval edges, the T_crit = 117 switch, large T).

F_m(T) = gamma_lower(m + 1/2, T) / (2 T^(m + 1/2)), F_m(0) = 1/(2m + 1),
evaluated with mpmath at 60 significant digits and rounded to double."""
import json
import os

import mpmath as mp

mp.mp.dps = 60
MMAX = 40


def boys(m, T):
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    a = mp.mpf(m) + mp.mpf(1) / 2
    return mp.gammainc(a, 0, T) / (2 * T ** a)


def main():
    Ts = [0.0] + [float(mp.power(mp.sqrt(10), i) / 1000) for i in range(12)]
    Ts += [1.0 / 7.0, 3.0 / 7.0 + 1e-9, 0.5, 2.0, 7.3, 19.99, 33.3, 55.0, 80.0, 116.9, 117.0, 117.2, 150.0, 400.0,
           1.0e3, 3.0e4]
    out = {"T": Ts, "mmax": MMAX, "F": [[float(boys(m, T)) for m in range(MMAX + 1)] for T in Ts]}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "boys_golden.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
