"""Boys-function fixtures F_m(T) = int_0^1 u^(2m) exp(-T u^2) du from the
definition, by mpmath quadrature at 40 significant digits (independent of any
code in this repository).  Writes tests/golden/boys_points.json.

    python tests/golden/gen_boys_fixtures.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
TS = [0.0, 1e-12, 1e-6, 0.01, 0.3, 1.0, 2.5, 7.0, 12.0, 19.5, 25.0, 29.34071349, 33.0, 44.9, 45.1, 60.0, 120.0,
      800.0]
MS = list(range(0, 13))


def boys(m, t):
    t = mp.mpf(t)
    if t == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.quad(lambda u: u ** (2 * m) * mp.e ** (-t * u * u), [0, mp.mpf(1) / 4, mp.mpf(1) / 2, 1])


def main():
    recs = [{"m": m, "T": t, "F": mp.nstr(boys(m, t), 30)} for t in TS for m in MS]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_points.json")
    with open(path, "w") as fh:
        json.dump(recs, fh, indent=0)
    print(f"{len(recs)} records -> {path}")


if __name__ == "__main__":
    main()
