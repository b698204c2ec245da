"""Write tests/golden/xc_points.json: exchange-correlation energy densities and
their first partial derivatives at fixed spin-density points, in 50-digit
arithmetic (mpmath), from the published formulas:

  Dirac/Slater exchange;  Vosko, Wilk, Nusair, Can. J. Phys. 58, 1200 (1980),
  VWN5 and VWN-RPA fits;  Becke, Phys. Rev. A 38, 3098 (1988), beta = 0.0042;
  Lee, Yang, Parr (1988) in the Miehlich, Savin, Stoll, Preuss form, Chem.
  Phys. Lett. 157, 200 (1989), eq. 2;  Perdew, Burke, Ernzerhof, PRL 77, 3865
  (1996) on Perdew-Wang 92 correlation (PRB 45, 13244).
Mixing coefficients and fit constants are the ones CalChem uses
(crates/calchem-dft/src/xc.rs:159-223, 257-310, 566-573, 642-667).  Only the
energy density is transcribed; derivatives come from mpmath.diff, so the
fixtures are independent of any hand-derived derivative formula.

Run:  python tests/golden/gen_xc_fixtures.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50
PI = mp.pi
THIRD = mp.mpf(1) / 3


def slater(ra, rb, gaa, gab, gbb):
    c = mp.mpf(3) / 4 * mp.cbrt(6 / PI)
    return -c * (ra ** (4 * THIRD) + rb ** (4 * THIRD))


def _fz(z):
    return ((1 + z) ** (4 * THIRD) + (1 - z) ** (4 * THIRD) - 2) / (2 ** (4 * THIRD) - 2)


def _vwn_eps(x, A, x0, b, c):
    X = lambda t: t * t + b * t + c
    Q = mp.sqrt(4 * c - b * b)
    at = mp.atan(Q / (2 * x + b))
    return A * (mp.log(x * x / X(x)) + 2 * b / Q * at
                - b * x0 / X(x0) * (mp.log((x - x0) ** 2 / X(x)) + 2 * (b + 2 * x0) / Q * at))


VWN5 = dict(P=(mp.mpf("0.0310907"), mp.mpf("-0.10498"), mp.mpf("3.72744"), mp.mpf("12.9352")),
            F=(mp.mpf("0.01554535"), mp.mpf("-0.32500"), mp.mpf("7.06042"), mp.mpf("18.0578")),
            A=(-1 / (6 * PI ** 2), mp.mpf("-0.00475840"), mp.mpf("1.13107"), mp.mpf("13.0045")))
VWNRPA = dict(P=(mp.mpf("0.0310907"), mp.mpf("-0.409286"), mp.mpf("13.0720"), mp.mpf("42.7198")),
              F=(mp.mpf("0.01554535"), mp.mpf("-0.743294"), mp.mpf("20.1231"), mp.mpf("101.578")),
              A=(-1 / (6 * PI ** 2), mp.mpf("-0.228344"), mp.mpf("1.06835"), mp.mpf("11.4813")))


def vwn(pars):
    def f(ra, rb, gaa, gab, gbb):
        rho = ra + rb
        rs = mp.cbrt(3 / (4 * PI * rho))
        x = mp.sqrt(rs)
        z = (ra - rb) / rho
        eP, eF, eA = (_vwn_eps(x, *pars[k]) for k in "PFA")
        fdd0 = 4 / (9 * (mp.cbrt(2) - 1))
        ec = eP + eA * _fz(z) / fdd0 * (1 - z ** 4) + (eF - eP) * _fz(z) * z ** 4
        return rho * ec
    return f


def b88(ra, rb, gaa, gab, gbb):
    beta = mp.mpf("0.0042")
    tot = 0
    for r, g in ((ra, gaa), (rb, gbb)):
        if r == 0:
            continue
        x = mp.sqrt(g) / r ** (4 * THIRD)
        tot += -beta * r ** (4 * THIRD) * x * x / (1 + 6 * beta * x * mp.asinh(x))
    return tot


def lyp(ra, rb, gaa, gab, gbb):
    a, b, c, d = mp.mpf("0.04918"), mp.mpf("0.132"), mp.mpf("0.2533"), mp.mpf("0.349")
    rho = ra + rb
    g = gaa + 2 * gab + gbb
    rm = rho ** (-THIRD)
    om = mp.exp(-c * rm) / (1 + d * rm) * rho ** (-11 * THIRD)
    de = c * rm + d * rm / (1 + d * rm)
    CF = mp.mpf(3) / 10 * (3 * PI ** 2) ** (2 * THIRD)
    brace = (ra * rb * (2 ** (11 * THIRD) * CF * (ra ** (8 * THIRD) + rb ** (8 * THIRD))
                        + (mp.mpf(47) / 18 - 7 * de / 18) * g
                        - (mp.mpf(5) / 2 - de / 18) * (gaa + gbb)
                        - (de - 11) / 9 * (ra / rho * gaa + rb / rho * gbb))
             - mp.mpf(2) / 3 * rho ** 2 * g
             + (mp.mpf(2) / 3 * rho ** 2 - ra ** 2) * gbb
             + (mp.mpf(2) / 3 * rho ** 2 - rb ** 2) * gaa)
    return -a * 4 / (1 + d * rm) * ra * rb / rho - a * b * om * brace


def pbe_x(ra, rb, gaa, gab, gbb):
    kappa, mu = mp.mpf("0.804"), mp.mpf("0.2195149727645171")
    tot = 0
    for r, gs in ((ra, gaa), (rb, gbb)):
        if r == 0:
            continue
        n, gn = 2 * r, 4 * gs
        kf = mp.cbrt(3 * PI ** 2 * n)
        s2 = gn / (4 * kf ** 2 * n ** 2)
        fx = 1 + kappa - kappa / (1 + mu * s2 / kappa)
        tot += mp.mpf(1) / 2 * (-mp.mpf(3) / 4 * mp.cbrt(3 / PI) * n ** (4 * THIRD)) * fx
    return tot


def _pw(rs, A, a1, b1, b2, b3, b4):
    x = mp.sqrt(rs)
    return -2 * A * (1 + a1 * rs) * mp.log(1 + 1 / (2 * A * (b1 * x + b2 * rs + b3 * x ** 3 + b4 * rs ** 2)))


PW0 = tuple(mp.mpf(s) for s in ("0.0310907", "0.21370", "7.5957", "3.5876", "1.6382", "0.49294"))
PW1 = tuple(mp.mpf(s) for s in ("0.01554535", "0.20548", "14.1189", "6.1977", "3.3662", "0.62517"))
PWA = tuple(mp.mpf(s) for s in ("0.0168869", "0.11125", "10.357", "3.6231", "0.88026", "0.49671"))


def pbe_c(ra, rb, gaa, gab, gbb):
    beta, gam = mp.mpf("0.06672455060314922"), (1 - mp.log(2)) / PI ** 2
    rho = ra + rb
    z = (ra - rb) / rho
    rs = mp.cbrt(3 / (4 * PI * rho))
    fdd0 = 4 / (9 * (mp.cbrt(2) - 1))
    e0, e1, ac = _pw(rs, *PW0), _pw(rs, *PW1), -_pw(rs, *PWA)
    ec = e0 + ac * _fz(z) / fdd0 * (1 - z ** 4) + (e1 - e0) * _fz(z) * z ** 4
    g = gaa + 2 * gab + gbb
    phi = ((1 + z) ** (2 * THIRD) + (1 - z) ** (2 * THIRD)) / 2
    kf = mp.cbrt(3 * PI ** 2 * rho)
    ks = mp.sqrt(4 * kf / PI)
    t2 = g / (4 * phi ** 2 * ks ** 2 * rho ** 2)
    A = beta / gam / (mp.exp(-ec / (gam * phi ** 3)) - 1)
    H = gam * phi ** 3 * mp.log(1 + beta / gam * t2 * (1 + A * t2) / (1 + A * t2 + A ** 2 * t2 ** 2))
    return rho * (ec + H)


def combo(*terms):
    return lambda *a: sum(c * f(*a) for c, f in terms)


FUNCS = {
    "svwn5": combo((1, slater), (1, vwn(VWN5))),
    "blyp": combo((1, slater), (1, b88), (1, lyp)),
    "pbe": combo((1, pbe_x), (1, pbe_c)),
    "b3lyp": combo((mp.mpf("0.80"), slater), (mp.mpf("0.72"), b88), (mp.mpf("0.19"), vwn(VWNRPA)),
                   (mp.mpf("0.81"), lyp)),
    "b3lyp5": combo((mp.mpf("0.80"), slater), (mp.mpf("0.72"), b88), (mp.mpf("0.19"), vwn(VWN5)),
                    (mp.mpf("0.81"), lyp)),
    "pbe0": combo((mp.mpf("0.75"), pbe_x), (1, pbe_c)),
}

# The five points of crates/calchem-dft/src/xc.rs:1228-1236 plus a closed-shell
# core and a closed-shell tail point.
POINTS = [
    ("0.37", "0.19", "0.41", "0.12", "0.23"),
    ("0.25", "0.25", "0.18", "0.18", "0.18"),
    ("14.2", "13.8", "120.0", "110.0", "105.0"),
    ("1.2e-4", "0.7e-4", "3.5e-7", "1.1e-7", "2.0e-7"),
    ("0.42", "0", "0.31", "0", "0"),
    ("150.0", "150.0", "4.0e5", "4.0e5", "4.0e5"),
    ("3.0e-6", "3.0e-6", "2.0e-11", "2.0e-11", "2.0e-11"),
]


def main():
    out = []
    for name, f in FUNCS.items():
        for p in POINTS:
            x = [mp.mpf(s) for s in p]
            rec = {"func": name, "point": [float(v) for v in x], "f": mp.nstr(f(*x), 25)}
            polarized = x[1] == 0
            partials = []
            for k in range(5):
                if polarized and k in (1, 3, 4):
                    partials.append(None)  # one-sided at rho_b = 0: not fixed here
                    continue
                h = max(abs(x[k]) * mp.mpf("1e-12"), mp.mpf("1e-30"))
                partials.append(mp.nstr(mp.diff(lambda t: f(*(x[:k] + [t] + x[k + 1:])), x[k], h=h), 25))
            rec["d"] = partials
            out.append(rec)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "xc_points.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {len(out)} records to {path}")


if __name__ == "__main__":
    main()
