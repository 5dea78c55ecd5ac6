"""Writes tables_mpmath.json: 1D Gauss-Legendre rules on [0, 1] (q = 1..13)
and Lagrange tables B, Dm on the reference's node order [0, 1, 1/p, ...]
(p = 1..9), computed with mpmath at 50 digits -- the golden vectors for the
1D tables of SumFactorizationMap (src/SumFactorizationMaps.jl:50-55).

Gauss nodes: Newton on the Legendre polynomial P_q (mpmath.legendre) from
mpmath's own float guesses; weights 2 / ((1 - z^2) P_q'(z)^2).  Lagrange
values and derivatives: product formula in exact rational nodes.

Run:  python tests/golden/make_tables_mpmath.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50


def legendre_and_derivative(q, z):
    p0, p1 = mp.mpf(1), z
    if q == 0:
        return p0, mp.mpf(0)
    for k in range(2, q + 1):
        p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
    dp = q * (z * p1 - p0) / (z * z - 1)
    return p1, dp


def gauss01(q):
    pts, wts = [], []
    for i in range(q):
        z = mp.cos(mp.pi * (i + mp.mpf(3) / 4) / (q + mp.mpf(1) / 2))
        for _ in range(200):
            p, dp = legendre_and_derivative(q, z)
            dz = p / dp
            z -= dz
            if abs(dz) < mp.mpf(10) ** -45:
                break
        p, dp = legendre_and_derivative(q, z)
        w = 2 / ((1 - z * z) * dp * dp)
        pts.append((z + 1) / 2)
        wts.append(w / 2)
    order = sorted(range(q), key=lambda k: pts[k])
    return [pts[k] for k in order], [wts[k] for k in order]


def nodes(p):
    return [mp.mpf(0), mp.mpf(1)] + [mp.mpf(k) / p for k in range(1, p)]


def lagrange(p, x):
    t = nodes(p)
    n = p + 1
    vals, ders = [], []
    for j in range(n):
        den = mp.mpf(1)
        for m in range(n):
            if m != j:
                den *= (t[j] - t[m])
        val = mp.mpf(1)
        for m in range(n):
            if m != j:
                val *= (x - t[m])
        der = mp.mpf(0)
        for k in range(n):
            if k == j:
                continue
            pr = mp.mpf(1)
            for m in range(n):
                if m != j and m != k:
                    pr *= (x - t[m])
            der += pr
        vals.append(val / den)
        ders.append(der / den)
    return vals, ders


def s(v):
    return mp.nstr(v, 25, strip_zeros=False)


def main():
    out = {"gauss": {}, "lagrange": {}}
    for q in range(1, 14):
        pts, wts = gauss01(q)
        out["gauss"][str(q)] = {"points": [s(v) for v in pts], "weights": [s(v) for v in wts]}
    for p in range(1, 10):
        for q in range(max(1, p), p + 5):
            pts, _ = gauss01(q)
            B, Dm = [], []
            for x in pts:
                v, d = lagrange(p, x)
                B.append([s(a) for a in v])
                Dm.append([s(a) for a in d])
            out["lagrange"]["%d_%d" % (p, q)] = {"B": B, "Dm": Dm}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables_mpmath.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
