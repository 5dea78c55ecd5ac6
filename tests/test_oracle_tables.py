"""The 1D tables of the oracle (Legendre-Vandermonde construction) and of
the library (product formula, sf_tables_1d -- host code, no GPU) against the
50-digit mpmath golden vectors, plus the quadrature and basis invariants."""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tables_mpmath.json")))
PQ = [tuple(int(v) for v in k.split("_")) for k in GOLD["lagrange"]]


def gold_gauss(q):
    g = GOLD["gauss"][str(q)]
    return np.array([float(v) for v in g["points"]]), np.array([float(v) for v in g["weights"]])


def table_tols(p, gD):
    """Rounding floor of a table evaluated at double-precision Gauss points:
    an ulp of the point times the slope of the entry (|phi'| for B, about
    p |phi'| for phi')."""
    s = float(np.max(np.abs(gD)))
    return 4e-16 * s + 1e-15, 4e-15 * s * max(1, p)


def gold_lagrange(p, q):
    g = GOLD["lagrange"]["%d_%d" % (p, q)]
    return (np.array([[float(v) for v in r] for r in g["B"]]),
            np.array([[float(v) for v in r] for r in g["Dm"]]))


@pytest.mark.parametrize("q", range(1, 14))
def test_gauss_points_oracle_vs_gold(brute, q):
    x, w = brute.gauss01(q)
    gx, gw = gold_gauss(q)
    assert np.max(np.abs(x - gx)) < 1e-15
    assert np.max(np.abs(w - gw)) < 1e-15


@pytest.mark.parametrize("q", range(1, 14))
def test_gauss_points_library_vs_gold(pkg, q):
    _, _, x, w = pkg.tables_1d(1, q)
    gx, gw = gold_gauss(q)
    assert np.max(np.abs(x - gx)) < 1e-15
    assert np.max(np.abs(w - gw)) < 2e-15


@pytest.mark.parametrize("q", range(1, 14))
def test_gauss_exactness(pkg, q):
    # exact for x^k, k <= 2q - 1 on [0, 1]
    _, _, x, w = pkg.tables_1d(1, q)
    for k in range(2 * q):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14
    # and for x^(2q) the error is the closed form of the Gauss remainder
    # int - sum = (q!)^4 / ((2q + 1) ((2q)!)^2) on [0, 1]
    from math import factorial
    E = factorial(q) ** 4 / ((2 * q + 1) * factorial(2 * q) ** 2)
    err = 1.0 / (2 * q + 1) - np.dot(w, x ** (2 * q))
    assert abs(err - E) < 2e-15 + 1e-6 * E


@pytest.mark.parametrize("p,q", PQ)
def test_lagrange_oracle_vs_gold(brute, p, q):
    B, Dm, _, _ = brute.tables(p, q)
    gB, gD = gold_lagrange(p, q)
    tB, tD = table_tols(p, gD)
    assert np.max(np.abs(B - gB)) < 10 * tB
    assert np.max(np.abs(Dm - gD)) < 10 * tD


@pytest.mark.parametrize("p,q", PQ)
def test_lagrange_library_vs_gold(pkg, p, q):
    B, Dm, _, _ = pkg.tables_1d(p, q)
    gB, gD = gold_lagrange(p, q)
    tB, tD = table_tols(p, gD)
    assert np.max(np.abs(B - gB)) < tB
    assert np.max(np.abs(Dm - gD)) < tD


@pytest.mark.parametrize("p", range(1, 10))
def test_lagrange_invariants(pkg, p):
    q = p + 2
    B, Dm, x, _ = pkg.tables_1d(p, q)
    # partition of unity / derivative of a constant
    assert np.allclose(B.sum(1), 1.0, atol=1e-14)
    assert np.allclose(Dm.sum(1), 0.0, atol=1e-12)
    # exact reproduction of x^k, k <= p, and of its derivative, on the
    # reference's node order [0, 1, 1/p, ...]
    t = np.array([0.0, 1.0] + [k / p for k in range(1, p)])
    for k in range(p + 1):
        assert np.allclose(B @ t ** k, x ** k, atol=1e-13)
        dk = k * x ** (k - 1) if k > 0 else 0 * x
        assert np.allclose(Dm @ t ** k, dk, atol=1e-11)


def test_gauss_points_for_degree(pkg):
    # Gridap: npoints = ceil((degree + 1) / 2); test/Demo_StepByStep.jl:53
    # documents degree 4 -> 3 points; degree 6 -> 4 (test/Demo.jl:23)
    assert pkg.gauss_points_for_degree(4) == 3
    assert pkg.gauss_points_for_degree(6) == 4
    for deg in range(0, 30):
        assert pkg.gauss_points_for_degree(deg) == -(-(deg + 1) // 2)


def test_sumfactorization_map_mirror(pkg):
    m = pkg.SumFactorizationMap(3, (4, 4, 4), (6, 6, 6))
    D, SB, SQ = pkg.get_dimensional_parameters(m)
    assert (D, SB, SQ) == (3, (5, 5, 5), (4, 4, 4))
    # gradMats[d][k] = B0 if d == k else A0 (src/SumFactorizationMaps.jl:58)
    for d in range(3):
        for k in range(3):
            want = m.B0 if d == k else m.A0
            assert m.gradMats[d][k] is want
    with pytest.raises(ValueError):
        pkg.SumFactorizationMap(3, (4, 4, 3), (6, 6, 6))
