"""CPU checks of the drop-in boundary: libsfhip.so loads, exports every
entry point include/sfhip.h declares, and validates arguments before any
device work; host-side mesh / numbering / RNG logic."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sfhip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding declares the same set
    bound = sorted(n for n, _, _ in pkg._lib.SIGNATURES)
    assert bound == names


def test_version_and_error_string(pkg):
    L = pkg.load_library()
    assert b"gfx950" in L.sf_version()
    assert isinstance(L.sf_last_error(), bytes)


def test_plan_create_validates_before_device(pkg):
    L = pkg.load_library()
    h = ctypes.c_void_p()
    assert L.sf_plan_create(ctypes.byref(h), 4, 3, 5, 10, 10, 0) == pkg._lib.SF_ERR_ARG
    assert b"dim" in L.sf_last_error()
    assert L.sf_plan_create(ctypes.byref(h), 3, 0, 5, 10, 10, 0) == pkg._lib.SF_ERR_ARG
    assert L.sf_plan_create(ctypes.byref(h), 3, 10, 5, 10, 10, 0) == pkg._lib.SF_ERR_ARG
    assert L.sf_plan_create(ctypes.byref(h), 3, 3, 5, -1, 10, 0) == pkg._lib.SF_ERR_ARG
    assert L.sf_plan_create(ctypes.byref(h), 3, 3, 5, 1, 10, 7) == pkg._lib.SF_ERR_ARG
    # (n, q) outside the instantiated list
    assert L.sf_plan_create(ctypes.byref(h), 3, 3, 9, 1, 10, 0) == pkg._lib.SF_ERR_UNSUPPORTED
    assert L.sf_plan_destroy(None) == 0
    assert L.sf_apply(None, None, None, None) == pkg._lib.SF_ERR_ARG


def test_tables_argument_errors(pkg):
    L = pkg.load_library()
    assert L.sf_tables_1d(0, 3, None, None, None, None) == pkg._lib.SF_ERR_ARG
    assert L.sf_tables_1d(3, 0, None, None, None, None) == pkg._lib.SF_ERR_ARG
    assert L.sf_tables_1d(3, 4, None, None, None, None) == 0


def test_splitmix_reproducible_and_uniform(pkg):
    a = pkg.uniform(0, 100000, -1, 1)
    b = pkg.uniform(0, 100000, -1, 1)
    assert np.array_equal(a, b)
    assert a.min() >= -1 and a.max() < 1
    assert abs(a.mean()) < 0.01 and abs(a.var() - 1 / 3) < 0.01
    assert not np.array_equal(a, pkg.uniform(1, 100000, -1, 1))
    assert not np.array_equal(pkg.uniform(0, 10, stream=1), pkg.uniform(0, 10, stream=2))
    # prefix property of the counter-based stream
    assert np.array_equal(pkg.uniform(7, 10), pkg.uniform(7, 1000)[:10])
    z = pkg.normal(3, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    # first SplitMix64 outputs are fixed bits (any host reproduces them)
    v = pkg.splitmix64(0, 2)
    assert v.dtype == np.uint64 and v[0] != v[1]


@pytest.mark.parametrize("D,m,p,dir_", [(2, 3, 2, True), (3, 2, 3, True), (3, 2, 2, False)])
def test_fespace_numbering(pkg, D, m, p, dir_):
    model = pkg.CartesianModel(D, m, perturb=0.2, seed=5)
    space = pkg.FESpace(model, p, dirichlet=dir_)
    nL = p * m + 1
    assert space.n_free == ((nL - 2) if dir_ else nL) ** D
    cd = space.cell_dofs
    assert cd.shape == (m ** D, (p + 1) ** D)
    # every free dof appears, ids in range
    ids = cd[cd >= 0]
    assert ids.max() == space.n_free - 1
    assert np.array_equal(np.unique(ids), np.arange(space.n_free))
    # a cell's ids are distinct
    for c in range(cd.shape[0]):
        v = cd[c][cd[c] >= 0]
        assert len(np.unique(v)) == len(v)
    # lattice multiplicity: interior-of-cell nodes appear once, vertices 2^D
    counts = np.bincount(ids, minlength=space.n_free)
    assert counts.max() <= 2 ** D and counts.min() >= 1
    # perturbed vertices: boundary fixed, interior moved by <= 0.2 h
    V = model.vertices
    h = 1.0 / m
    g = np.round(V / h)
    bnd = np.any((g == 0) | (g == m), axis=1)
    assert np.all(np.abs(V[bnd] - g[bnd] * h) < 1e-15)
    assert np.all(np.abs(V - g * h) <= 0.2 * h + 1e-15)
