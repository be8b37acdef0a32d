"""GPU: the MD-code entry points of wrappers/gromacs.cxx (FMM_Init,
FMM_Partition, FMM_Coulomb, Ewald_Coulomb, Direct_Coulomb), single rank,
called through ctypes exactly as an MD code calls them.

Bar: FMM_Coulomb against Direct_Coulomb on the same arrays within the FMM's
own truncation error (P = 10, theta = 0.4: 1e-5 potential, 1e-4 gradient on
open boundaries); periodic FMM (2 image levels, with the dipole correction
both apply) against Ewald_Coulomb within 1e-3 / 1e-2; FMM_Partition returns
a permutation with x and q moved consistently; FMM_Coulomb's accumulation
follows the reference (incoming p, f divided by q, then the field added).
"""
import ctypes
import os

import numpy as np
import pytest

import exafmm_hip as X

pytestmark = pytest.mark.gpu

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)


@pytest.fixture
def lib(monkeypatch):
    monkeypatch.setenv("EXAFMM_HIP_P", "10")
    L = ctypes.CDLL(X.LIB_PATH)
    L.FMM_Init.argtypes = [ctypes.c_int]
    L.FMM_Partition.argtypes = [ip, ip, dp, dp, ctypes.c_double]
    for name in ("FMM_Coulomb", "Direct_Coulomb"):
        getattr(L, name).argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_double]
    L.Ewald_Coulomb.argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double]
    yield L
    L.FMM_Finalize()


def P(a, t):
    return a.ctypes.data_as(t)


def system(n, seed, cycle):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-cycle / 2, cycle / 2, (n, 3)).ravel().copy()
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x, q


def rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_coulomb_open_boundary_against_direct(lib):
    n, cyc = 8000, 2 * np.pi
    x, q = system(n, 1, cyc)
    print(f"[gpu] FMM_Coulomb open n={n}", flush=True)
    lib.FMM_Init(0)
    p = np.zeros(n)
    f = np.zeros(3 * n)
    lib.FMM_Coulomb(n, P(x, dp), P(q, dp), P(p, dp), P(f, dp), cyc)
    pd = np.zeros(n)
    fd = np.zeros(3 * n)
    lib.Direct_Coulomb(n, P(x, dp), P(q, dp), P(pd, dp), P(fd, dp), cyc)
    assert rel(p, pd) < 1e-5
    assert rel(f, fd) < 1e-4
    # Accumulation: incoming (p, f) are divided by q before the field is added.
    p2 = np.full(n, 3.0)
    f2 = np.full(3 * n, -2.0)
    lib.FMM_Coulomb(n, P(x, dp), P(q, dp), P(p2, dp), P(f2, dp), cyc)
    assert np.allclose(p2, 3.0 / q + p, rtol=1e-12, atol=1e-12 * np.abs(3.0 / q).max())
    assert np.allclose(f2, -2.0 / np.repeat(q, 3) + f, rtol=1e-12, atol=1e-12 * np.abs(2.0 / q).max())


def test_partition_permutes_consistently(lib):
    n, cyc = 5000, 2.0
    x, q = system(n, 2, cyc)
    # Some bodies just outside the box (within the range vec.h's unwrap
    # inverts): wrapped for the tree, unwrapped on return.
    x[:30] = np.where(x[:30] > 0, x[:30] + 0.5, x[:30] - 0.5)
    x0, q0 = x.copy(), q.copy()
    idx = np.arange(n, dtype=np.int32)
    nn = ctypes.c_int(n)
    lib.FMM_Init(0)
    lib.FMM_Partition(ctypes.byref(nn), P(idx, ip), P(x, dp), P(q, dp), cyc)
    assert nn.value == n
    assert np.array_equal(np.sort(idx), np.arange(n))
    assert not np.array_equal(idx, np.arange(n))  # tree order differs from input order
    assert np.array_equal(q, q0[idx])
    assert np.allclose(x.reshape(-1, 3), x0.reshape(-1, 3)[idx], rtol=0, atol=1e-12)
    # FMM_Coulomb on the partitioned arrays matches direct summation of the
    # same system: FMM_Coulomb wraps positions into the box and
    # Direct_Coulomb does not (as in the reference), so the direct sum gets
    # the wrapped positions.
    p = np.zeros(n)
    f = np.zeros(3 * n)
    lib.FMM_Coulomb(n, P(x, dp), P(q, dp), P(p, dp), P(f, dp), cyc)
    xw = np.where(x < -cyc / 2, x + cyc, np.where(x > cyc / 2, x - cyc, x))
    pd = np.zeros(n)
    fd = np.zeros(3 * n)
    lib.Direct_Coulomb(n, P(xw, dp), P(q, dp), P(pd, dp), P(fd, dp), cyc)
    assert rel(p, pd) < 1e-5 and rel(f, fd) < 1e-4


def test_periodic_coulomb_against_ewald(lib):
    m, cyc = 16, 2 * np.pi
    h = cyc / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(5)
    x = ((g + 0.5) * h + rng.uniform(-0.25, 0.25, g.shape) * h - cyc / 2).ravel().copy()
    q = np.where(g.sum(1) % 2 == 0, 1.0, -1.0)
    n = q.shape[0]
    print(f"[gpu] FMM_Coulomb periodic n={n}", flush=True)
    lib.FMM_Init(2)
    p = np.zeros(n)
    f = np.zeros(3 * n)
    lib.FMM_Coulomb(n, P(x, dp), P(q, dp), P(p, dp), P(f, dp), cyc)
    pe = np.zeros(n)
    fe = np.zeros(3 * n)
    lib.Ewald_Coulomb(n, P(x, dp), P(q, dp), P(pe, dp), P(fe, dp), 11, 10 / cyc, 0.25 / np.pi, cyc / 2, cyc)
    assert abs(np.sum(q * p) - np.sum(q * pe)) / abs(np.sum(q * pe)) < 1e-3
    assert rel(f, fe) < 1e-2
