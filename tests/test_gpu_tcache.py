"""The lattice M2L's T(d) table kept across evaluations.

T(d) depends on the offset key alone, so a row computed for one tree serves
every later tree that uses the same offset.  Bar: an evaluation on a handle
whose table already holds rows from other trees (and has to add rows for new
offsets, or start afresh when a row limit is reached, or when the precision
changes) gives results identical, bit for bit, to the same evaluation on a
fresh handle: the rows hold the same values wherever they sit in the table,
and every sum runs in the same order.
"""
import numpy as np
import pytest

import exafmm_hip as X

pytestmark = pytest.mark.gpu


def cube(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-scale, scale, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def run(f, x):
    f.set_bodies(x)
    f.evaluate()
    f.synchronize()
    return f.results().copy(), f.stats()


def fresh(x, **kw):
    f = X.FMM(**kw)
    try:
        return run(f, x)[0]
    finally:
        f.close()


# Three trees with different offset sets: a uniform cube, a denser cube (one
# level deeper) and a clustered (adaptive) distribution.
def trees():
    rng = np.random.default_rng(7)
    c = rng.normal(0, 0.15, (30000, 3))
    q = rng.uniform(-0.5, 0.5, 30000)
    clustered = np.concatenate([c, ((q - q.mean()) / 30000)[:, None]], axis=1)
    return [cube(20000, 1), cube(120000, 2), clustered]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rows_kept_across_trees_match_fresh_handles(precision):
    kw = dict(P=10, theta=0.4, ncrit=64, precision=precision)
    xs = trees()
    ref = [fresh(x, **kw) for x in xs]
    f = X.FMM(**kw)
    try:
        for rep in range(2):  # second pass: every row already in the table
            for x, r in zip(xs, ref):
                got, st = run(f, x)
                assert st["num_m2l_lattice"] > 0
                assert np.array_equal(got, r)
    finally:
        f.close()


def test_row_limit_restarts_the_table(monkeypatch):
    kw = dict(P=8, theta=0.5, ncrit=64, precision="fp64")
    xs = trees()
    ref = [fresh(x, **kw) for x in xs]
    monkeypatch.setenv("EXAFMM_HIP_TCACHE_ROWS", "40")  # fewer rows than one tree uses
    f = X.FMM(**kw)
    try:
        for x, r in zip(xs + xs, ref + ref):
            assert np.array_equal(run(f, x)[0], r)
    finally:
        f.close()


def test_precision_switch_rebuilds_the_table():
    x = cube(20000, 3)
    r64 = fresh(x, P=10, precision="fp64")
    rmx = fresh(x, P=10, precision="mixed")
    f = X.FMM(P=10, precision="fp64")
    try:
        assert np.array_equal(run(f, x)[0], r64)
        f.set_config(precision="mixed")
        assert np.array_equal(run(f, x)[0], rmx)
        f.set_config(precision="fp64")
        assert np.array_equal(run(f, x)[0], r64)
    finally:
        f.close()
