"""The device-driven traversal (levels enqueued without host round trips).

A handle's first traversal runs the per-level passes (each level's sizes read
back to the host) and records the frontier capacity; later traversals run the
device-driven passes, and fall back to the per-level passes when an output
would not fit.  Bar: the lists (in their emitted, target-grouped order) and
the results of a later evaluation are identical, bit for bit, to those of a
fresh handle -- for open and periodic boundaries, mutual traversal, the
adaptive builder, batched forests, a tree that outgrows the recorded
capacities, and a deeper tree that fits them but needs more levels than the
last traversal had.
"""
import numpy as np
import pytest

import exafmm_hip as X

pytestmark = pytest.mark.gpu


def cube(n, seed, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def plummer(n, seed):
    rng = np.random.default_rng(seed)
    r = 1.0 / np.sqrt(rng.uniform(0.01, 1.0, n) ** (-2.0 / 3.0) - 1.0)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    q = np.full(n, 1.0 / n)
    return np.concatenate([v * r[:, None], q[:, None]], axis=1)


def snapshot(f, x):
    f.set_bodies(x)
    f.evaluate()
    f.synchronize()
    L = f.lists()
    return f.results().copy(), {k: v.copy() for k, v in L.items()}


def same(a, b):
    ra, la = a
    rb, lb = b
    assert np.array_equal(ra, rb)
    for k in la:
        assert np.array_equal(la[k], lb[k]), k


CASES = {
    "open": (dict(P=6, ncrit=32), lambda: cube(30000, 1)),
    "periodic": (dict(P=6, ncrit=32, images=2, cycle=2.0), lambda: cube(20000, 2)),
    "mutual": (dict(P=6, ncrit=32, mutual=1), lambda: cube(20000, 3)),
    "adaptive": (dict(P=6, ncrit=32, builder="adaptive"), lambda: plummer(30000, 4)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_later_traversals_match_a_fresh_handle(name):
    kw, make = CASES[name]
    x = make()
    f = X.FMM(**kw)
    try:
        first = snapshot(f, x)   # per-level passes
        second = snapshot(f, x)  # device-driven passes
        third = snapshot(f, x)
        same(first, second)
        same(first, third)
    finally:
        f.close()


def test_outgrowing_the_recorded_capacity_falls_back():
    # A small tree records small capacities; a larger, denser one must not fit
    # them on its first device-driven traversal.
    small, big = cube(3000, 5), plummer(60000, 6)
    ref = X.FMM(P=6, ncrit=32, builder="adaptive")
    try:
        want = snapshot(ref, big)
    finally:
        ref.close()
    f = X.FMM(P=6, ncrit=32, builder="adaptive")
    try:
        snapshot(f, small)
        snapshot(f, small)
        same(snapshot(f, big), want)
        same(snapshot(f, big), want)
    finally:
        f.close()


def test_batched_forest_repeats():
    systems = [cube(2000 + 500 * s, 10 + s, -1 - s, 1 + s) for s in range(6)]
    f = X.FMM(P=6, ncrit=32)
    try:
        a = f.evaluate_batch(systems)
        b = f.evaluate_batch(systems)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
    finally:
        f.close()


def test_deeper_tree_within_the_recorded_capacity():
    # A large shallow tree records large capacities and its level count; a
    # smaller, deeper one then fits them but needs more levels than were
    # enqueued, so the device-driven passes continue after the read-back.
    wide, deep = cube(60000, 7), plummer(20000, 8)
    ref = X.FMM(P=6, ncrit=32, builder="adaptive")
    try:
        want = snapshot(ref, deep)
    finally:
        ref.close()
    f = X.FMM(P=6, ncrit=32, builder="adaptive")
    try:
        snapshot(f, wide)
        snapshot(f, wide)
        shallow_levels = f.stats()["num_levels"]
        got = snapshot(f, deep)
        assert f.stats()["num_levels"] > shallow_levels
        same(got, want)
        same(snapshot(f, deep), want)
    finally:
        f.close()
