"""GPU batched evaluation: many independent systems in one forest.

Bar: each system's potential and gradient from `evaluate_batch` within 1e-13
(relative L2) of a single-system evaluation of the same bodies on the same
handle configuration (the forest gives every system the tree its own build
makes; only the grouping of M2L work into tiles and segments, and so the
rounding of partial sums, can differ), in FP64 and in the mixed form; and the
same interaction counts as the single-system runs added up.
"""
import numpy as np
import pytest

import exafmm_hip as X
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def close(a, b, tol):
    """Relative L2 within tol; exact when the reference is zero (a lone body)."""
    if not np.any(b):
        return np.array_equal(a, b)
    return rel_l2(a, b) < tol


def cube(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def sphere_surface(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([v + rng.uniform(-3, 3, 3), q[:, None]], axis=1)


def systems(sizes, seed=0):
    out = []
    for i, n in enumerate(sizes):
        out.append(cube(n, seed + i) if i % 2 == 0 else sphere_surface(n, seed + i))
    return out


@pytest.mark.parametrize("precision,P,theta,ncrit", [("fp64", 8, 0.5, 64), ("mixed", 8, 0.5, 32), ("fp64", 10, 0.4, 64)])
def test_batch_matches_single_systems(precision, P, theta, ncrit):
    sizes = [1, 37, 300, 4096, 9000, 20000, 5000, 65536]
    sy = systems(sizes, seed=11)
    print(f"[gpu] batch {precision} P={P} sizes={sizes}", flush=True)
    f = X.FMM(P=P, theta=theta, ncrit=ncrit, precision=precision)
    out = f.evaluate_batch(sy)
    sb = f.stats()
    npairs = f.count_p2p_bodies()
    m2l, p2p, pb = 0, 0, 0
    for x, r in zip(sy, out):
        f.set_bodies(x)
        f.evaluate()
        ref = f.results()
        s = f.stats()
        m2l += s["num_m2l_pairs"]
        p2p += s["num_p2p_pairs"]
        pb += f.count_p2p_bodies()
        assert r.shape == ref.shape
        assert close(r[:, 0], ref[:, 0], 1e-13), len(x)
        assert close(r[:, 1:], ref[:, 1:], 1e-13), len(x)
    assert sb["num_m2l_pairs"] == m2l and sb["num_p2p_pairs"] == p2p and npairs == pb
    f.close()


def test_batch_repeat_and_single_system():
    """A batch of one system is that system's evaluation; a batch evaluated
    twice gives identical results (deterministic)."""
    sy = systems([20000, 3000, 12000], seed=3)
    f = X.FMM(P=6, theta=0.5, ncrit=64)
    a = f.evaluate_batch(sy)
    b = f.evaluate_batch(sy)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    one = f.evaluate_batch(sy[:1])[0]
    f.set_bodies(sy[0])
    f.evaluate()
    assert rel_l2(one, f.results()) < 1e-14
    f.close()


def test_batch_rejects_unsupported():
    f = X.FMM(P=6, images=2)
    with pytest.raises(X.ExafmmError):
        f.evaluate_batch(systems([100, 200]))
    f.close()
    f = X.FMM(P=6)
    with pytest.raises(X.ExafmmError):
        f.evaluate_batch([np.zeros((0, 4))])
    f.close()


def test_autotune_picks_parameters_and_matches_untuned_run():
    """With autotuning on, the first batch of a signature times candidate
    configurations and keeps the fastest; the results equal an untuned run
    with the chosen parameters pinned (set_tuning) bit for bit, and a second
    call reuses the choice."""
    sy = systems([3000, 20000, 45000, 9000, 70000], seed=5)
    print("[gpu] batch autotune", flush=True)
    f = X.FMM(P=8, theta=0.5, ncrit=64)
    f.set_autotune(True)
    a = f.evaluate_batch(sy)
    t = f.tuning()
    assert t["ncrit"] in (32, 64, 128) and t["p2p_waves_per_cu"] in (8, 12, 16)
    assert t["p2p_targets_per_lane"] in (1, 3, 4, 5) and t["m2l_groups_per_cu"] in (0, 2)
    b = f.evaluate_batch(sy)
    assert f.tuning() == t
    g = X.FMM(P=8, theta=0.5, ncrit=t["ncrit"])
    g.set_tuning(**t)
    assert g.tuning() == t
    c = g.evaluate_batch(sy)
    for u, v, w in zip(a, b, c):
        assert np.array_equal(u, w) and np.array_equal(v, w)
    # A single system of its own signature.
    f.set_bodies(sy[2])
    f.evaluate()
    ts = f.tuning()
    g2 = X.FMM(P=8, theta=0.5, ncrit=ts["ncrit"])
    g2.set_tuning(**ts)
    g2.set_bodies(sy[2])
    g2.evaluate()
    assert np.array_equal(f.results(), g2.results())
    for h in (f, g, g2):
        h.close()
