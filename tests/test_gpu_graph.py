"""GPU re-evaluation with new charges on a kept geometry, replayed from a HIP
graph (exafmm_hip_evaluate_charges).

Bar: every call -- the first one (ordinary launches, then recorded) and the
graph replays after it, with other work of the handle between them -- gives
results bit-identical to a full evaluation of the same bodies with those
charges on a fresh handle, in FP64 and in the mixed form, for a uniform cube,
an adaptive Plummer tree, a periodic lattice and a batched forest; the
result is linear in the charges (a property the graph cannot fake: it is
checked against two other replays); calls that change the geometry drop the
graph and the next call records a new one; unsupported settings are
rejected.
"""
import numpy as np
import pytest

import exafmm_hip as X
from helpers import plummer, rel_l2, unit_cube

pytestmark = pytest.mark.gpu


def charges(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.uniform(-0.5, 0.5, n)
    return (q - q.mean()) / n


def fresh(x, q, **kw):
    g = X.FMM(**kw)
    g.set_bodies(np.concatenate([x[:, :3], q[:, None]], axis=1))
    g.evaluate()
    r = g.results()
    g.close()
    return r


CASES = {
    "cube-fp64": (lambda: unit_cube(16384, 3), dict(P=10, theta=0.4, ncrit=64)),
    "cube-mixed": (lambda: unit_cube(16384, 3), dict(P=10, theta=0.4, ncrit=64, precision="mixed")),
    "plummer-adaptive": (lambda: plummer(16384, 4), dict(P=8, theta=0.4, ncrit=32, builder="adaptive")),
    "lattice-periodic": (lambda: np.concatenate([(np.stack(np.meshgrid(*[np.arange(12)] * 3, indexing="ij"), -1)
                                                  .reshape(-1, 3) + 0.5) * (2 * np.pi / 12),
                                                 np.zeros((1728, 1))], axis=1),
                         dict(P=10, theta=0.35, ncrit=32, images=2, cycle=2 * np.pi)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_replays_match_fresh_evaluations(name):
    make, kw = CASES[name]
    x = make()
    n = x.shape[0]
    print(f"[gpu] evaluate_charges {name} n={n}", flush=True)
    f = X.FMM(**kw)
    f.set_bodies(np.concatenate([x[:, :3], charges(n, 1)[:, None]], axis=1))
    f.evaluate()
    tgt = np.arange(0, n, 97, dtype=np.int32)
    outs = []
    for k in range(4):
        q = charges(n, 10 + k)
        f.evaluate_charges(q)
        r = f.results()
        assert f.stats()["t_total"] > 0
        ref = fresh(x, q, **kw)
        assert np.array_equal(r, ref), (name, k, rel_l2(r, ref))
        outs.append((q, r))
        # Other work of the handle between replays (none of it drops the graph).
        f.direct(tgt)
        f.stats()
    # Linearity in the charges across three replays (to the rounding of the
    # form: FP64, or the FP32 pair terms and tables of the mixed form).
    (q0, r0), (q1, r1) = outs[0], outs[1]
    f.evaluate_charges(2.0 * q0 - 3.0 * q1)
    tol = 1e-6 if kw.get("precision") == "mixed" else 1e-12
    assert rel_l2(f.results(), 2.0 * r0 - 3.0 * r1) < tol
    f.close()


def test_geometry_change_drops_the_graph():
    x = unit_cube(8192, 5)
    y = unit_cube(8192, 6)
    kw = dict(P=8, theta=0.4, ncrit=64)
    f = X.FMM(**kw)
    f.set_bodies(x)
    f.evaluate()
    q = charges(8192, 7)
    f.evaluate_charges(q)
    f.evaluate_charges(q)  # a replay
    assert np.array_equal(f.results(), fresh(x, q, **kw))
    f.set_bodies(y)  # new positions: the kept geometry and its graph go
    with pytest.raises(X.ExafmmError):
        f.evaluate_charges(q)
    f.evaluate()
    f.evaluate_charges(q)
    f.evaluate_charges(q)
    assert np.array_equal(f.results(), fresh(y, q, **kw))
    # A configuration change re-traverses and records again.
    f.set_config(theta=0.5)
    f.evaluate()
    f.evaluate_charges(q)
    f.evaluate_charges(q)
    assert np.array_equal(f.results(), fresh(y, q, P=8, theta=0.5, ncrit=64))
    f.close()


def test_batched_forest_replays():
    sizes = [3000, 20000, 9000, 45000]
    systems = [unit_cube(m, 20 + i) for i, m in enumerate(sizes)]
    kw = dict(P=8, theta=0.5, ncrit=64)
    print("[gpu] evaluate_charges batched forest", flush=True)
    f = X.FMM(**kw)
    f.evaluate_batch(systems)
    g = X.FMM(**kw)
    for k in range(3):
        qs = [charges(m, 30 + 7 * k + i) for i, m in enumerate(sizes)]
        f.evaluate_charges(qs)
        got = f.results_batch()
        want = g.evaluate_batch([np.concatenate([s[:, :3], q[:, None]], axis=1) for s, q in zip(systems, qs)])
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
    f.close()
    g.close()


def test_unsupported_and_bad_calls():
    x = unit_cube(4096, 9)
    f = X.FMM(P=6, theta=0.4, ncrit=32, useRopt=1)
    f.set_bodies(x)
    with pytest.raises(X.ExafmmError):
        f.evaluate_charges(x[:, 3])  # nothing evaluated yet
    f.evaluate()
    with pytest.raises(X.ExafmmError):
        f.evaluate_charges(x[:, 3])  # useRopt: the radii depend on the charges
    f.close()
    f = X.FMM(P=6, theta=0.4, ncrit=32)
    f.set_bodies(x)
    f.evaluate()
    with pytest.raises(X.ExafmmError):
        f.evaluate_charges(x[:100, 3])  # wrong body count
    f.evaluate_charges(x[:, 3])
    assert np.array_equal(f.results(), fresh(x, x[:, 3], P=6, theta=0.4, ncrit=32))
    f.close()
