"""Handle state across calls: stale per-evaluation state must not leak.

* A split evaluate() on a system without M2L pairs queues the lattice M2L's
  scaled source multipoles that nothing consumes; a later stepwise run on a
  larger tree must rebuild them for its own cells (reading the stale ones
  would index past them).  Bar: bit-identical results to a fresh handle.
* set_config invalidates the passes that depend on a changed parameter:
  after a change of theta the traversal refuses to run until the upward
  pass has been rerun, and the rerun equals a fresh handle bit for bit.
* Bounds stay on the device for the builder: bounds a caller passes to
  build_tree replace those an evaluation computed, and get_bounds returns
  the computed ones; every tree equals a fresh handle's bit for bit.
"""
import numpy as np
import pytest

import exafmm_hip as X
from helpers import unit_cube

pytestmark = pytest.mark.gpu


def stepwise(f, x):
    f.set_bodies(x)
    f.get_bounds()
    f.build_tree()
    f.upward_pass()
    f.traverse()
    f.downward_pass()
    return f.results()


def test_no_m2l_evaluate_then_stepwise_on_larger_tree():
    tiny = unit_cube(40, 3)
    big = unit_cube(16384, 4)
    print("[gpu] handle state: tiny evaluate then stepwise 16k", flush=True)
    f = X.FMM(P=10, ncrit=64)
    f.set_bodies(tiny)
    f.evaluate()
    assert f.stats()["num_m2l_pairs"] == 0
    a = stepwise(f, big)
    assert f.stats()["num_m2l_pairs"] > 0
    g = X.FMM(P=10, ncrit=64)
    b = stepwise(g, big)
    assert np.array_equal(a, b)
    # And an evaluate() on the same handle afterwards.
    f.set_bodies(big)
    f.evaluate()
    g.set_bodies(big)
    g.evaluate()
    assert np.array_equal(f.results(), g.results())
    f.close()
    g.close()


def test_set_config_invalidates_dependent_passes():
    x = unit_cube(16384, 5)
    print("[gpu] handle state: set_config invalidation", flush=True)
    f = X.FMM(P=8, ncrit=64, theta=0.4)
    f.set_bodies(x)
    f.get_bounds()
    f.build_tree()
    f.upward_pass()
    f.set_config(theta=0.5)
    with pytest.raises(X.ExafmmError):
        f.traverse()
    f.upward_pass()
    f.traverse()
    f.downward_pass()
    a = f.results()
    g = X.FMM(P=8, ncrit=64, theta=0.5)
    b = stepwise(g, x)
    assert np.array_equal(a, b)
    # A precision change invalidates the traversal only.
    f.set_config(precision="mixed")
    with pytest.raises(X.ExafmmError):
        f.results()
    f.traverse()
    f.downward_pass()
    assert np.all(np.isfinite(f.results()))
    f.close()
    g.close()


def test_caller_bounds_replace_device_bounds():
    """Bounds given to build_tree replace the ones an evaluation computed on
    the device (whose host copy may still be in flight), and get_bounds
    brings the computed ones back: each tree equals a fresh handle's."""
    x = unit_cube(16384, 6)
    print("[gpu] handle state: caller bounds after device bounds", flush=True)
    f = X.FMM(P=8, ncrit=64)
    f.set_bodies(x)
    f.evaluate()
    f.build_tree((np.full(3, -1.5), np.full(3, 1.25)))
    cw = f.cells()
    g = X.FMM(P=8, ncrit=64)
    g.set_bodies(x)
    g.build_tree((np.full(3, -1.5), np.full(3, 1.25)))
    cg = g.cells()
    for k in ("iparent", "ichild", "nchild", "ibody", "nbody", "icell", "X", "R"):
        assert np.array_equal(cw[k], cg[k]), k
    lo, hi = f.get_bounds()
    assert np.array_equal(lo, x[:, :3].min(0)) and np.array_equal(hi, x[:, :3].max(0))
    f.build_tree()
    cd = f.cells()
    h = X.FMM(P=8, ncrit=64)
    h.set_bodies(x)
    h.get_bounds()
    h.build_tree()
    ch = h.cells()
    for k in ("iparent", "ichild", "nchild", "ibody", "nbody", "icell", "X", "R"):
        assert np.array_equal(cd[k], ch[k]), k
    assert not np.array_equal(cd["X"][:1], cw["X"][:1])  # the caller's box was used
    for o in (f, g, h):
        o.close()
