"""Evaluation schedule and pass forms at a size where every form is used.

At N = 2^18 (uniform tree of six levels, 4096 parents on the last L2L
level) the L2L runs both as the row-block-split kernel (levels of fewer
than 1024 parents) and as the whole-operator kernel, P2P claims target leaves
from a resident grid with its target count read on the device, and the
lattice M2L set-up starts before P2P.  Bars:
* the resident P2P grid and one wave per leaf (EXAFMM_HIP_P2P_WAVES=0, which
  reads the target count back on the host) give identical results, bit for
  bit: every target's sums run in the same order in both;
* the operator-form L2L/M2M (matrix cores) and the per-cell kernels
  (EXAFMM_HIP_UPDOWN=cells) agree to 1e-12 in relative L2 norm (rounding
  of two arrangements of the same sums);
* the FMM result against the direct sum on 64 targets: relative L2 error of
  the potential below 1e-4 and of the force below 5e-4 at P = 10, theta = 0.4.
"""
import numpy as np
import pytest

import exafmm_hip as X

pytestmark = pytest.mark.gpu

N = 1 << 18


def cube(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def evaluate(x, monkeypatch, **env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    f = X.FMM(P=10, theta=0.4, ncrit=64)
    try:
        f.set_bodies(x)
        f.evaluate()
        f.evaluate()  # the second evaluation runs the device-driven traversal
        f.synchronize()
        r = f.results().copy()
        t = np.arange(0, N, N // 64)
        d = f.direct(t)
        return r, t, d
    finally:
        f.close()
        for k in env:
            monkeypatch.delenv(k)


@pytest.fixture(scope="module")
def bodies():
    return cube(N, 4242)


def test_resident_p2p_grid_matches_one_wave_per_leaf(bodies, monkeypatch):
    print(f"[case] N={N} P=10 resident vs one-wave P2P", flush=True)
    a, _, _ = evaluate(bodies, monkeypatch)
    b, _, _ = evaluate(bodies, monkeypatch, EXAFMM_HIP_P2P_WAVES="0")
    assert np.array_equal(a, b)


def test_operator_form_updown_matches_per_cell(bodies, monkeypatch):
    print(f"[case] N={N} P=10 operator vs per-cell L2L/M2M", flush=True)
    a, _, _ = evaluate(bodies, monkeypatch)
    b, _, _ = evaluate(bodies, monkeypatch, EXAFMM_HIP_UPDOWN="cells")
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-12


def test_fmm_against_direct(bodies, monkeypatch):
    print(f"[case] N={N} P=10 FMM vs direct on 64 targets", flush=True)
    r, t, d = evaluate(bodies, monkeypatch)
    err = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert err(r[t, 0], d[:, 0]) < 1e-4
    assert err(r[t, 1:], d[:, 1:]) < 5e-4
