"""Oracle pins for the periodic path: traversePeriodic, the dipole correction
and the Ewald summation (traversal.h:127-185, up_down_pass.h:151-168,
ewald.h:25-181).

* Bit-exact against oracle/ref_ewald.cxx dumps of the compiled reference
  (tests/golden/*_ewald.npz): periodic FMM results before and after the
  dipole correction, the dipole, and the Ewald results.
* Closed forms: the Madelung constant of the rock-salt lattice
  (1.747564594633...) from both Ewald and the periodic FMM; Ewald results
  independent of the splitting parameter alpha; the real-space sum equal to
  a brute-force minimum-image sum.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, rel_l2

CASES = ["lattice12_p10_img3_ewald", "lattice12_p6_img2_ewald", "cube2k_p10_img4_ewald", "lattice12_p6_cart_img3_ewald"]
MADELUNG_NACL = 1.747564594633182


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def caller_order(o):
    b = o.bodies()
    out = np.empty_like(b["trg"])
    out[b["ibody"]] = b["trg"]
    return out


@pytest.mark.parametrize("name", CASES)
def test_periodic_fmm_dipole_ewald_bitexact(name):
    z = load(name)
    x = z["input"]
    cyc = float(z["cycle"])
    cart = int(z.get("basis", 0)) == 1
    o = O.Oracle(P=int(z["P"]), theta=float(z["theta"]), ncrit=int(z["ncrit"]), nspawn=int(z["nspawn"]),
                 images=int(z["images"]), cycle=cyc, basis=int(z.get("basis", 0)))
    o.set_bodies(x)
    o.run()
    # Spherical: bit-exact.  Cartesian: the derivative tensor comes from the
    # Taylor recurrence instead of the reference's unrolled closed forms, so
    # it agrees to rounding.
    tol = 1e-13 if cart else 0.0
    assert rel_l2(caller_order(o), z["fmm_raw"]) <= tol
    d = o.dipole()
    assert np.array_equal(d, z["dipole"])
    o.dipole_correction(d, x.shape[0], cyc)
    assert rel_l2(caller_order(o), z["fmm"]) <= tol
    e = o.ewald(int(z["ksize"]), float(z["alpha"]), float(z["sigma"]), float(z["cutoff"]), cyc)
    assert np.array_equal(e, z["ewald"])
    # FMM vs Ewald as examples/ewald.cxx:136-148 reports them (the
    # reference's own agreement on these inputs: 4e-6..3.4e-4 in the
    # potential sum, 1.6e-4..1.2e-3 in the gradient).
    q = x[:, 3]
    ps, pe = np.sum(q * z["fmm"][:, 0]), np.sum(q * e[:, 0])
    assert abs(ps - pe) / abs(pe) < 1e-3
    assert rel_l2(z["fmm"][:, 1:], e[:, 1:]) < 5e-3


def rocksalt(m):
    cyc = 2 * np.pi
    h = cyc / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * h
    q = np.where(g.sum(1) % 2 == 0, 1.0, -1.0)
    return np.concatenate([x, q[:, None]], axis=1), h


def test_madelung_constant_ewald_and_fmm():
    x, h = rocksalt(8)
    cyc = 2 * np.pi
    o = O.Oracle(P=10, theta=0.35, ncrit=32, images=4, cycle=cyc)
    o.set_bodies(x)
    o.run()
    e = o.ewald(11, 10 / cyc, 0.25 / np.pi, cyc / 2, cyc)
    madelung = -e[:, 0] * x[:, 3] * h
    # Truncation of the splitting at alpha = 10 / cycle, |k| <= 11: ~5e-8.
    assert np.allclose(madelung, MADELUNG_NACL, rtol=1e-7, atol=0)
    assert np.abs(e[:, 1:]).max() < 1e-6  # every site is a centre of symmetry
    d = o.dipole()
    o.dipole_correction(d, x.shape[0], cyc)
    f = caller_order(o)
    # Periodic FMM at P = 10, theta = 0.35: 2.2e-4 (measured).
    assert np.allclose(-f[:, 0] * x[:, 3] * h, MADELUNG_NACL, rtol=5e-4, atol=0)
    assert np.abs(f[:, 1:]).max() < 5e-3


def test_ewald_independent_of_alpha():
    z = load("cube2k_p10_img4_ewald")
    x = z["input"]
    cyc = 2 * np.pi
    o = O.Oracle(P=4, ncrit=32, images=1, cycle=cyc)
    o.set_bodies(x)
    o.run(3)  # bounds + tree
    a = o.ewald(11, 10 / cyc, 0.25 / np.pi, cyc / 2, cyc)
    b = o.ewald(11, 8 / cyc, 0.25 / np.pi, cyc / 2, cyc)
    # The split is exact; only the truncations (erfc(8/2) ~ 1.5e-8 in real
    # space, exp(-k^2/4alpha^2) ~ 1e-8 at |k| = 11) differ.
    assert rel_l2(a[:, 0], b[:, 0]) < 1e-6
    assert rel_l2(a[:, 1:], b[:, 1:]) < 1e-6


def test_ewald_real_part_brute_force():
    """With no waves (ksize = 0) and the self term added back, the oracle's
    real-space part over the tree equals a minimum-image brute-force sum.
    The reference fixes each source leaf's image by the wrapped offset of the
    cell centres and prunes source cells whose wrapped centre distance is not
    below 3 cutoff; with cutoff = 2 (3 cutoff > the largest wrapped distance
    pi sqrt(3)) and leaves 2 pi / 8 wide (cutoff + leaf < pi) both rules
    pick the minimum image for every pair inside the cut-off."""
    from scipy.special import erfc
    rng = np.random.default_rng(5)
    cyc = 2 * np.pi
    n = 600
    x = np.concatenate([rng.uniform(0, cyc, (n, 3)), rng.uniform(-1, 1, (n, 1))], axis=1)
    alpha, cutoff = 10 / cyc, 2.0
    o = O.Oracle(P=4, ncrit=8, images=1, cycle=cyc)
    o.set_bodies(x)
    o.run(3)
    c = o.cells()
    assert c["nchild"][0] > 0 and np.all(c["Rbuild"][c["nchild"] == 0] * 2 < np.pi - cutoff)
    e = o.ewald(0, alpha, 0.25 / np.pi, cutoff, cyc)
    e[:, 0] += 2 / np.sqrt(np.pi) * x[:, 3] * alpha
    ref = np.zeros((n, 4))
    for i in range(n):
        dX = x[i, :3] - x[:, :3]
        dX -= cyc * np.round(dX / cyc)
        r2 = np.sum(dX * dX, axis=1)
        m = (r2 > 0) & (r2 < cutoff * cutoff)
        r = np.sqrt(r2[m])
        qj = x[m, 3]
        ref[i, 0] = np.sum(qj * erfc(alpha * r) / r)
        g = qj * (2 * alpha / np.sqrt(np.pi) * np.exp(-alpha * alpha * r * r) + erfc(alpha * r) / r) / (r * r)
        ref[i, 1:] = -np.sum(dX[m] * g[:, None], axis=0)
    assert rel_l2(e[:, 0], ref[:, 0]) < 1e-12
    assert rel_l2(e[:, 1:], ref[:, 1:]) < 1e-12
