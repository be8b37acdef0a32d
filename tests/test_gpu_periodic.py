"""GPU periodic path: traversePeriodic, dipole correction and Ewald against
the oracle and the reference fixtures.

Bar (FP64): identical interaction-pair multisets (the 27 images of the tree
are traversed non-mutually, traversal.h:250-259), results within 1e-12
(relative L2) of the oracle and of the reference's dumps; the Ewald sum
within 1e-12 of the reference's ewald.h output (the GPU's erfc/exp/sincos
and summation order differ from glibc's by a few ulps).
"""
import os

import numpy as np
import pytest

import exafmm_hip as X
from helpers import GOLDEN, case_params, load_case, rel_l2, run_oracle

pytestmark = pytest.mark.gpu

EWALD = ["lattice12_p10_img3_ewald", "lattice12_p6_img2_ewald", "cube2k_p10_img4_ewald",
         "lattice12_p6_cart_img3_ewald"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def pair_multiset(i, j):
    return np.sort(i.astype(np.int64) * 10 ** 7 + j)


@pytest.mark.parametrize("name", EWALD)
def test_periodic_fmm_dipole_ewald(name):
    z = load(name)
    x = z["input"]
    cyc = float(z["cycle"])
    P, images = int(z["P"]), int(z["images"])
    print(f"[gpu] periodic {name} N={x.shape[0]} P={P} images={images}", flush=True)
    basis = int(z.get("basis", 0))
    f = X.FMM(P=P, theta=float(z["theta"]), ncrit=int(z["ncrit"]), nspawn=int(z["nspawn"]), images=images, cycle=cyc,
              basis=basis)
    f.set_bodies(x)
    f.evaluate()
    o = run_oracle(x, P=P, theta=float(z["theta"]), ncrit=int(z["ncrit"]), nspawn=int(z["nspawn"]), images=images,
                   cycle=cyc, basis=basis)
    gl, ol = f.lists(), o.lists()
    for a, b in (("p2p_i", "p2p_j"), ("m2l_i", "m2l_j")):
        assert np.array_equal(pair_multiset(gl[a], gl[b]), pair_multiset(ol[a], ol[b])), a
    r = f.results()
    raw = z["fmm_raw"]
    assert rel_l2(r[:, 0], raw[:, 0]) < 1e-12
    assert rel_l2(r[:, 1:], raw[:, 1:]) < 1e-12
    d = f.dipole()
    scale = np.abs(x[:, :3] * x[:, 3:4]).sum()
    assert np.abs(d - z["dipole"]).max() <= 1e-13 * scale
    f.dipole_correction(z["dipole"], x.shape[0], cyc)
    r = f.results()
    assert rel_l2(r[:, 0], z["fmm"][:, 0]) < 1e-12
    assert rel_l2(r[:, 1:], z["fmm"][:, 1:]) < 1e-12
    e = f.ewald(int(z["ksize"]), float(z["alpha"]), float(z["sigma"]), float(z["cutoff"]), cyc)
    assert rel_l2(e[:, 0], z["ewald"][:, 0]) < 1e-12
    assert rel_l2(e[:, 1:], z["ewald"][:, 1:]) < 1e-12
    # The three parts add up to the whole (accumulate semantics).
    parts = np.zeros_like(e)
    for p in (1, 2, 4):
        f.ewald(int(z["ksize"]), float(z["alpha"]), float(z["sigma"]), float(z["cutoff"]), cyc, parts=p, trg=parts)
    assert rel_l2(parts, e) < 1e-14
    f.close()


@pytest.mark.parametrize("name", ["cube2k_u_p6_periodic", "mass_cube2k_u_p8_img3", "cube4k_a_p10_img3",
                                  "cube2k_u_p6_cart_img3"])
def test_periodic_fmm_matches_reference(name):
    z = load_case(name)
    p = case_params(z)
    builder = "adaptive" if p["adaptive"] else "uniform"
    print(f"[gpu] periodic {name} builder={builder} images={p['images']} mass={p['mass']}", flush=True)
    f = X.FMM(P=p["P"], theta=p["theta"], ncrit=p["ncrit"], nspawn=p["nspawn"], useRmax=p["useRmax"],
              useRopt=p["useRopt"], images=p["images"], builder=builder, mass=p["mass"], basis=p["basis"])
    f.set_bodies(z["input"])
    f.evaluate()
    o = run_oracle(z["input"], **p)
    c, oc = f.cells(), o.cells()
    scale = np.abs(oc["L"]).max()
    assert np.abs(c["L"] - oc["L"]).max() <= 1e-12 * scale
    trg = f.results()
    ob = o.bodies()
    otrg = np.empty_like(ob["trg"])
    otrg[ob["ibody"]] = ob["trg"]
    assert rel_l2(trg[:, 0], otrg[:, 0]) < 1e-12
    assert rel_l2(trg[:, 1:], otrg[:, 1:]) < 1e-12
    assert rel_l2(trg[z["sample_ibody"]], z["sample_fmm"]) < 1e-12
    f.close()


@pytest.mark.parametrize("basis,P,tol_pot,tol_acc", [("spherical", 10, 1e-4, 1e-3), ("cartesian", 6, 1e-3, 1e-2)])
def test_periodic_fmm_vs_ewald_lattice32(basis, P, tol_pot, tol_acc):
    """The check of examples/ewald.cxx:136-148 at 32^3 bodies of the jittered
    rock-salt lattice (theta 0.35, 3 image levels; spherical P = 10 and
    Cartesian P = 6): potential sum and gradient of the GPU periodic FMM
    against the GPU Ewald sum."""
    m = 32
    cyc = 2 * np.pi
    h = cyc / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(4)
    x = np.concatenate([(g + 0.5) * h + rng.uniform(-0.25, 0.25, g.shape) * h,
                        np.where(g.sum(1) % 2 == 0, 1.0, -1.0)[:, None]], axis=1)
    print(f"[gpu] lattice 32^3 FMM vs Ewald {basis} P={P}", flush=True)
    f = X.FMM(P=P, theta=0.35, ncrit=64, images=3, cycle=cyc, basis=basis)
    f.set_bodies(x)
    f.evaluate()
    d = f.dipole()
    f.dipole_correction(d, x.shape[0], cyc)
    r = f.results()
    e = f.ewald(11, 10 / cyc, 0.25 / np.pi, cyc / 2, cyc)
    q = x[:, 3]
    ps, pe = np.sum(q * r[:, 0]), np.sum(q * e[:, 0])
    print(f"[gpu] pot-sum rel {abs(ps - pe) / abs(pe):.2e} acc rel {rel_l2(r[:, 1:], e[:, 1:]):.2e}", flush=True)
    assert abs(ps - pe) / abs(pe) < tol_pot
    assert rel_l2(r[:, 1:], e[:, 1:]) < tol_acc
    f.close()
