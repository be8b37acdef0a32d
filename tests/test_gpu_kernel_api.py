"""The per-cell kernel API (include/kernel.h:6-14) on the GPU.

oracle/kernel_chain.cxx drives kernel::P2M, M2M, mutual M2L across a
periodic shift, L2L, L2P, mutual P2P and self P2P (with and without
softening) over a two-level cell set.  Compiled against the reference's
kernels it wrote the fixtures tests/golden/kernel_chain_*.npz; compiled
against exafmm-dev_amd/host/kernel.h (oracle/_ref/kernel_chain_hip_*) the
same calls run through the C-ABI's per-cell operators on the device.

Bar: every multipole and local expansion and every body result within 1e-12
(relative L2 over the array) of the reference's; and the chain's far-field
potential (L2P) agrees with its near-field sum (P2P) to the truncation error
the reference's own chain shows (the intent of examples/kernel.cxx).
Variants: spherical P = 10, spherical P = 10 MASS, Cartesian P = 6.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VARIANTS = ["p10", "p10_mass", "p6_cart"]
KEYS = ["M", "L_parents", "L_children", "trg_l2p", "trg_p2p_targets", "trg_p2p_sources", "trg_p2p_self",
        "trg_p2p_normalised"]


def fixture(v):
    return dict(np.load(os.path.join(GOLDEN, f"kernel_chain_{v}.npz"), allow_pickle=False))


def chain_error(d):
    """L2P against the normalised P2P of the same targets: (potential, gradient)."""
    a, r = d["trg_l2p"], d["trg_p2p_normalised"]
    return rel_l2(a[:, 0], r[:, 0]), rel_l2(a[:, 1:], r[:, 1:])


@pytest.mark.parametrize("v", VARIANTS)
def test_chain_fixture_converges(v):
    """The reference's own chain: L2P reproduces P2P to the truncation error
    of its order (P = 10: ~1e-9 potential at this separation; P = 6 ~1e-6)."""
    ep, eg = chain_error(fixture(v))
    bound = (1e-8, 1e-6) if v.startswith("p10") else (1e-5, 1e-3)
    assert ep < bound[0] and eg < bound[1], (ep, eg)


@pytest.mark.gpu
@pytest.mark.parametrize("v", VARIANTS)
def test_kernel_api_matches_reference(v, tmp_path):
    exe = os.path.join(ROOT, "oracle", "_ref", f"kernel_chain_hip_{v}")
    assert os.path.exists(exe), f"{exe} missing: build with __graft_entry__.build() (make -C oracle drivers)"
    out = tmp_path / "chain.bin"
    print(f"[gpu] kernel chain {v}", flush=True)
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = O.read_kernel_chain_dump(str(out))
    z = fixture(v)
    for k in KEYS:
        assert g[k].shape == z[k].shape, k
        assert np.all(np.isfinite(g[k])), k
        err = rel_l2(g[k], z[k])
        assert err <= 1e-12, (k, err)
    # The device chain converges as the reference's does.
    ge, ze = chain_error(g), chain_error(z)
    for a, b in zip(ge, ze):
        assert abs(a - b) <= 1e-3 * b + 1e-14, (ge, ze)
