"""The oracle against fixtures dumped by the compiled reference (tests/golden).

Pins: drand48 and the dataset generators bit for bit; both tree builders
(cell structure, keys, centres, radii before and after the upward pass, body
order) bit for bit; interaction-pair counts exactly; multipole/local
coefficients and the FMM results of the sampled bodies.
"""
import numpy as np
import pytest

import oracle as O
from helpers import case_params, load_case, rel_l2

CASES = ["cube4k_u_p10", "cube16k_u_p10", "cube16k_a_p10", "plummer16k_a_p10", "plummer16k_u_p10",
         "sphere16k_a_p8", "mass_plummer16k_a_p10", "mass_plummer8k_u_p8", "cube4k_u_p4", "cube4k_a_p12",
         "octant8k_a_p6", "cube2k_u_p6_periodic", "mass_cube2k_u_p8_img3", "cube4k_a_p10_img3", "unit16k_u_p10", "small100_u_p10", "tiny40_u_p10",
         "cube4k_u_p10_mut", "cube16k_u_p10_mut", "plummer16k_u_p10_mut", "cube16k_a_p10_mut",
         "mass_plummer16k_a_p10_mut", "cube16k_u_p6_cart", "cube16k_a_p6_cart", "cube4k_u_p4_cart",
         "sphere8k_a_p8_cart", "mass_plummer16k_a_p6_cart", "cube4k_u_p6_cart_mut", "cube2k_u_p6_cart_img3",
         "mass_plummer16k_a_p4", "mass_plummer16k_a_p8", "mass_plummer16k_a_p12"]


def test_drand48_and_generators(golden_dir):
    g = np.load(f"{golden_dir}/generators.npz", allow_pickle=False)
    # POSIX drand48 from srand48(seed), glibc's sequence.
    assert np.array_equal(O.drand48_seq(0, 16), g["drand48_seed0"])
    assert np.array_equal(O.drand48_seq(42, 16), g["drand48_seed42"])
    for d in "cspol":
        ref = g["dataset_" + d]
        assert np.array_equal(O.reference_dataset(d, ref.shape[0] if d != "l" else 1000, 0, False), ref), d
    assert np.array_equal(O.reference_dataset("p", 1000, 0, True), g["dataset_p_mass"])


def test_drand48_closed_form():
    # X1 = (a X0 + c) mod 2^48 with X0 = 0x1234ABCD330E for srand48(0x1234ABCD).
    x0 = (0x1234ABCD << 16) | 0x330E
    x1 = (0x5DEECE66D * x0 + 0xB) % (1 << 48)
    assert O.drand48_seq(0x1234ABCD, 1)[0] == x1 / float(1 << 48)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    z = load_case(name)
    o = O.Oracle(**case_params(z))
    o.set_bodies(z["input"])
    o.run()
    c = o.cells()
    xmin, xmax = o.bounds()
    assert np.array_equal(xmin, z["xmin"]) and np.array_equal(xmax, z["xmax"])
    for k in ("iparent", "ichild", "nchild", "ibody", "nbody", "icell"):
        assert np.array_equal(c[k], z["c_" + k]), k
    assert np.array_equal(c["X"], z["c_X"])
    assert np.array_equal(c["Rbuild"], z["c_Rbuild"])
    assert np.array_equal(c["R"], z["c_R"])
    b = o.bodies()
    assert np.array_equal(b["ibody"], z["ibody"])
    if not z["adaptive"]:
        assert np.array_equal(b["key"], z["key"])
    lst = o.lists()
    if int(z["images"]) == 0 and not z["mutual"]:
        assert len(lst["p2p_i"]) == z["num_p2p"]
        assert len(lst["m2l_i"]) == z["num_m2l"]
    elif int(z["images"]) == 0:
        # A mutual call covers both directions; each leaf's self pair is one call.
        nleaf = int(np.sum(c["nchild"] == 0))
        assert len(lst["m2l_i"]) == 2 * z["num_m2l"]
        assert len(lst["p2p_i"]) == 2 * z["num_p2p"] - nleaf
    # Results of the sampled bodies.
    pos = np.empty(len(b["ibody"]), np.int64)
    pos[b["ibody"]] = np.arange(len(b["ibody"]))
    fmm = b["trg"][pos[z["sample_ibody"]]]
    assert rel_l2(fmm, z["sample_fmm"]) < 1e-13
    if "trg" in z:
        assert rel_l2(b["trg"], z["trg"]) < 1e-13
        assert np.abs(c["M"] - z["M"]).max() <= 1e-13 * np.abs(z["M"]).max()
        assert np.abs(c["L"] - z["L"]).max() <= 1e-13 * np.abs(z["L"]).max()


@pytest.mark.parametrize("name", ["cube16k_u_p10", "plummer16k_a_p10", "cube2k_u_p6_periodic"])
def test_oracle_direct_matches_reference(name):
    z = load_case(name)
    x = z["input"]
    t = x[z["sample_ibody"]]
    d = O.direct(t, x, images=int(z["images"]), cycle=2 * np.pi)
    assert rel_l2(d, z["sample_direct"]) < 1e-14


@pytest.mark.parametrize("name", ["cube16k_u_p10", "plummer16k_a_p10", "mass_plummer16k_a_p10", "cube16k_u_p6_cart",
                                  "cube4k_u_p10_mut", "cube2k_u_p6_periodic"])
def test_parallel_oracle_is_bit_identical_to_serial(name):
    """The oracle's parallel mode (target cells' kernel calls deferred and run
    in emission order, passes depth by depth) against its serial recursion:
    every expansion and result bit for bit."""
    z = load_case(name)
    out = []
    for threads in (1, 0):
        o = O.Oracle(**dict(case_params(z), threads=threads))
        o.set_bodies(z["input"])
        o.run()
        c, b = o.cells(), o.bodies()
        out.append((c["M"], c["L"], b["trg"]))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
