"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs in the development container only: it executes the reference harness
binaries built by `make -C oracle ref` (oracle/ref_harness.cxx compiled
against the reference sources) and stores their inputs and outputs as small
.npz files.  The fixtures are data -- inputs and expected outputs -- and are
what the CPU tests pin the oracle against.

Usage: python tests/golden/make_golden.py [fixture names]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.dirname(os.path.abspath(__file__))

# name, binary, mode args, harness options (theta ncrit nspawn useRmax useRopt mutual images), what to keep
CASES = [
    ("cube4k_u_p10", "harness_u_p10", ["gen", "c", "4096"], [0.4, 64, 1000, 1, 0, 0, 0], "full"),
    ("cube16k_u_p10", "harness_u_p10", ["gen", "c", "16384"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("cube16k_a_p10", "harness_a_p10", ["gen", "c", "16384"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("plummer16k_a_p10", "harness_a_p10", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("plummer16k_u_p10", "harness_u_p10", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("sphere16k_a_p8", "harness_a_p8", ["gen", "s", "16384"], [0.5, 32, 1000, 1, 0, 0, 0], "tree"),
    ("mass_plummer16k_a_p10", "harness_a_p10_mass", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 1, 0, 0], "tree"),
    ("mass_plummer8k_u_p8", "harness_u_p8_mass", ["gen", "p", "8192"], [0.4, 64, 1000, 1, 1, 0, 0], "tree"),
    ("cube4k_u_p4", "harness_u_p4", ["gen", "c", "4096"], [0.4, 64, 1000, 1, 0, 0, 0], "full"),
    ("cube4k_a_p12", "harness_a_p12", ["gen", "c", "4096"], [0.4, 32, 1000, 1, 0, 0, 0], "tree"),
    ("octant8k_a_p6", "harness_a_p6", ["gen", "o", "8192"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("cube2k_u_p6_periodic", "harness_u_p6", ["gen", "c", "2048"], [0.4, 64, 1000, 1, 0, 0, 2], "tree"),
    ("mass_cube2k_u_p8_img3", "harness_u_p8_mass", ["gen", "c", "2048"], [0.4, 64, 1000, 1, 0, 0, 3], "tree"),
    ("cube4k_a_p10_img3", "harness_a_p10", ["gen", "c", "4096"], [0.4, 64, 1000, 1, 0, 0, 3], "tree"),
    # Inputs of the BASELINE-style generator (U[-1,1]^3, zero-mean charges / N).
    ("unit16k_u_p10", "harness_u_p10", ["file"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("small100_u_p10", "harness_u_p10", ["gen", "c", "100"], [0.4, 64, 1000, 1, 0, 0, 0], "full"),
    ("tiny40_u_p10", "harness_u_p10", ["gen", "c", "40"], [0.4, 64, 1000, 1, 0, 0, 0], "full"),
    # Mutual traversal (the reference's default -m 1).
    ("cube4k_u_p10_mut", "harness_u_p10", ["gen", "c", "4096"], [0.4, 64, 1000, 1, 0, 1, 0], "full"),
    ("cube16k_u_p10_mut", "harness_u_p10", ["gen", "c", "16384"], [0.4, 64, 1000, 1, 0, 1, 0], "tree"),
    ("plummer16k_u_p10_mut", "harness_u_p10", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 0, 1, 0], "tree"),
    ("cube16k_a_p10_mut", "harness_a_p10", ["gen", "c", "16384"], [0.4, 64, 1000, 1, 0, 1, 0], "tree"),
    ("mass_plummer16k_a_p10_mut", "harness_a_p10_mass", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 1, 1, 0], "tree"),
    # Cartesian (Taylor) basis, kernels/LaplaceCartesianCPU.cxx.
    ("cube16k_u_p6_cart", "harness_u_p6_cart", ["gen", "c", "16384"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("cube16k_a_p6_cart", "harness_a_p6_cart", ["gen", "c", "16384"], [0.4, 64, 1000, 1, 0, 0, 0], "tree"),
    ("cube4k_u_p4_cart", "harness_u_p4_cart", ["gen", "c", "4096"], [0.4, 64, 1000, 1, 0, 0, 0], "full"),
    ("sphere8k_a_p8_cart", "harness_a_p8_cart", ["gen", "s", "8192"], [0.5, 32, 1000, 1, 0, 0, 0], "tree"),
    ("mass_plummer16k_a_p6_cart", "harness_a_p6_cart_mass", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 1, 0, 0],
     "tree"),
    ("cube4k_u_p6_cart_mut", "harness_u_p6_cart", ["gen", "c", "4096"], [0.4, 64, 1000, 1, 0, 1, 0], "full"),
    ("cube2k_u_p6_cart_img3", "harness_u_p6_cart", ["gen", "c", "2048"], [0.4, 64, 1000, 1, 0, 0, 3], "tree"),
    # The reference Plummer set (dataset.h:108-135, MASS charges 1/N, useRopt 1)
    # at the expansion orders of the mixed-precision sweep.
    ("mass_plummer16k_a_p4", "harness_a_p4_mass", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 1, 0, 0], "tree"),
    ("mass_plummer16k_a_p8", "harness_a_p8_mass", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 1, 0, 0], "tree"),
    ("mass_plummer16k_a_p12", "harness_a_p12_mass", ["gen", "p", "16384"], [0.4, 64, 1000, 1, 1, 0, 0], "tree"),
]

# The per-cell kernel API chain (oracle/kernel_chain.cxx) on the reference's
# kernels: name, binary.
KERNEL_CHAIN_CASES = [
    ("kernel_chain_p10", "kernel_chain_p10"),
    ("kernel_chain_p10_mass", "kernel_chain_p10_mass"),
    ("kernel_chain_p6_cart", "kernel_chain_p6_cart"),
]


# Periodic cases for oracle/ref_ewald.cxx: name, binary, input, options
# (theta ncrit nspawn images ksize alpha sigma cutoff cycle).
CYCLE = 2 * np.pi
EWALD_CASES = [
    ("lattice12_p10_img3_ewald", "ref_ewald_p10", ("lattice", 12, 1),
     [0.35, 64, 1000, 3, 11, 10 / CYCLE, 0.25 / np.pi, CYCLE / 2, CYCLE]),
    ("lattice12_p6_img2_ewald", "ref_ewald_p6", ("lattice", 12, 2),
     [0.35, 64, 1000, 2, 11, 10 / CYCLE, 0.25 / np.pi, 10.0, CYCLE]),
    ("cube2k_p10_img4_ewald", "ref_ewald_p10", ("cube", 2048, 3),
     [0.4, 32, 1000, 4, 11, 10 / CYCLE, 0.25 / np.pi, CYCLE / 2, CYCLE]),
    ("lattice12_p6_cart_img3_ewald", "ref_ewald_p6_cart", ("lattice", 12, 5),
     [0.35, 64, 1000, 3, 11, 10 / CYCLE, 0.25 / np.pi, CYCLE / 2, CYCLE]),
]


def jittered_lattice(m, seed):
    """m^3 bodies on a cubic lattice in [0, 2 pi)^3, each moved by a uniform
    jitter of +-0.25 spacing, with alternating +-1 charges (neutral for even m)."""
    h = CYCLE / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    x = (g + 0.5) * h + rng.uniform(-0.25, 0.25, g.shape) * h
    q = np.where(g.sum(1) % 2 == 0, 1.0, -1.0)
    return np.concatenate([x, q[:, None]], axis=1)


def periodic_cube(n, seed):
    """n bodies uniform in [0, 2 pi)^3 with zero-mean charges."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, CYCLE, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q -= q.mean()
    return np.concatenate([x, q[:, None]], axis=1)


def read_ewald_dump(path):
    b = open(path, "rb").read()
    assert b[:4] == b"EXEW"
    n = int(np.frombuffer(b, np.int32, 1, 4)[0])
    off = 8
    dip = np.frombuffer(b, np.float64, 3, off).copy()
    off += 24
    out = dict(dipole=dip)
    for k in ("fmm_raw", "fmm", "ewald"):
        out[k] = np.frombuffer(b, np.float64, 4 * n, off).reshape(n, 4).copy()
        off += 32 * n
    return out


def unit_cube(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def main(only=None):
    def wanted(name):
        return only is None or name in only

    with tempfile.TemporaryDirectory() as td:
        for name, binary in KERNEL_CHAIN_CASES:
            if not wanted(name):
                continue
            out = os.path.join(td, name + ".bin")
            print(binary, flush=True)
            subprocess.run([os.path.join(REF, binary), out], check=True)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **O.read_kernel_chain_dump(out))
        for name, binary, mode, opts, keep in CASES:
            if not wanted(name):
                continue
            out = os.path.join(td, name + ".bin")
            args = [os.path.join(REF, binary)]
            if mode[0] == "file":
                inp = os.path.join(td, name + ".in")
                x = unit_cube(16384, 7)
                with open(inp, "wb") as f:
                    np.array([x.shape[0]], np.int32).tofile(f)
                    x.tofile(f)
                args += ["file", inp, out]
            else:
                args += mode + [out]
            args += [str(v) for v in opts]
            print(" ".join(args[1:]), flush=True)
            subprocess.run(args, check=True)
            r = O.read_ref_dump(out)
            c = r["cells"]
            d = dict(theta=opts[0], ncrit=opts[1], nspawn=opts[2], useRmax=opts[3],
                     useRopt=opts[4], mutual=opts[5], images=opts[6], P=r["P"], num_p2p=r["num_p2p"], num_m2l=r["num_m2l"],
                     err_pot=r["err_pot"], err_acc=r["err_acc"], xmin=r["xmin"], xmax=r["xmax"],
                     ibody=r["ibody"], key=r["key"], sample_ibody=r["sample_ibody"],
                     sample_direct=r["sample_direct"], adaptive="_a_" in name, mass=name.startswith("mass"),
                     basis=1 if "_cart" in name else 0,
                     c_iparent=c["iparent"], c_ichild=c["ichild"], c_nchild=c["nchild"], c_ibody=c["ibody"],
                     c_nbody=c["nbody"], c_icell=c["icell"], c_X=c["X"], c_Rbuild=c["Rbuild"], c_R=c["R"])
            if mode[0] == "file":
                d["input"] = r["input"]
            else:
                # Inputs come from the reference generators, which the oracle
                # reproduces bit for bit (pinned by generators.npz).
                d["dist"] = mode[1]
                d["n"] = int(mode[2])
                regen = O.reference_dataset(mode[1], int(mode[2]), 0, name.startswith("mass"))
                assert np.array_equal(regen, r["input"]), name
            # FMM results of the sampled bodies (tree order -> caller order).
            pos = np.empty(r["n"], np.int64)
            pos[r["ibody"]] = np.arange(r["n"])
            d["sample_fmm"] = r["trg"][pos[r["sample_ibody"]]]
            if keep == "full":
                d["trg"] = r["trg"]
                d["M"] = c["M"]
                d["L"] = c["L"]
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
        for name, binary, (kind, size, seed), opts in EWALD_CASES:
            if not wanted(name):
                continue
            x = jittered_lattice(size, seed) if kind == "lattice" else periodic_cube(size, seed)
            inp = os.path.join(td, name + ".in")
            out = os.path.join(td, name + ".bin")
            with open(inp, "wb") as f:
                np.array([x.shape[0]], np.int32).tofile(f)
                x.tofile(f)
            args = [os.path.join(REF, binary), inp, out] + [repr(float(v)) if isinstance(v, float) else str(v)
                                                            for v in opts]
            print(" ".join(args[1:]), flush=True)
            subprocess.run(args, check=True)
            r = read_ewald_dump(out)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), input=x, theta=opts[0], ncrit=opts[1],
                                nspawn=opts[2], images=opts[3], ksize=opts[4], alpha=opts[5], sigma=opts[6],
                                cutoff=opts[7], cycle=opts[8], P=int(binary.split("_p")[1].split("_")[0]),
                                basis=1 if binary.endswith("_cart") else 0, **r)
    if not wanted("generators"):
        return
    # Generator fixtures: drand48 and the first bodies of every distribution.
    g = dict(drand48_seed0=O.drand48_seq(0, 16), drand48_seed42=O.drand48_seq(42, 16))
    for dist in "cspol":
        path = os.path.join(OUT, "gen_tmp.bin")
        subprocess.run([os.path.join(REF, "harness_u_p4"), "gen", dist, "1000", path], check=True,
                       stdout=subprocess.DEVNULL)
        g["dataset_" + dist] = O.read_ref_dump(path)["input"]
        os.remove(path)
    path = os.path.join(OUT, "gen_tmp.bin")
    subprocess.run([os.path.join(REF, "harness_u_p4_mass"), "gen", "p", "1000", path], check=True,
                   stdout=subprocess.DEVNULL)
    g["dataset_p_mass"] = O.read_ref_dump(path)["input"]
    os.remove(path)
    np.savez_compressed(os.path.join(OUT, "generators.npz"), **g)


if __name__ == "__main__":
    # Optional: the fixture names to (re)write, e.g. kernel_chain_p10; all if none.
    main(set(sys.argv[1:]) or None)
