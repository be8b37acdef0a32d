"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs in the development container only: it executes the reference harness
binaries built by `make -C oracle ref` (oracle/ref_harness.cxx compiled
against the reference sources) and stores their inputs and outputs as small
.npz files.  The fixtures are data -- inputs and expected outputs -- and are
what the CPU tests pin the oracle against.

Usage: python tests/golden/make_golden.py
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
]


def unit_cube(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def main():
    with tempfile.TemporaryDirectory() as td:
        for name, binary, mode, opts, keep in CASES:
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
    main()
