"""Benchmark of the FMM evaluation path on one MI355X.

Workload (one step): a full Laplace FMM evaluation -- bounds, uniform
Morton-tree build, upward pass, dual-tree traversal lists, P2P + M2L,
downward pass -- of N = 1,048,576 bodies uniform in [-1,1]^3 with zero-mean
charges divided by N, spherical P = 10, theta = 0.4, ncrit = 64, in FP64.
The bodies are resident in HBM before the timed region; each step rebuilds
everything from them.

Prints one JSON line (metric, value = particles/s, roofline of the dominant
kernel measured with HIP events on the library's stream, and the reference
CPU path timed on this host).
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "exafmm-dev_amd"))

import exafmm_hip  # noqa: E402

METRIC = "FMM particles/s at N=1M P=10, mixed vs FP64 GPU; P2P Ginter/s; rel L2 err"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and MFMA)
PEAK_HBM_GBS = 8000.0


def m2l_flops_per_pair(P):
    # Complex multiply-adds of LaplaceSphericalCPU.cxx:156-190, 8 flops each.
    return 8 * P * (P + 1) ** 2 * (P + 2) // 12


def unit_cube(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def cpu_baseline(n, P):
    """Time the reference's own driver (examples/serial.cxx built from the
    reference sources into oracle/_ref/) on this host, one thread."""
    exe = os.path.join(ROOT, "oracle", "_ref", f"serial_p{P}")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OMP_NUM_THREADS="1")
    cmd = [exe, "-n", str(n), "-c", "64", "-t", "0.4", "-o", "0", "-d", "c", "-T", "1", "-v", "1"]
    t0 = time.time()
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600).stdout
    wall = time.time() - t0
    m = re.search(r"Total FMM\s*:\s*([0-9.eE+-]+)", out)
    if not m:
        return None
    t = float(m.group(1))
    acc = re.search(r"Rel. L2 Error \(acc\)\s*:\s*([0-9.eE+-]+)", out)
    return {
        "value": n / t,
        "unit": "particles/s",
        "cores": 1,
        "kind": "reference",
        "sample": (f"reference examples/serial.cxx (OpenMP build, uniform tree, mutual traversal), "
                   f"N={n} cube, P={P}, theta=0.4, ncrit=64, useRopt=0, 1 thread; Total FMM {t:.2f} s "
                   f"(wall {wall:.1f} s), its rel-L2 acc error {float(acc.group(1)) if acc else float('nan'):.2e}"),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines `traffic <kernel> <bytes>`
    in profiles/*.md), or None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(pdir, name)) as fh:
            for line in fh:
                m = re.match(r"\s*traffic\s+(\S+)\s+(\d+)\s*$", line)
                if m and m.group(1) == kernel:
                    return int(m.group(2))
    return None


def copy_probe():
    """One device-to-device copy of 1 GiB (reads 1 GiB, writes 1 GiB): the
    known-size kernel that scales the HBM byte counters."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    b = a.clone()
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 8, "ok": bool(b[-1].item() == 1.0)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--P", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--with-mixed", action="store_true", default=True)
    ap.add_argument("--no-mixed", dest="with_mixed", action="store_false")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X; multi-GPU sharding is design-only "
              "(see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    n, P = args.n, args.P
    x = unit_cube(n, 2024)
    rng = np.random.default_rng(7)
    tgt = rng.choice(n, 128, replace=False).astype(np.int32)

    def run(precision):
        f = exafmm_hip.FMM(P=P, theta=0.4, ncrit=64, nspawn=1000, useRmax=1, useRopt=0, builder="uniform",
                           precision=precision)
        f.set_bodies(x)  # host -> HBM once; steps read the resident copy
        for _ in range(args.warmup):
            f.evaluate()
        f.synchronize()
        stats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f.evaluate()
            stats.append(f.stats())
        f.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        stats[-1]["num_p2p_bodies"] = f.count_p2p_bodies()  # outside the timed steps
        # Accuracy against FP64 direct summation on 128 random targets.
        res = f.results()[tgt]
        ref = f.direct(tgt)
        err_pot = float(np.sqrt(np.sum((res[:, 0] - ref[:, 0]) ** 2) / np.sum(ref[:, 0] ** 2)))
        err_acc = float(np.sqrt(np.sum((res[:, 1:] - ref[:, 1:]) ** 2) / np.sum(ref[:, 1:] ** 2)))
        f.close()
        med = {k: float(np.median([d[k] for d in stats])) for k in stats[0] if k.startswith("t_")}
        return ms, stats[-1], med, err_pot, err_acc

    ms, s, med, err_pot, err_acc = run("fp64")
    value = n / (ms * 1e-3)
    mixed = None
    if args.precision == "mixed" or args.with_mixed:
        mms, ms_stats, mmed, mpot, macc = run("mixed")
        mixed = {"value": n / (mms * 1e-3), "unit": "particles/s", "ms_per_step": mms,
                 "speedup_vs_fp64": ms / mms, "rel_l2_err": {"pot": mpot, "acc": macc, "targets": 128},
                 "phases_ms": {k: round(v, 4) for k, v in mmed.items()},
                 "dtype": "f32 P2P (compensated) and f32 M2L tables; f64 centres, offsets, expansions"}

    # Per-kernel rates (algorithmic work / HIP-event time of that kernel).
    t_m2l, t_p2p = med["t_m2l"], med["t_p2p"]
    t_mfma, t_pairs = med["t_m2l_mfma"], med["t_m2l_pairs"]
    n_lat = s["num_m2l_lattice"]
    n_irr = s["num_m2l_pairs"] - n_lat
    fpp = m2l_flops_per_pair(P)
    p2p_pairs = s["num_p2p_bodies"]
    kern = {
        "k_m2l_mfma<%d,double>" % P: (n_lat * fpp, t_mfma, "mfma", PEAK_FP64_TFLOPS),
        "k_m2l<%d,double>" % P: (n_irr * fpp, t_pairs, "mfma", PEAK_FP64_TFLOPS),
        "k_p2p_f64": (20 * p2p_pairs, t_p2p, "mfma", PEAK_FP64_TFLOPS),
    }
    kernels_tflops = {k: round(w / (t * 1e-3) / 1e12, 4) for k, (w, t, _, _) in kern.items() if t > 0}
    dom_name = max((k for k in kern if kern[k][1] > 0), key=lambda k: kern[k][1])
    w, t, bound, peak = kern[dom_name]
    ach = w / (t * 1e-3) / 1e12
    roofline = {"bound": bound, "achieved": round(ach, 4), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 5), "traffic": measured_traffic(dom_name) if n == 1 << 20 and P == 10 else None,
                "kernel": dom_name, "kernel_ms": round(t, 4)}

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, P)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (U[-1,1]^3 bodies, zero-mean uniform charges / N, seeded)",
        "config": {"workload": "plain FP64 GPU FMM evaluation, Laplace spherical", "N": n, "P": P, "theta": 0.4,
                   "ncrit": 64, "tree": "uniform Morton (build_tree_omp.h)", "precision": "fp64"},
        "p2p_ginter_per_s": p2p_pairs / (t_p2p * 1e-3) / 1e9,
        "m2l": {"pairs": s["num_m2l_pairs"], "lattice_pairs": n_lat, "offsets": s["num_m2l_offsets"],
                "tiles": s["num_m2l_tiles"]},
        "rel_l2_err": {"pot": err_pot, "acc": err_acc, "targets": 128},
        "phases_ms": {k: round(v, 4) for k, v in med.items()},
        "counts": {"cells": s["num_cells"], "p2p_cell_pairs": s["num_p2p_pairs"],
                   "m2l_cell_pairs": s["num_m2l_pairs"], "p2p_body_pairs": p2p_pairs},
        "kernels_tflops": kernels_tflops,
        "roofline": roofline,
        "mixed": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
