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
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 (vector and MFMA)
PEAK_HBM_GBS = 8000.0


def p2p_kernel_name(tuning, form):
    """Kernel name of the P2P variant the library ran (get_tuning:
    p2p_targets_per_lane = 1 for the one-target-per-lane kernel)."""
    k = tuning["p2p_targets_per_lane"]
    return f"k_p2p_{form}" if k == 1 else f"k_p2p_{form}_rb<{k}>"


def m2l_flops_per_pair(P):
    # Complex multiply-adds of LaplaceSphericalCPU.cxx:156-190, 8 flops each.
    return 8 * P * (P + 1) ** 2 * (P + 2) // 12


def unit_cube(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


# How oracle/Makefile builds the two CPU baselines.
REF_FLAGS = "g++ -O3 -fopenmp -DOPENMP -DSpherical -DFP64 -DEXPANSION=10 (examples/serial.cxx + kernels/*CPU.cxx)"
ORACLE_FLAGS = "g++ -O2 -std=c++17 -fopenmp -ffp-contract=off (oracle/oracle.cpp)"


def host_info():
    """CPU model and hardware threads of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "threads_available": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


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
        "compiler_flags": REF_FLAGS,
        "host": host_info(),
    }


def cpu_baseline_all_cores(x, P):
    """The same evaluation by the CPU restatement of the reference path
    (oracle/), its kernel calls spread over every host thread (target cells
    in parallel, each in the reference's order; passes level by level): the
    all-cores CPU figure beside the reference's single-core one."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle
    except Exception:
        return None
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    o = oracle.Oracle(P=P, theta=0.4, ncrit=64, nspawn=1000, useRmax=1, useRopt=0, threads=threads)
    o.set_bodies(x)
    t0 = time.perf_counter()
    o.run()
    t = time.perf_counter() - t0
    n = x.shape[0]
    return {"value": n / t, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": (f"CPU restatement of the reference path (oracle/oracle.cpp, bit-identical trees and lists, "
                       f"results to 1e-12), N={n} cube, P={P}, theta=0.4, ncrit=64, non-mutual, {threads} threads; "
                       f"bounds to downward pass {t:.2f} s"),
            "compiler_flags": ORACLE_FLAGS, "host": host_info()}


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


def rel_err(res, ref):
    pot = float(np.sqrt(np.sum((res[:, 0] - ref[:, 0]) ** 2) / np.sum(ref[:, 0] ** 2)))
    acc = float(np.sqrt(np.sum((res[:, 1:] - ref[:, 1:]) ** 2) / np.sum(ref[:, 1:] ** 2)))
    return pot, acc


def bench_batched(args):
    """256 independent systems, N log-uniform in [4096, 262144], half uniform
    cube, half uniform on the unit sphere surface, zero-mean charges; P = 8,
    theta = 0.5; ncrit, the P2P grid and targets per lane and the contraction
    grid chosen by the library's autotuner (exafmm_hip_set_autotune).  value =
    total bodies / device time of one batched evaluation (HIP events); the
    wall time per call including the host copies of inputs and results is
    reported beside it; compared with one system of the same total N (uniform
    cube, tuned the same way)."""
    rng = np.random.default_rng(2024)
    sizes = np.floor(np.exp(rng.uniform(np.log(4096), np.log(262144), 256))).astype(np.int64)
    systems = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            x = rng.uniform(-1, 1, (n, 3))
        else:
            v = rng.normal(size=(n, 3))
            x = v / np.linalg.norm(v, axis=1)[:, None]
        q = rng.uniform(-0.5, 0.5, n)
        q = (q - q.mean()) / n
        systems.append(np.concatenate([x, q[:, None]], axis=1))
    total = int(sizes.sum())
    steps = max(1, args.steps)
    # The library's autotuner picks ncrit, the P2P grid and targets per lane
    # and the contraction grid on the first call of this input signature.
    f = exafmm_hip.FMM(P=8, theta=0.5, ncrit=64)
    f.set_autotune(True)
    t0 = time.perf_counter()
    out = f.evaluate_batch(systems)
    tune_s = time.perf_counter() - t0
    tuned = f.tuning()
    ncrit = tuned["ncrit"]
    dev, wall = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        out = f.evaluate_batch(systems)
        wall.append((time.perf_counter() - t0) * 1e3)
        dev.append(f.stats()["t_total"])
    ms = float(np.median(dev))
    ms_wall = float(np.median(wall))
    # New charges on the same systems: the forest's numeric passes replayed
    # from a HIP graph (exafmm_hip_evaluate_charges), results left on the
    # device (wall time includes the charge upload and the device work).
    qs = np.concatenate([sy[:, 3] for sy in systems])
    f.evaluate_charges(qs * 0.5)
    rdev, rwall = [], []
    for k in range(steps):
        t0 = time.perf_counter()
        f.evaluate_charges(qs * (1.0 + 0.01 * k))
        rwall.append((time.perf_counter() - t0) * 1e3)
        rdev.append(f.stats()["t_total"])
    f.evaluate_charges(qs)
    out = f.results_batch()
    rms = float(np.median(rdev))
    # Per-system error against FP64 direct summation on 64 sampled targets.
    errs = []
    for x, r in zip(systems, out):
        idx = rng.choice(x.shape[0], 64, replace=False)
        ref = f.direct_arrays(x[idx], x)
        errs.append(rel_err(r[idx], ref))
    errs = np.array(errs)
    f.close()
    # One system of the same total N.
    g = exafmm_hip.FMM(P=8, theta=0.5, ncrit=ncrit)
    g.set_autotune(True)  # the single system gets its own tuning
    xs = unit_cube(total, 99)
    g.set_bodies(xs)
    g.evaluate()
    sdev = []
    for _ in range(steps):
        g.evaluate()
        sdev.append(g.stats()["t_total"])
    g.evaluate_charges(xs[:, 3] * 0.5)
    srdev = []
    for k in range(steps):
        g.evaluate_charges(xs[:, 3] * (1.0 + 0.01 * k))
        srdev.append(g.stats()["t_total"])
    g.close()
    sms = float(np.median(sdev))
    srms = float(np.median(srdev))
    line = {
        "metric": "batched FMM particles/s: 256 independent systems, N log-uniform in [4096, 262144], P=8, theta=0.5",
        "value": total / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": steps,
        "ms_per_step": ms, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (seeded; half uniform cube, half uniform on the unit sphere surface; zero-mean charges)",
        "config": {"workload": "batched evaluation, one forest of per-system uniform trees", "systems": len(sizes),
                   "total_N": total, "N_min": int(sizes.min()), "N_max": int(sizes.max()), "P": 8, "theta": 0.5,
                   "ncrit": ncrit},
        "autotune": {"chosen": tuned, "first_call_s_incl_tuning": round(tune_s, 3)},
        "ms_wall_incl_host_copies": round(ms_wall, 3),
        "single_system_same_N": {"value": total / (sms * 1e-3), "ms": sms, "workload": "uniform cube, same N, ncrit"},
        "ratio_vs_single_system": (total / (ms * 1e-3)) / (total / (sms * 1e-3)),
        "charges_only_replay": {
            "workload": "same 256 systems, new charges: the forest's numeric passes replayed as one HIP graph launch",
            "ms_device": round(rms, 4), "ms_wall_incl_host_copy": round(float(np.median(rwall)), 4),
            "value": total / (rms * 1e-3), "single_system_same_N_ms": round(srms, 4),
            "ratio_vs_single_system": srms / rms},
        "rel_l2_err_per_system": {"pot_max": float(errs[:, 0].max()), "acc_max": float(errs[:, 1].max()),
                                  "acc_median": float(np.median(errs[:, 1])), "targets_per_system": 64},
    }
    print(json.dumps(line), flush=True)


def plummer(n, seed):
    """Plummer sphere, scale radius 1, radius cut at 10, isotropic; charges 1/N."""
    rng = np.random.default_rng(seed)
    out = np.empty((0, 3))
    while out.shape[0] < n:
        m = int((n - out.shape[0]) * 1.05) + 16
        u = rng.uniform(0, 1, m)
        r = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        v = rng.normal(size=(m, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        out = np.concatenate([out, (v * r[:, None])[r < 10]], axis=0)
    x = out[:n]
    return np.concatenate([x, np.full((n, 1), 1.0 / n)], axis=1)


def bench_plummer(args):
    """N = 4,194,304 Plummer bodies (adaptive tree), P in {4, 8, 12},
    theta = 0.4, FP64 and mixed; error vs FP64 direct on 1024 targets."""
    n = 1 << 22
    x = plummer(n, 5)
    rng = np.random.default_rng(9)
    tgt = rng.choice(n, 1024, replace=False).astype(np.int32)
    rows = {}
    for P in (4, 8, 12):
        for prec in ("fp64", "mixed"):
            f = exafmm_hip.FMM(P=P, theta=0.4, ncrit=64, builder="adaptive", precision=prec)
            f.set_bodies(x)
            f.evaluate()
            dev = []
            for _ in range(max(1, args.steps)):
                f.evaluate()
                dev.append(f.stats()["t_total"])
            res = f.results()[tgt]
            ref = f.direct(tgt)
            pot, acc = rel_err(res, ref)
            ms = float(np.median(dev))
            rows[f"P{P}_{prec}"] = {"ms": round(ms, 3), "particles_per_s": n / (ms * 1e-3), "rel_l2_pot": pot,
                                   "rel_l2_acc": acc}
            f.close()
    print(json.dumps({"metric": "FMM particles/s, Plummer N=4M, P in {4,8,12}, FP64 and mixed",
                      "config": {"N": n, "theta": 0.4, "ncrit": 64, "tree": "adaptive", "targets": 1024,
                                 "data": "synthetic Plummer (seeded), charges 1/N"},
                      "results": rows}), flush=True)


def bench_periodic(args):
    """N = 128^3 jittered rock-salt lattice in [0, 2 pi)^3, images = 3,
    theta = 0.35, spherical P = 10 and Cartesian P = 6; FMM (+ dipole
    correction) against the Ewald sum (alpha = 10/cycle, ksize = 11,
    cutoff = cycle/2): relative error of sum q phi and of the gradient."""
    m = 128
    cyc = 2 * np.pi
    h = cyc / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(4)
    x = np.concatenate([(g + 0.5) * h + rng.uniform(-0.25, 0.25, g.shape) * h,
                        np.where(g.sum(1) % 2 == 0, 1.0, -1.0)[:, None]], axis=1)
    rows = {}
    ewald = None
    for basis, P in (("spherical", 10), ("cartesian", 6)):
        f = exafmm_hip.FMM(P=P, theta=0.35, ncrit=64, images=3, cycle=cyc, basis=basis)
        f.set_bodies(x)
        f.evaluate()
        dev = []
        for _ in range(max(1, args.steps)):
            f.evaluate()
            dev.append(f.stats()["t_total"])
        d = f.dipole()
        f.dipole_correction(d, x.shape[0], cyc)
        r = f.results()
        if ewald is None:
            t0 = time.perf_counter()
            ewald = f.ewald(11, 10 / cyc, 0.25 / np.pi, cyc / 2, cyc)
            t_ew = time.perf_counter() - t0
        q = x[:, 3]
        ps, pe = float(np.sum(q * r[:, 0])), float(np.sum(q * ewald[:, 0]))
        acc = float(np.sqrt(np.sum((r[:, 1:] - ewald[:, 1:]) ** 2) / np.sum(ewald[:, 1:] ** 2)))
        ms = float(np.median(dev))
        rows[f"{basis}_P{P}"] = {"ms": round(ms, 3), "particles_per_s": x.shape[0] / (ms * 1e-3),
                                 "rel_err_sum_q_phi": abs(ps - pe) / abs(pe), "rel_l2_grad": acc}
        f.close()
    print(json.dumps({"metric": "periodic FMM particles/s, 128^3 lattice, images=3, vs Ewald",
                      "config": {"N": int(x.shape[0]), "theta": 0.35, "images": 3, "ewald_s": round(t_ew, 3),
                                 "ewald": {"ksize": 11, "alpha": 10 / cyc, "cutoff": cyc / 2}},
                      "results": rows}), flush=True)


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
    ap.add_argument("--no-charges-replay", dest="charges_replay", action="store_false",
                    help="skip the timing of new-charge re-evaluations replayed from a HIP graph")
    ap.add_argument("--no-isolated-rates", dest="isolated_rates", action="store_false",
                    help="skip the 3 extra evaluations that time the dominant kernels on an otherwise idle GPU")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--config", default="headline", choices=["headline", "batched", "plummer", "periodic"],
                    help="headline: the N=1M P=10 line; the others measure the batched, Plummer and periodic workloads")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.config != "headline":
        if args.gpus != 1:
            print("bench.py measures one GPU: run with --gpus 1.", file=sys.stderr)
            sys.exit(2)
        {"batched": bench_batched, "plummer": bench_plummer, "periodic": bench_periodic}[args.config](args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the evaluation is not sharded over several GPUs. Run with --gpus 1.",
              file=sys.stderr)
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
        tuning = f.tuning()
        # The same bodies with new charges (exafmm_hip_evaluate_charges): the
        # numeric passes only, replayed from the recorded HIP graph.
        charges = None
        if args.charges_replay:
            q0 = x[:, 3]
            f.evaluate_charges(q0 * 0.5)  # first call: ordinary launches, then recorded
            dev, wall = [], []
            for k in range(args.steps):
                t0 = time.perf_counter()
                f.evaluate_charges(q0 * (1.0 + 0.01 * k))
                wall.append((time.perf_counter() - t0) * 1e3)
                dev.append(f.stats()["t_total"])
            f.evaluate_charges(q0)  # leaves the results of the original charges
            cms = float(np.median(dev))
            charges = {"ms_device": round(cms, 4), "ms_wall_incl_host_copy": round(float(np.median(wall)), 4),
                       "value": n / (cms * 1e-3), "unit": "particles/s",
                       "workload": ("same bodies, new charges: P2M, M2M, M2L, L2L, L2P and P2P replayed from a "
                                    "HIP graph (tree, lists and M2L set-up kept)")}
        # Accuracy against FP64 direct summation on 128 random targets.
        res = f.results()[tgt]
        ref = f.direct(tgt)
        err_pot = float(np.sqrt(np.sum((res[:, 0] - ref[:, 0]) ** 2) / np.sum(ref[:, 0] ** 2)))
        err_acc = float(np.sqrt(np.sum((res[:, 1:] - ref[:, 1:]) ** 2) / np.sum(ref[:, 1:] ** 2)))
        f.close()
        med = {k: float(np.median([d[k] for d in stats])) for k in stats[0] if k.startswith("t_")}
        return ms, stats[-1], med, err_pot, err_acc, tuning, charges

    ms, s, med, err_pot, err_acc, fp64_tuning, fp64_charges = run("fp64")
    value = n / (ms * 1e-3)
    mixed = None
    mixed_med, mixed_tuning = None, None
    if args.precision == "mixed" or args.with_mixed:
        mms, ms_stats, mmed, mpot, macc, mixed_tuning, mixed_charges = run("mixed")
        mixed_med = mmed
        mixed = {"value": n / (mms * 1e-3), "unit": "particles/s", "ms_per_step": mms,
                 "speedup_vs_fp64": ms / mms, "rel_l2_err": {"pot": mpot, "acc": macc, "targets": 128},
                 "phases_ms": {k: round(v, 4) for k, v in mmed.items()},
                 "charges_only_replay": mixed_charges,
                 "dtype": "f32 P2P (compensated) and f32 M2L tables; f64 centres, offsets, expansions"}

    # Per-kernel rates (algorithmic work / HIP-event time of that kernel).
    # P2P and the per-pair M2L run on the vector ALU, the lattice M2L on the
    # matrix cores; on MI355X the FP64 vector and matrix peaks are equal.
    t_m2l, t_p2p = med["t_m2l"], med["t_p2p"]
    t_mfma, t_pairs = med["t_m2l_mfma"], med["t_m2l_pairs"]
    n_lat = s["num_m2l_lattice"]
    n_irr = s["num_m2l_pairs"] - n_lat
    fpp = m2l_flops_per_pair(P)
    p2p_pairs = s["num_p2p_bodies"]
    p2p_name = p2p_kernel_name(fp64_tuning, "f64")
    kern = {
        "k_m2l_mfma<%d,double>" % P: (n_lat * fpp, t_mfma, "mfma", PEAK_FP64_TFLOPS),
        "k_m2l<%d,double>" % P: (n_irr * fpp, t_pairs, "valu", PEAK_FP64_TFLOPS),
        p2p_name: (20 * p2p_pairs, t_p2p, "valu", PEAK_FP64_TFLOPS),
    }
    kernels_tflops = {k: round(w / (t * 1e-3) / 1e12, 4) for k, (w, t, _, _) in kern.items() if t > 0}
    dom_name = max((k for k in kern if kern[k][1] > 0), key=lambda k: kern[k][1])

    def rate(name, work, t, bound, peak):
        a = work / (t * 1e-3) / 1e12
        return {"bound": bound, "achieved": round(a, 4), "peak": peak, "unit": "TFLOP/s", "frac": round(a / peak, 5),
                "kernel_ms": round(t, 4)}

    w, t, bound, peak = kern[dom_name]
    roofline = rate(dom_name, w, t, bound, peak)
    roofline.update({"traffic": measured_traffic(dom_name) if n == 1 << 20 and P == 10 else None,
                     "kernel": dom_name,
                     "note": ("dominant kernel = the longest HIP-event duration in the timed steps, where the "
                              "contraction and P2P share the CUs; P2P runs on the FP64 vector ALU "
                              "(bound 'valu'), whose peak equals the FP64 matrix peak")})
    # Both heavy kernels of each form, in the timed steps.
    roofline["kernels"] = {k: dict(rate(k, *kern[k]), traffic=measured_traffic(k) if n == 1 << 20 and P == 10 else None)
                           for k in (p2p_name, "k_m2l_mfma<%d,double>" % P) if kern[k][1] > 0}
    if mixed is not None:
        mp2p = p2p_kernel_name(mixed_tuning, "mixed")
        # FP32 arithmetic: against the published FP32 vector / matrix peak.
        for k, work, tk, bound, peak in ((mp2p, 20 * p2p_pairs, mixed_med["t_p2p"], "valu", PEAK_FP32_TFLOPS),
                                         ("k_m2l_mfma<%d,float>" % P, n_lat * fpp, mixed_med["t_m2l_mfma"], "mfma",
                                          PEAK_FP32_TFLOPS)):
            if tk > 0:
                roofline["kernels"][k] = dict(rate(k, work, tk, bound, peak),
                                              traffic=measured_traffic(k) if n == 1 << 20 and P == 10 else None)
    # The same two kernels with the GPU to themselves: with EXAFMM_HIP_SERIAL=1
    # the contraction waits for P2P to finish (P2P keeps its grid; only the
    # light lattice set-up runs beside it); 3 evaluations after the timed
    # ones, outside the timed region.
    if args.isolated_rates:
        old_env = os.environ.get("EXAFMM_HIP_SERIAL")
        os.environ["EXAFMM_HIP_SERIAL"] = "1"
        try:
            f = exafmm_hip.FMM(P=P, theta=0.4, ncrit=64, nspawn=1000, useRmax=1, useRopt=0, builder="uniform")
            f.set_bodies(x)
            f.evaluate()
            iso = []
            for _ in range(3):
                f.evaluate()
                iso.append(f.stats())
            f.synchronize()
            f.close()
        finally:
            if old_env is None:
                del os.environ["EXAFMM_HIP_SERIAL"]
            else:
                os.environ["EXAFMM_HIP_SERIAL"] = old_env
        iso_t = {"k_m2l_mfma<%d,double>" % P: float(np.median([d["t_m2l_mfma"] for d in iso])),
                 p2p_name: float(np.median([d["t_p2p"] for d in iso]))}
        roofline["isolated"] = {}
        for k, ti in iso_t.items():
            if ti > 0:
                a = kern[k][0] / (ti * 1e-3) / 1e12
                roofline["isolated"][k] = {"kernel_ms": round(ti, 4), "achieved": round(a, 4),
                                           "frac": round(a / kern[k][3], 5)}

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, P)
    cpu_all = None if args.no_cpu_baseline else cpu_baseline_all_cores(x, P)
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
                "tiles": s["num_m2l_tiles"], "steps": s["num_m2l_steps"],
                "busy_columns": round(n_lat / (64 * s["num_m2l_steps"]), 4) if s["num_m2l_steps"] else None},
        "rel_l2_err": {"pot": err_pot, "acc": err_acc, "targets": 128},
        "phases_ms": {k: round(v, 4) for k, v in med.items()},
        # The same phases under the reference's logger timer names
        # (bound_box.h:77, build_tree_omp.h, up_down_pass.h:120,136,
        # traversal.h:242 -- its Traverse evaluates P2P and M2L --, serial.cxx:43).
        "phases_ms_reference_names": {"Get bounds": round(med["t_bounds"], 4), "Build tree": round(med["t_build"], 4),
                                      "Upward pass": round(med["t_upward"], 4),
                                      "Traverse": round(med["t_traverse"], 4),
                                      "Downward pass": round(med["t_downward"], 4),
                                      "Total FMM": round(med["t_total"], 4)},
        "charges_only_replay": fp64_charges,
        "counts": {"cells": s["num_cells"], "p2p_cell_pairs": s["num_p2p_pairs"],
                   "m2l_cell_pairs": s["num_m2l_pairs"], "p2p_body_pairs": p2p_pairs},
        "kernels_tflops": kernels_tflops,
        "roofline": roofline,
        "mixed": mixed,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
