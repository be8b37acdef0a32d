"""GPU mixed-precision form against the plain FP64 GPU path and FP64 direct.

The mixed form computes P2P in FP32 with per-target compensated (Kahan)
accumulation (kahan.h:35-41) and M2L with FP32 harmonic tables; cell
centres, offsets, expansions and accumulators stay FP64.  It shares the
tree and the interaction lists with the FP64 path, so the bar is:

* identical interaction-pair sets;
* results within 1e-6 (potential) and 5e-7 (gradient) (relative L2 over all
  bodies) of the FP64 FMM -- the FP32 rounding of single terms is ~6e-8, the
  compensated P2P sums keep the accumulation error near that level, and the
  FP32 MFMA M2L sums up to 110 products per step in FP32 (potential, whose
  zero-mean sources cancel most, measured 4-7e-7; gradient ~1e-7);
* error against FP64 direct summation no worse than the FP64 FMM's own
  truncation error times 1.02 plus that FP32 floor (3e-7 absolute), which
  keeps every case below the 1e-6 force-error bar wherever the FP64 FMM is.
"""
import numpy as np
import pytest

import exafmm_hip as X
from helpers import case_params, load_case, plummer, rel_l2, unit_cube

pytestmark = pytest.mark.gpu

CASES = [("cube16k_u_p10", "uniform"), ("plummer16k_u_p10", "uniform"), ("cube16k_a_p10", "adaptive"),
         ("mass_plummer16k_a_p10", "adaptive"), ("sphere16k_a_p8", "adaptive"), ("cube4k_a_p12", "adaptive"),
         ("octant8k_a_p6", "adaptive"), ("cube4k_u_p4", "uniform"), ("cube16k_u_p10_mut", "uniform")]


def make(z, builder, precision):
    p = case_params(z)
    print(f"[gpu] mixed N={z['input'].shape[0]} P={p['P']} builder={builder} precision={precision}", flush=True)
    f = X.FMM(P=p["P"], theta=p["theta"], ncrit=p["ncrit"], nspawn=p["nspawn"], useRmax=p["useRmax"],
              useRopt=p["useRopt"], images=p["images"], builder=builder, mass=p["mass"], mutual=int(p["mutual"]),
              precision=precision)
    f.set_bodies(z["input"])
    f.evaluate()
    return f


@pytest.mark.parametrize("name,builder", CASES)
def test_mixed_against_fp64(name, builder):
    z = load_case(name)
    f64 = make(z, builder, "fp64")
    fmx = make(z, builder, "mixed")
    a, b = f64.lists(), fmx.lists()
    for k in ("p2p_i", "p2p_j", "m2l_i", "m2l_j"):
        assert np.array_equal(a[k], b[k]), k
    r64, rmx = f64.results(), fmx.results()
    assert np.all(np.isfinite(rmx))
    assert rel_l2(rmx[:, 0], r64[:, 0]) < 1e-6
    assert rel_l2(rmx[:, 1:], r64[:, 1:]) < 5e-7
    # Error against FP64 direct on the reference's sampled bodies.
    ref = z["sample_direct"]
    idx = z["sample_ibody"]
    for sl in (slice(0, 1), slice(1, 4)):
        e64 = rel_l2(r64[idx][:, sl], ref[:, sl])
        emx = rel_l2(rmx[idx][:, sl], ref[:, sl])
        assert emx <= 1.02 * e64 + 3e-7, (sl, emx, e64)
    f64.close()
    fmx.close()


@pytest.mark.parametrize("kind", ["cube", "plummer"])
def test_mixed_scale_invariance(kind):
    """Shrinking all coordinates by 2^-12 (deep, tiny cells whose unscaled
    FP32 harmonics would overflow) changes the normalised results by exact
    powers of two only up to FP32 rounding."""
    x = unit_cube(20000, 11) if kind == "cube" else plummer(20000, 11)
    xs = x.copy()
    xs[:, :3] *= 2.0 ** -12
    print(f"[gpu] mixed scale invariance {kind}", flush=True)
    out = []
    for xx in (x, xs):
        f = X.FMM(P=10, ncrit=64, builder="adaptive", precision="mixed")
        f.set_bodies(xx)
        f.evaluate()
        out.append(f.results())
        f.close()
    a, b = out
    assert np.all(np.isfinite(b))
    assert rel_l2(b[:, 0] * 2.0 ** -12, a[:, 0]) < 1e-6
    assert rel_l2(b[:, 1:] * 2.0 ** -24, a[:, 1:]) < 1e-6


def test_mixed_direct_error_bound():
    """Mixed-form error against FP64 direct on 128 random targets at N=65536,
    P=10, theta=0.4: within 2% of the FP64 FMM's error plus the FP32 floor."""
    x = unit_cube(65536, 2024)
    tgt = np.random.default_rng(7).choice(x.shape[0], 128, replace=False).astype(np.int32)
    errs = {}
    for prec in ("fp64", "mixed"):
        print(f"[gpu] direct error bound {prec}", flush=True)
        f = X.FMM(P=10, ncrit=64, precision=prec)
        f.set_bodies(x)
        f.evaluate()
        ref = f.direct(tgt)
        res = f.results()[tgt]
        errs[prec] = (rel_l2(res[:, 0], ref[:, 0]), rel_l2(res[:, 1:], ref[:, 1:]))
        f.close()
    for i in range(2):
        assert errs["mixed"][i] <= 1.02 * errs["fp64"][i] + 3e-7, errs


def test_mixed_close_pairs_across_leaves():
    """Close pairs in different leaves: 2,048 pairs of bodies straddle the
    uniform tree's leaf faces 1e-6 leaf widths apart, among 16,384 uniform
    bodies.  Their separation is far below one FP32 ulp of the neighbourhood's
    extent, so the mixed P2P must add the positions' lo parts for them (an
    unmasked stage: the partner sits in another leaf); without that the force
    on those bodies would be off by ~10%.  Bar: the mixed form within 1e-6 /
    5e-7 of the FP64 FMM, and its error against direct summation on the
    straddling bodies within 1.02x of the FP64 form's plus 1e-6 (their forces
    are dominated by the exact pair term, so both errors sit at the FP32 /
    FP64 rounding of that term)."""
    rng = np.random.default_rng(21)
    n0, npair = 16384, 2048
    x = unit_cube(n0, 22)
    x[0, :3] = -1.0  # the bounding box is [-1, 1]^3
    x[1, :3] = 1.0
    n = n0 + 2 * npair
    # Uniform build (build_tree_omp.h:216-231): L levels, cell size d (float)
    # from R0 = 1.00001, keys measured from the lower bound -1.
    L = 1 + int(np.log(n // 64) / np.log(2) / 3)
    w = float(np.float32(2 * 1.00001 / (1 << L)))
    face = -1.0 + rng.integers(1, (1 << L), npair) * w  # interior leaf faces along x
    eps = 1e-6 * w
    y = rng.uniform(-0.9, 0.9, (npair, 2))
    q = rng.uniform(-0.5, 0.5, 2 * npair) / n0
    a = np.column_stack([face - eps / 2, y, q[:npair]])
    b = np.column_stack([face + eps / 2, y, q[npair:]])
    bodies = np.concatenate([x, a, b])
    bodies[:, 3] -= bodies[:, 3].mean()
    out = {}
    for prec in ("fp64", "mixed"):
        print(f"[gpu] mixed close pairs {prec}", flush=True)
        f = X.FMM(P=10, theta=0.4, ncrit=64, precision=prec)
        f.set_bodies(bodies)
        f.evaluate()
        out[prec] = f.results()
        if prec == "fp64":
            ref = f.direct(np.arange(n0, n, dtype=np.int32))
        f.close()
    r64, rmx = out["fp64"], out["mixed"]
    assert np.all(np.isfinite(rmx))
    assert rel_l2(rmx[:, 0], r64[:, 0]) < 1e-6
    assert rel_l2(rmx[:, 1:], r64[:, 1:]) < 5e-7
    for sl in (slice(0, 1), slice(1, 4)):
        e64 = rel_l2(r64[n0:][:, sl], ref[:, sl])
        emx = rel_l2(rmx[n0:][:, sl], ref[:, sl])
        print(f"[gpu] close pairs {sl}: FP64 {e64:.3e} mixed {emx:.3e}", flush=True)
        assert emx <= 1.02 * e64 + 1e-6, (sl, emx, e64)
