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
