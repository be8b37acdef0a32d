"""The allocation-free entry point exafmm_hip_direct_arrays_static.

It stages targets, sources and results in the library's static device array
instead of per-call device buffers; it runs the same kernel as
exafmm_hip_direct_arrays, so results must be bit-identical to it.  Two calls
in a row, the second smaller than the first (the array keeps the first
call's contents, which must not leak into the second), then a request beyond
the array's capacity, which must fail with EXAFMM_HIP_EINVAL and leave the
handle usable.
"""
import numpy as np
import pytest

import exafmm_hip as X
from helpers import unit_cube

pytestmark = pytest.mark.gpu

CAPACITY = 1 << 19  # EXAFMM_HIP_STATIC_SCRATCH_BODIES


@pytest.mark.parametrize("images", [0, 1])
def test_static_direct_matches_allocating_direct(images):
    f = X.FMM(P=6, images=images, cycle=2.5)
    for n_src, n_tgt, seed in ((60000, 700, 1), (9000, 130, 2)):
        src = unit_cube(n_src, seed)
        tgt = src[np.random.default_rng(seed).choice(n_src, n_tgt, replace=False)]
        print(f"[gpu] static direct images={images} nt={n_tgt} ns={n_src}", flush=True)
        a = f.direct_arrays(tgt, src)
        b = f.direct_arrays_static(tgt, src)
        assert np.array_equal(a, b)
        assert np.all(np.isfinite(b))
    f.close()


def test_static_direct_rejects_oversize_and_stays_usable():
    f = X.FMM(P=6)
    src = unit_cube(CAPACITY - 2 * 64 + 1, 3)
    tgt = src[:64]
    with pytest.raises(X.ExafmmError):
        f.direct_arrays_static(tgt, src)
    src = src[:5000]
    a = f.direct_arrays(tgt, src)
    b = f.direct_arrays_static(tgt, src)
    assert np.array_equal(a, b)
    f.close()
