"""Shared helpers of the test suite (inputs, fixture loading, comparisons)."""
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_case(name):
    """A golden case with its input bodies restored."""
    z = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    if "input" not in z:
        z["input"] = O.reference_dataset(str(z["dist"]), int(z["n"]), 0, bool(z["mass"]))
    return z


def case_params(z):
    return dict(P=int(z["P"]), theta=float(z["theta"]), ncrit=int(z["ncrit"]), nspawn=int(z["nspawn"]),
                useRmax=int(z["useRmax"]), useRopt=int(z["useRopt"]), images=int(z["images"]),
                adaptive=bool(z["adaptive"]), mass=bool(z["mass"]), mutual=bool(z["mutual"]),
                basis="cartesian" if int(z.get("basis", 0)) == 1 else "spherical")


def unit_cube(n, seed):
    """Bodies U[-1,1]^3 with uniform charges in [-0.5, 0.5], mean-shifted and divided by n."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-0.5, 0.5, n)
    q = (q - q.mean()) / n
    return np.concatenate([x, q[:, None]], axis=1)


def plummer(n, seed, a=1.0, rmax=10.0):
    """Plummer sphere (scale radius a, radial cut-off rmax), equal charges 1/n."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, 4))
    k = 0
    while k < n:
        m = (n - k) * 2
        u = rng.uniform(0, 1, m)
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[(r < rmax) & np.isfinite(r)][: n - k]
        cz = rng.uniform(-1, 1, r.shape[0])
        ph = rng.uniform(0, 2 * np.pi, r.shape[0])
        sz = np.sqrt(1 - cz * cz)
        out[k:k + r.shape[0], 0] = r * sz * np.cos(ph)
        out[k:k + r.shape[0], 1] = r * sz * np.sin(ph)
        out[k:k + r.shape[0], 2] = r * cz
        k += r.shape[0]
    out[:, 3] = 1.0 / n
    return out


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def run_oracle(xyzq, **kw):
    o = O.Oracle(**kw)
    o.set_bodies(xyzq)
    o.run()
    return o
