"""Python binding of libexafmm_hip (the C-ABI of include/exafmm_hip.h).

`FMM` mirrors the pass-level API of the reference driver examples/serial.cxx:
get_bounds (BoundBox::getBounds), build_tree (BuildTree::buildTree),
upward_pass / downward_pass (UpDownPass), traverse
(Traversal::dualTreeTraversal), direct (Traversal::direct + normalize), and
evaluate for the whole sequence.  Every call goes to the HIP library; there is
no CPU fallback: a missing library or a failing call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexafmm_hip.so")

# Every symbol declared in include/exafmm_hip.h.
EXPORTS = (
    "exafmm_hip_create", "exafmm_hip_destroy", "exafmm_hip_set_config", "exafmm_hip_last_error",
    "exafmm_hip_set_bodies", "exafmm_hip_direct_arrays",
    "exafmm_hip_set_bodies_device", "exafmm_hip_get_bounds", "exafmm_hip_build_tree", "exafmm_hip_upward_pass",
    "exafmm_hip_traverse", "exafmm_hip_downward_pass", "exafmm_hip_evaluate", "exafmm_hip_get_cells",
    "exafmm_hip_get_bodies", "exafmm_hip_get_results", "exafmm_hip_results_device", "exafmm_hip_get_list_counts",
    "exafmm_hip_get_lists", "exafmm_hip_direct", "exafmm_hip_get_stats", "exafmm_hip_synchronize",
    "exafmm_hip_get_dipole", "exafmm_hip_dipole_correction", "exafmm_hip_ewald", "exafmm_hip_count_p2p_bodies",
    "exafmm_hip_evaluate_batch", "exafmm_hip_direct_arrays_static",
    "exafmm_hip_cell_p2p", "exafmm_hip_cell_p2p_self", "exafmm_hip_cell_p2m", "exafmm_hip_cell_m2m",
    "exafmm_hip_cell_m2l", "exafmm_hip_cell_l2l", "exafmm_hip_cell_l2p",
    "exafmm_hip_set_autotune", "exafmm_hip_get_tuning", "exafmm_hip_set_tuning",
    "exafmm_hip_evaluate_charges", "exafmm_hip_evaluate_charges_device",
)


class Config(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int),
        ("theta", ctypes.c_double),
        ("ncrit", ctypes.c_int),
        ("nspawn", ctypes.c_int),
        ("useRmax", ctypes.c_int),
        ("useRopt", ctypes.c_int),
        ("mutual", ctypes.c_int),
        ("images", ctypes.c_int),
        ("cycle", ctypes.c_double),
        ("builder", ctypes.c_int),
        ("mass", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("device", ctypes.c_int),
        ("basis", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("t_bounds", "t_build", "t_upward", "t_lists", "t_p2p", "t_m2l", "t_downward", "t_total",
                 "t_traverse")] + [
        ("num_p2p_pairs", ctypes.c_int64), ("num_m2l_pairs", ctypes.c_int64), ("num_p2p_bodies", ctypes.c_int64),
        ("num_cells", ctypes.c_int), ("num_levels", ctypes.c_int), ("num_m2l_lattice", ctypes.c_int64),
        ("num_m2l_offsets", ctypes.c_int), ("num_m2l_tiles", ctypes.c_int), ("t_m2l_mfma", ctypes.c_double),
        ("t_m2l_pairs", ctypes.c_double), ("num_m2l_steps", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ExafmmError(RuntimeError):
    pass


_lib = None


def load():
    """Load libexafmm_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ExafmmError(f"{LIB_PATH} is missing: build it with `make -C exafmm-dev_amd`")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    L.exafmm_hip_create.argtypes = [P(Config), P(vp)]
    L.exafmm_hip_destroy.argtypes = [vp]
    L.exafmm_hip_last_error.argtypes = [vp]
    L.exafmm_hip_last_error.restype = ctypes.c_char_p
    L.exafmm_hip_set_bodies.argtypes = [vp, ip, vp]
    L.exafmm_hip_set_bodies_device.argtypes = [vp, ip, vp]
    L.exafmm_hip_get_bounds.argtypes = [vp, vp, vp]
    L.exafmm_hip_build_tree.argtypes = [vp, vp, vp, P(ip)]
    for f in ("upward_pass", "traverse", "downward_pass", "evaluate", "synchronize"):
        getattr(L, "exafmm_hip_" + f).argtypes = [vp]
    L.exafmm_hip_get_cells.argtypes = [vp] + [vp] * 6
    L.exafmm_hip_get_bodies.argtypes = [vp, vp, vp, vp]
    L.exafmm_hip_get_results.argtypes = [vp, vp]
    L.exafmm_hip_results_device.argtypes = [vp, P(vp)]
    L.exafmm_hip_get_list_counts.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int64)]
    L.exafmm_hip_get_lists.argtypes = [vp, vp, vp, vp, vp]
    L.exafmm_hip_direct.argtypes = [vp, ip, vp, vp]
    L.exafmm_hip_direct_arrays.argtypes = [vp, ip, vp, ip, vp, vp]
    L.exafmm_hip_direct_arrays_static.argtypes = [vp, ip, vp, ip, vp, vp]
    L.exafmm_hip_cell_p2p.argtypes = [vp, ip, vp, vp, ip, vp, vp, dp, vp, ip]
    L.exafmm_hip_cell_p2p_self.argtypes = [vp, ip, vp, vp, dp]
    L.exafmm_hip_cell_p2m.argtypes = [vp, ip, vp, vp, vp]
    L.exafmm_hip_cell_m2m.argtypes = [vp, ip, vp, vp, vp, vp]
    L.exafmm_hip_cell_m2l.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ip]
    L.exafmm_hip_cell_l2l.argtypes = [vp, vp, vp, vp, vp, vp]
    L.exafmm_hip_cell_l2p.argtypes = [vp, ip, vp, vp, vp, vp]
    L.exafmm_hip_set_autotune.argtypes = [vp, ip]
    L.exafmm_hip_get_tuning.argtypes = [vp, P(ip), P(ip), P(ip), P(ip)]
    L.exafmm_hip_set_tuning.argtypes = [vp, ip, ip, ip, ip]
    L.exafmm_hip_set_config.argtypes = [vp, P(Config)]
    L.exafmm_hip_get_stats.argtypes = [vp, P(Stats)]
    L.exafmm_hip_count_p2p_bodies.argtypes = [vp, P(ctypes.c_int64)]
    L.exafmm_hip_evaluate_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    L.exafmm_hip_evaluate_charges.argtypes = [vp, ip, vp]
    L.exafmm_hip_evaluate_charges_device.argtypes = [vp, ip, vp]
    L.exafmm_hip_get_dipole.argtypes = [vp, vp, vp]
    L.exafmm_hip_dipole_correction.argtypes = [vp, vp, ip, dp]
    L.exafmm_hip_ewald.argtypes = [vp, ip, dp, dp, dp, dp, ip, vp]
    for f in EXPORTS:
        if f != "exafmm_hip_last_error":
            getattr(L, f).restype = ctypes.c_int
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class FMM:
    """One particle system on one GPU (one handle, one HIP stream)."""

    def __init__(self, P=10, theta=0.4, ncrit=64, nspawn=1000, useRmax=1, useRopt=0, mutual=0, images=0,
                 cycle=2 * np.pi, builder="uniform", mass=False, precision="fp64", device=0, basis="spherical"):
        self.lib = load()
        b = {"uniform": 0, "adaptive": 1}[builder] if isinstance(builder, str) else int(builder)
        prec = {"fp64": 0, "mixed": 1}[precision] if isinstance(precision, str) else int(precision)
        bas = {"spherical": 0, "cartesian": 1}[basis] if isinstance(basis, str) else int(basis)
        self.cfg = Config(P, theta, ncrit, nspawn, useRmax, useRopt, mutual, images, cycle, b, int(mass), prec, device,
                          bas)
        h = ctypes.c_void_p()
        rc = self.lib.exafmm_hip_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if rc != 0:
            raise ExafmmError(f"exafmm_hip_create failed with status {rc}")
        self.h = h
        self.P = P
        self.nterm = P * (P + 1) * (P + 2) // 6 if bas == 1 else P * (P + 1) // 2
        self.n = 0

    def set_config(self, **changes):
        """Change run-time parameters (theta, ncrit, precision, images, ...) as
        exafmm_hip_set_config does; P, basis and device stay fixed."""
        names = {"builder": {"uniform": 0, "adaptive": 1}, "precision": {"fp64": 0, "mixed": 1},
                 "basis": {"spherical": 0, "cartesian": 1}}
        for k, v in changes.items():
            if k not in dict(Config._fields_):
                raise KeyError(k)
            if isinstance(v, str):
                v = names[k][v]
            setattr(self.cfg, k, v)
        self._check(self.lib.exafmm_hip_set_config(self.h, ctypes.byref(self.cfg)), "set_config")

    def close(self):
        if getattr(self, "h", None):
            self.lib.exafmm_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.exafmm_hip_last_error(self.h)
            raise ExafmmError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    # -- bodies ------------------------------------------------------------
    def set_bodies(self, xyzq):
        a = np.ascontiguousarray(xyzq, dtype=np.float64)
        if a.ndim != 2 or a.shape[1] != 4:
            raise ValueError("bodies must be an (n, 4) array of x, y, z, q")
        self.n = a.shape[0]
        self._check(self.lib.exafmm_hip_set_bodies(self.h, self.n, _ptr(a)), "set_bodies")

    def set_bodies_device(self, n, dev_ptr):
        self.n = int(n)
        self._check(self.lib.exafmm_hip_set_bodies_device(self.h, self.n, ctypes.c_void_p(dev_ptr)),
                    "set_bodies_device")

    # -- passes ------------------------------------------------------------
    def get_bounds(self):
        a, b = np.zeros(3), np.zeros(3)
        self._check(self.lib.exafmm_hip_get_bounds(self.h, _ptr(a), _ptr(b)), "get_bounds")
        return a, b

    def build_tree(self, bounds=None):
        nc = ctypes.c_int()
        if bounds is None:
            rc = self.lib.exafmm_hip_build_tree(self.h, None, None, ctypes.byref(nc))
        else:
            a = np.ascontiguousarray(bounds[0], np.float64)
            b = np.ascontiguousarray(bounds[1], np.float64)
            rc = self.lib.exafmm_hip_build_tree(self.h, _ptr(a), _ptr(b), ctypes.byref(nc))
        self._check(rc, "build_tree")
        return nc.value

    def upward_pass(self):
        self._check(self.lib.exafmm_hip_upward_pass(self.h), "upward_pass")

    def traverse(self):
        self._check(self.lib.exafmm_hip_traverse(self.h), "traverse")

    def downward_pass(self):
        self._check(self.lib.exafmm_hip_downward_pass(self.h), "downward_pass")

    def evaluate(self):
        self._check(self.lib.exafmm_hip_evaluate(self.h), "evaluate")

    def synchronize(self):
        self._check(self.lib.exafmm_hip_synchronize(self.h), "synchronize")

    # -- readback ----------------------------------------------------------
    def cells(self):
        s = self.stats()
        nc = s["num_cells"]
        ints = np.zeros((nc, 5), np.int32)
        icell = np.zeros(nc, np.uint64)
        X = np.zeros((nc, 3))
        R = np.zeros(nc)
        M = np.zeros((nc, self.nterm), np.complex128)
        L = np.zeros((nc, self.nterm), np.complex128)
        self._check(self.lib.exafmm_hip_get_cells(self.h, _ptr(ints), _ptr(icell), _ptr(X), _ptr(R), _ptr(M),
                                                  _ptr(L)), "get_cells")
        return dict(iparent=ints[:, 0].copy(), ichild=ints[:, 1].copy(), nchild=ints[:, 2].copy(),
                    ibody=ints[:, 3].copy(), nbody=ints[:, 4].copy(), icell=icell, X=X, R=R, M=M, L=L)

    def bodies(self):
        xyzq = np.zeros((self.n, 4))
        ib = np.zeros(self.n, np.int32)
        trg = np.zeros((self.n, 4))
        self._check(self.lib.exafmm_hip_get_bodies(self.h, _ptr(xyzq), _ptr(ib), _ptr(trg)), "get_bodies")
        return dict(xyzq=xyzq, ibody=ib, trg=trg)

    def results(self):
        """(n, 4) potential and gradient in the caller's body order."""
        out = np.zeros((self.n, 4))
        self._check(self.lib.exafmm_hip_get_results(self.h, _ptr(out)), "get_results")
        return out

    def lists(self):
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        self._check(self.lib.exafmm_hip_get_list_counts(self.h, ctypes.byref(a), ctypes.byref(b)), "list_counts")
        pi = np.zeros(a.value, np.int32)
        pj = np.zeros(a.value, np.int32)
        mi = np.zeros(b.value, np.int32)
        mj = np.zeros(b.value, np.int32)
        self._check(self.lib.exafmm_hip_get_lists(self.h, _ptr(pi), _ptr(pj), _ptr(mi), _ptr(mj)), "get_lists")
        return dict(p2p_i=pi, p2p_j=pj, m2l_i=mi, m2l_j=mj)

    def direct(self, target_ibody):
        t = np.ascontiguousarray(target_ibody, np.int32)
        out = np.zeros((t.shape[0], 4))
        self._check(self.lib.exafmm_hip_direct(self.h, t.shape[0], _ptr(t), _ptr(out)), "direct")
        return out

    def direct_arrays(self, targets, sources):
        t = np.ascontiguousarray(targets, np.float64)
        s = np.ascontiguousarray(sources, np.float64)
        out = np.zeros((t.shape[0], 4))
        self._check(self.lib.exafmm_hip_direct_arrays(self.h, t.shape[0], _ptr(t), s.shape[0], _ptr(s), _ptr(out)),
                    "direct_arrays")
        return out

    def direct_arrays_static(self, targets, sources):
        """direct_arrays through the allocation-free entry point (staging in
        the library's static device array; 2 nt + ns bodies at most)."""
        t = np.ascontiguousarray(targets, np.float64)
        s = np.ascontiguousarray(sources, np.float64)
        out = np.zeros((t.shape[0], 4))
        self._check(self.lib.exafmm_hip_direct_arrays_static(self.h, t.shape[0], _ptr(t), s.shape[0], _ptr(s),
                                                             _ptr(out)), "direct_arrays_static")
        return out

    def dipole(self, X0=(0.0, 0.0, 0.0)):
        """UpDownPass::getDipole: sum of (X - X0) q over the bodies."""
        x0 = np.ascontiguousarray(X0, np.float64)
        out = np.zeros(3)
        self._check(self.lib.exafmm_hip_get_dipole(self.h, _ptr(x0), _ptr(out)), "get_dipole")
        return out

    def dipole_correction(self, dipole, num_bodies, cycle):
        """UpDownPass::dipoleCorrection on the device-resident results."""
        d = np.ascontiguousarray(dipole, np.float64)
        self._check(self.lib.exafmm_hip_dipole_correction(self.h, _ptr(d), int(num_bodies), float(cycle)),
                    "dipole_correction")

    def ewald(self, ksize, alpha, sigma, cutoff, cycle, parts=7, trg=None):
        """Ewald summation (wave 1, real 2, self 4) added to trg (n x 4,
        caller order; zeros when None), which is returned."""
        n = self.n
        out = np.zeros((n, 4)) if trg is None else np.ascontiguousarray(trg, np.float64)
        self._check(self.lib.exafmm_hip_ewald(self.h, int(ksize), float(alpha), float(sigma), float(cutoff),
                                            float(cycle), int(parts), _ptr(out)), "ewald")
        return out

    def stats(self):
        s = Stats()
        self._check(self.lib.exafmm_hip_get_stats(self.h, ctypes.byref(s)), "get_stats")
        return s.as_dict()

    def evaluate_batch(self, systems):
        """Evaluate independent systems (a list of n_s x 4 arrays x, y, z, q)
        in one batched pass; returns the list of n_s x 4 results."""
        arrs = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 4) for x in systems]
        n = np.array([a.shape[0] for a in arrs], dtype=np.int32)
        x = np.ascontiguousarray(np.concatenate(arrs, axis=0))
        out = np.empty_like(x)
        self._check(self.lib.exafmm_hip_evaluate_batch(self.h, len(arrs), n.ctypes.data, x.ctypes.data,
                                                       out.ctypes.data), "evaluate_batch")
        self.n = x.shape[0]
        self._batch = n
        return np.split(out, np.cumsum(n)[:-1])

    def evaluate_charges(self, q):
        """Re-evaluate the last evaluated bodies with new charges q (caller
        order; for a batch, the systems' charges as a list or concatenated):
        the numeric passes only, replayed from a HIP graph after the first
        call.  Results as after evaluate (results(), or results_batch())."""
        if isinstance(q, (list, tuple)):
            q = np.concatenate([np.asarray(v, np.float64).reshape(-1) for v in q])
        a = np.ascontiguousarray(q, np.float64).reshape(-1)
        self._check(self.lib.exafmm_hip_evaluate_charges(self.h, a.shape[0], _ptr(a)), "evaluate_charges")

    def results_batch(self):
        """results() split per system of the last evaluate_batch."""
        return np.split(self.results(), np.cumsum(self._batch)[:-1])

    def set_autotune(self, enable=True):
        """Tune ncrit, the P2P grid and targets per lane, and the contraction
        grid per input signature (the first evaluation of a signature times
        the candidates; later ones reuse the winner)."""
        self._check(self.lib.exafmm_hip_set_autotune(self.h, int(bool(enable))), "set_autotune")

    def tuning(self):
        """The parameters in use: ncrit, P2P waves per CU, P2P targets per
        lane, contraction workgroups per CU (0: one per work item)."""
        v = [ctypes.c_int(0) for _ in range(4)]
        self._check(self.lib.exafmm_hip_get_tuning(self.h, *[ctypes.byref(x) for x in v]), "get_tuning")
        return dict(zip(("ncrit", "p2p_waves_per_cu", "p2p_targets_per_lane", "m2l_groups_per_cu"),
                        (x.value for x in v)))

    def set_tuning(self, ncrit, p2p_waves_per_cu, p2p_targets_per_lane, m2l_groups_per_cu):
        """Pin the parameters that tuning() reports (e.g. the ones a tuned
        run chose)."""
        self._check(self.lib.exafmm_hip_set_tuning(self.h, int(ncrit), int(p2p_waves_per_cu),
                                                   int(p2p_targets_per_lane), int(m2l_groups_per_cu)), "set_tuning")

    def count_p2p_bodies(self):
        """Target-source body pairs of the current P2P list."""
        n = ctypes.c_int64(0)
        self._check(self.lib.exafmm_hip_count_p2p_bodies(self.h, ctypes.byref(n)), "count_p2p_bodies")
        return int(n.value)


def rel_l2(a, b):
    """Relative L2 difference sqrt(sum (a-b)^2 / sum b^2) (include/verify.h)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
