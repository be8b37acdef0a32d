"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product path never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class Params(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int),
        ("theta", ctypes.c_double),
        ("ncrit", ctypes.c_int),
        ("nspawn", ctypes.c_int),
        ("useRmax", ctypes.c_int),
        ("useRopt", ctypes.c_int),
        ("images", ctypes.c_int),
        ("cycle", ctypes.c_double),
        ("adaptive", ctypes.c_int),
        ("mass", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("mutual", ctypes.c_int),
        ("basis", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(Params)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_bodies.argtypes = [vp, ctypes.c_int, dp]
        L.orc_run.argtypes = [vp, ctypes.c_int]
        for f in ("orc_num_cells", "orc_num_bodies", "orc_nterm"):
            getattr(L, f).argtypes = [vp]
        L.orc_num_p2p.argtypes = [vp]
        L.orc_num_p2p.restype = ctypes.c_long
        L.orc_num_m2l.argtypes = [vp]
        L.orc_num_m2l.restype = ctypes.c_long
        L.orc_get_bounds.argtypes = [vp, dp, dp]
        L.orc_get_cells.argtypes = [vp] + [ctypes.c_void_p] * 7
        L.orc_get_bodies.argtypes = [vp] + [ctypes.c_void_p] * 4
        L.orc_get_lists.argtypes = [vp] + [ctypes.c_void_p] * 4
        L.orc_direct.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int]
        L.orc_reference_dataset.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_drand48_seq.argtypes = [ctypes.c_long, ctypes.c_int, dp]
        L.orc_get_dipole.argtypes = [vp, dp, dp]
        L.orc_dipole_correction.argtypes = [vp, dp, ctypes.c_int, ctypes.c_double]
        L.orc_ewald.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _vp(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One FMM evaluation on the CPU with the reference's serial semantics."""

    STAGES_ALL = 31

    def __init__(self, P=10, theta=0.4, ncrit=64, nspawn=1000, useRmax=1, useRopt=0, images=0,
                 cycle=2 * np.pi, adaptive=False, mass=False, threads=0, mutual=False, basis="spherical"):
        b = {"spherical": 0, "cartesian": 1}[basis] if isinstance(basis, str) else int(basis)
        self.params = Params(P, theta, ncrit, nspawn, useRmax, useRopt, images, cycle, int(adaptive), int(mass),
                             threads, int(mutual), b)
        self.h = lib().orc_create(ctypes.byref(self.params))
        self.P = P

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_bodies(self, xyzq):
        xyzq = np.ascontiguousarray(xyzq, dtype=np.float64)
        assert xyzq.ndim == 2 and xyzq.shape[1] == 4
        lib().orc_set_bodies(self.h, xyzq.shape[0], _dp(xyzq))

    def run(self, stages=STAGES_ALL):
        lib().orc_run(self.h, stages)

    def bounds(self):
        a = np.zeros(3)
        b = np.zeros(3)
        lib().orc_get_bounds(self.h, _dp(a), _dp(b))
        return a, b

    def cells(self):
        nc = lib().orc_num_cells(self.h)
        nt = lib().orc_nterm(self.h)
        ints = np.zeros((nc, 5), np.int32)
        icell = np.zeros(nc, np.uint64)
        X = np.zeros((nc, 3))
        R = np.zeros(nc)
        Rb = np.zeros(nc)
        M = np.zeros((nc, nt), np.complex128)
        L = np.zeros((nc, nt), np.complex128)
        lib().orc_get_cells(self.h, _vp(ints), _vp(icell), _vp(X), _vp(R), _vp(Rb), _vp(M), _vp(L))
        return dict(iparent=ints[:, 0].copy(), ichild=ints[:, 1].copy(), nchild=ints[:, 2].copy(),
                    ibody=ints[:, 3].copy(), nbody=ints[:, 4].copy(), icell=icell, X=X, R=R, Rbuild=Rb, M=M, L=L)

    def bodies(self):
        n = lib().orc_num_bodies(self.h)
        xyzq = np.zeros((n, 4))
        ib = np.zeros(n, np.int32)
        key = np.zeros(n, np.uint64)
        trg = np.zeros((n, 4))
        lib().orc_get_bodies(self.h, _vp(xyzq), _vp(ib), _vp(key), _vp(trg))
        return dict(xyzq=xyzq, ibody=ib, key=key, trg=trg)

    def dipole(self, X0=(0.0, 0.0, 0.0)):
        x0 = np.ascontiguousarray(X0, np.float64)
        out = np.zeros(3)
        lib().orc_get_dipole(self.h, _dp(x0), _dp(out))
        return out

    def dipole_correction(self, dipole, num_bodies, cycle):
        d = np.ascontiguousarray(dipole, np.float64)
        lib().orc_dipole_correction(self.h, _dp(d), int(num_bodies), float(cycle))

    def ewald(self, ksize, alpha, sigma, cutoff, cycle):
        """Ewald potential and gradient of every body, in caller order."""
        n = lib().orc_num_bodies(self.h)
        out = np.zeros((n, 4))
        rc = lib().orc_ewald(self.h, ksize, alpha, sigma, cutoff, cycle, _dp(out))
        if rc != 0:
            raise ValueError("orc_ewald: invalid arguments or no tree")
        res = np.empty_like(out)
        res[self.bodies()["ibody"]] = out
        return res

    def lists(self):
        n1 = lib().orc_num_p2p(self.h)
        n2 = lib().orc_num_m2l(self.h)
        a = np.zeros(n1, np.int32)
        b = np.zeros(n1, np.int32)
        c = np.zeros(n2, np.int32)
        d = np.zeros(n2, np.int32)
        lib().orc_get_lists(self.h, _vp(a), _vp(b), _vp(c), _vp(d))
        return dict(p2p_i=a, p2p_j=b, m2l_i=c, m2l_j=d)


def direct(targets, sources, images=0, cycle=2 * np.pi, threads=0):
    """Direct sum onto targets (n x 4 xyzq); returns n x 4 (pot, grad) normalised."""
    t = np.ascontiguousarray(targets, np.float64)
    s = np.ascontiguousarray(sources, np.float64)
    out = np.zeros((t.shape[0], 4))
    lib().orc_direct(t.shape[0], _dp(t), s.shape[0], _dp(s), images, cycle, _dp(out), threads)
    return out


def reference_dataset(dist, n, seed=0, mass=False):
    """dataset.h generators (drand48-exact): dist in c, s, o, p, l."""
    out = np.zeros((n, 4))
    cnt = lib().orc_reference_dataset(dist.encode()[:1], n, seed, int(mass), _dp(out))
    return out[:cnt]


def drand48_seq(seed, n):
    out = np.zeros(n)
    lib().orc_drand48_seq(seed, n, _dp(out))
    return out


def read_ref_dump(path):
    """Read a dump written by oracle/ref_harness.cxx (format version 2)."""
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:4] == b"EXFX", "not a reference dump"
    off = 4
    ver, n, nc, P, nt = np.frombuffer(buf, np.int32, 5, off)
    off += 20
    assert ver == 2
    hdr = np.frombuffer(buf, np.float64, 14, off)
    off += 14 * 8
    out = dict(n=int(n), ncells=int(nc), P=int(P), nterm=int(nt), num_p2p=hdr[0], num_m2l=hdr[1],
               err_pot=hdr[2], err_acc=hdr[3], t_build=hdr[4], t_up=hdr[5], t_trav=hdr[6], t_down=hdr[7],
               xmin=hdr[8:11].copy(), xmax=hdr[11:14].copy())
    out["input"] = np.frombuffer(buf, np.float64, 4 * n, off).reshape(n, 4).copy()
    off += 32 * n
    bdt = np.dtype([("X", "<f8", 3), ("SRC", "<f8"), ("IBODY", "<i4"), ("ICELL", "<u8"), ("TRG", "<f8", 4)])
    bodies = np.frombuffer(buf, bdt, n, off)
    off += bdt.itemsize * n
    out["xyzq"] = np.concatenate([bodies["X"], bodies["SRC"][:, None]], axis=1)
    out["ibody"] = bodies["IBODY"].copy()
    out["key"] = bodies["ICELL"].copy()
    out["trg"] = bodies["TRG"].copy()
    cdt = np.dtype([("IPARENT", "<i4"), ("ICHILD", "<i4"), ("NCHILD", "<i4"), ("IBODY", "<i4"), ("NBODY", "<i4"),
                    ("ICELL", "<u8"), ("X", "<f8", 3), ("Rbuild", "<f8"), ("R", "<f8"),
                    ("M", "<f8", 2 * nt), ("Ltrav", "<f8", 2 * nt), ("L", "<f8", 2 * nt)])
    cells = np.frombuffer(buf, cdt, nc, off)
    off += cdt.itemsize * nc
    out["cells"] = dict(iparent=cells["IPARENT"].copy(), ichild=cells["ICHILD"].copy(),
                        nchild=cells["NCHILD"].copy(), ibody=cells["IBODY"].copy(), nbody=cells["NBODY"].copy(),
                        icell=cells["ICELL"].copy(), X=cells["X"].copy(), Rbuild=cells["Rbuild"].copy(),
                        R=cells["R"].copy(),
                        M=cells["M"].copy().view(np.complex128), Ltrav=cells["Ltrav"].copy().view(np.complex128),
                        L=cells["L"].copy().view(np.complex128))
    ns = int(np.frombuffer(buf, np.int32, 1, off)[0])
    off += 4
    sdt = np.dtype([("IBODY", "<i4"), ("TRG", "<f8", 4)])
    samp = np.frombuffer(buf, sdt, ns, off)
    out["sample_ibody"] = samp["IBODY"].copy()
    out["sample_direct"] = samp["TRG"].copy()
    return out


def read_kernel_chain_dump(path):
    """Read a dump of oracle/kernel_chain.cxx (the per-cell kernel API chain):
    M of the 18 cells, L of the two parents after the mutual M2L and of the 8
    target children after L2L, and the bodies' TRG after L2P, after the mutual
    P2P (targets, then sources), after the self P2P, and the normalised P2P."""
    with open(path, "rb") as f:
        b = f.read()
    nt, nb = (int(v) for v in np.frombuffer(b, np.int32, 2, 0))
    off = 8
    out = {}

    def take(shape):
        nonlocal off
        k = int(np.prod(shape))
        v = np.frombuffer(b, np.float64, k, off).reshape(shape).copy()
        off += 8 * k
        return v

    out["M"] = take((18, nt, 2))
    out["L_parents"] = take((2, nt, 2))
    out["L_children"] = take((8, nt, 2))
    n = 8 * nb
    for k in ("trg_l2p", "trg_p2p_targets", "trg_p2p_sources", "trg_p2p_self", "trg_p2p_normalised"):
        out[k] = take((n, 4))
    assert off == len(b), "unexpected dump length"
    return out
