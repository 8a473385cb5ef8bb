"""ctypes wrapper of liboracle.so -- the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only as the checker or the timed CPU baseline; the product (libdem.so) never calls it.
Build: `make -C oracle` (done by __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

c_uint = ctypes.c_uint
c_int = ctypes.c_int
c_double = ctypes.c_double
P = ctypes.c_void_p

ORC_EMPTY = 0xFFFFFFFF


class OrcMat(ctypes.Structure):
    _fields_ = [(n, c_double) for n in (
        "estarp", "ufricp", "vpois", "dampnp", "damptp", "upp",
        "estarw", "ufricw", "vpoisw", "dampnw", "damptw", "upw")]


class OrcWalls(ctypes.Structure):
    _fields_ = [("kind", c_int), ("pad", c_int), ("size", c_double * 3), ("radius", c_double)]


class OrcSetup(ctypes.Structure):
    _fields_ = [
        ("g", c_uint * 3), ("ncells", c_uint),
        ("cell", c_double * 3), ("origin", c_double * 3),
        ("skin", c_double), ("ann2", c_double), ("dt", c_double),
        ("width", c_int), ("pad", c_int),
        ("mat", OrcMat),
        ("walls", OrcWalls),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing; run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_reduced_material": ([c_double] * 12 + [P], None),
        "orc_calc_hash": ([c_int, P, P, P, P, P], None),
        "orc_radix_sort": ([P, P, c_int, c_int], None),
        "orc_find_cell_start": ([P, P, c_int, c_uint], None),
        "orc_neighbor_array_h": ([c_int, P, P, c_double, P, P, P, P, c_int, P, P, P], c_int),
        "orc_prefix_sum": ([c_int, P, P, P], c_uint),
        "orc_pair_list": ([c_int, c_int, c_int, P, P, P, P, P, P, P, P, P, P, P, P, P], None),
        "orc_law_pair": ([P, P, P, P, P, P, c_double, c_double, c_double, c_double, P, c_double,
                          c_double, P, P, P, P, P, P], c_int),
        "orc_law_wall": ([P, P, P, P, c_double, c_double, P, c_double, c_double, P, P, P, P, P],
                         c_int),
        "orc_force_pairs": ([c_uint, P, P, P, P, P, P, P, P, P, P, P, P, c_double, P, P, P, P, P],
                            None),
        "orc_sum_forces": ([c_int, c_int, P, P, P, P, P, P, P, P, P, P], None),
        "orc_wall_gravity": ([c_int, P, P, P, P, P, P, c_double, c_uint, P, P, P, P, P], None),
        "orc_update": ([c_int, P, P, P, P, P, P, P, P, c_double, c_double, c_double, P, P],
                       None),
        "orc_set_float_length": ([c_int], None),
        "orc_contact_total": ([c_int, P], c_uint),
        "orc_rebuild": ([P, P], c_int),
        "orc_step": ([P, P], c_int),
        "orc_run": ([P, P, c_int], c_int),
        "orc_state_create": ([c_int, c_int, c_uint, c_uint], P),
        "orc_state_destroy": ([P], None),
        "orc_state_field": ([P, c_int], P),
        "orc_state_scalars": ([P, P], None),
        "orc_state_set_iall": ([P, c_uint], None),
        "orc_state_set_n": ([P, c_int], None),
        "orc_brute_pairs": ([c_int, P, P, c_double, P, ctypes.c_longlong], ctypes.c_longlong),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ptr(a: np.ndarray):
    return a.ctypes.data_as(P)


def mat_from(material) -> OrcMat:
    """OrcMat from a dem_amd.Material (same 12 doubles, same order)."""
    m = OrcMat()
    for k in ("estarp", "ufricp", "vpois", "dampnp", "damptp", "upp"):
        setattr(m, k, getattr(material.p, k))
    for k in ("estarw", "ufricw", "vpoisw", "dampnw", "damptw", "upw"):
        setattr(m, k, getattr(material.w, k))
    return m


def walls_from(walls) -> OrcWalls:
    w = OrcWalls()
    w.kind = walls.kind
    for k in range(3):
        w.size[k] = walls.size[k]
    w.radius = walls.radius
    return w


_FIELDS = {
    # name: (index, dtype, per-particle/per-pair, columns)
    "pos": (0, np.float64, "n", 3), "pdis": (1, np.float64, "n", 3), "angv": (2, np.float64, "n", 3),
    "rad": (3, np.float64, "n", 1), "rmass": (4, np.float64, "n", 1), "inert": (5, np.float64, "n", 1),
    "qi": (6, np.float64, "n", 3), "disptw": (7, np.float64, "n", 3), "fricpw": (8, np.float64, "n", 1),
    "forcei": (9, np.float64, "n", 3), "torquei": (10, np.float64, "n", 3),
    "contacti": (11, np.uint32, "n", 1), "hash": (12, np.uint32, "n", 2),
    "sorted": (13, np.uint32, "n", 2), "cell_start": (14, np.uint32, "c", 1),
    "anei": (15, np.uint32, "nw", 1), "aref": (16, np.uint32, "nw", 1),
    "njgi": (17, np.uint32, "n", 1), "njli": (18, np.uint32, "n", 1), "sjgi": (19, np.uint32, "n", 1),
    "lp": (20, np.uint32, "p", 1), "ln": (21, np.uint32, "p", 1), "dispt": (22, np.float64, "p", 3),
    "fricp": (23, np.float64, "p", 1), "forcepair": (24, np.float64, "p", 3),
    "torqpairi": (25, np.float64, "p", 3), "torqpairj": (26, np.float64, "p", 3),
    "contactpair": (27, np.uint32, "p", 1),
}


class OracleSim:
    """Reference-layout CPU simulation of one instance (dem_amd.Instance + state arrays)."""

    def __init__(self, inst, pos, pdis, angv, rad, rmass, inert, width: int = 64,
                 pair_cap: int | None = None, float_length: bool = False):
        L = lib()
        self.L = L
        # the reference's float length() in vprojec and the rebuild trigger (a process-wide
        # switch of the C oracle, set before each rebuild and step of this simulation)
        self.float_length = bool(float_length)
        n = int(rad.shape[0])
        self.n = n
        self.width = width
        self.pair_cap = pair_cap if pair_cap is not None else max(16 * n, 1024)
        s = OrcSetup()
        for k in range(3):
            s.g[k] = inst.gsize[k]
            s.cell[k] = inst.cell[k]
            s.origin[k] = inst.origin[k]
        s.ncells = inst.gsize[0] * inst.gsize[1] * inst.gsize[2]
        s.skin = inst.skin
        ann = 0.5 * inst.skin
        s.ann2 = ann * ann
        s.dt = inst.dt
        s.width = width
        s.mat = mat_from(inst.mat)
        s.walls = walls_from(inst.walls)
        self.setup = s
        self.ncells = s.ncells
        self.st = L.orc_state_create(n, width, self.pair_cap, s.ncells)
        for name, a in (("pos", pos), ("pdis", pdis), ("angv", angv), ("rad", rad),
                        ("rmass", rmass), ("inert", inert)):
            self.field(name)[...] = np.asarray(a, dtype=np.float64).reshape(self.field(name).shape)

    def __del__(self):
        try:
            if self.st:
                self.L.orc_state_destroy(self.st)
                self.st = None
        except Exception:
            pass

    def field(self, name: str) -> np.ndarray:
        idx, dt, kind, cols = _FIELDS[name]
        rows = {"n": self.n, "p": self.pair_cap, "c": self.ncells, "nw": self.n * self.width}[kind]
        p = self.L.orc_state_field(self.st, idx)
        ct = ctypes.c_double if dt == np.float64 else ctypes.c_uint
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(rows * cols,))
        return arr.reshape((rows, cols)) if cols > 1 else arr

    def scalars(self) -> dict:
        out = (c_double * 7)()
        self.L.orc_state_scalars(self.st, out)
        keys = ("n", "nold", "iall", "npairs", "overflow", "pcntot", "maxhz")
        return {k: (out[i] if k == "maxhz" else int(out[i])) for i, k in enumerate(keys)}

    def set_iall(self, v: int):
        self.L.orc_state_set_iall(self.st, v)

    def set_n(self, n: int):
        """Activate particles [0, n) (layer insertion); forces a rebuild at the next step."""
        self.L.orc_state_set_n(self.st, int(n))

    def rebuild(self) -> int:
        self.L.orc_set_float_length(1 if self.float_length else 0)
        return self.L.orc_rebuild(ctypes.byref(self.setup), self.st)

    def step(self, k: int = 1) -> int:
        self.L.orc_set_float_length(1 if self.float_length else 0)
        rc = self.L.orc_run(ctypes.byref(self.setup), self.st, int(k))
        self.L.orc_set_float_length(0)
        return rc

    def pairs(self):
        """Current pair list (i, j), i < j, list order."""
        np_ = self.scalars()["npairs"]
        return np.stack([self.field("lp")[:np_], self.field("ln")[:np_]], axis=1).copy()


def brute_pairs(pos, rad, skin):
    L = lib()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rad = np.ascontiguousarray(rad, dtype=np.float64)
    n = rad.shape[0]
    cap = max(64 * n, 1024)
    out = np.empty((cap, 2), dtype=np.uint32)
    c = L.orc_brute_pairs(n, ptr(pos), ptr(rad), skin, ptr(out), cap)
    if c < 0:
        raise RuntimeError("brute-force pair capacity exceeded")
    return out[:c].copy()


def law_pair(xi, xj, vi, vj, wi, wj, ri, rj, mi, mj, hin, dt, mat, float_length=False):
    """Oracle pair law for one pair; returns (contact, F_i, T_i, T_j, h_out, |Ft|).
    float_length: the reference's float length() in vprojec."""
    L = lib()
    L.orc_set_float_length(1 if float_length else 0)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (xi, xj, vi, vj, wi, wj, hin)]
    hout = np.zeros(3)
    fout = c_double()
    fi = np.zeros(3)
    ti = np.zeros(3)
    tj = np.zeros(3)
    m = mat_from(mat) if not isinstance(mat, OrcMat) else mat
    c = L.orc_law_pair(ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]), ptr(a[5]), ri, rj, mi,
                       mj, ptr(a[6]), 0.0, dt, ctypes.byref(m), ptr(hout), ctypes.byref(fout),
                       ptr(fi), ptr(ti), ptr(tj))
    L.orc_set_float_length(0)
    return c, fi, ti, tj, hout, fout.value


def law_wall(walls, x, v, om, r, m, hin, dt, mat):
    """Oracle wall law for one sphere against every wall of `walls` (a dem_amd.Walls or an
    OrcWalls); returns (walls hit, F, T, h_out, |Ft|) -- the force and torque summed over
    the walls the sphere overlaps, the history of the deepest one."""
    L = lib()
    a = [np.ascontiguousarray(q, dtype=np.float64) for q in (x, v, om, hin)]
    w = walls if isinstance(walls, OrcWalls) else walls_from(walls)
    mm = mat_from(mat) if not isinstance(mat, OrcMat) else mat
    hout = np.zeros(3)
    fout = c_double()
    fw = np.zeros(3)
    tw = np.zeros(3)
    hits = L.orc_law_wall(ctypes.byref(w), ptr(a[0]), ptr(a[1]), ptr(a[2]), r, m, ptr(a[3]), 0.0,
                          dt, ctypes.byref(mm), ptr(hout), ctypes.byref(fout), ptr(fw), ptr(tw))
    return hits, fw, tw, hout, fout.value
