"""dem_amd -- Python binding of libdem.so, the gfx950 DEM packing step.

The compute path is the HIP library; this module only marshals arguments through ctypes
(C ABI declared in include/dem.h) and composes the seeded workloads the benchmark and tests
use.  There is no CPU fallback: if libdem.so is missing or no GPU is visible, the calls that
need it raise.

Reduced units throughout (reference dempacking.cu:146-153): length = particle diameter,
mass = rho*pi*d^3/6, force = m*g, time = sqrt(d/g).  `pdis` and `angv` are per-step
displacements, as in the reference integrator (neiglist.cu:356-360).
"""
from __future__ import annotations

import ctypes
import functools
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdem.so")

c_uint = ctypes.c_uint
c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_ull = ctypes.c_ulonglong

DEM_OK = 0
DEM_ERR_INVALID = 1
DEM_ERR_HIP = 2
DEM_ERR_CAPACITY = 3
DEM_ERR_NO_DEVICE = 4
DEM_ERR_IO = 5

WALLS_BOX = 0
WALLS_CYLINDER = 1

OPT_RECORD_FORCES = 1
OPT_REBUILD_INTERVAL = 2
OPT_SKIN = 4
OPT_PAIR_BLOCK = 5
OPT_PAIR_GRID = 6
OPT_PARTICLE_BLOCK = 7
OPT_ADJ_ROLL_REG = 8
OPT_FUSED_STEP = 9
OPT_TIME_STEPS = 10
OPT_PIPELINE = 11
OPT_FUSED_FAST = 12
OPT_XCD_CHUNK = 13
OPT_PIPELINE_BATCH = 14
OPT_ROLL_FLOOR = 15
OPT_FLOAT_LENGTH = 16

# neiglist.cu:27-28: cua (unit diameter) and cub (candidate cutoff); the list skin is cub - cua.
CUA = 1.0
CUB = 1.55
SKIN = CUB - CUA


class Matproperties(ctypes.Structure):
    _fields_ = [(n, c_double) for n in ("estarp", "ufricp", "vpois", "dampnp", "damptp", "upp")]


class Wallproperties(ctypes.Structure):
    _fields_ = [(n, c_double) for n in ("estarw", "ufricw", "vpoisw", "dampnw", "damptw", "upw")]


class Material(ctypes.Structure):
    _fields_ = [("p", Matproperties), ("w", Wallproperties)]


class Walls(ctypes.Structure):
    _fields_ = [("kind", c_int), ("pad", c_int), ("size", c_double * 3), ("radius", c_double)]


class Instance(ctypes.Structure):
    _fields_ = [
        ("n", c_uint),
        ("gsize", c_uint * 3),
        ("cell", c_double * 3),
        ("origin", c_double * 3),
        ("walls", Walls),
        ("mat", Material),
        ("dt", c_double),
        ("skin", c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("steps", c_ull),
        ("rebuilds", c_ull),
        ("n", c_uint),
        ("npairs", c_uint),
        ("pair_capacity", c_uint),
        ("ninstances", c_uint),
        ("last_rebuild_flag", c_uint),
        ("overflow", c_uint),
        ("maxhz", c_double),
        ("pcntot", c_uint),
        ("step_path", c_uint),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("skin", c_double), ("cell", c_double), ("interval", c_int),
        ("pair_block", c_uint), ("pair_grid_per_cu", c_uint), ("particle_block", c_uint),
        ("ms_per_step", c_double), ("score_ms", c_double), ("trials", c_int), ("step_path", c_int),
    ]


class Deck(ctypes.Structure):
    _fields_ = [
        ("iforcemode", c_int),
        ("emod", c_double), ("vpois", c_double), ("denp", c_double),
        ("diam", c_double),
        ("ntot", c_int),
        ("dt", c_double), ("tstop", c_uint),
        ("dmt1", c_double), ("dmt2", c_double), ("thick", c_double),
        ("zht1", c_double), ("zht2", c_double),
        ("vel0", c_double), ("nrings", c_int),
        ("ufricp", c_double), ("ufricw", c_double),
        ("dampnp", c_double), ("damptp", c_double), ("dampnw", c_double), ("damptw", c_double),
        ("cdamp", c_double),
        ("stiffness", c_double), ("dampcoeffn", c_double), ("dampcoefft", c_double),
        ("upp", c_double), ("upw", c_double),
    ]


class Obs(ctypes.Structure):
    """Smooth-observable region (dem_obs): side walls moved in by `margin`, zlo < z < zhi,
    logistic edges of width `eps`; soft contacts ramp over overlaps in [0, eps_c]."""
    _fields_ = [("margin", c_double), ("zlo", c_double), ("zhi", c_double), ("eps", c_double),
                ("eps_c", c_double)]


def make_obs(zhi: float, margin: float = 2.0, zlo: float = 2.0, eps: float = 0.5,
             eps_c: float = 1e-4) -> Obs:
    o = Obs()
    o.margin, o.zlo, o.zhi, o.eps, o.eps_c = margin, zlo, zhi, eps, eps_c
    return o


class DemError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: status {code}")
        self.code = code


_lib = None

# Exported symbols of libdem.so (every function include/dem.h declares).
EXPORTS = [
    "dem_last_error", "dem_status_string", "setMaterialProperties", "setWallGeometry",
    "calcHash", "RadixSort", "RadixSortNoAlloc", "findCellStart", "neighborarray", "prefixsum", "pairlist",
    "calculateforcepair", "sumallforcei", "calculateforceiw", "updateposition",
    "outputporositycn", "dem_create", "dem_destroy", "dem_set_option", "dem_set_state",
    "dem_get_state", "dem_get_wall_history", "dem_save_checkpoint", "dem_load_checkpoint",
    "dem_read_restart", "dem_write_restart_resume", "dem_read_sidecar", "dem_step", "dem_step_graph",
    "dem_request_rebuild", "dem_rebuild", "dem_autotune", "dem_set_active", "dem_active_count", "dem_get_stats", "dem_get_instance_stats",
    "dem_sync", "dem_last_ms", "dem_time_kernels", "dem_get_hash", "dem_get_sorted", "dem_get_cell_start",
    "dem_get_pairs", "dem_get_contacts", "dem_get_forces", "dem_device_info", "dem_read_deck",
    "dem_deck_material", "dem_write_restart", "dem_damping_for_restitution",
    "dem_restitution_for_damping", "dem_gen_diameters", "dem_gen_insert", "dem_gen_lattice",
    "dem_particle_props", "dem_make_grid", "dem_dev_alloc", "dem_dev_free", "dem_memcpy",
    "dem_memzero", "dem_device_sync", "dem_law_pair_eval", "dem_contact_filter", "dem_law_pair_vjp",
    "dem_law_walls_eval", "dem_law_walls_vjp", "dem_set_material", "dem_observe",
    "dem_adj_forward", "dem_adj_backward", "dem_adj_state_grad", "dem_adj_info",
    "dem_adj_release", "dem_get_step_timing", "dem_time_rebuild", "dem_checkpoint_layout",
    "dem_law_pair_eval_floor", "dem_law_pair_vjp_floor", "dem_law_walls_eval_floor",
    "dem_law_walls_vjp_floor",
]


def lib() -> ctypes.CDLL:
    """Load libdem.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    sig = {
        "dem_last_error": ([], c_int),
        "dem_status_string": ([c_int], ctypes.c_char_p),
        "dem_create": ([ctypes.POINTER(P), c_int, P, c_uint], c_int),
        "dem_destroy": ([P], c_int),
        "dem_set_option": ([P, c_int, c_double], c_int),
        "dem_set_state": ([P, P, P, P, P, P, P, c_int], c_int),
        "dem_get_state": ([P, P, P, P, P, c_int], c_int),
        "dem_get_wall_history": ([P, P, P], c_int),
        "dem_save_checkpoint": ([P, ctypes.c_char_p], c_int),
        "dem_load_checkpoint": ([P, ctypes.c_char_p], c_int),
        "dem_read_restart": ([ctypes.c_char_p, P, P, P], c_int),
        "dem_read_sidecar": ([ctypes.c_char_p, P], c_int),
        "dem_write_restart_resume": ([ctypes.c_char_p, c_int, c_int, c_int], c_int),
        "dem_step": ([P, c_int], c_int),
        "dem_step_graph": ([P, c_int], c_int),
        "dem_request_rebuild": ([P], c_int),
        "dem_rebuild": ([P], c_int),
        "dem_autotune": ([P, c_int, P], c_int),
        "dem_set_active": ([P, c_uint], c_int),
        "dem_active_count": ([P], c_uint),
        "dem_get_stats": ([P, P], c_int),
        "dem_get_instance_stats": ([P, P, P], c_int),
        "dem_sync": ([P], c_int),
        "dem_last_ms": ([P], c_double),
        "dem_time_kernels": ([P, c_int, P, P], c_int),
        "dem_get_step_timing": ([P, P], c_int),
        "dem_time_rebuild": ([P, P], c_int),
        "dem_checkpoint_layout": ([P], c_int),
        "dem_get_hash": ([P, P], c_int),
        "dem_get_sorted": ([P, P], c_int),
        "dem_get_cell_start": ([P, P, c_uint], c_int),
        "dem_get_pairs": ([P, P, P, P, c_uint], c_int),
        "dem_get_forces": ([P, P, P, P], c_int),
        "dem_get_contacts": ([P, P], c_int),
        "dem_device_info": ([P, ctypes.c_char_p, c_int], c_int),
        "dem_dev_alloc": ([ctypes.c_size_t], P),
        "dem_dev_free": ([P], c_int),
        "dem_memcpy": ([P, P, ctypes.c_size_t, c_int], c_int),
        "dem_memzero": ([P, ctypes.c_size_t], c_int),
        "dem_device_sync": ([], c_int),
        "dem_law_pair_eval": ([P, P, P], c_int),
        "dem_contact_filter": ([P, c_int, P], c_int),
        "dem_law_pair_vjp": ([P, P, P, P, P], c_int),
        "dem_law_walls_eval": ([P, P, P, P], c_int),
        "dem_law_walls_vjp": ([P, P, P, P, P, P], c_int),
        "dem_law_pair_eval_floor": ([P, P, c_double, P], c_int),
        "dem_law_pair_vjp_floor": ([P, P, c_double, P, P, P], c_int),
        "dem_law_walls_eval_floor": ([P, P, P, c_double, P], c_int),
        "dem_law_walls_vjp_floor": ([P, P, P, c_double, P, P, P], c_int),
        "dem_set_material": ([P, c_int, P], c_int),
        "dem_observe": ([P, P, P, P], c_int),
        "dem_adj_forward": ([P, c_int, c_int, P, P, P], c_int),
        "dem_adj_backward": ([P, P, P, P], c_int),
        "dem_adj_state_grad": ([P, P, P, P], c_int),
        "dem_adj_info": ([P, P], c_int),
        "dem_adj_release": ([P], c_int),
        "dem_read_deck": ([ctypes.c_char_p, P], c_int),
        "dem_deck_material": ([P, P], c_int),
        "dem_write_restart": ([ctypes.c_char_p], c_int),
        "dem_damping_for_restitution": ([c_double], c_double),
        "dem_restitution_for_damping": ([c_double], c_double),
        "dem_gen_diameters": ([c_int, c_uint, c_double, c_double, c_double, c_ull, P], c_int),
        "dem_gen_insert": ([P, c_double, c_double, c_uint, P, c_ull, P], c_int),
        "dem_gen_lattice": ([P, c_double, c_double, c_double, c_uint, c_ull, P], c_int),
        "dem_particle_props": ([c_uint, P, P, P, P], c_int),
        "dem_make_grid": ([P, c_double, c_double, P, P], c_int),
        # drop-in layer (reference signatures, particleSystem.cu)
        "setMaterialProperties": ([P, P], None),
        "setWallGeometry": ([P], None),
        "calcHash": ([P, P, P, P, P, c_int], None),
        "RadixSort": ([P, P, c_uint, c_uint], None),
        "RadixSortNoAlloc": ([P, P, c_uint, c_uint], c_int),
        "findCellStart": ([P, P, c_int, c_uint], None),
        "neighborarray": ([P, P, P, P, P, P, c_int, c_int, P, P, P], None),
        "prefixsum": ([c_uint, P, P, P, P], None),
        "pairlist": ([P, P, P, P, P, P, P, P, c_int, c_int, c_uint, P, P, P, P, P], None),
        "calculateforcepair": ([P, P, P, P, P, P, P, P, P, P, P, P, c_double, P, P, P, P, c_uint], None),
        "sumallforcei": ([c_int, c_int, P, P, P, P, P, P, P, P, P, P], None),
        "calculateforceiw": ([P, P, P, P, P, P, c_double, c_uint, c_int, P, P, P, P], None),
        "updateposition": ([P, c_int, P, P, P, P, P, P, P, P, c_double, c_double, P, P], None),
        "outputporositycn": ([c_uint, P, c_double, P, P], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


def _check(code: int, what: str):
    if code != DEM_OK:
        raise DemError(code, what)


# ----------------------------------------------------------------------------------------
# Material / units
# ----------------------------------------------------------------------------------------

def reduced_material(E: float, nu: float, rho: float, diam: float, mu: float, mu_w: float,
                     e: float | None = None, damp: float | None = None, mu_r: float = 0.0,
                     mu_rw: float | None = None) -> Material:
    """Material in reduced units (reference dempacking.cu:141-149).  `mu_r` is the
    rolling-friction arm in units of d.  Either restitution `e` (mapped to the dashpot
    coefficient by the head-on collision ODE) or the coefficient `damp` itself."""
    er = 6.0 * E / (3.14159265 * rho * 9.81 * diam)
    m = Material()
    m.p.estarp = 1.33333333 * er / (2.0 * (1.0 - nu * nu))
    m.w.estarw = 1.33333333 / ((1.0 - nu * nu) / er + (1.0 - nu * nu) / er)
    if damp is None:
        damp = damping_for_restitution(e if e is not None else 1.0)
    m.p.ufricp, m.p.vpois, m.p.dampnp, m.p.damptp, m.p.upp = mu, nu, damp, damp, mu_r
    m.w.ufricw, m.w.vpoisw, m.w.dampnw, m.w.damptw = mu_w, nu, damp, damp
    m.w.upw = mu_r if mu_rw is None else mu_rw
    return m


@functools.lru_cache(maxsize=4096)
def damping_for_restitution(e: float) -> float:
    g = lib().dem_damping_for_restitution(float(e))
    if g < 0:
        raise ValueError(f"restitution {e} out of range")
    return g


def restitution_for_damping(g: float) -> float:
    return lib().dem_restitution_for_damping(float(g))


def reduced_dt(dt_seconds: float, diam_m: float) -> float:
    """Reduced time step: t_unit = sqrt(d/g) (dempacking.cu:153)."""
    return dt_seconds / math.sqrt(diam_m / 9.81)


def make_walls(kind: int, size=(0.0, 0.0, 0.0), radius: float = 0.0) -> Walls:
    w = Walls()
    w.kind = kind
    w.size[0], w.size[1], w.size[2] = size
    w.radius = radius
    return w


def make_instance(n: int, walls: Walls, height: float, cell: float, mat: Material, dt: float,
                  skin: float = SKIN) -> Instance:
    """Instance with the reference's grid pattern (dempacking.cu:329-344)."""
    L = lib()
    g = (c_uint * 3)()
    o = (c_double * 3)()
    _check(L.dem_make_grid(ctypes.byref(walls), height, cell, g, o), "dem_make_grid")
    inst = Instance()
    inst.n = n
    for k in range(3):
        inst.gsize[k] = g[k]
        inst.cell[k] = cell
        inst.origin[k] = o[k]
    inst.walls = walls
    inst.mat = mat
    inst.dt = dt
    inst.skin = skin
    return inst


# ----------------------------------------------------------------------------------------
# Workloads (seeded; the configurations the benchmark and tests run)
# ----------------------------------------------------------------------------------------

@dataclass
class Workload:
    name: str
    instances: list
    pos: np.ndarray
    pdis: np.ndarray
    angv: np.ndarray
    rad: np.ndarray
    rmass: np.ndarray
    inert: np.ndarray
    rebuild_interval: int = 0
    steps: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.rad.shape[0])


def _props(diam: np.ndarray):
    L = lib()
    diam = np.ascontiguousarray(diam, dtype=np.float64).reshape(-1)
    n = diam.shape[0]
    rad = np.empty(n)
    rmass = np.empty(n)
    inert = np.empty(n)
    _check(L.dem_particle_props(n, _ptr(diam), _ptr(rad), _ptr(rmass), _ptr(inert)), "props")
    return rad, rmass, inert


def gen_diameters(kind: int, n: int, p0: float, p1: float = 0.0, p2: float = 0.0, seed: int = 1):
    d = np.empty(n)
    _check(lib().dem_gen_diameters(kind, n, p0, p1, p2, seed, _ptr(d)), "dem_gen_diameters")
    return d


def gen_insert(walls: Walls, z0: float, z1: float, diam: np.ndarray, seed: int):
    diam = np.ascontiguousarray(diam, dtype=np.float64).reshape(-1)
    pos = np.empty((diam.shape[0], 3))
    _check(lib().dem_gen_insert(ctypes.byref(walls), z0, z1, diam.shape[0], _ptr(diam), seed,
                                _ptr(pos)), "dem_gen_insert")
    return pos


def gen_lattice(walls: Walls, z0: float, a: float, jitter: float, n: int, seed: int):
    pos = np.empty((n, 3))
    _check(lib().dem_gen_lattice(ctypes.byref(walls), z0, a, jitter, n, seed, _ptr(pos)),
           "dem_gen_lattice")
    return pos


# Material of configs 0-2 and 4: E = 1e7 Pa, nu = 0.29, rho = 2500 kg/m^3, d = 1 mm,
# sliding friction 0.3, rolling friction 0.002 d, restitution 0.6, dt = 1e-6 s.
MAT0 = dict(E=1e7, nu=0.29, rho=2500.0, diam=1e-3, mu=0.3, mu_w=0.3, e=0.6, mu_r=0.002)
DT0 = 1e-6


def config0(n: int = 2000, seed: int = 1) -> Workload:
    """2,000 monodisperse spheres poured from a seeded jittered lattice (jitter 0.05 d) into a
    20d x 20d x 60d box; rebuild every 20 steps."""
    walls = make_walls(WALLS_BOX, (20.0, 20.0, 60.0))
    mat = reduced_material(**MAT0)
    dt = reduced_dt(DT0, MAT0["diam"])
    diam = np.ones(n)
    pos = gen_lattice(walls, 1.0, 1.2, 0.05, n, seed)
    rad, rmass, inert = _props(diam)
    cell = 1.0 + SKIN
    inst = make_instance(n, walls, 60.0, cell, mat, dt)
    z = np.zeros((n, 3))
    return Workload("plain-2k-monodisperse-box", [inst], pos, z.copy(), z.copy(), rad, rmass,
                    inert, rebuild_interval=20, steps=20000)


def config1(n: int = 500_000, seed: int = 2, phi0: float = 0.25) -> Workload:
    """500K lognormal spheres (mean 1, sigma 0.1, clipped to [0.7, 1.3]) inserted at random
    without overlap into a cylinder of radius 40 mean-d with a flat floor, loose fill phi0."""
    radius = 40.0
    diam = gen_diameters(1, n, 0.1, 0.7, 1.3, seed)
    vol = float(np.sum(math.pi / 6.0 * diam ** 3))
    height = vol / phi0 / (math.pi * radius * radius)
    walls = make_walls(WALLS_CYLINDER, (2 * radius, 2 * radius, height), radius)
    pos = gen_insert(walls, 0.0, height, diam, seed + 100)
    rad, rmass, inert = _props(diam)
    mat = reduced_material(**MAT0)
    dt = reduced_dt(DT0, MAT0["diam"])
    cell = float(diam.max()) + SKIN
    inst = make_instance(n, walls, height, cell, mat, dt)
    z = np.zeros((n, 3))
    return Workload("plain-500k-lognormal-cylinder", [inst], pos, z.copy(), z.copy(), rad, rmass,
                    inert, rebuild_interval=0, steps=50000, meta={"height": height, "phi0": phi0})


def config2(n: int = 2_000_000, seed: int = 3, phi0: float = 0.3) -> Workload:
    """2M bimodal spheres (d = 1 and 0.5, 30:70 by volume) inserted at random into a
    120d x 120d box (height sized for fill phi0, at most 200d); friction 0.5."""
    diam = gen_diameters(2, n, 1.0, 0.5, 0.3, seed)
    vol = float(np.sum(math.pi / 6.0 * diam ** 3))
    height = min(200.0, vol / phi0 / (120.0 * 120.0))
    walls = make_walls(WALLS_BOX, (120.0, 120.0, 200.0))
    pos = gen_insert(walls, 0.0, height, diam, seed + 100)
    rad, rmass, inert = _props(diam)
    m = dict(MAT0)
    m.update(mu=0.5, mu_w=0.5)
    mat = reduced_material(**m)
    dt = reduced_dt(DT0, MAT0["diam"])
    inst = make_instance(n, walls, 200.0, 1.0 + SKIN, mat, dt)
    z = np.zeros((n, 3))
    return Workload("plain-2m-bimodal-box", [inst], pos, z.copy(), z.copy(), rad, rmass, inert,
                    rebuild_interval=0, steps=20000, meta={"height": height})


def config3(ninst: int = 256, seed: int = 4, nmin: int = 1000, nmax: int = 64000,
            rebuild_interval: int = 20) -> Workload:
    """Batched form: `ninst` independent monodisperse packings, N log-uniform in [nmin, nmax],
    friction U[0.1, 0.6], restitution U[0.3, 0.9], E log-uniform [1e6, 1e8] Pa, boxes sized
    for a 3x-loose initial fill (phi ~ 0.2)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), ninst)).astype(np.int64)
    mus = rng.uniform(0.1, 0.6, ninst)
    es = rng.uniform(0.3, 0.9, ninst)
    Es = np.exp(rng.uniform(math.log(1e6), math.log(1e8), ninst))
    dt = reduced_dt(DT0, 1e-3)

    def one(k):
        n = int(ns[k])
        phi = 0.2
        side = max(6.0, (n * math.pi / 6.0 / phi / 2.0) ** (1.0 / 3.0))   # H = 2 * side
        h = 2.0 * side
        walls = make_walls(WALLS_BOX, (side, side, h))
        pos = gen_insert(walls, 0.0, h, np.ones(n), seed * 1000 + k)
        mat = reduced_material(E=float(Es[k]), nu=0.29, rho=2500.0, diam=1e-3, mu=float(mus[k]),
                               mu_w=float(mus[k]), e=float(es[k]), mu_r=0.002)
        return make_instance(n, walls, h, 1.0 + SKIN, mat, dt), pos

    # the C generator releases the GIL, so instances are generated in parallel
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(one, range(ninst)))
    insts = [r[0] for r in res]
    pos = np.concatenate([r[1] for r in res])
    n = pos.shape[0]
    rad, rmass, inert = _props(np.ones(n))
    z = np.zeros((n, 3))
    return Workload("batched-256-monodisperse-boxes", insts, pos, z.copy(), z.copy(), rad, rmass,
                    inert, rebuild_interval=rebuild_interval, steps=10000,
                    meta={"ns": ns.tolist(), "mu": mus.tolist(), "e": es.tolist(), "E": Es.tolist()})


def config4(n: int = 20_000, seed: int = 5, phi0: float = 0.3, side: float = 30.0) -> Workload:
    """Differentiable form: `n` lognormal spheres (as config 1) inserted at random into a
    `side` x `side` box at loose fill phi0; `settle_bed` pours them into a bed, and the
    5,000-step window (checkpoints every 100 steps) starts from that bed after a tap."""
    diam = gen_diameters(1, n, 0.1, 0.7, 1.3, seed)
    vol = float(np.sum(math.pi / 6.0 * diam ** 3))
    height = vol / phi0 / (side * side)
    walls = make_walls(WALLS_BOX, (side, side, height))
    pos = gen_insert(walls, 0.0, height, diam, seed + 100)
    rad, rmass, inert = _props(diam)
    mat = reduced_material(**MAT0)
    dt = reduced_dt(DT0, MAT0["diam"])
    cell = float(diam.max()) + SKIN
    inst = make_instance(n, walls, height, cell, mat, dt)
    z = np.zeros((n, 3))
    return Workload("adjoint-20k-lognormal-box", [inst], pos, z.copy(), z.copy(), rad, rmass,
                    inert, rebuild_interval=0, steps=5000,
                    meta={"height": height, "phi0": phi0, "every": 100})


def settle_bed(w: Workload, steps: int = 100_000, interval: int = 50, tap: float = 0.5,
               seed: int = 6, graph: bool = True) -> Workload:
    """Pour `w` into a bed on the GPU (hipGraph steps, or the host loop with graph=False:
    the same steps bit for bit; fixed rebuild interval: no particle moves half the skin in
    `interval` steps of free fall from the fill height), then 'tap' it: every particle gets
    a random velocity of scale `tap` (in sqrt(g d)) and contact histories restart at zero.
    Returns the tapped bed as a new workload."""
    ctx = Context(w.instances)
    ctx.set_state(w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
    ctx.set_option(OPT_REBUILD_INTERVAL, interval)
    # graph replays in chunks: each call sizes the pair buffers for the list it starts from
    chunk = max(interval, (10_000 // interval) * interval)
    left = (steps // interval) * interval
    while left > 0:
        k = min(chunk, left)
        if graph:
            ctx.step_graph(k)
        else:
            ctx.step(k)
        left -= k
    st = ctx.state()
    ctx.close()
    dt = w.instances[0].dt
    rng = np.random.default_rng(seed)
    pdis = st["pdis"] + rng.normal(scale=tap * dt, size=st["pdis"].shape)
    maxhz = float(st["pos"][:, 2].max())
    meta = dict(w.meta)
    meta.update(settle_steps=steps, tap=tap, bed_top=maxhz)
    return Workload(w.name + "-settled", w.instances, st["pos"], pdis, st["angv"].copy(),
                    w.rad, w.rmass, w.inert, rebuild_interval=w.rebuild_interval, steps=w.steps,
                    meta=meta)


# ----------------------------------------------------------------------------------------
# Context
# ----------------------------------------------------------------------------------------

class Context:
    """One dem_ctx: `instances` independent packings stepped together."""

    def __init__(self, instances, pair_capacity: int = 0):
        L = lib()
        self._L = L
        arr = (Instance * len(instances))(*instances)
        self._inst = arr
        self.ninst = len(instances)
        self.n = int(sum(i.n for i in instances))
        self.ncells = int(sum(i.gsize[0] * i.gsize[1] * i.gsize[2] for i in instances))
        h = c_void_p()
        _check(L.dem_create(ctypes.byref(h), len(instances), ctypes.cast(arr, c_void_p),
                            pair_capacity), "dem_create")
        self._h = h

    def close(self):
        if self._h:
            self._L.dem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, opt: int, value: float):
        _check(self._L.dem_set_option(self._h, opt, float(value)), "dem_set_option")

    def set_state(self, pos, pdis, angv, rad, rmass, inert):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (pos, pdis, angv, rad, rmass, inert)]
        self._keep = arrs
        _check(self._L.dem_set_state(self._h, *[_ptr(a) for a in arrs], 0), "dem_set_state")

    def load(self, w: Workload):
        self.set_state(w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
        if w.rebuild_interval:
            self.set_option(OPT_REBUILD_INTERVAL, w.rebuild_interval)

    def state(self):
        n = self.n
        pos, pdis, angv, qi = (np.empty((n, 3)) for _ in range(4))
        _check(self._L.dem_get_state(self._h, _ptr(pos), _ptr(pdis), _ptr(angv), _ptr(qi), 0),
               "dem_get_state")
        return {"pos": pos, "pdis": pdis, "angv": angv, "qi": qi}

    def wall_history(self):
        d = np.empty((self.n, 3))
        f = np.empty(self.n)
        _check(self._L.dem_get_wall_history(self._h, _ptr(d), _ptr(f)), "dem_get_wall_history")
        return d, f

    def save_checkpoint(self, path):
        """Full-state checkpoint (state, pair list and histories, grids, counters) to `path`."""
        _check(self._L.dem_save_checkpoint(self._h, os.fsencode(path)), "dem_save_checkpoint")

    def load_checkpoint(self, path):
        """Replace this context's state by a checkpoint made for the same instances."""
        _check(self._L.dem_load_checkpoint(self._h, os.fsencode(path)), "dem_load_checkpoint")

    def step(self, nsteps: int = 1):
        _check(self._L.dem_step(self._h, int(nsteps)), "dem_step")

    def step_graph(self, nsteps: int):
        _check(self._L.dem_step_graph(self._h, int(nsteps)), "dem_step_graph")

    def rebuild(self):
        _check(self._L.dem_rebuild(self._h), "dem_rebuild")

    def autotune(self, steps: int = 64) -> dict:
        """Run the project autotuner (advances the simulation by its trial steps)."""
        r = TuneResult()
        _check(self._L.dem_autotune(self._h, int(steps), ctypes.byref(r)), "dem_autotune")
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def set_active(self, n: int):
        _check(self._L.dem_set_active(self._h, int(n)), "dem_set_active")

    def active_count(self) -> int:
        return int(self._L.dem_active_count(self._h))

    def request_rebuild(self):
        _check(self._L.dem_request_rebuild(self._h), "dem_request_rebuild")

    def stats(self) -> Stats:
        s = Stats()
        _check(self._L.dem_get_stats(self._h, ctypes.byref(s)), "dem_get_stats")
        return s

    def instance_stats(self):
        z = np.empty(self.ninst)
        c = np.empty(self.ninst, dtype=np.uint32)
        _check(self._L.dem_get_instance_stats(self._h, _ptr(z), _ptr(c)), "dem_get_instance_stats")
        return z, c

    def last_ms(self) -> float:
        return float(self._L.dem_last_ms(self._h))

    def time_kernels(self, nsteps: int):
        """Per-kernel mean device times (ms) over `nsteps` steps: pair, particle, rebuild."""
        out = np.zeros(4)
        nr = c_int(0)
        _check(self._L.dem_time_kernels(self._h, int(nsteps), _ptr(out), ctypes.byref(nr)),
               "dem_time_kernels")
        return {"pair_ms": out[0], "particle_ms": out[1], "rebuild_ms": out[2],
                "total_ms": out[3], "rebuilds": nr.value}

    def step_timing(self) -> dict:
        """Kernel times dem_step accumulated under OPT_TIME_STEPS (ms), then reset."""
        o = np.zeros(10)
        _check(self._L.dem_get_step_timing(self._h, _ptr(o)), "dem_get_step_timing")
        return {"split": {"pair_walls_ms": o[0], "particle_ms": o[1], "steps": int(o[2]),
                          "pairs_sum": o[3]},
                "fused": {"walls_ms": o[4], "step_ms": o[5], "steps": int(o[6]), "pairs_sum": o[7]},
                "rebuild_ms": o[8], "rebuilds": int(o[9])}

    REBUILD_STAGES = ("hash", "sort", "cell_range", "reorder", "nbr_count", "scans", "nbr_fill",
                      "remap", "transpose")

    def time_rebuild(self) -> dict:
        """One list rebuild now, timed stage by stage (ms), with the sizes it ran at."""
        o = np.zeros(15)
        _check(self._L.dem_time_rebuild(self._h, _ptr(o)), "dem_time_rebuild")
        out = {k: float(o[i]) for i, k in enumerate(self.REBUILD_STAGES)}
        out.update(passes=int(o[9]), npairs=int(o[10]), npairs_old=int(o[11]), ncells=int(o[12]),
                   n=int(o[13]), phases_ms=float(o[14]))
        return out

    def sync(self):
        _check(self._L.dem_sync(self._h), "dem_sync")

    def hash(self):
        kv = np.empty((self.n, 2), dtype=np.uint32)
        _check(self._L.dem_get_hash(self._h, _ptr(kv)), "dem_get_hash")
        return kv

    def sorted(self):
        kv = np.empty((self.n, 2), dtype=np.uint32)
        _check(self._L.dem_get_sorted(self._h, _ptr(kv)), "dem_get_sorted")
        return kv

    def cell_start(self):
        cs = np.empty(self.ncells, dtype=np.uint32)
        _check(self._L.dem_get_cell_start(self._h, _ptr(cs), self.ncells), "dem_get_cell_start")
        return cs

    def pairs(self):
        np_ = self.stats().npairs
        ids = np.empty((np_, 2), dtype=np.uint32)
        hist = np.empty((np_, 4))
        out = np.empty((np_, 10))
        _check(self._L.dem_get_pairs(self._h, _ptr(ids), _ptr(hist), _ptr(out), np_),
               "dem_get_pairs")
        return ids, hist, out

    def contacts(self):
        c = np.empty(self.n, dtype=np.uint32)
        _check(self._L.dem_get_contacts(self._h, _ptr(c)), "dem_get_contacts")
        return c

    def forces(self):
        f = np.empty((self.n, 3))
        t = np.empty((self.n, 3))
        c = np.empty(self.n, dtype=np.uint32)
        _check(self._L.dem_get_forces(self._h, _ptr(f), _ptr(t), _ptr(c)), "dem_get_forces")
        return f, t, c

    # ---- adjoint form ----
    def set_material(self, k: int, mat: Material):
        _check(self._L.dem_set_material(self._h, int(k), ctypes.byref(mat)), "dem_set_material")

    def observe(self, obs: Obs):
        """Smooth (phi, CN) per instance of the current state."""
        phi = np.empty(self.ninst)
        cn = np.empty(self.ninst)
        _check(self._L.dem_observe(self._h, ctypes.byref(obs), _ptr(phi), _ptr(cn)), "dem_observe")
        return phi, cn

    def adj_forward(self, nsteps: int, every: int, obs: Obs):
        """Checkpointed forward of `nsteps` steps; returns (phi, CN) per instance at the end."""
        phi = np.empty(self.ninst)
        cn = np.empty(self.ninst)
        self._obs = obs
        _check(self._L.dem_adj_forward(self._h, int(nsteps), int(every), ctypes.byref(obs),
                                       _ptr(phi), _ptr(cn)), "dem_adj_forward")
        return phi, cn

    def adj_backward(self, gphi, gcn):
        """Adjoint sweep; returns theta_bar[ninst, 10]: d/d(E*, mu, gamma_n, gamma_t, mu_r) of
        the pair law, then of the wall law."""
        gp = np.ascontiguousarray(np.broadcast_to(np.asarray(gphi, dtype=np.float64), (self.ninst,)))
        gc = np.ascontiguousarray(np.broadcast_to(np.asarray(gcn, dtype=np.float64), (self.ninst,)))
        th = np.empty((self.ninst, 10))
        _check(self._L.dem_adj_backward(self._h, _ptr(gp), _ptr(gc), _ptr(th)), "dem_adj_backward")
        return th

    def adj_state_grad(self):
        xb, vb, wb = (np.empty((self.n, 3)) for _ in range(3))
        _check(self._L.dem_adj_state_grad(self._h, _ptr(xb), _ptr(vb), _ptr(wb)),
               "dem_adj_state_grad")
        return xb, vb, wb

    def adj_info(self) -> dict:
        out = np.zeros(6)
        _check(self._L.dem_adj_info(self._h, _ptr(out)), "dem_adj_info")
        return {"forward_ms": out[0], "backward_ms": out[1], "segments": int(out[2]),
                "max_pairs": int(out[3]), "checkpoint_bytes": int(out[4]),
                "tape_bytes": int(out[5])}

    def adj_release(self):
        _check(self._L.dem_adj_release(self._h), "dem_adj_release")


def packing_stats(pos, rad, contacti, walls, margin: float = 2.0, top_margin: float = 3.0) -> dict:
    """Packing fraction and mean coordination number of a bed.

    Reference formulas (dempacking.cu:768-770): CN = sum(contacti) / N and
    phi = sum(pi/6 d^3) / (floor area * maxhz), maxhz = highest centre (the reference has
    N*pi/6 for unit spheres).  Inner-region estimate: spheres whose centres lie at least
    `margin` from the walls and floor and `top_margin` below maxhz; phi_inner = their volume /
    region volume, cn_inner = their mean contact count."""
    pos = np.asarray(pos)
    rad = np.asarray(rad)
    c = np.asarray(contacti, dtype=np.float64)
    vol = math.pi / 6.0 * (2.0 * rad) ** 3
    maxhz = max(float(pos[:, 2].max()), 0.0)
    if walls.kind == WALLS_BOX:
        area = walls.size[0] * walls.size[1]
        inside = (np.abs(pos[:, 0]) < walls.size[0] / 2 - margin) & \
                 (np.abs(pos[:, 1]) < walls.size[1] / 2 - margin)
        area_in = (walls.size[0] - 2 * margin) * (walls.size[1] - 2 * margin)
    else:
        area = math.pi * walls.radius ** 2
        inside = np.hypot(pos[:, 0], pos[:, 1]) < walls.radius - margin
        area_in = math.pi * (walls.radius - margin) ** 2
    z0, z1 = margin, maxhz - top_margin
    inside &= (pos[:, 2] > z0) & (pos[:, 2] < z1)
    region = area_in * max(z1 - z0, 0.0)
    return {
        "phi_ref": float(vol.sum() / (area * maxhz)) if maxhz > 0 else 0.0,
        "cn_ref": float(c.sum() / (len(c) + 1e-30)),
        "phi_inner": float(vol[inside].sum() / region) if region > 0 else 0.0,
        "cn_inner": float(c[inside].mean()) if inside.any() else 0.0,
        "n_inner": int(inside.sum()),
        "maxhz": maxhz,
    }


class DeviceArray:
    """A device allocation from libdem (for driving the drop-in entry points without torch)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().dem_dev_alloc(self.nbytes)
        if not self.ptr:
            raise DemError(DEM_ERR_HIP, "dem_dev_alloc")
        self.zero()

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(lib().dem_memcpy(d.ptr, a.ctypes.data_as(c_void_p), d.nbytes, 1), "dem_memcpy")
        return d

    def zero(self):
        _check(lib().dem_memzero(self.ptr, self.nbytes), "dem_memzero")

    def host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().dem_memcpy(out.ctypes.data_as(c_void_p), self.ptr, self.nbytes, 2), "dem_memcpy")
        return out

    @property
    def p(self):
        return c_void_p(self.ptr)

    def __del__(self):
        try:
            if self.ptr:
                lib().dem_dev_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def read_deck(path: str) -> Deck:
    d = Deck()
    _check(lib().dem_read_deck(path.encode(), ctypes.byref(d)), "dem_read_deck")
    return d


def deck_material(d: Deck) -> Material:
    m = Material()
    _check(lib().dem_deck_material(ctypes.byref(d), ctypes.byref(m)), "dem_deck_material")
    return m


def write_restart(path: str):
    _check(lib().dem_write_restart(path.encode()), "dem_write_restart")


class Sidecar(ctypes.Structure):
    """dem_sidecar: the deck side-car's knobs (include/dem.h)."""
    _fields_ = [
        ("walls", c_int), ("particles", c_int), ("box", c_double * 3), ("radius", c_double),
        ("height", c_double), ("diameters", c_int), ("insertion", c_int),
        ("dpar", c_double * 3), ("ipar", c_double * 3), ("seed", c_ull),
        ("restitution", c_double), ("rebuild_interval", c_int), ("pad", c_int),
    ]


DECKS = os.path.join(HERE, "decks")


def read_sidecar(path: str) -> Sidecar:
    sc = Sidecar()
    _check(lib().dem_read_sidecar(os.fsencode(path), ctypes.byref(sc)), "dem_read_sidecar")
    return sc


def read_restart(path: str):
    """(nrestart, npregas, nparticledat) from restart.dat (dempacking.cu:361-366), or None when
    the file does not start with three integers (the end-of-run file: start fresh)."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    code = lib().dem_read_restart(path.encode(), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    if code == DEM_ERR_INVALID:
        return None
    _check(code, "dem_read_restart")
    return a.value, b.value, c.value


def write_restart_resume(path: str, nrestart: int, npregas: int, nparticledat: int):
    _check(lib().dem_write_restart_resume(path.encode(), nrestart, npregas, nparticledat),
           "dem_write_restart_resume")


def __getattr__(name):
    # `adjoint` (torch.autograd.Function, calibration) is imported on first use
    if name == "adjoint":
        import importlib
        return importlib.import_module(__name__ + ".adjoint")
    raise AttributeError(name)
