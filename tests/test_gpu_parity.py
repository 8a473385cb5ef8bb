"""GPU parity: libdem.so (HIP, gfx950) against the CPU oracle on the same seeded inputs.

Bars: hashes, sort order, cell starts and pair sets bit-exact (pair sets compared after
canonical (min, max) ordering); forces, torques, histories and states within 1e-10 normwise
relative error (||gpu - oracle||_inf <= 1e-10 * ||oracle||_inf).
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / scale)


def dense_workload(dem, n=3000, seed=11, box=(14.0, 14.0, 60.0), spacing=0.995, poly=False,
                   cyl=False, mu=0.4, mu_r=0.01, e=0.5):
    """Slightly overlapping jittered lattice with random velocities and spins, so that the
    normal, tangential, rolling and wall terms are all active from the first step."""
    rng = np.random.default_rng(seed)
    if cyl:
        walls = dem.make_walls(dem.WALLS_CYLINDER, (box[0], box[0], box[2]), radius=box[0] / 2)
    else:
        walls = dem.make_walls(dem.WALLS_BOX, box)
    if poly:
        spacing = min(spacing, 0.99)
    pos = dem.gen_lattice(walls, 0.498, spacing, 0.01, n, seed)
    diam = np.ones(n) if not poly else rng.uniform(0.97, 1.03, n)
    rad = diam / 2
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    mat = dem.reduced_material(E=1e6, nu=0.29, rho=2500.0, diam=1e-3, mu=mu, mu_w=mu, e=e, mu_r=mu_r)
    dt = 2e-4
    pdis = rng.normal(scale=2e-5, size=(n, 3))
    angv = rng.normal(scale=4e-5, size=(n, 3))
    inst = dem.make_instance(n, walls, box[2], float(diam.max()) + dem.SKIN, mat, dt)
    return dem.Workload("dense", [inst], pos, pdis, angv, rad, rmass, inert)


def make_pair(dem, orc, w, width=64, interval=0):
    ctx = dem.Context(w.instances)
    ctx.set_state(w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
    ctx.set_option(dem.OPT_RECORD_FORCES, 1)
    if interval:
        ctx.set_option(dem.OPT_REBUILD_INTERVAL, interval)
    sim = orc.OracleSim(w.instances[0], w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert, width=width)
    return ctx, sim


def oracle_steps(sim, k, interval, start):
    """Advance the oracle k steps on the fixed rebuild schedule (rebuild when step % K == 0)."""
    for s in range(start, start + k):
        sim.set_iall(1 if s % interval == 0 else 0)
        assert sim.step(1) == 0


def canon(pairs):
    p = np.sort(np.asarray(pairs, dtype=np.int64), axis=1)
    return set(map(tuple, p.tolist()))


def small_cyl(dem, n=20000, seed=21):
    """Config-1-like (lognormal, cylinder) at a size the oracle checks in seconds."""
    radius = 12.0
    diam = dem.gen_diameters(1, n, 0.1, 0.7, 1.3, seed)
    vol = float(np.sum(math.pi / 6 * diam ** 3))
    h = vol / 0.3 / (math.pi * radius * radius)
    walls = dem.make_walls(dem.WALLS_CYLINDER, (2 * radius, 2 * radius, h), radius)
    pos = dem.gen_insert(walls, 0.0, h, diam, seed + 1)
    rad = diam / 2
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    mat = dem.reduced_material(**dem.MAT0)
    inst = dem.make_instance(n, walls, h, float(diam.max()) + dem.SKIN, mat,
                             dem.reduced_dt(1e-6, 1e-3))
    z = np.zeros((n, 3))
    return dem.Workload("cyl", [inst], pos, z, z.copy(), rad, rmass, inert)


def small_bimodal(dem, n=20000, seed=31):
    diam = dem.gen_diameters(2, n, 1.0, 0.5, 0.3, seed)
    walls = dem.make_walls(dem.WALLS_BOX, (24.0, 24.0, 60.0))
    vol = float(np.sum(math.pi / 6 * diam ** 3))
    h = vol / 0.3 / (24.0 * 24.0)
    pos = dem.gen_insert(walls, 0.0, h, diam, seed + 1)
    rad = diam / 2
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    m = dict(dem.MAT0)
    m.update(mu=0.5, mu_w=0.5)
    inst = dem.make_instance(n, walls, 60.0, 1.0 + dem.SKIN, dem.reduced_material(**m),
                             dem.reduced_dt(1e-6, 1e-3))
    z = np.zeros((n, 3))
    return dem.Workload("bimodal", [inst], pos, z, z.copy(), rad, rmass, inert)


@pytest.mark.parametrize("which", ["config0", "cyl", "bimodal", "dense"])
def test_rebuild_bit_exact(dem, orc, gpu_ok, which):
    w = {"config0": lambda: dem.config0(), "cyl": lambda: small_cyl(dem),
         "bimodal": lambda: small_bimodal(dem), "dense": lambda: dense_workload(dem)}[which]()
    say(f"[rebuild] {which} n={w.n}")
    ctx, sim = make_pair(dem, orc, w, width=128)
    ctx.rebuild()
    assert sim.rebuild() == 0
    assert np.array_equal(ctx.hash(), sim.field("hash"))
    assert np.array_equal(ctx.sorted(), sim.field("sorted"))
    assert np.array_equal(ctx.cell_start(), sim.field("cell_start"))
    ids, _, _ = ctx.pairs()
    assert ctx.stats().npairs == sim.scalars()["npairs"]
    assert canon(ids) == canon(sim.pairs())


def crowded_cells(dem, seed=41):
    """A lattice of 404 spheres plus 600 spheres outside the grid, which the hash clamps into
    two boundary cells (300 each, in a row of spheres 0.16 d apart): cells of hundreds of
    particles (the stable in-cell ranking, long equal-cell runs in a wave, stencil rows past
    64 slots) and a last wave of 44 live lanes (1,004 spheres)."""
    w = dense_workload(dem, n=404, seed=seed)
    inst = w.instances[0]
    hx, hy = inst.walls.size[0] / 2, inst.walls.size[1] / 2
    k = np.arange(300, dtype=np.float64)
    above = np.stack([np.full(300, 0.3), np.full(300, -0.2), inst.walls.size[2] + 2.0 + 0.163 * k], 1)
    beyond = np.stack([hx + 1.0 + 0.163 * k, np.full(300, hy - 0.4), np.full(300, 3.3)], 1)
    pos = np.concatenate([w.pos, above, beyond])
    n = len(pos)
    rad = np.full(n, 0.5)
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    z = np.zeros((n, 3))
    inst2 = dem.make_instance(n, inst.walls, inst.walls.size[2], 1.0 + dem.SKIN, inst.mat, inst.dt)
    return dem.Workload("crowded", [inst2], pos, z, z.copy(), rad, rmass, inert)


def test_rebuild_bit_exact_crowded_cells(dem, orc, gpu_ok):
    w = crowded_cells(dem)
    say(f"[rebuild] crowded n={w.n}")
    ctx, sim = make_pair(dem, orc, w, width=128)
    ctx.rebuild()
    assert sim.rebuild() == 0
    cs = sim.field("cell_start")
    filled = np.diff(np.append(cs[cs != 0xFFFFFFFF], w.n))
    assert filled.max() >= 300   # the clamped spheres do crowd into single cells
    assert np.array_equal(ctx.hash(), sim.field("hash"))
    assert np.array_equal(ctx.sorted(), sim.field("sorted"))
    assert np.array_equal(ctx.cell_start(), cs)
    ids, _, _ = ctx.pairs()
    assert ctx.stats().npairs == sim.scalars()["npairs"]
    assert canon(ids) == canon(sim.pairs())
    # a second rebuild from the cell-sorted slots (equal cells now in consecutive lanes)
    ctx.rebuild()
    assert np.array_equal(ctx.sorted(), sim.field("sorted"))
    assert canon(ctx.pairs()[0]) == canon(sim.pairs())


@pytest.mark.parametrize("poly,cyl", [(False, False), (True, False), (False, True)])
def test_first_step_forces_and_state(dem, orc, gpu_ok, poly, cyl):
    w = dense_workload(dem, poly=poly, cyl=cyl)
    say(f"[step1] poly={poly} cyl={cyl} n={w.n}")
    ctx, sim = make_pair(dem, orc, w)
    ctx.step(1)
    assert sim.step(1) == 0
    f, t, c = ctx.forces()
    assert np.array_equal(c, sim.field("contacti"))
    assert c.sum() > w.n, "workload must be in contact"
    assert rel_err(f, sim.field("forcei")) <= TOL
    assert rel_err(t, sim.field("torquei")) <= TOL
    st = ctx.state()
    for k in ("pdis", "angv"):
        assert rel_err(st[k], sim.field(k)) <= TOL, k
    assert rel_err(st["pos"] - w.pos, sim.field("pos") - w.pos) <= TOL
    dw, fw = ctx.wall_history()
    assert rel_err(dw, sim.field("disptw")) <= TOL
    assert rel_err(fw, sim.field("fricpw")) <= TOL
    s = ctx.stats()
    assert s.pcntot == sim.scalars()["pcntot"]
    assert s.maxhz == pytest.approx(sim.scalars()["maxhz"], rel=1e-14)


def test_pair_level_parity_after_rebuilds(dem, orc, gpu_ok):
    """Per-pair force, torques and carried history match the oracle after several rebuilds
    (orientation of the GPU's slot-ordered pair mapped onto the oracle's index order)."""
    w = dense_workload(dem, n=2500, seed=5)
    K = 10
    ctx, sim = make_pair(dem, orc, w, interval=K)
    ctx.step(35)
    oracle_steps(sim, 35, K, 0)
    ids, hist, out = ctx.pairs()
    P = sim.scalars()["npairs"]
    lp, ln = sim.field("lp")[:P], sim.field("ln")[:P]
    index = {(int(i), int(j)): k for k, (i, j) in enumerate(zip(lp, ln))}
    F_o, Ti_o, Tj_o = sim.field("forcepair"), sim.field("torqpairi"), sim.field("torqpairj")
    H_o, fr_o = sim.field("dispt"), sim.field("fricp")
    g = np.zeros((len(ids), 13))
    o = np.zeros_like(g)
    for q, (a, b) in enumerate(ids):
        if a < b:
            k = index[(int(a), int(b))]
            o[q] = np.concatenate([F_o[k], Ti_o[k], Tj_o[k], H_o[k], [fr_o[k]]])
        else:
            k = index[(int(b), int(a))]
            o[q] = np.concatenate([-F_o[k], Tj_o[k], Ti_o[k], -H_o[k], [fr_o[k]]])
        g[q] = np.concatenate([out[q, 0:3], out[q, 4:7], out[q, 7:10], hist[q, 0:3], hist[q, 3:4]])
    assert len(ids) == P
    assert np.count_nonzero(o[:, 9:12]) > 100, "history must be exercised"
    for sl, name in ((slice(0, 3), "F"), (slice(3, 6), "Ta"), (slice(6, 9), "Tb"),
                     (slice(9, 12), "hist"), (slice(12, 13), "fricp")):
        assert rel_err(g[:, sl], o[:, sl]) <= TOL, name
    flags = out[:, 3]
    assert int(flags.sum()) == int(sim.field("contactpair")[:P].sum())


def test_trajectory_parity_fixed_interval(dem, orc, gpu_ok):
    w = dense_workload(dem, n=3000, seed=3)
    K = 20
    ctx, sim = make_pair(dem, orc, w, interval=K)
    ctx.step(200)
    oracle_steps(sim, 200, K, 0)
    st = ctx.state()
    for k in ("pdis", "angv"):
        assert rel_err(st[k], sim.field(k)) <= TOL, k
    assert rel_err(st["pos"] - w.pos, sim.field("pos") - w.pos) <= TOL


def test_timed_rebuild_is_a_rebuild(dem, orc, gpu_ok):
    """dem_time_rebuild (two rebuilds in a row, timed by stage and by phase) leaves the context
    as one rebuild at that step does: the trajectory equals the oracle's with a rebuild forced
    at the same step, and the fixed-interval schedule counts from it."""
    w = dense_workload(dem, n=3000, seed=3)
    K = 20
    ctx, sim = make_pair(dem, orc, w, interval=K)
    ctx.step(50)
    rb = ctx.time_rebuild()
    ctx.step(50)
    rebuild_at = {0, 20, 40, 50, 70, 90}
    for s in range(100):
        sim.set_iall(1 if s in rebuild_at else 0)
        assert sim.step(1) == 0
    st = ctx.state()
    for k in ("pdis", "angv"):
        assert rel_err(st[k], sim.field(k)) <= TOL, k
    assert rel_err(st["pos"] - w.pos, sim.field("pos") - w.pos) <= TOL
    assert rb["n"] == w.n and rb["npairs"] > 0 and rb["npairs_old"] > 0
    assert rb["ncells"] == int(np.prod(w.instances[0].gsize[:]))
    assert all(rb[k] >= 0.0 for k in ctx.REBUILD_STAGES) and rb["phases_ms"] > 0.0
    assert ctx.stats().rebuilds == 7   # 0, 20, 40, the two timed ones, 70, 90


def test_trajectory_parity_displacement_trigger(dem, orc, gpu_ok):
    """Reference rebuild policy (|q| > 0.5*skin triggers a rebuild at the next step): both
    sides must rebuild at the same steps."""
    w = dense_workload(dem, n=2000, seed=8)
    w.pdis *= 200.0       # fast enough to trip the trigger within the run
    ctx, sim = make_pair(dem, orc, w)
    rebuild_steps_o = []
    for s in range(120):
        if sim.scalars()["iall"]:
            rebuild_steps_o.append(s)
        assert sim.step(1) == 0
    ctx.step(120)
    st = ctx.state()
    assert ctx.stats().rebuilds == len(rebuild_steps_o) and len(rebuild_steps_o) >= 2
    for k in ("pdis", "angv"):
        assert rel_err(st[k], sim.field(k)) <= TOL, k


def test_graph_equals_host_loop(dem, gpu_ok):
    w = dense_workload(dem, n=4000, seed=13)
    K = 10
    a = dem.Context(w.instances)
    b = dem.Context(w.instances)
    for c in (a, b):
        c.load(w)
        c.set_option(dem.OPT_REBUILD_INTERVAL, K)
    a.step(50)
    b.step_graph(30)
    b.step_graph(20)
    sa, sb = a.state(), b.state()
    for k in ("pos", "pdis", "angv"):
        assert np.array_equal(sa[k], sb[k]), k
    assert a.stats().npairs == b.stats().npairs


@pytest.mark.parametrize("mode", ["trigger", "interval", "graph", "batched", "odd-graph"])
def test_fused_and_split_paths_identical(dem, gpu_ok, mode):
    """The fused per-particle step (one-pass and chunked contact tests), the pair-kernel +
    gather path and the auto mode (which switches between them) give the same states,
    histories, contacts and pair outputs bit for bit, under both rebuild policies, in graphs
    (even and odd group counts, even and odd K) and in the batched form."""
    if mode == "batched":
        ws = [dense_workload(dem, n=n, seed=60 + i, box=(bx, bx, 60.0))
              for i, (n, bx) in enumerate([(1500, 12.0), (2600, 16.0)])]
        insts = [w.instances[0] for w in ws]
        cat = lambda k: np.concatenate([getattr(w, k) for w in ws])
        w = dem.Workload("b", insts, cat("pos"), cat("pdis"), cat("angv"), cat("rad"),
                         cat("rmass"), cat("inert"))
    else:
        w = dense_workload(dem, n=4000, seed=51)
    if mode == "trigger":
        w.pdis = w.pdis * 100.0   # fast enough for displacement-triggered rebuilds
    res = []
    # fused (one-pass contact tests), split, auto (timed probes switch paths mid-run), fused
    # with every particle on the chunked path, fused and split with one contiguous range of
    # blocks per XCD, and fused with blocks of 64 and 256 particles (a partner inside the
    # block is read from the block's LDS rows in the plain form, so the block edges move)
    for fused, fast, chunk, pblk in ((1, 64, 8, 128), (0, 64, 8, 128), (2, 64, 8, 128),
                                     (1, 0, 8, 128), (1, 64, 0, 128), (0, 64, 0, 128),
                                     (1, 64, 8, 64), (1, 64, 8, 256)):
        c = dem.Context(w.instances)
        c.set_option(dem.OPT_FUSED_STEP, fused)
        c.set_option(dem.OPT_FUSED_FAST, fast)
        c.set_option(dem.OPT_XCD_CHUNK, chunk)
        c.set_option(dem.OPT_PARTICLE_BLOCK, pblk)
        c.load(w)
        c.set_option(dem.OPT_RECORD_FORCES, 1)
        if mode == "trigger":
            c.step(200)
        elif mode in ("interval", "batched"):
            c.set_option(dem.OPT_REBUILD_INTERVAL, 10)
            c.step(60)
        elif mode == "graph":
            c.set_option(dem.OPT_REBUILD_INTERVAL, 10)
            c.step_graph(30)
            c.step_graph(20)
            c.step(5)
        else:
            c.set_option(dem.OPT_REBUILD_INTERVAL, 7)
            c.step_graph(7)
            c.step_graph(21)
            c.step(3)
        st = c.state()
        dh, fh = c.wall_history()
        f, t, cnt = c.forces()
        ids, hist, out = c.pairs()
        res.append((st, dh, fh, f, t, cnt, ids, hist, out, c.stats().rebuilds))
        c.close()
    a = res[0]
    for b in res[1:]:
        for k in ("pos", "pdis", "angv", "qi"):
            assert np.array_equal(a[0][k], b[0][k]), k
        for i in range(1, 9):
            assert np.array_equal(a[i], b[i]), i
        assert a[9] == b[9] and a[9] >= 2
    assert int(a[5].sum()) > 0


def test_queued_steps_equal_synchronous(dem, gpu_ok):
    """Under the displacement trigger dem_step queues steps ahead of the host (a queued step
    after a trigger does nothing and is issued again after the rebuild); the steps run, the
    rebuilds and the results equal the synchronous loop's bit for bit, whatever the depth and
    however many steps share one event."""
    w = dense_workload(dem, n=4000, seed=52)
    w.pdis = w.pdis * 100.0   # frequent triggers
    res = []
    for depth, batch in ((1, 1), (2, 1), (3, 1), (8, 1), (8, 4), (6, 3), (8, 8), (5, 2), (16, 8),
                         (16, 16)):
        c = dem.Context(w.instances)
        c.set_option(dem.OPT_PIPELINE, depth)
        c.set_option(dem.OPT_PIPELINE_BATCH, batch)
        c.load(w)
        for k in (1, 7, 50, 142):   # calls of several lengths
            c.step(k)
        st = c.state()
        s = c.stats()
        res.append((st, s.steps, s.rebuilds, c.contacts(), c.wall_history()[0]))
        c.close()
    a = res[0]
    assert a[2] >= 4, a[2]
    for b in res[1:]:
        for k in ("pos", "pdis", "angv", "qi"):
            assert np.array_equal(a[0][k], b[0][k]), k
        assert a[1] == b[1] == 200 and a[2] == b[2]
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])


@pytest.mark.parametrize("flen", [0, 1])
def test_batched_equals_plain(dem, gpu_ok, flen):
    """Three instances of uneven size and material in one batched context equal three plain
    contexts bit for bit (with the FP64 length and with the float-length option)."""
    ws = [dense_workload(dem, n=n, seed=40 + i, mu=mu, e=e, box=(bx, bx, 60.0))
          for i, (n, mu, e, bx) in enumerate([(1500, 0.2, 0.4, 12.0), (3100, 0.5, 0.8, 16.0),
                                               (700, 0.3, 0.6, 9.0)])]
    if flen:
        for w in ws:
            w.pdis *= 100.0   # sliding fast enough for vprojec's rescale
    K = 10
    insts = [w.instances[0] for w in ws]
    big = dem.Context(insts)
    cat = lambda k: np.concatenate([getattr(w, k) for w in ws])
    big.set_state(cat("pos"), cat("pdis"), cat("angv"), cat("rad"), cat("rmass"), cat("inert"))
    big.set_option(dem.OPT_REBUILD_INTERVAL, K)
    big.set_option(dem.OPT_FLOAT_LENGTH, flen)
    big.step(40)
    sb = big.state()
    off = 0
    for w in ws:
        c = dem.Context(w.instances)
        c.load(w)
        c.set_option(dem.OPT_REBUILD_INTERVAL, K)
        c.set_option(dem.OPT_FLOAT_LENGTH, flen)
        c.step(40)
        s = c.state()
        for k in ("pos", "pdis", "angv"):
            assert np.array_equal(sb[k][off:off + w.n], s[k]), k
        off += w.n
    z, cnt = big.instance_stats()
    assert np.all(cnt > 0)


def test_full_size_batched_256_equals_per_instance(dem, gpu_ok):
    """The batched configuration at full size (256 instances, N log-uniform in [1K, 64K],
    per-instance materials, sum N = 4.5M): 10 graph groups of 20 steps in one batched context
    end, instance by instance, bit-identical to each instance stepped in a plain context of
    its own on the same fixed rebuild interval (the batched kernels read per-instance
    constants; the slot order inside an instance is the plain context's)."""
    K, steps = 20, 200
    w = dem.config3(rebuild_interval=K)
    rng = np.random.default_rng(4)   # random motion, so that contacts form within the window
    w.pdis = rng.normal(scale=2e-3, size=w.pdis.shape)
    w.angv = rng.normal(scale=1e-3, size=w.angv.shape)
    say(f"[c3] {len(w.instances)} instances, n={w.n}")
    big = dem.Context(w.instances)
    big.load(w)
    big.set_option(dem.OPT_REBUILD_INTERVAL, K)
    big.step_graph(steps)
    sb = big.state()
    z, cnt = big.instance_stats()
    big.close()
    assert int(np.count_nonzero(cnt)) > 200, "most instances must carry contacts"
    off = 0
    for k, inst in enumerate(w.instances):
        sl = slice(off, off + inst.n)
        c = dem.Context([inst])
        c.set_state(w.pos[sl], w.pdis[sl], w.angv[sl], w.rad[sl], w.rmass[sl], w.inert[sl])
        c.set_option(dem.OPT_REBUILD_INTERVAL, K)
        c.step_graph(steps)
        s = c.state()
        c.close()
        for key in ("pos", "pdis", "angv", "qi"):
            assert np.array_equal(sb[key][sl], s[key]), (k, key)
        off += inst.n
    say(f"[c3] all {len(w.instances)} instances bit-identical")


def sparse_workload(dem, n):
    """n spheres more than a cutoff apart, the lowest pressed 0.001 d into the floor, the rest
    just above it, with spins and sideways velocities: wall contacts from the first step and
    never a candidate pair (P = 0, the reference's zero-pair case)."""
    walls = dem.make_walls(dem.WALLS_BOX, (24.0, 24.0, 30.0))
    k = np.arange(n, dtype=np.float64)
    pos = np.stack([-10.0 + 4.0 * k, 9.0 - 4.5 * k, 0.499 + 0.3 * k], axis=1)
    rng = np.random.default_rng(n)
    pdis = rng.normal(scale=2e-5, size=(n, 3))
    angv = rng.normal(scale=4e-5, size=(n, 3))
    rad = np.full(n, 0.5)
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    mat = dem.reduced_material(E=1e6, nu=0.29, rho=2500.0, diam=1e-3, mu=0.4, mu_w=0.4, e=0.5,
                               mu_r=0.01)
    inst = dem.make_instance(n, walls, 30.0, 1.0 + dem.SKIN, mat, 2e-4)
    return dem.Workload("sparse", [inst], pos, pdis, angv, rad, rmass, inert)


@pytest.mark.parametrize("n", [1, 5])
@pytest.mark.parametrize("mode", ["fused", "split", "graph"])
def test_workloads_without_pairs(dem, orc, gpu_ok, n, mode):
    """One sphere, and spheres farther apart than the cutoff: empty pair lists through every
    rebuild kernel and both step paths (and a captured graph), wall contacts and free fall
    equal to the oracle."""
    say(f"[sparse] n={n} mode={mode}")
    w = sparse_workload(dem, n)
    K = 10
    ctx, sim = make_pair(dem, orc, w, interval=K if mode == "graph" else 0)
    ctx.set_option(dem.OPT_FUSED_STEP, 0 if mode == "split" else 1)
    steps = 200
    if mode == "graph":
        ctx.step_graph(steps)
        oracle_steps(sim, steps, K, 0)
    else:
        ctx.step(steps)
        for _ in range(steps):
            assert sim.step(1) == 0
    st = ctx.state()
    assert ctx.stats().npairs == 0 and sim.scalars()["npairs"] == 0
    for key in ("pdis", "angv"):
        assert rel_err(st[key], sim.field(key)) <= TOL, key
    assert rel_err(st["pos"] - w.pos, sim.field("pos") - w.pos) <= TOL
    f, t, c = ctx.forces()
    assert np.all(c == 0)
    assert rel_err(f, sim.field("forcei")) <= TOL and rel_err(t, sim.field("torquei")) <= TOL
    assert abs(st["pos"][0, 2] - 0.5) < 0.01, "the lowest sphere rests on the floor"


@pytest.mark.parametrize("mode", ["fused", "split", "graph"])
def test_float_length_option(dem, orc, gpu_ok, mode):
    """DEM_OPT_FLOAT_LENGTH: the reference's float length() in vprojec and in the rebuild
    trigger.  Fast spheres (histories past vprojec's 0.001 threshold, several triggers): the
    trajectory and the rebuild steps equal the oracle run with the same option, and differ
    from the FP64 default; the adjoint form refuses the option."""
    say(f"[float-length] mode={mode}")
    w = dense_workload(dem, n=2000, seed=8)
    w.pdis *= 200.0
    K = 10
    graph = mode == "graph"
    ctx, sim = make_pair(dem, orc, w, interval=K if graph else 0)
    sim.float_length = True
    ctx.set_option(dem.OPT_FLOAT_LENGTH, 1)
    ctx.set_option(dem.OPT_FUSED_STEP, 0 if mode == "split" else 1)
    steps = 120
    if graph:
        ctx.step_graph(steps)
        oracle_steps(sim, steps, K, 0)
    else:
        rebuilds = 0
        for _ in range(steps):
            rebuilds += 1 if sim.scalars()["iall"] else 0
            assert sim.step(1) == 0
        ctx.step(steps)
        assert ctx.stats().rebuilds == rebuilds and rebuilds >= 2
    st = ctx.state()
    for key in ("pdis", "angv"):
        assert rel_err(st[key], sim.field(key)) <= TOL, key
    assert rel_err(st["pos"] - w.pos, sim.field("pos") - w.pos) <= TOL
    ref = dem.Context(w.instances)
    ref.load(w)
    ref.set_option(dem.OPT_FUSED_STEP, 0 if mode == "split" else 1)
    if graph:
        ref.set_option(dem.OPT_REBUILD_INTERVAL, K)
        ref.step_graph(steps)
    else:
        ref.step(steps)
    assert not np.array_equal(ref.state()["pdis"], st["pdis"]), "the option must change the bits"
    with pytest.raises(dem.DemError) as ei:
        ctx.adj_forward(10, 5, dem.make_obs(zhi=10.0))
    assert ei.value.code == dem.DEM_ERR_INVALID
    for bad in (0.5, 2.0, -1.0):
        with pytest.raises(dem.DemError):
            ctx.set_option(dem.OPT_FLOAT_LENGTH, bad)


def test_pair_capacity_grows(dem, orc, gpu_ok):
    w = dense_workload(dem, n=1500, seed=17)
    ctx = dem.Context(w.instances, pair_capacity=64)
    ctx.load(w)
    ctx.set_option(dem.OPT_RECORD_FORCES, 1)
    ctx.step(3)
    sim = orc.OracleSim(w.instances[0], w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
    assert sim.step(3) == 0
    assert ctx.stats().pair_capacity >= ctx.stats().npairs > 64
    assert rel_err(ctx.state()["pdis"], sim.field("pdis")) <= TOL


def test_dropin_reference_sequence(dem, orc, gpu_ok):
    """The reference-named entry points, called in dempacking.cu:535-739 order on
    caller-owned device buffers, reproduce the oracle (which follows the same layout)."""
    import ctypes
    L = dem.lib()
    D = dem.DeviceArray
    w = dense_workload(dem, n=2000, seed=23, mu_r=0.0)
    inst = w.instances[0]
    n = w.n
    W = 30
    S = n * W
    f64, u32 = np.float64, np.uint32
    pos, pdis, angv = (D.from_host(getattr(w, k)) for k in ("pos", "pdis", "angv"))
    rad, rmass, inert = (D.from_host(getattr(w, k)) for k in ("rad", "rmass", "inert"))
    C = inst.gsize[0] * inst.gsize[1] * inst.gsize[2]
    hash0, hash1 = D(2 * n, u32), D(2 * n, u32)
    cstart = D(C, u32)
    anei, aref = D(S, u32), D(S, u32)
    njgi, njli, sjgi, oldsjgi = D(n, u32), D(n, u32), D(n, u32), D(n, u32)
    lp, ln, oldln, cpair = D(S, u32), D(S, u32), D(S, u32), D(S, u32)
    dispt, olddispt, fp, tpi, tpj = (D(3 * S, f64) for _ in range(5))
    fricp, oldfricp = D(S, f64), D(S, f64)
    disptw, fricpw = D(3 * n, f64), D(n, f64)
    forcei, torquei, qi = D(3 * n, f64), D(3 * n, f64), D(3 * n, f64)
    contacti = D(n, u32)
    G = (ctypes.c_uint * 3)(*inst.gsize)
    cell = (ctypes.c_double * 3)(*inst.cell)
    orig = (ctypes.c_double * 3)(*inst.origin)
    ws = (ctypes.c_double * 3)(*inst.walls.size)
    L.setMaterialProperties(ctypes.byref(inst.mat.p), ctypes.byref(inst.mat.w))
    L.setWallGeometry(None)
    iall = ctypes.c_uint(1)
    maxhz = ctypes.c_double(0.0)
    P = ctypes.c_uint(0)
    pcn = ctypes.c_uint(0)
    sim = orc.OracleSim(inst, w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert, width=W)
    nsteps = 25
    for it in range(nsteps):
        forcei.zero(); torquei.zero(); contacti.zero()
        if iall.value:
            qi.zero()
            L.calcHash(pos.p, hash0.p, G, cell, orig, n)
            L.RadixSort(hash0.p, hash1.p, n, 32)
            L.findCellStart(hash0.p, cstart.p, n, C)
            if it == 0:
                sim_h = orc.OracleSim(inst, w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert, width=W)
                assert sim_h.rebuild() == 0
                assert np.array_equal(hash0.host().reshape(n, 2), sim_h.field("sorted"))
                assert np.array_equal(cstart.host(), sim_h.field("cell_start"))
            L.neighborarray(pos.p, hash0.p, cstart.p, G, cell, orig, W, n, anei.p, njgi.p, njli.p)
            if it == 0:   # whole rows, the empty tails included
                assert np.array_equal(anei.host().reshape(n, W), sim_h.field("anei").reshape(n, W))
                assert np.array_equal(njgi.host(), sim_h.field("njgi"))
                assert np.array_equal(njli.host(), sim_h.field("njli"))
            L.prefixsum(n, njgi.p, sjgi.p, oldsjgi.p, ctypes.byref(P))
            L.pairlist(sjgi.p, oldsjgi.p, oldln.p, anei.p, njgi.p, njli.p, olddispt.p,
                       oldfricp.p, W, W, n, aref.p, lp.p, ln.p, dispt.p, fricp.p)
            iall.value = 0
        L.calculateforcepair(pos.p, pdis.p, angv.p, rad.p, rmass.p, lp.p, ln.p, oldln.p,
                             dispt.p, fricp.p, olddispt.p, oldfricp.p, inst.dt, fp.p, tpi.p,
                             tpj.p, cpair.p, P.value)
        L.sumallforcei(W, n, njgi.p, njli.p, fp.p, tpi.p, tpj.p, cpair.p, aref.p,
                       forcei.p, torquei.p, contacti.p)
        L.calculateforceiw(ws, pos.p, pdis.p, angv.p, rad.p, rmass.p, inst.dt, P.value, n,
                           disptw.p, fricpw.p, forcei.p, torquei.p)
        L.updateposition(ws, n, pos.p, pdis.p, angv.p, rmass.p, inert.p, forcei.p,
                         torquei.p, qi.p, inst.dt, 0.3, ctypes.byref(iall), ctypes.byref(maxhz))
        assert L.dem_last_error() == 0
        assert sim.step(1) == 0
        assert P.value == sim.scalars()["npairs"]
        assert iall.value == sim.scalars()["iall"]
    L.outputporositycn(n, contacti.p, maxhz.value, ws, ctypes.byref(pcn))
    assert pcn.value == sim.scalars()["pcntot"]
    assert maxhz.value == sim.scalars()["maxhz"]
    Pn = P.value
    assert np.array_equal(lp.host()[:Pn], sim.field("lp")[:Pn])
    assert np.array_equal(ln.host()[:Pn], sim.field("ln")[:Pn])
    assert np.array_equal(aref.host().reshape(n, W), sim.field("aref").reshape(n, W))
    assert np.array_equal(contacti.host(), sim.field("contacti"))
    for t, k in ((pos, "pos"), (pdis, "pdis"), (angv, "angv")):
        g = t.host().reshape(n, 3)
        ref = sim.field(k)
        if k == "pos":
            assert rel_err(g - w.pos, ref - w.pos) <= TOL
        else:
            assert rel_err(g, ref) <= TOL, k
    assert rel_err(fp.host()[:3 * Pn], sim.field("forcepair")[:Pn].ravel()) <= TOL
    assert rel_err(dispt.host()[:3 * Pn], sim.field("dispt")[:Pn].ravel()) <= TOL
    assert int(cpair.host()[:Pn].sum()) > n, "drop-in workload must be in contact"


def test_invalid_arguments_and_capacity_errors(dem, gpu_ok):
    """Error paths return a status instead of ending the process (the reference's gpuErrchk
    exit()s, particleSystem.cu:15-23): bad instance descriptions, option values, step counts
    and active counts give DEM_ERR_INVALID and leave the context usable; a neighbour row too
    short for a particle's candidates gives DEM_ERR_CAPACITY through dem_last_error, which the
    read clears, and no entry is written past the row."""
    import ctypes
    w = dense_workload(dem, n=800, seed=71)
    inst = w.instances[0]

    def variant(f):
        b = type(inst)()
        ctypes.pointer(b)[0] = inst
        f(b)
        return b

    edits = (lambda b: setattr(b, "n", 0), lambda b: setattr(b, "dt", 0.0),
             lambda b: setattr(b, "skin", -0.1), lambda b: b.gsize.__setitem__(0, 0),
             lambda b: b.cell.__setitem__(1, 0.0))
    for f in edits:
        with pytest.raises(dem.DemError) as ei:
            dem.Context([variant(f)])
        assert ei.value.code == dem.DEM_ERR_INVALID
    c = dem.Context([inst])
    c.load(w)
    c.step(5)
    for opt, val in ((dem.OPT_PAIR_BLOCK, 100), (dem.OPT_PARTICLE_BLOCK, 32),
                     (dem.OPT_FUSED_STEP, 3), (dem.OPT_FUSED_STEP, 0.5), (dem.OPT_PIPELINE, 0),
                     (dem.OPT_PIPELINE, 17), (dem.OPT_PIPELINE_BATCH, 0),
                     (dem.OPT_PIPELINE_BATCH, 17), (dem.OPT_XCD_CHUNK, -1),
                     (dem.OPT_FUSED_FAST, 65), (dem.OPT_REBUILD_INTERVAL, -1),
                     (dem.OPT_REBUILD_INTERVAL, float("nan")), (dem.OPT_REBUILD_INTERVAL, 1e20),
                     (dem.OPT_REBUILD_INTERVAL, 2.5),
                     (dem.OPT_SKIN, 0.0), (dem.OPT_PAIR_GRID, 0), (99, 1.0)):
        with pytest.raises(dem.DemError) as ei:
            c.set_option(opt, val)
        assert ei.value.code == dem.DEM_ERR_INVALID, (opt, val)
    calls = (lambda: c.step(-1), lambda: c.step_graph(10), lambda: c.set_active(0),
             lambda: c.set_active(inst.n + 1), lambda: c.set_active(inst.n - 100))
    for f in calls:
        with pytest.raises(dem.DemError) as ei:
            f()
        assert ei.value.code == dem.DEM_ERR_INVALID
    c.set_option(dem.OPT_REBUILD_INTERVAL, 20)
    with pytest.raises(dem.DemError) as ei:
        c.step_graph(30)
    assert ei.value.code == dem.DEM_ERR_INVALID
    c.step(10)
    assert c.stats().steps == 15 and np.all(np.isfinite(c.state()["pos"]))
    two = dem.Context([inst, inst])
    with pytest.raises(dem.DemError) as ei:
        two.set_active(10)
    assert ei.value.code == dem.DEM_ERR_INVALID
    # drop-in neighbour rows of width 2 on a compressed lattice (about 12 candidates each)
    L = dem.lib()
    D = dem.DeviceArray
    n = w.n
    C = inst.gsize[0] * inst.gsize[1] * inst.gsize[2]
    pos = D.from_host(w.pos)
    hash0, hash1 = D(2 * n, np.uint32), D(2 * n, np.uint32)
    cstart = D(C, np.uint32)
    guard = D(2 * n + 64, np.uint32)   # the rows, then a guard band that must stay zero
    njgi, njli = D(n, np.uint32), D(n, np.uint32)
    G = (ctypes.c_uint * 3)(*inst.gsize)
    cell = (ctypes.c_double * 3)(*inst.cell)
    orig = (ctypes.c_double * 3)(*inst.origin)
    L.dem_last_error()   # start from a clear error word
    L.calcHash(pos.p, hash0.p, G, cell, orig, n)
    L.RadixSort(hash0.p, hash1.p, n, 32)
    L.findCellStart(hash0.p, cstart.p, n, C)
    L.neighborarray(pos.p, hash0.p, cstart.p, G, cell, orig, 2, n, guard.p, njgi.p, njli.p)
    assert L.dem_last_error() == dem.DEM_ERR_CAPACITY
    assert L.dem_last_error() == 0
    g = guard.host()
    assert np.all(g[2 * n:] == 0), "neighborarray wrote past its rows"
    assert int(njgi.host().sum() + njli.host().sum()) > 4 * n


def test_full_size_500k_rebuild_and_step(dem, orc, gpu_ok):
    """The 500K lognormal cylinder workload at full size, its loose fill compressed to 55% of
    its height and given random displacements and spins so that most spheres are in
    contact: lists bit-exact, one step's forces and state within 1e-10 of the oracle, and
    the pair forces cancel (Newton's third law): the total force over all spheres is
    gravity plus the wall forces, which the oracle's wall law gives for the spheres that
    touch a wall."""
    w = dem.config1()
    n = w.n
    say(f"[500k] n={n}")
    rng = np.random.default_rng(51)
    w.pos[:, 2] = w.pos[:, 2] * 0.55
    w.pdis = rng.normal(scale=1e-4, size=(n, 3))
    w.angv = rng.normal(scale=1e-4, size=(n, 3))
    ctx, sim = make_pair(dem, orc, w, width=64)
    ctx.rebuild()
    assert sim.rebuild() == 0
    assert np.array_equal(ctx.hash(), sim.field("hash"))
    assert np.array_equal(ctx.sorted(), sim.field("sorted"))
    ids, _, _ = ctx.pairs()
    assert canon(ids) == canon(sim.pairs())
    sim.set_iall(0)
    ctx.step(1)
    assert sim.step(1) == 0
    f, t, c = ctx.forces()
    assert rel_err(f, sim.field("forcei")) <= TOL
    assert rel_err(t, sim.field("torquei")) <= TOL
    assert np.array_equal(c, sim.field("contacti"))
    assert int(c.sum()) > n, "the compressed bed must be in contact"
    st = ctx.state()
    assert rel_err(st["pdis"], sim.field("pdis")) <= TOL
    # internal forces cancel: total force = gravity + walls
    inst = w.instances[0]
    R = inst.walls.radius
    touch = np.nonzero((w.pos[:, 2] < w.rad) |
                       (np.hypot(w.pos[:, 0], w.pos[:, 1]) + w.rad > R))[0]
    fw = np.zeros(3)
    for i in touch:
        hits, fwi, _, _, _ = orc.law_wall(inst.walls, w.pos[i], w.pdis[i], w.angv[i], w.rad[i],
                                          w.rmass[i], np.zeros(3), inst.dt, inst.mat)
        fw += fwi
    resid = f.sum(axis=0) - fw - np.array([0.0, 0.0, -w.rmass.sum()])
    scale = float(np.max(np.linalg.norm(f, axis=1)))
    say(f"[500k] {len(touch)} spheres on a wall; pair-force residual {np.max(np.abs(resid)) / scale:.2e}")
    assert len(touch) > 0
    assert np.max(np.abs(resid)) <= 1e-10 * scale


def pair_keys(pairs, n):
    """Canonical (min, max) pairs as sorted int64 keys (fast set equality for large lists)."""
    p = np.sort(np.asarray(pairs, dtype=np.int64), axis=1)
    return np.sort(p[:, 0] * n + p[:, 1])


def test_full_size_2m_bimodal_box(dem, orc, gpu_ok):
    """The 2M bimodal box workload (d = 1 and 0.5, friction 0.5) at full size, the pair-list
    imbalance case: the loose fill compressed to 80% of its height and given random
    displacements and spins, so that contacts of every size combination and the walls are
    active.  Lists bit-exact after the first rebuild; forces, torques and state within 1e-10
    of the oracle after one step and after six (rebuilds every 2 steps, so histories are
    remapped twice at full size); the fused step (one-pass and chunked tests) and the split
    path bit-identical."""
    w = dem.config2()
    n = w.n
    say(f"[2m] n={n}")
    rng = np.random.default_rng(41)
    w.pos[:, 2] = w.pos[:, 2] * 0.8
    w.pdis = rng.normal(scale=5e-3, size=(n, 3))
    w.angv = rng.normal(scale=5e-3, size=(n, 3))
    K = 2
    ctx, sim = make_pair(dem, orc, w, width=128, interval=K)
    ctx.rebuild()
    assert sim.rebuild() == 0
    assert np.array_equal(ctx.hash(), sim.field("hash"))
    assert np.array_equal(ctx.sorted(), sim.field("sorted"))
    ids, _, _ = ctx.pairs()
    P = sim.scalars()["npairs"]
    assert len(ids) == P
    po = np.stack([sim.field("lp")[:P], sim.field("ln")[:P]], axis=1)
    assert np.array_equal(pair_keys(ids, n), pair_keys(po, n))
    say(f"[2m] lists equal, P={P} ({P / n:.2f} per sphere)")
    sim.set_iall(0)
    ctx.step(1)
    assert sim.step(1) == 0
    f, t, c = ctx.forces()
    assert rel_err(f, sim.field("forcei")) <= TOL
    assert rel_err(t, sim.field("torquei")) <= TOL
    assert np.array_equal(c, sim.field("contacti"))
    assert int(c.sum()) > n // 10, "the compressed bed must be in contact"
    ctx.step(5)
    oracle_steps(sim, 5, K, 1)
    f, t, c = ctx.forces()
    st = ctx.state()
    assert rel_err(f, sim.field("forcei")) <= TOL
    assert rel_err(t, sim.field("torquei")) <= TOL
    assert np.array_equal(c, sim.field("contacti"))
    for k in ("pdis", "angv"):
        assert rel_err(st[k], sim.field(k)) <= TOL, k
    assert rel_err(st["pos"] - w.pos, sim.field("pos") - w.pos) <= TOL
    assert ctx.stats().rebuilds == 3
    say("[2m] oracle parity after 6 steps")
    del sim
    ref = (st, f, t, c)
    ctx.close()
    for fused, fast in ((1, 0), (0, 64)):
        b = dem.Context(w.instances)
        b.set_state(w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
        b.set_option(dem.OPT_RECORD_FORCES, 1)
        b.set_option(dem.OPT_REBUILD_INTERVAL, K)
        b.set_option(dem.OPT_FUSED_STEP, fused)
        b.set_option(dem.OPT_FUSED_FAST, fast)
        b.rebuild()
        b.step(6)
        sb = b.state()
        fb, tb, cb = b.forces()
        for k in ("pos", "pdis", "angv", "qi"):
            assert np.array_equal(sb[k], ref[0][k]), (fused, fast, k)
        assert np.array_equal(fb, ref[1]) and np.array_equal(tb, ref[2])
        assert np.array_equal(cb, ref[3])
        b.close()


def test_full_pour_packing_and_coordination(dem, orc, gpu_ok):
    """Pour of the 2,000-sphere configuration (jittered lattice into a 20d x 20d box, rebuild
    every 20 steps) run until settled on the GPU (hipGraph path) and on the oracle: packing
    fraction and mean coordination number agree within 0.5%."""
    w = dem.config0()
    K = w.rebuild_interval
    nsteps = 80000
    ctx = dem.Context(w.instances)
    ctx.load(w)
    ctx.step_graph(nsteps)
    sim = orc.OracleSim(w.instances[0], w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
    say(f"[pour] gpu done {ctx.last_ms():.0f} ms; oracle running {nsteps} steps")
    oracle_steps(sim, nsteps, K, 0)
    st = ctx.state()
    walls = w.instances[0].walls
    g = dem.packing_stats(st["pos"], w.rad, ctx.contacts(), walls)
    o = dem.packing_stats(sim.field("pos"), w.rad, sim.field("contacti"), walls)
    say(f"[pour] gpu {g}\n[pour] oracle {o}")
    assert o["cn_inner"] > 3.0 and o["phi_inner"] > 0.5, "bed must be packed"
    for k in ("phi_ref", "cn_ref", "phi_inner", "cn_inner"):
        assert g[k] == pytest.approx(o[k], rel=5e-3), k
    say(f"[pour] max |dpos| = {np.max(np.abs(st['pos'] - sim.field('pos'))):.3e}")


def test_layer_insertion_parity(dem, orc, gpu_ok):
    """Layer-by-layer pour (dempacking.cu:378-462): particles join in index order and force a
    rebuild; history of existing pairs is kept.  GPU (dem_set_active) vs oracle."""
    walls = dem.make_walls(dem.WALLS_BOX, (10.0, 10.0, 40.0))
    per = 64
    nl = 5
    n = per * nl
    rng = np.random.default_rng(7)
    layer = np.array([[-4.2 + 1.2 * i, -4.2 + 1.2 * j] for i in range(8) for j in range(8)])
    pos = np.zeros((n, 3))
    for L in range(nl):
        pos[L * per:(L + 1) * per, :2] = layer + rng.uniform(-0.05, 0.05, (per, 2))
        pos[L * per:(L + 1) * per, 2] = 8.0
    pdis = np.zeros((n, 3))
    pdis[:, 2] = -4e-3
    rad = np.full(n, 0.5)
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    mat = dem.reduced_material(E=1e6, nu=0.29, rho=2500.0, diam=1e-3, mu=0.4, mu_w=0.4, e=0.5,
                               mu_r=0.002)
    inst = dem.make_instance(n, walls, 40.0, 1.55, mat, 1e-3)
    ctx = dem.Context([inst])
    ctx.set_state(pos, pdis, np.zeros((n, 3)), rad, rmass, inert)
    sim = orc.OracleSim(inst, pos, pdis, np.zeros((n, 3)), rad, rmass, inert)
    active = 0
    for it in range(1, 4001):
        if active < n and it % 400 == 1:
            active = min(n, active + per)
            ctx.set_active(active)
            sim.set_n(active)
        ctx.step(1)
        assert sim.step(1) == 0
    assert ctx.stats().rebuilds >= nl
    st = ctx.state()
    assert np.array_equal(ctx.contacts(), sim.field("contacti"))
    assert sim.field("contacti").sum() > n
    for k in ("pdis", "angv"):
        assert rel_err(st[k], sim.field(k)) <= TOL, k
    assert rel_err(st["pos"] - pos, sim.field("pos") - pos) <= TOL


def test_gpusphere_driver_on_shipped_deck(dem, gpu_ok, tmp_path):
    """The gpusphere driver on the reference's own hopp3d.inp (first 2,500 steps): insertion
    schedule, ppor.dat / particle.dat formats, restart.dat bytes."""
    import os
    import re
    import shutil
    import subprocess
    golden = os.path.join(os.path.dirname(__file__), "golden")
    shutil.copy(os.path.join(golden, "hopp3d.inp"), tmp_path / "hopp3d.inp")
    exe = os.path.join(dem.HERE, "gpusphere")
    r = subprocess.run([exe, "--tstop", "2500", "--quiet"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    assert "program comes to end!" in r.stdout
    lines = (tmp_path / "ppor.dat").read_text().splitlines()
    assert len(lines) == 2
    pat = re.compile(r"^\s{0,7}\d{1,8}\s{1,8}\d{1,8}\s{1,8}\d{1,8}(\s+-?\d+\.\d{6}){4}$")
    for ln in lines:
        assert pat.match(ln), ln
    it, n = zip(*[(int(ln.split()[0]), int(ln.split()[1])) for ln in lines])
    # layers of 47 x 47 = 2209 enter at steps 1, 501, 1001, 1501 (time > N * 0.5 / nplayer)
    assert it == (1000, 2000) and n == (4418, 8836)
    assert (tmp_path / "restart.dat").read_bytes() == open(os.path.join(golden, "restart.dat"), "rb").read()
    assert (tmp_path / "jobfinished").read_text() == "job finished\n"
    pd = (tmp_path / "particle.dat").read_text().splitlines()
    assert pd[0] == 'ZONE T="ZONE001"' and len(pd) == 1 + 2209
    assert (tmp_path / "view.dat").read_text().startswith("number of particles=10000\n")


def test_gpusphere_with_sidecar_runs_configuration_0(dem, gpu_ok, tmp_path):
    """The driver on the shipped configuration-0 deck and side-car (jittered lattice,
    restitution 0.6, rebuild every 20 steps) runs the same steps as the context API on the
    package's configuration-0 workload: its checkpoint after step 200 holds the same state bit
    for bit."""
    import os
    import shutil
    import subprocess
    for f in ("config0.inp", "config0.side"):
        shutil.copy(os.path.join(dem.DECKS, f), tmp_path / f)
    exe = os.path.join(dem.HERE, "gpusphere")
    r = subprocess.run([exe, "config0.inp", "--side", "config0.side", "--tstop", "400",
                        "--checkpoint-every", "200", "--quiet"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    pre = (tmp_path / "preflow.dat").read_text().splitlines()
    assert pre[0] == "     200,    2000,       0" and len(pre) == 2001
    w = dem.config0()
    a = dem.Context(w.instances)
    a.load(w)
    a.step(200)
    ref = a.state()
    b = dem.Context(w.instances)
    b.load_checkpoint(str(tmp_path / "preflow.bin"))
    assert b.stats().steps == 200
    got = b.state()
    for k in ("pos", "pdis", "angv", "qi"):
        assert np.array_equal(got[k], ref[k]), k
    for x, y in zip(b.wall_history(), a.wall_history()):
        assert np.array_equal(x, y)
    # the loaded context steps with the driver's instance table (materials from the deck and
    # the side-car's restitution): once the lattice reaches the floor and the spheres touch,
    # any difference from the package's material would show
    a.step(16000)
    b.step(16000)
    assert a.contacts().sum() > 500, "the spheres must be in contact by then"
    got, ref = b.state(), a.state()
    for k in ("pos", "pdis", "angv", "qi"):
        assert np.array_equal(got[k], ref[k]), k


def test_gpusphere_ppor_values_equal_the_oracle(dem, orc, gpu_ok, tmp_path):
    """ppor.dat of the driver's full configuration-0 run (20,000 steps: the lattice falls,
    its lowest layer reaches the floor near step 13,000 and the contacts grow) against the
    oracle run of the same workload: at every 1,000-step mark the particle count and
    contacts / 2 are equal and the
    porosity, packing fraction, mean coordination number and bed height columns are the same
    text ("%15.6f", dempacking.cu:758-777), computed from the oracle's contact total and
    maximum height with the reference's formulas (dempacking.cu:768-770); particle.dat's
    zone of step 1 and view.dat are the same text as the oracle's state and the deck
    formatted alike."""
    import os
    import shutil
    import subprocess
    for f in ("config0.inp", "config0.side"):
        shutil.copy(os.path.join(dem.DECKS, f), tmp_path / f)
    exe = os.path.join(dem.HERE, "gpusphere")
    nsteps = 20000
    r = subprocess.run([exe, "config0.inp", "--side", "config0.side", "--tstop", str(nsteps),
                        "--quiet"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "ppor.dat").read_text().splitlines()
    assert len(lines) == nsteps // 1000
    w = dem.config0()
    n = w.n
    sim = orc.OracleSim(w.instances[0], w.pos, w.pdis, w.angv, w.rad, w.rmass, w.inert)
    pi = math.atan(1.0) * 4.0
    area = w.instances[0].walls.size[0] * w.instances[0].walls.size[1]
    expect = []
    pos1 = None
    for it in range(1, nsteps + 1):
        sim.set_iall(1 if (it - 1) % w.rebuild_interval == 0 else 0)
        assert sim.step(1) == 0
        if it == 1:
            pos1 = sim.field("pos").copy()
        if it % 1000 == 0:
            sc = sim.scalars()
            avercn = sc["pcntot"] / (n * 1.0 + 1.0e-30)
            porosity5 = 1.0 - n * pi / 6 / (area * sc["maxhz"])
            expect.append("%8d %8d %8d %15.6f %15.6f %15.6f %15.6f" % (
                it, n, sc["pcntot"] // 2, porosity5, 1.0 - porosity5, avercn, sc["maxhz"]))
    assert lines == expect
    # particle.dat: the Tecplot zone of step 1 (dempacking.cu:779-792), positions and radii in
    # metres, the same text as the oracle's state formatted alike
    d = dem.read_deck(str(tmp_path / "config0.inp"))
    zone = ["ZONE T=\"ZONE001\""] + [
        "%15.6g%15.6g%15.6g%15.6g%8d" % (pos1[i, 0] * d.diam, pos1[i, 1] * d.diam,
                                       pos1[i, 2] * d.diam, w.rad[i] * d.diam, i)
        for i in range(n)]
    assert (tmp_path / "particle.dat").read_text().splitlines() == zone
    # view.dat: the run description of dempacking.cu:227-234 from the deck values
    view = ["number of particles=%4d" % n, "diameters =%6.3f" % d.diam,
            "time step=%f bed width=%6.3f thickniss=%6.3f" % (d.dt, d.dmt1 / d.diam,
                                                             d.dmt2 / d.diam),
            "glabal damping=%10.4f bed hbot=%10.4f htop=%10.4f" % (
                d.cdamp, d.zht1 / d.diam, d.zht1 / d.diam + d.zht2 / d.diam)]
    assert (tmp_path / "view.dat").read_text().splitlines() == view
    last = lines[-1].split()
    assert int(last[2]) > 500, "the spheres must be in contact by the end of the run"
    say(f"[ppor] {lines[-1]}")


@pytest.mark.parametrize("config", ["config1", "config2"])
def test_gpusphere_sidecar_cylinder_and_bimodal(dem, gpu_ok, tmp_path, config):
    """The driver on the shipped configuration-1 (cylinder, lognormal, random insertion) and
    configuration-2 (bimodal, random insertion capped at the container height) decks and
    side-cars, with the particle count cut to 20,000 by one more side-car line (and the
    configuration-2 box narrowed to 20 d to keep a bed several diameters deep): 1,000 steps,
    all particles in from the first step, the radii of the distribution in particle.dat, a
    packing fraction from each sphere's own volume in ppor.dat."""
    import os
    import shutil
    import subprocess
    shutil.copy(os.path.join(dem.DECKS, config + ".inp"), tmp_path / "deck.inp")
    side = open(os.path.join(dem.DECKS, config + ".side")).read() + "particles 20000\n"
    if config == "config2":   # 20,000 spheres would fill the 120 d box to under one diameter
        side += "walls box 20 20 200\n"
    (tmp_path / "deck.side").write_text(side)
    exe = os.path.join(dem.HERE, "gpusphere")
    r = subprocess.run([exe, "deck.inp", "--side", "deck.side", "--tstop", "1000", "--quiet"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    ppor = (tmp_path / "ppor.dat").read_text().split()
    it, n, half_contacts = int(ppor[0]), int(ppor[1]), int(ppor[2])
    phi, maxhz = float(ppor[4]), float(ppor[6])
    assert (it, n) == (1000, 20000) and half_contacts >= 0
    assert 0.05 < phi < 0.7 and maxhz > 5.0
    zone = (tmp_path / "particle.dat").read_text().splitlines()
    assert zone[0] == 'ZONE T="ZONE001"' and len(zone) == 1 + 20000
    r_mm = np.array([float(ln.split()[3]) for ln in zone[1:]]) * 1e3   # radius in mm
    if config == "config1":
        assert 0.35 <= r_mm.min() and r_mm.max() <= 0.65 and r_mm.std() > 0.02
    else:
        assert set(np.round(r_mm, 6)) == {0.25, 0.5}


def test_autotune_then_graph(dem, gpu_ok):
    """The autotuner's picks are from its candidate sets, and a graph run at the chosen fixed
    rebuild interval never raises the displacement trigger (no particle outran the skin)."""
    ws = [dense_workload(dem, n=n, seed=60 + i, box=(bx, bx, 60.0)) for i, (n, bx) in
          enumerate([(1200, 11.0), (2600, 15.0)])]
    insts = [w.instances[0] for w in ws]
    cat = lambda k: np.concatenate([getattr(w, k) for w in ws])
    big = dem.Context(insts)
    big.set_state(cat("pos"), cat("pdis"), cat("angv"), cat("rad"), cat("rmass"), cat("inert"))
    r = big.autotune(24)
    say(f"[tune] {r}")
    assert r["skin"] in (0.3, 0.45, 0.55, 0.7, 0.9)
    assert r["pair_block"] in (64, 128, 256) and r["particle_block"] in (64, 128, 256)
    assert r["pair_grid_per_cu"] in (4, 8, 16, 32) and r["interval"] >= 1
    K = min(int(r["interval"]), 50)
    big.set_option(dem.OPT_REBUILD_INTERVAL, K)
    big.step_graph(4 * K)
    assert big.stats().last_rebuild_flag == 0
    st = big.state()
    assert np.all(np.isfinite(st["pos"]))
