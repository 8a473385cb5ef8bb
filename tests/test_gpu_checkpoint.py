"""Full-state checkpoints (dem_save_checkpoint / dem_load_checkpoint) and the gpusphere
driver's resume through restart.dat and preflow.bin (the reference's unfinished restart,
dempacking.cu:358-367, 794-875).  The bar is the reproducibility the reference's resume was
meant to give: the steps run after a load are the uninterrupted run's steps, so states, pair
lists, contact and wall histories and counters must be identical bit for bit."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compressed_lattice(dem, n=800, seed=5, interval=0, vscale=3e-5):
    """Compressed jittered lattice with random displacements (scale `vscale` per step) and
    spins: pair, wall and rolling contacts and list rebuilds from the first steps."""
    w = dem.config0(n=n, seed=seed)
    w.pos[:, 2] = 0.5 + (w.pos[:, 2] - w.pos[:, 2].min()) * (0.997 / 1.2)
    w.pos[:, :2] *= 0.997 / 1.2
    rng = np.random.default_rng(seed)
    w.pdis[:] = rng.normal(scale=vscale, size=w.pdis.shape)
    w.angv[:] = rng.normal(scale=3e-5, size=w.angv.shape)
    w.rebuild_interval = interval
    return w


def snapshot(ctx):
    st = ctx.state()
    ids, hist, _ = ctx.pairs()
    dw, fw = ctx.wall_history()
    s = ctx.stats()
    return dict(st, ids=ids, hist=hist, disptw=dw, fricpw=fw, contacts=ctx.contacts(),
                steps=s.steps, rebuilds=s.rebuilds, npairs=s.npairs, flag=s.last_rebuild_flag)


def assert_same(a, b):
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


def advance(ctx, k, graph):
    if graph:
        ctx.step_graph(k)
    else:
        ctx.step(k)


@pytest.mark.parametrize("mode", ["trigger", "interval", "graph"])
def test_resume_repeats_the_run(dem, gpu_ok, tmp_path, mode):
    """Save in the middle of a list's life, continue; a fresh context that loads the file and
    runs the same steps ends in the same state bit for bit (and saving leaves the run as it was)."""
    say(f"[ckpt] mode={mode}")
    interval = {"trigger": 0, "interval": 20, "graph": 10}[mode]
    graph = mode == "graph"
    k1, k2 = (130, 160) if graph else (137, 163)
    # under the displacement trigger, fast enough that lists are rebuilt after the save too
    w = compressed_lattice(dem, interval=interval, vscale=1e-3 if mode == "trigger" else 3e-5)
    path = str(tmp_path / "state.bin")
    a = dem.Context(w.instances)
    a.load(w)
    advance(a, k1, graph)
    a.save_checkpoint(path)
    r1 = a.stats().rebuilds
    advance(a, k2, graph)
    ref = snapshot(a)
    a.close()
    say(f"[ckpt] rebuilds {r1} at the save, {ref['rebuilds']} at the end")
    assert ref["contacts"].sum() > 0 and ref["rebuilds"] > r1 >= 2
    b = dem.Context(w.instances)
    b.load(w)                 # another state first: the load must replace all of it
    b.step(3)
    b.load_checkpoint(path)
    assert b.stats().steps == k1
    advance(b, k2, graph)
    assert_same(ref, snapshot(b))
    # and the uninterrupted run itself, without any save
    c = dem.Context(w.instances)
    c.load(w)
    advance(c, k1 + k2, graph)
    assert_same(ref, snapshot(c))


def test_checkpoint_before_any_step_into_a_fresh_context(dem, gpu_ok, tmp_path):
    """A checkpoint taken before the first step (no list yet) loaded into a context that was
    never given a state: the file alone initialises it, and the run that follows is the
    original's."""
    w = compressed_lattice(dem, n=700, seed=8)
    path = str(tmp_path / "start.bin")
    a = dem.Context(w.instances)
    a.load(w)
    a.save_checkpoint(path)
    a.step(120)
    ref = snapshot(a)
    b = dem.Context(w.instances)
    b.load_checkpoint(path)
    assert b.stats().steps == 0
    b.step(120)
    assert_same(ref, snapshot(b))


def test_resume_with_waiting_layers(dem, gpu_ok, tmp_path):
    """A checkpoint taken while layers still wait (dem_set_active) keeps their state; the
    resumed run inserts them on the same schedule and ends as the uninterrupted one."""
    walls = dem.make_walls(dem.WALLS_BOX, (10.0, 10.0, 40.0))
    per, nl = 64, 5
    n = per * nl
    rng = np.random.default_rng(9)
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
    path = str(tmp_path / "layers.bin")

    def run(ctx, first, last, active, save_at=None):
        for it in range(first, last + 1):
            if active < n and it % 400 == 1:
                active = min(n, active + per)
                ctx.set_active(active)
            ctx.step(1)
            if it == save_at:
                ctx.save_checkpoint(path)
        return active

    a = dem.Context([inst])
    a.set_state(pos, pdis, np.zeros((n, 3)), rad, rmass, inert)
    run(a, 1, 2000, 0, save_at=1000)
    ref = snapshot(a)
    assert ref["contacts"].sum() > 0 and ref["rebuilds"] >= 5
    b = dem.Context([inst])
    b.set_state(pos, pdis, np.zeros((n, 3)), rad, rmass, inert)
    b.load_checkpoint(path)
    assert b.active_count() == 3 * per
    run(b, 1001, 2000, b.active_count())
    assert_same(ref, snapshot(b))


def test_resume_batched_keeps_materials(dem, gpu_ok, tmp_path):
    """Batched context whose instance 1 got other materials (dem_set_material): the file
    carries the instance table, so a fresh context resumes with them."""
    ws = [compressed_lattice(dem, n=n, seed=s, interval=0) for n, s in ((500, 1), (700, 2))]
    insts = [x.instances[0] for x in ws]
    cat = lambda k: np.concatenate([getattr(x, k) for x in ws])
    m = dem.reduced_material(E=3e6, nu=0.29, rho=2500.0, diam=1e-3, mu=0.55, mu_w=0.2, e=0.4,
                             mu_r=0.01)
    path = str(tmp_path / "batched.bin")
    a = dem.Context(insts)
    a.set_material(1, m)
    a.set_state(cat("pos"), cat("pdis"), cat("angv"), cat("rad"), cat("rmass"), cat("inert"))
    a.step(90)
    a.save_checkpoint(path)
    a.step(110)
    ref = snapshot(a)
    b = dem.Context(insts)
    b.set_state(cat("pos"), cat("pdis"), cat("angv"), cat("rad"), cat("rmass"), cat("inert"))
    b.load_checkpoint(path)
    b.step(110)
    assert_same(ref, snapshot(b))


def test_damaged_or_foreign_checkpoints_are_refused(dem, gpu_ok, tmp_path):
    """Truncated or corrupted files give DEM_ERR_IO and a context of other sizes
    DEM_ERR_INVALID, before anything is changed: the context still loads a good file."""
    w = compressed_lattice(dem, n=600)
    good = tmp_path / "good.bin"
    a = dem.Context(w.instances)
    a.load(w)
    a.step(40)
    a.save_checkpoint(str(good))
    a.step(25)
    ref = snapshot(a)
    raw = good.read_bytes()
    (tmp_path / "cut.bin").write_bytes(raw[: len(raw) // 2])
    flipped = bytearray(raw)
    flipped[len(raw) // 3] ^= 0x10
    (tmp_path / "flip.bin").write_bytes(bytes(flipped))
    b = dem.Context(w.instances)
    b.load(w)
    for name, code in (("cut.bin", dem.DEM_ERR_IO), ("flip.bin", dem.DEM_ERR_IO),
                       ("missing.bin", dem.DEM_ERR_IO)):
        with pytest.raises(dem.DemError) as ei:
            b.load_checkpoint(str(tmp_path / name))
        assert ei.value.code == code, name
    other = compressed_lattice(dem, n=640)
    c = dem.Context(other.instances)
    c.load(other)
    with pytest.raises(dem.DemError) as ei:
        c.load_checkpoint(str(good))
    assert ei.value.code == dem.DEM_ERR_INVALID
    b.load_checkpoint(str(good))
    b.step(25)
    assert_same(ref, snapshot(b))


def test_gpusphere_resume_through_restart_dat(dem, gpu_ok, tmp_path):
    """gpusphere on the shipped deck, 2,500 steps: once uninterrupted, once stopped by the
    CPU-time limit at step 1000 (checkpoint, restart.dat "0 0 1", exit status 3) and run again,
    which resumes at step 1001.  ppor.dat and particle.dat of the two are identical."""
    golden = os.path.join(os.path.dirname(__file__), "golden")
    exe = os.path.join(dem.HERE, "gpusphere")
    full, split = tmp_path / "full", tmp_path / "split"
    for d in (full, split):
        d.mkdir()
        shutil.copy(os.path.join(golden, "hopp3d.inp"), d / "hopp3d.inp")
    r = subprocess.run([exe, "--tstop", "2500", "--quiet"], cwd=full, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    say("[ckpt] driver: uninterrupted run done")
    r = subprocess.run([exe, "--tstop", "2500", "--quiet", "--max-cpu-seconds", "0"], cwd=split,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert (split / "restart.dat").read_bytes() == b"       0        0        1"
    pre = (split / "preflow.dat").read_text().splitlines()
    assert pre[0] == "    1000,    4418,       0" and len(pre) == 1 + 4418
    assert len(pre[1].split()) == 9
    assert dem.read_restart(str(split / "restart.dat")) == (0, 0, 1)
    say("[ckpt] driver: stopped after step 1000")
    r = subprocess.run([exe, "--tstop", "2500", "--quiet"], cwd=split, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    assert "resuming at step 1001 with 4418 particles" in r.stdout
    for f in ("ppor.dat", "particle.dat", "restart.dat", "jobfinished"):
        assert (split / f).read_bytes() == (full / f).read_bytes(), f
    assert (full / "restart.dat").read_bytes() == open(os.path.join(golden, "restart.dat"), "rb").read()
    # periodic checkpoints (--checkpoint-every) leave the run unchanged; the last one is of
    # step 2100 (all 10,000 particles in), and the finished run's restart.dat is the end-of-run
    # text again
    every = tmp_path / "every"
    every.mkdir()
    shutil.copy(os.path.join(golden, "hopp3d.inp"), every / "hopp3d.inp")
    r = subprocess.run([exe, "--tstop", "2500", "--quiet", "--checkpoint-every", "700"], cwd=every,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    for f in ("ppor.dat", "particle.dat", "restart.dat", "jobfinished"):
        assert (every / f).read_bytes() == (full / f).read_bytes(), f
    assert (every / "preflow.dat").read_text().startswith("    2100,   10000,       0\n")
    assert (every / "preflow.bin").stat().st_size > 0
