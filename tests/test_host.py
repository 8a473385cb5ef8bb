"""Host-side logic of libdem.so (no GPU calls): deck, restart file, generators, exported ABI."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(dem):
    L = ctypes.CDLL(dem.LIB_PATH)
    hdr = open(os.path.join(ROOT, "include", "dem.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?!typedef|enum|struct|#)[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(",
                         hdr, flags=re.M):
        names.add(m.group(1))
    names -= {"if", "while", "for", "sizeof"}
    assert len(names) >= 40
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(dem.EXPORTS) <= names


def test_deck_shipped_layout(dem):
    """hopp3d.inp as shipped, read positionally as dempacking.cu:112-125."""
    d = dem.read_deck(os.path.join(GOLDEN, "hopp3d.inp"))
    assert d.iforcemode == 1
    assert (d.emod, d.vpois, d.denp, d.diam) == (1e7, 0.29, 2500.0, 0.01)
    assert (d.ntot, d.dt, d.tstop) == (10000, 0.001, 40000)
    assert (d.dmt1, d.dmt2, d.thick, d.zht1, d.zht2) == (0.5, 0.5, 0.5, 0.0, 0.6)
    assert (d.vel0, d.nrings, d.ufricp, d.ufricw) == (4.0, 9, 0.4, 0.4)
    assert (d.dampnp, d.damptp, d.dampnw, d.damptw, d.cdamp) == (0.3, 0.3, 0.3, 0.3, 0.3)
    assert (d.stiffness, d.dampcoeffn, d.dampcoefft, d.upp, d.upw) == (10000.0, 100.0, 100.0, 0.0, 0.0)


def test_deck_without_iforcemode_line(dem, tmp_path):
    """A deck lacking the leading iforcemode line (26 values) is accepted with mode 0."""
    txt = open(os.path.join(GOLDEN, "hopp3d.inp")).read().split("\n", 1)[1]
    p = tmp_path / "hopp3d.inp"
    p.write_text(txt)
    d = dem.read_deck(str(p))
    assert d.iforcemode == 0 and d.ntot == 10000 and d.emod == 1e7


def test_deck_truncated_is_an_error(dem, tmp_path):
    p = tmp_path / "hopp3d.inp"
    p.write_text("1\n10000000 0.29 2500.0\n0.01\n")
    with pytest.raises(dem.DemError) as ei:
        dem.read_deck(str(p))
    assert ei.value.code == dem.DEM_ERR_IO


def test_deck_missing_file(dem):
    with pytest.raises(dem.DemError) as ei:
        dem.read_deck("/nonexistent/hopp3d.inp")
    assert ei.value.code == dem.DEM_ERR_IO


def test_restart_file_bytes(dem, tmp_path):
    p = tmp_path / "restart.dat"
    dem.write_restart(str(p))
    assert p.read_bytes() == open(os.path.join(GOLDEN, "restart.dat"), "rb").read()
    assert len(p.read_bytes()) == 40


def test_restart_reader_end_of_run_file_means_fresh_start(dem):
    """The shipped end-of-run restart.dat does not start with three integers: the reference's
    fscanf("%d %d %d") (dempacking.cu:365) matches nothing, so no resume."""
    assert dem.read_restart(os.path.join(GOLDEN, "restart.dat")) is None


def test_restart_resume_round_trip(dem, tmp_path):
    p = tmp_path / "restart.dat"
    dem.write_restart_resume(str(p), 0, 7, 12)
    assert p.read_bytes() == b"       0        7       12"   # "%8d %8d %8d"
    assert dem.read_restart(str(p)) == (0, 7, 12)
    p.write_text("1 2 3\n")
    assert dem.read_restart(str(p)) == (1, 2, 3)
    with pytest.raises(dem.DemError) as ei:
        dem.read_restart(str(tmp_path / "missing.dat"))
    assert ei.value.code == dem.DEM_ERR_IO


def test_lognormal_diameters(dem):
    d = dem.gen_diameters(1, 200000, 0.1, 0.7, 1.3, seed=2)
    assert d.min() >= 0.7 and d.max() <= 1.3
    assert d.mean() == pytest.approx(1.0, abs=3e-3)
    assert np.log(d).std() == pytest.approx(0.1, rel=0.02)
    assert np.array_equal(d, dem.gen_diameters(1, 200000, 0.1, 0.7, 1.3, seed=2))


def test_bimodal_volume_fraction(dem):
    d = dem.gen_diameters(2, 100000, 1.0, 0.5, 0.3, seed=3)
    v = d ** 3
    assert v[d == 1.0].sum() / v.sum() == pytest.approx(0.3, abs=2e-3)
    assert set(np.unique(d)) == {0.5, 1.0}


@pytest.mark.parametrize("kind", [0, 1])
def test_insertion_is_non_overlapping_and_inside(dem, kind):
    n = 3000
    diam = dem.gen_diameters(1, n, 0.1, 0.7, 1.3, seed=5)
    if kind == 0:
        walls = dem.make_walls(dem.WALLS_BOX, (12.0, 10.0, 0.0))
    else:
        walls = dem.make_walls(dem.WALLS_CYLINDER, (0, 0, 0), radius=6.0)
    pos = dem.gen_insert(walls, 0.0, 60.0, diam, seed=7)
    r = diam / 2
    for i in range(n):
        d = np.linalg.norm(pos[i + 1:] - pos[i], axis=1)
        assert np.all(d >= r[i] + r[i + 1:] - 1e-12)
    if kind == 0:
        assert np.all(np.abs(pos[:, 0]) <= 6.0 - r + 1e-12) and np.all(np.abs(pos[:, 1]) <= 5.0 - r + 1e-12)
    else:
        assert np.all(np.hypot(pos[:, 0], pos[:, 1]) <= 6.0 - r + 1e-12)
    assert np.all(pos[:, 2] >= r - 1e-12) and np.all(pos[:, 2] <= 60.0 - r + 1e-12)


def test_grid_pattern(dem):
    """dempacking.cu:329-344: origin (-W/2, -T/2, 0), gridSize = (int)(W / cell) + 1."""
    w = dem.make_walls(dem.WALLS_BOX, (50.0, 50.0, 60.0))
    inst = dem.make_instance(10, w, 60.0, 1.55, dem.Material(), 1e-3)
    assert list(inst.gsize) == [33, 33, 39]
    assert list(inst.origin) == [-25.0, -25.0, 0.0]


def test_config0_workload(dem):
    w = dem.config0()
    assert w.n == 2000 and w.rebuild_interval == 20
    assert w.instances[0].dt == pytest.approx(9.9045e-5, rel=1e-4)
    p = w.pos
    for i in range(0, w.n, 7):
        d = np.linalg.norm(p - p[i], axis=1)
        d[i] = 9
        assert d.min() >= 1.0


def test_header_constants_match_the_binding(dem):
    """Option numbers and status codes of include/dem.h are the ones the Python binding uses."""
    hdr = open(os.path.join(ROOT, "include", "dem.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    opts = dict((m.group(1), int(m.group(2)))
                for m in re.finditer(r"\bDEM_OPT_([A-Z_]+)\s*=\s*(\d+)", hdr))
    assert len(opts) >= 13
    for name, val in opts.items():
        assert getattr(dem, "OPT_" + name) == val, name
    errs = dict((m.group(1), int(m.group(2)))
                for m in re.finditer(r"\bDEM_(OK|ERR_[A-Z_]+)\s*=\s*(\d+)", hdr))
    assert errs["OK"] == 0 and len(errs) >= 6
    for name, val in errs.items():
        assert getattr(dem, "DEM_" + name) == val, name


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message before touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_sidecar_defaults_and_shipped_configs(dem, tmp_path):
    """The deck side-car: an empty file keeps the reference's run; the shipped side-cars of
    configurations 0-2 parse to their documented settings."""
    p = tmp_path / "empty.side"
    p.write_text("# nothing but a comment\n\n")
    sc = dem.read_sidecar(str(p))
    assert (sc.walls, sc.particles, sc.height, sc.diameters, sc.insertion) == (-1, 0, 0.0, 0, 0)
    assert (sc.dpar[0], sc.seed, sc.restitution, sc.rebuild_interval) == (1.0, 1, -1.0, 0)
    c0 = dem.read_sidecar(os.path.join(dem.DECKS, "config0.side"))
    assert c0.insertion == 1 and list(c0.ipar) == [1.2, 0.05, 1.0]
    assert (c0.seed, c0.restitution, c0.rebuild_interval, c0.walls) == (1, 0.6, 20, -1)
    c1 = dem.read_sidecar(os.path.join(dem.DECKS, "config1.side"))
    assert (c1.walls, c1.radius, c1.diameters, list(c1.dpar)) == (
        dem.WALLS_CYLINDER, 40.0, 1, [0.1, 0.7, 1.3])
    assert (c1.insertion, c1.ipar[0], c1.seed, c1.rebuild_interval) == (2, 0.25, 2, 0)
    c2 = dem.read_sidecar(os.path.join(dem.DECKS, "config2.side"))
    assert (c2.diameters, list(c2.dpar), c2.insertion, c2.ipar[0], c2.height) == (
        2, [1.0, 0.5, 0.3], 2, 0.3, 200.0)
    for k in range(3):
        d = dem.read_deck(os.path.join(dem.DECKS, f"config{k}.inp"))
        assert d.ntot == (2000, 500000, 2000000)[k] and d.emod == 1e7 and d.diam == 0.001
    # configuration 0's deck gives configuration 0's material and time step bit for bit
    d0 = dem.read_deck(os.path.join(dem.DECKS, "config0.inp"))
    assert d0.dt == dem.reduced_dt(dem.DT0, 1e-3) and d0.upp / d0.diam == dem.MAT0["mu_r"]


def test_sidecar_rejects_unknown_keys_and_bad_values(dem, tmp_path):
    bad = ["walls sphere 3", "walls box 1 2", "walls cylinder -1", "height 0", "particles 2.5",
           "diameters lognormal 0.1 1.3 0.7", "diameters bimodal 1 0.5 1.5", "insertion random 0.7",
           "insertion lattice 0 0 0", "seed -1", "restitution 1.5", "rebuild interval 0",
           "rebuild sometimes", "colour blue", "height 10 20", "walls"]
    for i, line in enumerate(bad):
        p = tmp_path / f"bad{i}.side"
        p.write_text(line + "\n")
        with pytest.raises(dem.DemError) as ei:
            dem.read_sidecar(str(p))
        assert ei.value.code == dem.DEM_ERR_IO, line
    with pytest.raises(dem.DemError) as ei:
        dem.read_sidecar(str(tmp_path / "missing.side"))
    assert ei.value.code == dem.DEM_ERR_IO


def test_sidecar_rejects_lognormal_windows_without_mass(dem, tmp_path):
    """A clip window the lognormal cannot draw from (sigma 0 with 1 outside the window, or a
    window far out in the tail) is refused by the parser, and the generator refuses it too
    instead of redrawing forever."""
    for i, line in enumerate(["diameters lognormal 0 1.1 1.3", "diameters lognormal 0.1 1.5 2.0",
                              "diameters lognormal 0.1 1.2 1.2"]):
        p = tmp_path / f"w{i}.side"
        p.write_text(line + "\n")
        with pytest.raises(dem.DemError) as ei:
            dem.read_sidecar(str(p))
        assert ei.value.code == dem.DEM_ERR_IO, line
    # windows that do hold mass still parse (sigma 0 around 1, and the shipped window)
    for i, line in enumerate(["diameters lognormal 0 0.9 1.1", "diameters lognormal 0.1 0.7 1.3"]):
        p = tmp_path / f"ok{i}.side"
        p.write_text(line + "\n")
        assert dem.read_sidecar(str(p)).diameters == 1
    for args in ((0.0, 1.1, 1.3), (0.1, 1.2, 1.2)):
        with pytest.raises(dem.DemError) as ei:
            dem.gen_diameters(1, 10, *args)
        assert ei.value.code == dem.DEM_ERR_INVALID
    d = dem.gen_diameters(1, 1000, 0.0, 0.9, 1.1)
    assert np.all(d == 1.0)


def test_generators_take_any_diameter_array(dem):
    """Float32 and strided diameter arrays give the same radii, masses and insertion as the
    contiguous float64 ones (the binding converts before passing a pointer)."""
    d64 = dem.gen_diameters(1, 400, 0.1, 0.7, 1.3, seed=9).astype(np.float32).astype(np.float64)
    d32 = d64.astype(np.float32)
    strided = np.repeat(d64, 2)[::2]
    ref = dem._props(d64)
    for d in (d32, strided):
        got = dem._props(d)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    walls = dem.make_walls(dem.WALLS_BOX, (12.0, 12.0, 40.0))
    p_ref = dem.gen_insert(walls, 0.0, 20.0, d64, 3)
    assert np.array_equal(dem.gen_insert(walls, 0.0, 20.0, d32, 3), p_ref)
    assert np.array_equal(dem.gen_insert(walls, 0.0, 20.0, strided, 3), p_ref)


@pytest.mark.parametrize("box", ["20 3.9 30", "20 3 30", "3.99 20 30"])
def test_driver_refuses_layer_insertion_in_a_narrow_box(dem, tmp_path, box):
    """The reference's layer lattice has (unsigned)(w) - 3 rings per side: a side-car box
    under 4 d on either side is refused before any device work, instead of wrapping the ring
    count or dividing by zero."""
    import shutil
    import subprocess
    exe = os.path.join(dem.HERE, "gpusphere")
    assert os.path.exists(exe), "build the driver first"
    shutil.copy(os.path.join(GOLDEN, "hopp3d.inp"), tmp_path / "hopp3d.inp")
    side = tmp_path / "narrow.side"
    side.write_text(f"walls box {box}\nparticles 50\n")
    r = subprocess.run([exe, "--side", str(side)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0
    assert "at least 4 d" in (r.stdout + r.stderr)
