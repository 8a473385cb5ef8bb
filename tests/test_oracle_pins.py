"""CPU pins of the oracle functions the closed-form law tests leave open: the per-particle
force sum (A10, its partner branch -F / torqpairj in orc_sum_forces, reference
particleSystem.cu:243-277), the side walls and the cylinder of the wall law (calforceiw,
called at neiglist.cu:287-288), the Mindlin partial-slip regime and the history rotation
(vprojec, neiglist.cu:133-146) of the pair law.  Every expected value is computed here in
Python from the material constants, independently of the oracle's code.  No GPU."""
import math

import numpy as np
import pytest


def _mat(dem, **kw):
    m = dict(dem.MAT0)
    m.update(kw)
    return dem.reduced_material(**m)


def _far_walls_instance(dem, n, mat, dt=1e-4, box=(40.0, 40.0, 40.0), cell=None):
    walls = dem.make_walls(dem.WALLS_BOX, box)
    return dem.make_instance(n, walls, box[2], cell or (1.3 + dem.SKIN), mat, dt)


# ---------------------------------------------------------------------------------------
# A10: sum of pair forces
# ---------------------------------------------------------------------------------------

def _compressed_cluster(n_side, spacing, seed):
    """A cubic lattice of polydisperse spheres closer than their diameters (every lattice
    neighbour overlaps), centred at (0, 0, 20): far from every wall of a 40^3 box."""
    rng = np.random.default_rng(seed)
    g = np.arange(n_side) * spacing
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = pos - pos.mean(axis=0) + np.array([0.0, 0.0, 20.0])
    pos += rng.uniform(-0.01, 0.01, pos.shape)
    n = pos.shape[0]
    rad = 0.5 * rng.uniform(0.95, 1.05, n)
    rmass = 8 * rad ** 3
    inert = 0.4 * rmass * rad ** 2
    pdis = rng.normal(scale=2e-5, size=(n, 3))
    angv = rng.normal(scale=5e-5, size=(n, 3))
    return pos, pdis, angv, rad, rmass, inert


def _pair_sum_residual(sim, rmass):
    """|sum_i F_i + sum_i m_i z| relative to max_i |F_i| (walls far away: the total force
    is the pair sum plus gravity -m_i z)."""
    f = sim.field("forcei")
    tot = f.sum(axis=0) + np.array([0.0, 0.0, rmass.sum()])
    return float(np.max(np.abs(tot))) / float(np.max(np.linalg.norm(f, axis=1)))


def test_pair_forces_sum_to_zero(dem, orc):
    """Newton's third law through orc_sum_forces: with no wall contact the total force over
    all particles is gravity alone, to rounding, over steps with rebuilds and histories."""
    pos, pdis, angv, rad, rmass, inert = _compressed_cluster(8, 0.97, seed=3)
    inst = _far_walls_instance(dem, rad.shape[0], _mat(dem, mu_r=0.01))
    sim = orc.OracleSim(inst, pos, pdis, angv, rad, rmass, inert)
    for s in range(12):
        sim.set_iall(1 if s % 4 == 0 else 0)
        assert sim.step(1) == 0
        assert sim.scalars()["pcntot"] > 4 * rad.shape[0]     # a compressed, connected bed
        assert not sim.field("fricpw").any()                   # no wall was touched
        assert _pair_sum_residual(sim, rmass) < 1e-12


def test_pair_sum_check_catches_a_sign_slip(dem, orc):
    """The residual of the test above is sensitive: summing the oracle's own per-pair forces
    with +F on the second member as well (the slip the partner branch could make) leaves a
    residual of the order of the forces themselves."""
    pos, pdis, angv, rad, rmass, inert = _compressed_cluster(5, 0.97, seed=4)
    inst = _far_walls_instance(dem, rad.shape[0], _mat(dem))
    sim = orc.OracleSim(inst, pos, pdis, angv, rad, rmass, inert)
    assert sim.step(1) == 0
    npairs = sim.scalars()["npairs"]
    lp, ln = sim.field("lp")[:npairs], sim.field("ln")[:npairs]
    fp = sim.field("forcepair")[:npairs]
    good = np.zeros((rad.shape[0], 3))
    bad = np.zeros((rad.shape[0], 3))
    np.add.at(good, lp, fp)
    np.add.at(good, ln, -fp)
    np.add.at(bad, lp, fp)
    np.add.at(bad, ln, fp)
    scale = np.max(np.abs(fp))
    assert np.max(np.abs(good.sum(axis=0))) < 1e-12 * scale * npairs
    assert np.max(np.abs(bad.sum(axis=0))) > 1e-3 * scale
    # and the oracle's own sum equals the correct one per particle (same order, so bitwise
    # only up to association: compare to rounding)
    f = sim.field("forcei") + np.stack([np.zeros_like(rmass), np.zeros_like(rmass), rmass], 1)
    assert np.allclose(f, good, rtol=0, atol=1e-12 * scale)


def _mindlin_tangential(mat, dn, rs, ms, ht, vt, dt):
    """Tangential force of the pair law for a history ht (already in the tangent plane and
    including this step's slip vt): Mindlin partial slip mu Fcn (1 - (1 - xi)^1.5) along ht,
    dashpot gt sqrt(6 mu m* Fcn sqrt(1 - xi) / dt_max) vt / dt, Coulomb cap mu Fcn."""
    p = mat.p
    fcn = p.estarp * math.sqrt(rs) * dn ** 1.5
    dtmax = p.ufricp * (2.0 - p.vpois) / (2.0 * (1.0 - p.vpois)) * dn
    htm = np.linalg.norm(ht)
    if htm > dtmax:
        ht = ht * (dtmax / htm)
        htm = dtmax
    xi = htm / dtmax
    fct = p.ufricp * fcn * (1.0 - (1.0 - xi) ** 1.5)
    ct = p.damptp * math.sqrt(6.0 * p.ufricp * ms * fcn * math.sqrt(1.0 - xi) / dtmax)
    ft = (fct / htm if htm > 0 else 0.0) * ht + ct / dt * vt
    ftm = np.linalg.norm(ft)
    if ftm > p.ufricp * fcn:
        ft = ft * (p.ufricp * fcn / ftm)
    return fcn, ft, ht


@pytest.mark.parametrize("lower_on_top", [False, True])
def test_two_sphere_forces_and_torques_closed_form(dem, orc, lower_on_top):
    """Two unequal spheres in contact, one spinning about a horizontal axis: the oracle's
    per-particle totals (pair sum, then gravity) equal the closed form on both members --
    F on the first, -F on the second; torque r_i n x Ft + Mr on the first and
    r_j n x Ft - Mr on the second (n from first to second), Mr = -mu_r Fcn w_rel/|w_rel|.
    Both orientations, so a sign or a torqpairi / torqpairj slip in the partner branch
    shows on one of them."""
    mu_r = 0.01
    mat = _mat(dem, mu_r=mu_r)
    dt = 1e-4
    r = np.array([0.6, 0.4])
    dn = 2e-3
    d = r.sum() - dn
    spin = 4e-5   # per-step rotation of the spinning sphere about +x
    if lower_on_top:   # particle 0 (the pair's first member) above particle 1, 1 spins
        pos = np.array([[0.0, 0.0, 20.0 + d], [0.0, 0.0, 20.0]])
        angv = np.array([[0.0, 0.0, 0.0], [spin, 0.0, 0.0]])
    else:              # particle 0 below, 0 spins
        pos = np.array([[0.0, 0.0, 20.0], [0.0, 0.0, 20.0 + d]])
        angv = np.array([[spin, 0.0, 0.0], [0.0, 0.0, 0.0]])
    rmass = 8 * r ** 3
    inert = 0.4 * rmass * r ** 2
    inst = _far_walls_instance(dem, 2, mat, dt)
    sim = orc.OracleSim(inst, pos, np.zeros((2, 3)), angv, r, rmass, inert)
    assert sim.step(1) == 0
    assert sim.scalars()["npairs"] == 1 and list(sim.field("contacti")) == [1, 1]

    # closed form, first member a = particle 0, second b = particle 1
    ra, rb = r
    n = (pos[1] - pos[0]) / d
    rs = ra * rb / (ra + rb)
    ms = rmass[0] * rmass[1] / (rmass[0] + rmass[1])
    # slip of b's contact point relative to a's: -(ra wa + rb wb) x n (no translation)
    vt = -np.cross(ra * angv[0] + rb * angv[1], n)
    fcn, ft, ht = _mindlin_tangential(mat, dn, rs, ms, vt, vt, dt)
    assert 0.0 < np.linalg.norm(ft) < mat.p.ufricp * fcn      # partial slip, below the cap
    wrel = angv[0] - angv[1]
    mr = -mu_r * fcn * wrel / np.linalg.norm(wrel)
    F = -fcn * n + ft
    Ta = np.cross(ra * n, ft) + mr
    Tb = np.cross(rb * n, ft) - mr
    g0 = np.array([0.0, 0.0, -rmass[0]])
    g1 = np.array([0.0, 0.0, -rmass[1]])
    f = sim.field("forcei")
    t = sim.field("torquei")
    scale = np.linalg.norm(F)
    assert np.allclose(f[0], F + g0, rtol=0, atol=1e-12 * scale)
    assert np.allclose(f[1], -F + g1, rtol=0, atol=1e-12 * scale)
    tscale = np.linalg.norm(Ta) + np.linalg.norm(Tb)
    assert np.allclose(t[0], Ta, rtol=0, atol=1e-12 * tscale)
    assert np.allclose(t[1], Tb, rtol=0, atol=1e-12 * tscale)
    # the friction and rolling torques oppose the spin on the spinning sphere
    k = 1 if lower_on_top else 0
    assert t[k][0] * angv[k][0] < 0.0
    # the pair's stored history is the slip of this step; |Ft| is fricp
    assert np.allclose(sim.field("dispt")[0], ht, rtol=0, atol=1e-15)
    assert sim.field("fricp")[0] == pytest.approx(np.linalg.norm(ft), rel=1e-12)


# ---------------------------------------------------------------------------------------
# A9: Mindlin partial slip and the history rotation
# ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("hmag_frac", [0.05, 0.4, 0.9])
@pytest.mark.parametrize("dn", [5e-4, 6e-3])
def test_mindlin_partial_slip(dem, orc, dn, hmag_frac):
    """Zero relative velocity, tangential history h in the tangent plane with |h| below
    dt_max: |Ft| = mu Fcn (1 - (1 - xi)^1.5), xi = |h| / dt_max, along +h (the history is the
    second body's displacement relative to the first, which drags the first along).  The
    large overlap puts |h| above 0.001, where vprojec rescales by sqrt(pmag / rmag) = 1."""
    mat = _mat(dem, mu=0.4)
    p = mat.p
    rng = np.random.default_rng(int(dn * 1e6) + int(hmag_frac * 100))
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    t1 = np.cross(u, [0.3, -0.5, 0.8])
    t1 /= np.linalg.norm(t1)
    dtmax = p.ufricp * (2.0 - p.vpois) / (2.0 * (1.0 - p.vpois)) * dn
    h = hmag_frac * dtmax * t1
    ri, rj = 0.55, 0.45
    xi_ = np.array([1.0, -2.0, 3.0])
    xj_ = xi_ + (ri + rj - dn) * u
    z = np.zeros(3)
    c, fi, ti, tj, hout, fr = orc.law_pair(xi_, xj_, z, z, z, z, ri, rj, 8 * ri ** 3, 8 * rj ** 3,
                                           h, 1e-4, mat)
    assert c == 1
    rs = ri * rj / (ri + rj)
    fcn = p.estarp * math.sqrt(rs) * dn ** 1.5
    ft_expect = p.ufricp * fcn * (1.0 - (1.0 - hmag_frac) ** 1.5)
    ft = fi - np.dot(fi, u) * u
    assert np.linalg.norm(ft) == pytest.approx(ft_expect, rel=1e-10)
    assert np.allclose(ft / np.linalg.norm(ft), t1, atol=1e-10)
    assert np.dot(fi, u) == pytest.approx(-fcn, rel=1e-12)
    assert np.allclose(hout, h, rtol=0, atol=1e-12 * np.linalg.norm(h))
    assert fr == pytest.approx(ft_expect, rel=1e-10)
    # spring torques: ri n x Ft on the first, rj n x Ft on the second
    assert np.allclose(ti, np.cross(ri * u, ft), atol=1e-12 * np.linalg.norm(ti))
    assert np.allclose(tj, np.cross(rj * u, ft), atol=1e-12 * np.linalg.norm(tj))


def test_mindlin_partial_slip_with_dashpot(dem, orc):
    """History plus a slip this step, both in the tangent plane: the spring acts along the
    new history, the dashpot along the slip (closed form of the Mindlin-damped model)."""
    mat = _mat(dem, mu=0.5)
    dn = 3e-3
    ri = rj = 0.5
    u = np.array([0.0, 0.0, 1.0])
    h = np.array([2e-4, 0.0, 0.0])
    vt = np.array([0.0, 1e-5, 0.0])       # slip of j relative to i this step
    z = np.zeros(3)
    c, fi, ti, tj, hout, fr = orc.law_pair(z, (1.0 - dn) * u, z, vt, z, z, ri, rj, 1.0, 1.0, h,
                                           1e-4, mat)
    fcn, ft, ht = _mindlin_tangential(mat, dn, 0.25, 0.5, h + vt, vt, 1e-4)
    assert np.allclose(hout, ht, rtol=0, atol=1e-15)
    assert np.allclose(fi, -fcn * u + ft, rtol=0, atol=1e-12 * np.linalg.norm(fi))
    assert fr == pytest.approx(np.linalg.norm(ft), rel=1e-12)


def test_history_rotation_vprojec_literal(dem, orc):
    """A history left out of the new tangent plane is projected onto it and, when
    |h| >= 0.001 and |h| <= 100 |projection|, rescaled by sqrt(|h| / |projection|) -- the
    reference's literal vprojec (neiglist.cu:140-145), magnitude sqrt(|h| |projection|)."""
    mat = _mat(dem, mu=0.5)
    dn = 6e-3
    u = np.array([0.0, 0.0, 1.0])
    z = np.zeros(3)
    for h, scaled in ((np.array([1.5e-3, 0.0, 1.0e-3]), True),    # rescaled
                      (np.array([4e-4, 0.0, 3e-4]), False),        # |h| < 0.001: projection only
                      (np.array([1e-5, 0.0, 2e-3]), False)):       # |h| > 100 |proj|: projection only
        c, fi, ti, tj, hout, fr = orc.law_pair(z, (1.0 - dn) * u, z, z, z, z, 0.5, 0.5, 1.0, 1.0,
                                               h, 1e-4, mat)
        proj = h - np.dot(h, u) * u
        if scaled:
            proj = proj * math.sqrt(np.linalg.norm(h) / np.linalg.norm(proj))
        assert np.allclose(hout, proj, rtol=0, atol=1e-15), (h, hout, proj)


# ---------------------------------------------------------------------------------------
# A11: side walls and the cylinder
# ---------------------------------------------------------------------------------------

BOX = (10.0, 8.0, 30.0)


@pytest.mark.parametrize("wall,normal", [
    ("x-", (-1.0, 0.0, 0.0)), ("x+", (1.0, 0.0, 0.0)),
    ("y-", (0.0, -1.0, 0.0)), ("y+", (0.0, 1.0, 0.0))])
def test_side_wall_hertz_force(dem, orc, wall, normal):
    """A sphere pressed into one side wall of the box (far from the floor), at rest, with
    overlap dn: F = E*_w sqrt(r) dn^1.5 along the inward normal, no torque."""
    mat = _mat(dem)
    walls = dem.make_walls(dem.WALLS_BOX, BOX)
    r, dn = 0.45, 1.5e-3
    nrm = np.array(normal)
    half = np.array([BOX[0] / 2, BOX[1] / 2, 0.0])
    x = nrm * (half - (r - dn)) + np.array([0.0, 0.0, 10.0])
    x[2] = 10.0
    z = np.zeros(3)
    hits, fw, tw, hout, fo = orc.law_wall(walls, x, z, z, r, 8 * r ** 3, z, 1e-4, mat)
    assert hits == 1
    F = mat.w.estarw * math.sqrt(r) * dn ** 1.5
    assert np.allclose(fw, -F * nrm, rtol=0, atol=1e-10 * F)
    assert not tw.any() and fo == 0.0


@pytest.mark.parametrize("theta", [0.0, 0.7, 2.5, -1.9])
def test_cylinder_side_hertz_force(dem, orc, theta):
    """A sphere pressed into the cylinder's side at angle theta: F = E*_w sqrt(r) dn^1.5
    toward the axis, horizontal."""
    mat = _mat(dem)
    R = 12.0
    walls = dem.make_walls(dem.WALLS_CYLINDER, radius=R)
    r, dn = 0.6, 8e-4
    rho = R - r + dn
    e = np.array([math.cos(theta), math.sin(theta), 0.0])
    x = rho * e + np.array([0.0, 0.0, 7.0])
    z = np.zeros(3)
    hits, fw, tw, hout, fo = orc.law_wall(walls, x, z, z, r, 8 * r ** 3, z, 1e-4, mat)
    assert hits == 1
    F = mat.w.estarw * math.sqrt(r) * dn ** 1.5
    assert np.allclose(fw, -F * e, rtol=0, atol=1e-9 * F)
    assert not tw.any()


def test_side_wall_friction_and_torque(dem, orc):
    """A sphere on the x+ wall sliding up (+z) by v per step: friction along -z with the
    Mindlin magnitude of the slip, torque r n x Ft = r |Ft| y about the centre (the contact
    point is at +r x), rolling moment zero (no spin); the history is the wall's slip -v."""
    mat = _mat(dem, mu_w=0.4)
    walls = dem.make_walls(dem.WALLS_BOX, BOX)
    r, dn = 0.5, 2e-3
    x = np.array([BOX[0] / 2 - (r - dn), 0.0, 10.0])
    v = np.array([0.0, 0.0, 2e-6])
    z = np.zeros(3)
    dt = 1e-4
    hits, fw, tw, hout, fo = orc.law_wall(walls, x, v, z, r, 8 * r ** 3, z, dt, mat)
    assert hits == 1
    w = mat.w
    fcn = w.estarw * math.sqrt(r) * dn ** 1.5
    dtmax = w.ufricw * (2.0 - w.vpoisw) / (2.0 * (1.0 - w.vpoisw)) * dn
    slip = np.linalg.norm(v)
    xi = slip / dtmax
    assert xi < 1.0
    fct = w.ufricw * fcn * (1.0 - (1.0 - xi) ** 1.5)
    ct = w.damptw * math.sqrt(6.0 * w.ufricw * 8 * r ** 3 * fcn * math.sqrt(1.0 - xi) / dtmax)
    ftm = fct + ct / dt * slip
    ftm = min(ftm, w.ufricw * fcn)
    assert np.allclose(fw, [-fcn, 0.0, -ftm], rtol=0, atol=1e-10 * fcn)
    assert np.allclose(tw, [0.0, r * ftm, 0.0], rtol=0, atol=1e-10 * r * fcn)
    assert np.allclose(hout, -v, rtol=0, atol=1e-18)
    assert fo == pytest.approx(ftm, rel=1e-10)


def test_box_corner_two_walls_push(dem, orc):
    """In a corner (floor and x- wall) both walls push, each with its own Hertz force."""
    mat = _mat(dem)
    walls = dem.make_walls(dem.WALLS_BOX, BOX)
    r = 0.5
    d_floor, d_side = 1e-3, 4e-4
    x = np.array([-BOX[0] / 2 + (r - d_side), 0.0, r - d_floor])
    z = np.zeros(3)
    hits, fw, tw, hout, fo = orc.law_wall(walls, x, z, z, r, 1.0, z, 1e-4, mat)
    assert hits == 2
    k = mat.w.estarw * math.sqrt(r)
    assert np.allclose(fw, [k * d_side ** 1.5, 0.0, k * d_floor ** 1.5], rtol=1e-10, atol=0)


# ---------------------------------------------------------------------------------------
# Float-length compatibility: the reference's length() is sqrtf of the FP64 dot product
# rounded to float (cutil_math.h:1086-1089), in vprojec (neiglist.cu:140-142) and in the
# rebuild trigger (neiglist.cu:359,370)
# ---------------------------------------------------------------------------------------

def _float_len(v):
    d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    return float(np.sqrt(np.float32(d)))


def test_vprojec_rescale_with_float_length(dem, orc):
    """Two spheres on the z axis (normal exactly (0, 0, 1)), no relative motion, an incoming
    history with a normal component (|h| > 0.001, below the spring cap of the 0.5 overlap):
    the rotated history is (hx, hy, 0) * sqrt(pmag / rmag), with the magnitudes in float
    under the option and in FP64 without it."""
    mat = _mat(dem, mu=0.5, mu_w=0.5)
    z3 = np.zeros(3)
    xi, xj = np.zeros(3), np.array([0.0, 0.0, 0.5])
    hin = np.array([1.234567e-3, -7.654321e-4, 3.3e-4])
    r = np.array([hin[0], hin[1], 0.0])
    out = {}
    for flen in (False, True):
        c, _, _, _, hout, _ = orc.law_pair(xi, xj, z3, z3, z3, z3, 0.5, 0.5, 1.0, 1.0, hin,
                                           1e-4, mat, float_length=flen)
        assert c == 1
        if flen:
            pmag, rmag = _float_len(hin), _float_len(r)
        else:
            pmag, rmag = math.sqrt(float(hin @ hin)), math.sqrt(float(r @ r))
        pr = math.sqrt(pmag / rmag)
        want = np.array([r[0] * pr, r[1] * pr, 0.0])
        assert np.array_equal(hout, want), (flen, hout, want)
        out[flen] = hout
    assert not np.array_equal(out[False], out[True]), "float magnitudes must change the bits"


def test_rebuild_trigger_with_float_length(orc):
    """One particle displaced by |q| = 0.2749999999 since the last rebuild: under the FP64
    test |q| > 0.275 (half the skin) is false, under the reference's float length() the
    magnitude rounds to the float 0.27500000596 and the rebuild triggers."""
    import ctypes
    L = orc.lib()
    ann = 0.5 * 0.55
    for flen, want in ((0, 0), (1, 1)):
        L.orc_set_float_length(flen)
        pos, pdis, angv = np.zeros(3), np.zeros(3), np.zeros(3)
        qi = np.array([0.2749999999, 0.0, 0.0])
        f0 = np.zeros(3)
        one = np.ones(1)
        iall = ctypes.c_uint(7)
        mz = ctypes.c_double(0.0)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        L.orc_update(1, P(pos), P(pdis), P(angv), P(one), P(one), P(f0), P(f0), P(qi), 1e-4,
                     ann * ann, ann, ctypes.byref(iall), ctypes.byref(mz))
        assert iall.value == want, (flen, iall.value)
    L.orc_set_float_length(0)
    assert _float_len(np.array([0.2749999999, 0.0, 0.0])) > ann > 0.2749999999
