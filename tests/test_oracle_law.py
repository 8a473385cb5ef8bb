"""CPU pins of the oracle's contact laws and integrator against closed forms and an
independent ODE solution (scipy).  No GPU."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp


def _mat(dem, **kw):
    m = dict(dem.MAT0)
    m.update(kw)
    return dem.reduced_material(**m)


def test_reduced_material_shipped_deck(dem):
    """dempacking.cu:146-149 for hopp3d.inp (E = 1e7 Pa, nu = 0.29, rho = 2500, d = 0.01 m):
    E_red = 77874.1 and E* = 56683 (hand computation)."""
    import os
    d = dem.read_deck(os.path.join(os.path.dirname(__file__), "golden", "hopp3d.inp"))
    m = dem.deck_material(d)
    assert m.p.estarp == pytest.approx(56683.0, rel=2e-5)
    assert m.w.estarw == pytest.approx(m.p.estarp, rel=1e-12)


def test_hertz_static_force(dem, orc):
    mat = _mat(dem)
    # Textbook Hertz with E* = E/(2(1-nu^2)) for two identical spheres: F = 4/3 E* sqrt(R*) d^1.5
    E_red = 6.0 * 1e7 / (math.pi * 2500.0 * 9.81 * 1e-3)
    Estar = E_red / (2.0 * (1.0 - 0.29 ** 2))
    rng = np.random.default_rng(0)
    for dn in (1e-5, 3e-4, 2e-3):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        xi = np.zeros(3)
        xj = (1.0 - dn) * u
        z = np.zeros(3)
        c, fi, ti, tj, h, fr = orc.law_pair(xi, xj, z, z, z, z, 0.5, 0.5, 1.0, 1.0, z, 1e-4, mat)
        assert c == 1
        F = 4.0 / 3.0 * Estar * math.sqrt(0.25) * dn ** 1.5
        assert np.linalg.norm(fi) == pytest.approx(F, rel=2e-8)
        assert np.allclose(fi / np.linalg.norm(fi), -u, atol=1e-12)   # pushes i away from j
        assert np.all(ti == 0.0) and np.all(tj == 0.0) and fr == 0.0


def test_no_contact_is_all_zero(dem, orc):
    mat = _mat(dem)
    z = np.zeros(3)
    c, fi, ti, tj, h, fr = orc.law_pair(z, np.array([1.0 + 1e-9, 0, 0]), z, z, z, z, 0.5, 0.5, 1, 1,
                                        np.array([1e-4, 2e-4, 0.0]), 1e-4, mat)
    assert c == 0 and not fi.any() and not ti.any() and not tj.any() and not h.any() and fr == 0.0


def test_antisymmetry_bitwise(dem, orc):
    """Swapping the bodies negates force and history and swaps the torques, bit for bit."""
    mat = _mat(dem, mu_r=0.01)
    rng = np.random.default_rng(1)
    for _ in range(200):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        ri, rj = rng.uniform(0.35, 0.65, 2)
        xi = rng.normal(size=3)
        xj = xi + (ri + rj - rng.uniform(1e-6, 2e-3)) * u
        vi, vj, wi, wj = (rng.normal(scale=1e-4, size=3) for _ in range(4))
        hin = rng.normal(scale=1e-4, size=3)
        mi, mj = 8 * ri ** 3, 8 * rj ** 3
        a = orc.law_pair(xi, xj, vi, vj, wi, wj, ri, rj, mi, mj, hin, 1e-4, mat)
        b = orc.law_pair(xj, xi, vj, vi, wj, wi, rj, ri, mj, mi, -hin, 1e-4, mat)
        assert a[0] == b[0] == 1
        assert np.array_equal(a[1], -b[1])
        assert np.array_equal(a[2], b[3]) and np.array_equal(a[3], b[2])
        assert np.array_equal(a[4], -b[4]) and a[5] == b[5]


def test_coulomb_limit_when_sliding(dem, orc):
    mu = 0.3
    mat = _mat(dem, mu=mu)
    E_red = 6.0 * 1e7 / (math.pi * 2500.0 * 9.81 * 1e-3)
    dn = 5e-4
    xi = np.zeros(3)
    xj = np.array([0.0, 0.0, 1.0 - dn])
    z = np.zeros(3)
    vj = np.array([5e-3, 0.0, 0.0])     # large tangential slip of j along +x
    c, fi, ti, tj, h, fr = orc.law_pair(xi, xj, z, vj, z, z, 0.5, 0.5, 1.0, 1.0, z, 1e-4, mat)
    fcn = mat.p.estarp * math.sqrt(0.25) * dn ** 1.5
    assert fr == pytest.approx(mu * fcn, rel=1e-12)
    assert fi[0] == pytest.approx(mu * fcn, rel=1e-12)       # friction drags i along +x
    nu = 0.29
    dtmax = mu * (2 - nu) / (2 * (1 - nu)) * dn
    assert np.linalg.norm(h) == pytest.approx(dtmax, rel=1e-12)
    # torque on i: r n x Ft = 0.5 * (z x x) * |Ft| = +0.5 |Ft| y
    assert ti[1] == pytest.approx(0.5 * mu * fcn, rel=1e-12)
    assert tj[1] == pytest.approx(0.5 * mu * fcn, rel=1e-12)
    _ = E_red


def test_frictionless_has_no_tangential_force(dem, orc):
    mat = _mat(dem, mu=0.0, mu_r=0.0)
    z = np.zeros(3)
    c, fi, ti, tj, h, fr = orc.law_pair(z, np.array([0, 0, 0.999]), z, np.array([1e-3, 0, 0]), z, z,
                                        0.5, 0.5, 1.0, 1.0, np.array([1e-4, 0, 0]), 1e-4, mat)
    assert c == 1 and fi[0] == 0.0 and fi[1] == 0.0 and not h.any() and not ti.any()


def test_rolling_moment(dem, orc):
    ur = 0.002
    mat = _mat(dem, mu=0.0, mu_r=ur)
    dn = 1e-3
    z = np.zeros(3)
    wi = np.array([0.0, 0.0, 3e-4])
    c, fi, ti, tj, h, fr = orc.law_pair(z, np.array([0, 0, 1 - dn]), z, z, wi, z, 0.5, 0.5, 1, 1, z,
                                        1e-4, mat)
    fcn = mat.p.estarp * math.sqrt(0.25) * dn ** 1.5
    assert ti[2] == pytest.approx(-ur * fcn, rel=1e-12) and tj[2] == pytest.approx(ur * fcn, rel=1e-12)


def _collision_restitution_ode(g, v0, Estar, R, m):
    """Independent Hertz + dashpot head-on collision, physical form, scipy RK45."""
    k = Estar * math.sqrt(R)
    c0 = g * math.sqrt(6.0 * m * Estar)

    def rhs(t, y):
        d, v = y
        dp = max(d, 0.0)
        return [v, (-k * dp ** 1.5 - c0 * (R * dp) ** 0.25 * v) / m]

    def back(t, y):
        return y[0]
    back.terminal = True
    back.direction = -1
    sol = solve_ivp(rhs, [0, 100.0], [0.0, v0], events=back, rtol=1e-11, atol=1e-16, max_step=1e-3)
    return abs(sol.y_events[0][0][1]) / v0


@pytest.mark.parametrize("e", [0.3, 0.6, 0.9])
def test_restitution_map_against_scipy(dem, e):
    g = dem.damping_for_restitution(e)
    assert dem.restitution_for_damping(g) == pytest.approx(e, abs=2e-6)
    got = _collision_restitution_ode(g, 0.05, 5e4, 0.25, 0.5)
    assert got == pytest.approx(e, abs=2e-4)


@pytest.mark.parametrize("e", [0.3, 0.6, 0.9])
def test_restitution_map_derivative_against_scipy(dem, e):
    """d gamma / d e of the library's map (central difference, relative step 1e-4, as the
    adjoint's chain rule takes it) against the inverse slope of an independent tight-tolerance
    solve of the same collision in physical form."""
    g = dem.damping_for_restitution(e)
    dg = 1e-4 * g
    de_dg = (_collision_restitution_ode(g + dg, 0.05, 5e4, 0.25, 0.5)
             - _collision_restitution_ode(g - dg, 0.05, 5e4, 0.25, 0.5)) / (2 * dg)
    h = 1e-4 * e
    ours = (dem.damping_for_restitution(e + h) - dem.damping_for_restitution(e - h)) / (2 * h)
    assert ours == pytest.approx(1.0 / de_dg, rel=1e-3)


@pytest.mark.parametrize("e", [0.3, 0.6])
def test_discrete_collision_restitution(dem, orc, e):
    """Head-on collision integrated with the oracle law and the reference integrator update
    (pdis += F dt^2 / m; pos += pdis) reproduces the ODE restitution."""
    mat = _mat(dem, e=e, mu=0.0, mu_r=0.0)
    dt = 2e-5
    v0 = 0.05   # approach speed of each sphere (reduced)
    x = np.array([[0.0, 0, 0], [1.0 + 1e-6, 0, 0]])
    p = np.array([[v0 * dt, 0, 0], [-v0 * dt, 0, 0]])
    z = np.zeros(3)
    h = np.zeros(3)
    touched = False
    for _ in range(200000):
        c, fi, ti, tj, h, fr = orc.law_pair(x[0], x[1], p[0], p[1], z, z, 0.5, 0.5, 1.0, 1.0, h,
                                            dt, mat)
        touched |= bool(c)
        p[0] = p[0] + fi * dt * dt / 1.0
        p[1] = p[1] + (-fi) * dt * dt / 1.0
        x += p
        if touched and not c:
            break
    e_meas = (p[1, 0] - p[0, 0]) / (2 * v0 * dt)
    assert e_meas == pytest.approx(e, rel=2e-3)


def test_free_fall_series(dem, orc):
    """A12 closed form: with gravity only, z_n = z0 + n v0 - n(n+1)/2 dt^2 (reduced g = 1)."""
    walls = dem.make_walls(dem.WALLS_BOX, (20.0, 20.0, 200.0))
    mat = _mat(dem)
    dt = 1e-3
    inst = dem.make_instance(1, walls, 200.0, 1.55, mat, dt)
    pos = np.array([[0.1, -0.2, 150.0]])
    v0 = np.array([[1e-4, -2e-4, 3e-4]])
    r = np.array([0.5])
    z = np.zeros((1, 3))
    sim = orc.OracleSim(inst, pos, v0, z, r, 8 * r ** 3, 0.4 * 8 * r ** 5)
    nstep = 500
    assert sim.step(nstep) == 0
    zn = 150.0 + nstep * 3e-4 - nstep * (nstep + 1) / 2 * dt * dt
    assert sim.field("pos")[0, 2] == pytest.approx(zn, abs=1e-11)
    assert sim.field("pos")[0, 0] == pytest.approx(0.1 + nstep * 1e-4, abs=1e-12)
    assert sim.field("pdis")[0, 2] == pytest.approx(3e-4 - nstep * dt * dt, abs=1e-15)


def test_resting_sphere_on_floor(dem, orc):
    """A11 closed form: a sphere settles on the floor at the Hertz overlap
    (m g / (E*_w sqrt(r)))^(2/3)."""
    walls = dem.make_walls(dem.WALLS_BOX, (20.0, 20.0, 20.0))
    mat = dem.reduced_material(E=1e5, nu=0.29, rho=2500.0, diam=1e-3, mu=0.3, mu_w=0.3, e=0.2,
                               mu_r=0.0)
    dt = 2e-3
    inst = dem.make_instance(1, walls, 20.0, 1.55, mat, dt)
    r = np.array([0.5])
    m = 8 * r ** 3
    pos = np.array([[0.0, 0.0, 0.5005]])
    z = np.zeros((1, 3))
    sim = orc.OracleSim(inst, pos, z, z, r, m, 0.4 * m * r ** 2)
    assert sim.step(20000) == 0
    delta = (m[0] / (mat.w.estarw * math.sqrt(0.5))) ** (2.0 / 3.0)
    assert 0.5 - sim.field("pos")[0, 2] == pytest.approx(delta, rel=1e-4)
    assert sim.field("fricpw")[0] == pytest.approx(0.0, abs=1e-12)


def test_wall_history_slot_belongs_to_deepest_wall(dem, orc):
    """A sphere touching the floor and a side wall: both walls push, history is the deeper
    contact's (the floor here)."""
    walls = dem.make_walls(dem.WALLS_BOX, (4.0, 4.0, 20.0))
    mat = _mat(dem)
    inst = dem.make_instance(1, walls, 20.0, 1.55, mat, 1e-4)
    r = np.array([0.5])
    pos = np.array([[-2.0 + 0.4995, 0.0, 0.499]])   # side overlap 5e-4, floor overlap 1e-3
    v = np.array([[0.0, 1e-4, 0.0]])                 # slides along y on both walls
    z = np.zeros((1, 3))
    sim = orc.OracleSim(inst, pos, v, z, r, 8 * r ** 3, 0.4 * r ** 2)
    assert sim.step(1) == 0
    h = sim.field("disptw")[0]
    assert h[1] != 0.0 and h[2] == 0.0      # tangential history in the floor plane
