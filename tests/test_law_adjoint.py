"""CPU check of the hand-written reverse mode of the contact laws (law_adj.hpp) against
central finite differences of the forward laws (law.hpp), through libdem's host entry points.
Covers the branches: plain contact, capped tangential spring, Coulomb cap, the vprojec rescale
(|h| >= 0.001) and rolling resistance; box and cylinder walls, corners."""
import ctypes

import numpy as np
import pytest

P = ctypes.c_void_p


def ptr(a):
    return a.ctypes.data_as(P)


def pair_eval(L, x, mat):
    out = np.zeros(13)
    c = L.dem_law_pair_eval(ptr(x), ptr(mat), ptr(out))
    return c, out


def wall_eval(L, walls, x, mat):
    out = np.zeros(10)
    h = L.dem_law_walls_eval(ctypes.byref(walls), ptr(x), ptr(mat), ptr(out))
    return h, out


def fd_vjp(f, x, idx, ybar, eps_rel=1e-7):
    """Central-difference ybar . J over the inputs `idx`."""
    g = np.zeros(len(idx))
    for k, i in enumerate(idx):
        h = eps_rel * max(abs(x[i]), 1e-3)
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[k] = (ybar @ f(xp) - ybar @ f(xm)) / (2 * h)
    return g


def mat_fd(f, mat, k, ybar, eps_rel=1e-5):
    h = eps_rel * abs(mat[k])
    mp, mm = mat.copy(), mat.copy()
    mp[k] += h
    mm[k] -= h
    return (ybar @ f(mp) - ybar @ f(mm)) / (2 * h)


def pair_case(rng, kind):
    ra, rb = rng.uniform(0.4, 0.6, 2)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    xa = rng.normal(size=3)
    dn = rng.uniform(2e-4, 2e-3)
    xb = xa + (ra + rb - dn) * u
    va, vb = rng.normal(scale=2e-5, size=(2, 3))
    wa, wb = rng.normal(scale=4e-5, size=(2, 3))
    hin = rng.normal(scale=2e-5, size=3)
    mu = 0.4
    if kind == "capped":
        hin = rng.normal(scale=1e-3, size=3) * 0.3     # spring beyond dt_max
        mu = 0.05
    elif kind == "coulomb":
        vb = va + rng.normal(scale=3e-3, size=3)       # fast slip: damping exceeds the cap
        mu = 0.1
    elif kind == "rescale":
        hin = rng.normal(size=3)
        hin *= 2e-3 / np.linalg.norm(hin)
        mu = 0.9
        dn = 5e-3
        xb = xa + (ra + rb - dn) * u
    x = np.concatenate([xa, xb, va, vb, wa, wb, [ra, rb, 8 * ra ** 3, 8 * rb ** 3], hin, [2e-4]])
    mat = np.array([5.6e4, mu, 0.29, 0.35, 0.3, 0.01])
    return x, mat


@pytest.mark.parametrize("kind", ["plain", "capped", "coulomb", "rescale"])
def test_pair_vjp_matches_finite_differences(dem, kind):
    L = dem.lib()
    rng = np.random.default_rng({"plain": 1, "capped": 2, "coulomb": 3, "rescale": 4}[kind])
    for trial in range(6):
        x, mat = pair_case(rng, kind)
        c, out = pair_eval(L, x, mat)
        assert c == 1
        ybar = rng.normal(size=12) * np.r_[np.full(9, 1.0 / (np.abs(out[:9]).max() + 1e-30)),
                                           np.full(3, 1e3)]
        f = lambda z: pair_eval(L, z, mat)[1][:12]
        inbar = np.zeros(21)
        matbar = np.zeros(5)
        assert L.dem_law_pair_vjp(ptr(x), ptr(mat), ptr(ybar), ptr(inbar), ptr(matbar)) == 0
        idx = list(range(18)) + [22, 23, 24]
        g = fd_vjp(f, x, idx, ybar)
        scale = np.abs(g).max()
        assert np.max(np.abs(inbar - g)) <= 2e-5 * scale, (kind, trial, inbar - g)
        fm = lambda m: pair_eval(L, x, m)[1][:12]
        for k_out, k_mat in ((0, 0), (1, 1), (2, 3), (3, 4), (4, 5)):
            gm = mat_fd(fm, mat, k_mat, ybar)
            assert matbar[k_out] == pytest.approx(gm, rel=2e-5, abs=2e-5 * (abs(gm) + scale * 1e-6))


def test_pair_vjp_branch_coverage(dem):
    """The four fixtures really exercise their branches (checked on the forward outputs)."""
    L = dem.lib()
    rng = np.random.default_rng(9)
    x, mat = pair_case(rng, "capped")
    _, out = pair_eval(L, x, mat)
    dn = x[18] + x[19] - np.linalg.norm(x[3:6] - x[0:3])
    dtmax = mat[1] * (2 - mat[2]) / (2 * (1 - mat[2])) * dn
    assert np.linalg.norm(out[9:12]) == pytest.approx(dtmax, rel=1e-12)
    x, mat = pair_case(rng, "coulomb")
    _, out = pair_eval(L, x, mat)
    dn = x[18] + x[19] - np.linalg.norm(x[3:6] - x[0:3])
    rs = x[18] * x[19] / (x[18] + x[19])
    fcn = mat[0] * np.sqrt(rs) * dn ** 1.5
    assert out[12] == pytest.approx(mat[1] * fcn, rel=1e-12)


@pytest.mark.parametrize("geom", ["floor", "corner", "cylinder"])
def test_wall_vjp_matches_finite_differences(dem, geom):
    L = dem.lib()
    rng = np.random.default_rng({"floor": 5, "corner": 6, "cylinder": 7}[geom])
    if geom == "cylinder":
        walls = dem.make_walls(dem.WALLS_CYLINDER, (0, 0, 0), radius=5.0)
    else:
        walls = dem.make_walls(dem.WALLS_BOX, (6.0, 6.0, 20.0))
    for trial in range(6):
        r = rng.uniform(0.4, 0.6)
        if geom == "floor":
            xw = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), r - rng.uniform(2e-4, 2e-3)])
        elif geom == "corner":
            xw = np.array([-3.0 + r - rng.uniform(2e-4, 1e-3), 3.0 - r + rng.uniform(2e-4, 1e-3),
                           r - rng.uniform(2e-4, 2e-3)])
        else:
            th = rng.uniform(0, 2 * np.pi)
            rho = 5.0 - r + rng.uniform(2e-4, 2e-3)
            xw = np.array([rho * np.cos(th), rho * np.sin(th), 3.0])
        x = np.concatenate([xw, rng.normal(scale=2e-5, size=3), rng.normal(scale=4e-5, size=3),
                            [r, 8 * r ** 3], rng.normal(scale=2e-5, size=3), [2e-4]])
        mat = np.array([5.6e4, 0.4, 0.29, 0.35, 0.3, 0.01])
        hits, out = wall_eval(L, walls, x, mat)
        assert hits >= (2 if geom == "corner" else 1)
        ybar = rng.normal(size=9) * np.r_[np.full(6, 1.0 / (np.abs(out[:6]).max() + 1e-30)),
                                          np.full(3, 1e3)]
        f = lambda z: wall_eval(L, walls, z, mat)[1][:9]
        inbar = np.zeros(12)
        matbar = np.zeros(5)
        assert L.dem_law_walls_vjp(ctypes.byref(walls), ptr(x), ptr(mat), ptr(ybar), ptr(inbar),
                                   ptr(matbar)) == 0
        idx = list(range(9)) + [11, 12, 13]
        g = fd_vjp(f, x, idx, ybar)
        scale = np.abs(g).max()
        assert np.max(np.abs(inbar - g)) <= 2e-5 * scale, (geom, trial, inbar - g)


@pytest.mark.parametrize("regime", ["floored", "above"])
def test_pair_vjp_with_rolling_floor(dem, regime):
    """The rolling floor of the differentiable form (DEM_OPT_ROLL_FLOOR): below the floor the
    moment is -w_rel / roll_c (viscous, independent of mu_r and fcn), above it the reference's
    moment; the vjp is the exact reverse of that forward (central differences), and roll_c = 0
    is the plain law bit for bit."""
    L = dem.lib()
    rng = np.random.default_rng({"floored": 11, "above": 12}[regime])
    for trial in range(6):
        x, mat = pair_case(rng, "plain")
        dn = x[18] + x[19] - np.linalg.norm(x[3:6] - x[0:3])
        rs = x[18] * x[19] / (x[18] + x[19])
        fcn = mat[0] * np.sqrt(rs) * dn ** 1.5
        wrel = x[12:15] - x[15:18]
        wfl_target = np.linalg.norm(wrel) * (3.0 if regime == "floored" else 0.3)
        roll_c = wfl_target / (mat[5] * fcn)
        out0 = np.zeros(13)
        out1 = np.zeros(13)
        assert L.dem_law_pair_eval(ptr(x), ptr(mat), ptr(out0)) == 1
        assert L.dem_law_pair_eval_floor(ptr(x), ptr(mat), 0.0, ptr(out1)) == 1
        assert np.array_equal(out0, out1)
        outf = np.zeros(13)
        assert L.dem_law_pair_eval_floor(ptr(x), ptr(mat), roll_c, ptr(outf)) == 1
        if regime == "floored":   # the moment is -w_rel / roll_c
            ta_minus_tb = outf[3:6] - outf[6:9]
            # ta - tb = (ra - rb) n x ft + 2 mr
            n = (x[3:6] - x[0:3]) / (x[18] + x[19] - dn)
            ft = outf[0:3] - np.dot(outf[0:3], n) * n
            mr2 = 0.5 * (ta_minus_tb - np.cross((x[18] - x[19]) * n, ft))
            assert np.allclose(mr2, -wrel / roll_c, rtol=1e-9, atol=1e-14)
        else:
            assert np.array_equal(outf, out0)
        ybar = rng.normal(size=12) * np.r_[np.full(9, 1.0 / (np.abs(outf[:9]).max() + 1e-30)),
                                           np.full(3, 1e3)]
        f = lambda z: (lambda o: (L.dem_law_pair_eval_floor(ptr(z), ptr(mat), roll_c, ptr(o)), o)[1])(
            np.zeros(13))[:12]
        inbar = np.zeros(21)
        matbar = np.zeros(5)
        assert L.dem_law_pair_vjp_floor(ptr(x), ptr(mat), roll_c, ptr(ybar), ptr(inbar),
                                        ptr(matbar)) == 0
        idx = list(range(18)) + [22, 23, 24]
        g = fd_vjp(f, x, idx, ybar)
        scale = np.abs(g).max()
        assert np.max(np.abs(inbar - g)) <= 2e-5 * scale, (regime, trial, inbar - g)
        fm = lambda m: (lambda o: (L.dem_law_pair_eval_floor(ptr(x), ptr(m), roll_c, ptr(o)), o)[1])(
            np.zeros(13))[:12]
        for k_out, k_mat in ((0, 0), (1, 1), (2, 3), (3, 4), (4, 5)):
            gm = mat_fd(fm, mat, k_mat, ybar)
            assert matbar[k_out] == pytest.approx(gm, rel=2e-5, abs=2e-5 * (abs(gm) + scale * 1e-6))


def test_wall_vjp_with_rolling_floor(dem):
    """The wall law under the rolling floor (one inertia): vjp against central differences in
    the floored regime."""
    L = dem.lib()
    rng = np.random.default_rng(13)
    walls = dem.make_walls(dem.WALLS_BOX, (6.0, 6.0, 20.0))
    for trial in range(6):
        r = rng.uniform(0.4, 0.6)
        xw = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), r - rng.uniform(2e-4, 2e-3)])
        w = rng.normal(scale=4e-8, size=3)
        x = np.concatenate([xw, rng.normal(scale=2e-5, size=3), w, [r, 8 * r ** 3],
                            rng.normal(scale=2e-5, size=3), [2e-4]])
        mat = np.array([5.6e4, 0.4, 0.29, 0.35, 0.3, 0.01])
        dn = r - xw[2]
        fcn = mat[0] * np.sqrt(r) * dn ** 1.5
        roll_c = 3.0 * np.linalg.norm(w) / (mat[5] * fcn)
        out = np.zeros(10)
        assert L.dem_law_walls_eval_floor(ctypes.byref(walls), ptr(x), ptr(mat), roll_c,
                                          ptr(out)) == 1
        ybar = rng.normal(size=9) * np.r_[np.full(6, 1.0 / (np.abs(out[:6]).max() + 1e-30)),
                                          np.full(3, 1e3)]

        def f(z):
            o = np.zeros(10)
            L.dem_law_walls_eval_floor(ctypes.byref(walls), ptr(z), ptr(mat), roll_c, ptr(o))
            return o[:9]
        inbar = np.zeros(12)
        matbar = np.zeros(5)
        assert L.dem_law_walls_vjp_floor(ctypes.byref(walls), ptr(x), ptr(mat), roll_c, ptr(ybar),
                                         ptr(inbar), ptr(matbar)) == 0
        idx = list(range(9)) + [11, 12, 13]
        g = fd_vjp(f, x, idx, ybar)
        scale = np.abs(g).max()
        assert np.max(np.abs(inbar - g)) <= 2e-5 * scale, (trial, inbar - g)
