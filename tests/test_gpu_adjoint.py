"""GPU checks of the adjoint form (dem_adj_*): the checkpointed forward equals the plain
forward bit for bit, the gradient does not depend on the checkpoint interval, and adjoint
gradients of the smooth observables agree with central finite differences of the plain
forward -- per contact-law constant, per initial coordinate, in the calibration parameters
(mu, log E, e, mu_r) through the torch.autograd.Function, and per instance in the batched form.
The adjoint has no counterpart in the reference; finite differences of the parity-checked
forward are its check.  These checks run the exact reverse mode (OPT_ADJ_ROLL_REG = 0);
the regularised one is checked where it differs, in chattering rolling contact."""
import copy
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 200
EVERY = 50
# (section, field) of each contact-law constant, in the library's theta_bar order
LAW_FIELDS = [("p", "estarp"), ("p", "ufricp"), ("p", "dampnp"), ("p", "damptp"), ("p", "upp"),
              ("w", "estarw"), ("w", "ufricw"), ("w", "dampnw"), ("w", "damptw"), ("w", "upw")]


def small_workload(dem, n=600, seed=3, interval=20, vscale=2e-5):
    """Compressed jittered lattice in a 20d box: contacts, wall contacts and list rebuilds
    from the first steps; random displacements per step of scale `vscale`."""
    w = dem.config0(n=n, seed=seed)
    w.pos[:, 2] = 0.5 + (w.pos[:, 2] - w.pos[:, 2].min()) * (0.998 / 1.2)
    w.pos[:, :2] *= 0.998 / 1.2
    rng = np.random.default_rng(seed)
    w.pdis[:] = rng.normal(scale=vscale, size=w.pdis.shape)
    w.angv[:] = rng.normal(scale=2e-5, size=w.angv.shape)
    w.rebuild_interval = interval
    return w


def obs_small(dem):
    return dem.make_obs(zhi=2.2, margin=2.0, zlo=0.8, eps=0.5, eps_c=2e-3)


def run_plain(dem, w, mats, nsteps, obs):
    ctx = dem.Context(w.instances)
    for k, m in enumerate(mats):
        ctx.set_material(k, m)
    ctx.load(w)
    ctx.set_option(dem.OPT_ADJ_ROLL_REG, 0)
    ctx.step(nsteps)
    phi, cn = ctx.observe(obs)
    st = ctx.state()
    ctx.close()
    return phi, cn, st


def rel_close(a, b, rtol):
    scale = max(abs(a), abs(b), 1e-300)
    return abs(a - b) <= rtol * scale


def test_adj_forward_matches_plain(dem, gpu_ok):
    w = small_workload(dem)
    obs = obs_small(dem)
    phi0, cn0, st0 = run_plain(dem, w, [w.instances[0].mat], NSTEPS, obs)
    ctx = dem.Context(w.instances)
    ctx.load(w)
    ctx.set_option(dem.OPT_ADJ_ROLL_REG, 0)
    phi, cn = ctx.adj_forward(NSTEPS, EVERY, obs)
    assert phi[0] == phi0[0] and cn[0] == cn0[0]
    assert 0.2 < phi[0] < 0.8 and cn[0] > 1.0, (phi, cn)
    ctx.adj_backward([1.0], [0.0])
    st = ctx.state()   # the backward leaves the context at the forward's end state
    assert np.array_equal(st["pos"], st0["pos"]) and np.array_equal(st["pdis"], st0["pdis"])
    info = ctx.adj_info()
    assert info["segments"] == NSTEPS // EVERY and info["max_pairs"] > w.n
    # the context keeps stepping from there exactly as the plain one would
    ctx.step(10)
    phi2, _, st2 = run_plain(dem, w, [w.instances[0].mat], NSTEPS + 10, obs)
    assert np.array_equal(ctx.state()["pos"], st2["pos"])


def test_adj_independent_of_checkpoint_interval(dem, gpu_ok):
    w = small_workload(dem)
    obs = obs_small(dem)
    ref = None
    for every in (NSTEPS, 50, 7, 1):
        ctx = dem.Context(w.instances)
        ctx.load(w)
        ctx.set_option(dem.OPT_ADJ_ROLL_REG, 0)
        ctx.adj_forward(NSTEPS, every, obs)
        tb = ctx.adj_backward([0.7], [0.3])
        xb, vb, _ = ctx.adj_state_grad()
        if ref is None:
            ref = (tb, xb, vb)
            assert np.all(np.isfinite(tb)) and np.abs(tb).max() > 0
        else:
            assert np.array_equal(tb, ref[0]), every
            assert np.array_equal(xb, ref[1]) and np.array_equal(vb, ref[2]), every
        ctx.close()


# Central differences need small steps here: the dashpot term grows as overlap^(1/4) from
# contact onset (not Lipschitz at zero overlap), so a wide step straddles onset events and
# its difference quotient is off by up to tens of percent, while as the step shrinks the
# quotient converges onto the adjoint (relative steps 1e-5, 1e-6 and 1e-7 are used).
FD_RELS = (1e-5, 1e-6, 1e-7)


def agrees(adj, fds, rtol, floor):
    return any(rel_close(adj, fd, rtol) or abs(adj - fd) <= floor for fd in fds)


@pytest.mark.parametrize("interval", [20, 0])
def test_adj_law_constants_match_fd(dem, gpu_ok, interval):
    """d phi and d CN with respect to each of the ten contact-law constants; interval 0 is
    the reference's displacement trigger."""
    # the trigger case starts faster so that displacements reach half the skin
    w = small_workload(dem, interval=interval, vscale=2e-5 if interval else 1.5e-3)
    obs = obs_small(dem)
    base = w.instances[0].mat
    ctx = dem.Context(w.instances)
    ctx.load(w)
    ctx.set_option(dem.OPT_ADJ_ROLL_REG, 0)
    ctx.adj_forward(NSTEPS, EVERY, obs)
    tb_phi = ctx.adj_backward([1.0], [0.0])[0]
    tb_cn = ctx.adj_backward([0.0], [1.0])[0]
    assert ctx.stats().rebuilds >= 2
    ctx.close()
    vals0 = np.array([getattr(getattr(base, s), f) for s, f in LAW_FIELDS])
    bad = []
    for j, (sec, name) in enumerate(LAW_FIELDS):
        v0 = vals0[j]
        fd_phi, fd_cn = [], []
        for rel in FD_RELS:
            h = rel * abs(v0)
            vals = []
            for sgn in (1, -1):
                m = copy.deepcopy(base)
                setattr(getattr(m, sec), name, v0 + sgn * h)
                phi, cn, _ = run_plain(dem, w, [m], NSTEPS, obs)
                vals.append((phi[0], cn[0]))
            fd_phi.append((vals[0][0] - vals[1][0]) / (2 * h))
            fd_cn.append((vals[0][1] - vals[1][1]) / (2 * h))
        print(f"{sec}.{name}: phi adj {tb_phi[j]:.6e} fd {fd_phi} | cn adj {tb_cn[j]:.6e} "
              f"fd {fd_cn}", flush=True)
        # gradients far below the largest (in relative terms) sit at the FD noise floor
        floor_phi = 1e-5 * np.abs(tb_phi * vals0).max() / abs(v0)
        floor_cn = 1e-5 * np.abs(tb_cn * vals0).max() / abs(v0)
        if not agrees(tb_phi[j], fd_phi, 1e-3, floor_phi):
            bad.append(("phi", name, tb_phi[j], fd_phi))
        if not agrees(tb_cn[j], fd_cn, 1e-3, floor_cn):
            bad.append(("cn", name, tb_cn[j], fd_cn))
    assert not bad, bad


def test_adj_initial_state_matches_fd(dem, gpu_ok):
    w = small_workload(dem)
    obs = obs_small(dem)
    ctx = dem.Context(w.instances)
    ctx.load(w)
    ctx.set_option(dem.OPT_ADJ_ROLL_REG, 0)
    ctx.adj_forward(NSTEPS, EVERY, obs)
    ctx.adj_backward([1.0], [0.05])
    xb, vb, wb = ctx.adj_state_grad()
    ctx.close()
    rng = np.random.default_rng(11)
    # probe the particles with the largest sensitivities plus a few random ones
    g = np.abs(xb).sum(axis=1)
    picks = list(np.argsort(g)[-4:]) + list(rng.integers(0, w.n, 2))
    bad = []
    for i in picks:
        for k in range(3):
            for arr_name, grad, hs in (("pos", xb, (1e-6, 1e-7)), ("pdis", vb, (1e-9, 1e-10))):
                fds = []
                for h in hs:
                    vals = []
                    for sgn in (1, -1):
                        w2 = copy.copy(w)
                        a = getattr(w, arr_name).copy()
                        a[i, k] += sgn * h
                        setattr(w2, arr_name, a)
                        phi, cn, _ = run_plain(dem, w2, [w.instances[0].mat], NSTEPS, obs)
                        vals.append(phi[0] + 0.05 * cn[0])
                    fds.append((vals[0] - vals[1]) / (2 * h))
                if not agrees(grad[i, k], fds, 1e-3, 1e-5 * np.abs(grad).max()):
                    bad.append((arr_name, int(i), k, grad[i, k], fds))
    assert not bad, bad


def test_adj_batched_equals_single(dem, gpu_ok):
    """Three instances with different materials in one batched context give each instance's
    single-context gradient."""
    ws = [small_workload(dem, n=n, seed=s) for n, s in ((400, 3), (600, 4), (500, 5))]
    obs = obs_small(dem)
    mm = dem.adjoint.MaterialMap()
    thetas = [[0.3, math.log(1e7), 0.6, 0.002], [0.5, math.log(3e6), 0.4, 0.004],
              [0.2, math.log(2e7), 0.8, 0.001]]
    insts = []
    for w, th in zip(ws, thetas):
        inst = w.instances[0]
        inst.mat = mm.material(th)
        insts.append(inst)
    cat = lambda name: np.concatenate([getattr(w, name) for w in ws])
    big = dem.Workload("batched-3", insts, cat("pos"), cat("pdis"), cat("angv"), cat("rad"),
                       cat("rmass"), cat("inert"), rebuild_interval=20)
    ctx = dem.Context(insts)
    ctx.load(big)
    phi_b, cn_b = ctx.adj_forward(NSTEPS, EVERY, obs)
    tb_b = ctx.adj_backward([1.0, 2.0, 3.0], [0.1, 0.2, 0.3])
    ctx.close()
    for k, w in enumerate(ws):
        w.instances[0].mat = insts[k].mat
        c = dem.Context(w.instances)
        c.load(w)
        phi, cn = c.adj_forward(NSTEPS, EVERY, obs)
        tb = c.adj_backward([float(k + 1)], [0.1 * (k + 1)])
        c.close()
        assert phi[0] == phi_b[k] and cn[0] == cn_b[k]
        np.testing.assert_allclose(tb[0], tb_b[k], rtol=1e-12, atol=1e-300)


def test_autograd_function_matches_fd(dem, gpu_ok):
    """torch.autograd.Function in the calibration parameters (mu, log E, e, mu_r).  mu, log E
    and mu_r are differenced directly; e enters through gamma(e), a map defined by an ODE
    solve whose derivative is pinned on the CPU against an independent solve, so here the
    shared damping coefficient gamma (both laws, normal and tangential) is differenced and
    the chain factor d gamma / d e applied."""
    import torch
    A = dem.adjoint
    w = small_workload(dem)
    prob = A.PackingProblem(w, NSTEPS, EVERY, obs_small(dem), roll_reg=0.0)
    theta0 = np.array([[0.3, math.log(1e7), 0.6, 0.002]])
    theta = torch.tensor(theta0, requires_grad=True)
    phi, cn = A.packing_observables(theta, prob)
    (phi.sum() + 0.05 * cn.sum()).backward()
    g = theta.grad.numpy()[0]
    tb = prob.theta_bar[0]
    r = [A.fd_check(prob, theta0, rel=rel, which=(0, 1, 3)) for rel in FD_RELS]
    adj = r[0]["adj_phi"] + 0.05 * r[0]["adj_cn"]
    np.testing.assert_allclose(g, adj, rtol=1e-10)
    fds = [x["fd_phi"] + 0.05 * x["fd_cn"] for x in r]
    print("theta grads adj", adj, "fd", fds, flush=True)
    nat = np.array([theta0[0, 0], 1.0, theta0[0, 2], theta0[0, 3]])   # natural scales
    scale = np.abs(adj * nat).max()
    for j in (0, 1, 3):
        floor = 1e-5 * scale / nat[j]
        assert agrees(adj[j], [f[j] for f in fds], 1e-3, floor), (A.PARAMS[j], adj[j], fds)
    # restitution: FD in the shared gamma, then the chain factor
    base = prob.mmap.material(theta0[0])
    gam = base.p.dampnp
    fd_g = []
    for rel in FD_RELS:
        vals = []
        for sgn in (1, -1):
            m = copy.deepcopy(base)
            m.p.dampnp = m.p.damptp = m.w.dampnw = m.w.damptw = gam * (1 + sgn * rel)
            phi_, cn_, _ = run_plain(dem, w, [m], NSTEPS, obs_small(dem))
            vals.append(phi_[0] + 0.05 * cn_[0])
        fd_g.append((vals[0] - vals[1]) / (2 * rel * gam))
    adj_g = tb[2] + tb[3] + tb[7] + tb[8]
    assert agrees(adj_g, fd_g, 1e-3, 0.0), (adj_g, fd_g)
    assert adj[2] == pytest.approx(adj_g * prob.mmap.dgamma_de(0.6), rel=1e-12)


def test_calibration_reduces_misfit(dem, gpu_ok):
    """Four starts calibrated in one batched context toward the observables of hidden
    parameters: the summed misfit falls."""
    A = dem.adjoint
    hidden = [0.4, math.log(1e7), 0.5, 0.003]
    w = small_workload(dem, n=400)
    obs = obs_small(dem)
    target = A.PackingProblem(w, NSTEPS, EVERY, obs)
    tphi, tcn = target.evaluate([hidden])
    starts = np.array([[0.2, math.log(1e7), 0.7, 0.002], [0.6, math.log(1e7), 0.3, 0.004],
                       [0.3, math.log(1e7), 0.6, 0.001], [0.5, math.log(1e7), 0.4, 0.005]])
    insts, parts = [], []
    for k in range(4):
        wk = small_workload(dem, n=400)
        insts.append(wk.instances[0])
        parts.append(wk)
    cat = lambda name: np.concatenate([getattr(p, name) for p in parts])
    big = dem.Workload("calib-4", insts, cat("pos"), cat("pdis"), cat("angv"), cat("rad"),
                       cat("rmass"), cat("inert"), rebuild_interval=20)
    prob = A.PackingProblem(big, NSTEPS, EVERY, obs)
    res = A.calibrate(prob, float(tphi[0]), float(tcn[0]), starts, iters=8,
                      lr=(0.02, 0.05, 0.02, 0.0005), scale=(1e-2, 0.05))
    losses = [h["loss"] for h in res["history"]]
    print("calibration losses", losses, flush=True)
    assert losses[-1] < 0.6 * losses[0], losses


def test_config4_gradients_at_full_size(dem, gpu_ok):
    """Configuration 4 at its own size: 20K lognormal spheres settled in a 30d-wide box, 5,000
    steps, checkpoints every 100, with the differentiable form's rolling floor (k = 3: the
    forward's rolling moment is viscous where the reference's flips sign every step, and the
    adjoint is that forward's exact reverse mode).  Adjoint gradients of the inner-region
    packing fraction and coordination number in (mu, log E, e, mu_r) against central
    differences of the same forward, relative step 1e-4 (e through the dashpot coefficient,
    gamma(1 +- 1e-4), times d gamma / d e).  The bars are what this window supports: the
    response is piecewise smooth (contact onsets, Coulomb and spring caps), so central
    differences of phi do not settle as the step shrinks: against the adjoint, the quotients
    at steps 1e-3, 1e-4, 3e-5 and 1e-5 differ by up to 3.9e-3 for mu and by 2.5e-3 ... 3.4e-2
    for e (the two larger steps straddle contact onsets); log E and mu_r agree to about 1e-3
    at every step."""
    A = dem.adjoint
    w = dem.settle_bed(dem.config4(), tap=0.0)
    obs = dem.make_obs(zhi=w.meta["bed_top"] - 3.0, margin=2.0, zlo=2.0, eps=0.5, eps_c=1e-3)
    prob = A.PackingProblem(w, 5000, 100, obs, roll_floor=3.0)
    th0 = [0.3, math.log(1e7), 0.6, 0.002]
    rel = 1e-4
    r = A.fd_check(prob, [th0], rel=rel, which=(0, 1, 3))
    base = prob.mmap.material(th0)
    gam = base.p.dampnp
    vals = []
    for sgn in (1, -1):
        m = copy.deepcopy(base)
        m.p.dampnp = m.p.damptp = m.w.dampnw = m.w.damptw = gam * (1 + sgn * rel)
        prob.ctx.set_material(0, m)
        prob.ctx.load(w)
        prob.ctx.step(5000)
        vals.append(prob.ctx.observe(obs))
    dg = prob.mmap.dgamma_de(0.6)
    r["fd_phi"][2] = (vals[0][0][0] - vals[1][0][0]) / (2 * rel * gam) * dg
    r["fd_cn"][2] = (vals[0][1][0] - vals[1][1][0]) / (2 * rel * gam) * dg
    ga, gf = np.asarray(r["adj_phi"]), np.asarray(r["fd_phi"])
    ca, cf = np.asarray(r["adj_cn"]), np.asarray(r["fd_cn"])
    relerr = np.abs(ga - gf) / np.maximum(np.abs(ga), np.abs(gf))
    nat = np.array([th0[0], 1.0, th0[2], th0[3]])
    norm_phi = np.abs(ga - gf) * nat / np.max(np.abs(gf) * nat)
    norm_cn = np.abs(ca - cf) * nat / np.max(np.abs(cf) * nat)
    print(f"[config4] phi {r['phi']:.5f} adj {ga} fd {gf} rel {relerr} norm phi {norm_phi} "
          f"cn {norm_cn}", flush=True)
    assert np.all(np.isfinite(ga)) and np.all(np.isfinite(ca))
    assert relerr[1] <= 1.5e-3 and relerr[3] <= 1.5e-3, relerr   # log E, mu_r
    assert relerr[0] <= 6e-3, relerr                              # mu
    assert relerr[2] <= 5e-2, relerr                              # e
    assert np.all(norm_phi <= 2e-3), norm_phi
    assert np.all(norm_cn <= 5e-3), norm_cn
