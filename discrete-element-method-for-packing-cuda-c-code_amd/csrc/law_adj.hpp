// law_adj.hpp -- hand-written reverse mode (vector-Jacobian products) of the contact laws in
// law.hpp, host + device.
//
// Each *_adj function recomputes the forward intermediates of one contact, then propagates
// the adjoints of the outputs back to the inputs in reverse order, accumulating (+=) into the
// input adjoints.  Branches follow the forward pass (the derivative of the branch taken):
// contact on/off, the rare vprojec rescale, the tangential-spring cap, the Coulomb cap and the
// rolling-moment guard.  Inside the capped spring branch xi == 1 identically, so the terms
// through om = 1 - xi vanish there.  Material adjoints: E*, mu, gamma_n, gamma_t, mu_r (nu is
// a fixed property).
#pragma once
#include "law.hpp"

namespace dem {

struct LawParamsBar {
    double estar, mu, gn, gt, ur;
};

DEM_HD d3 add3(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEM_HD d3 sub3(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEM_HD void acc3(d3 &a, d3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
DEM_HD void acc3s(d3 &a, double s, d3 b) { a.x += s * b.x; a.y += s * b.y; a.z += s * b.z; }

// Reverse of law_core.  Output adjoints: fib (force on i), ftb_ext (tangential force, from
// the caller's torques), mrb (rolling moment), houtb (new history).  Accumulates into nb,
// dnb, vrelb, wrelb, hinb, Pb.
//
// roll_c regularises the rolling moment's linearisation.  The moment -mu_r fcn w/|w| has
// magnitude fixed by fcn, and its derivative across w is mu_r fcn / |w| (I - w w^T / |w|^2),
// which grows without bound as w -> 0.  Once |w| is below the change one step of that
// moment makes, w_c = mu_r fcn dt^2 (1/I_a + 1/I_b), the moment flips sign from step to
// step (chatter about rolling contact) and the exact linearisation amplifies perturbations
// by |1 - w_c / |w|| per step.  The caller passes roll_c = k dt^2 (1/I_a + 1/I_b) (one
// inertia for a wall), and the derivative uses max(|w|, roll_c mu_r fcn) = max(|w|, k w_c):
// exact away from chatter, at most critically damped (k = 1) within it.  roll_c = 0 is the
// exact reverse mode.
//
// roll_fwd: the forward itself floored |w| there (law_core's roll_c): the moment is then
// -w_rel / roll_c below the floor -- linear in w_rel, independent of mu_r and fcn -- and
// this is the exact reverse of that forward (the projector and the material terms only
// above the floor).
DEM_HD void law_core_adj(d3 n, double dn, double rs, double ms, d3 vrel, d3 wrel, d3 hin,
                         double dt, const LawParams &P, d3 fib, d3 ftb_ext, d3 mrb, d3 houtb,
                         d3 &nb, double &dnb, d3 &vrelb, d3 &wrelb, d3 &hinb, LawParamsBar &Pb,
                         double roll_c = 0.0, bool roll_fwd = false)
{
    // ---- forward recompute ----
    double vns = dot3(vrel, n);
    d3 vn = scale3(n, vns);
    d3 vt = sub3(vrel, vn);
    double sdn = sqrt(dn);
    double fcn = P.estar * sqrt(rs) * dn * sdn;
    double Bq = 6.0 * ms * P.estar * sqrt(rs * dn);
    double sB = sqrt(Bq);
    double cn = P.gn * sB;
    d3 q = cross3(n, hin);
    d3 mn = d3{-n.x, -n.y, -n.z};
    d3 r0 = cross3(mn, q);
    double pmag = sqrt(dot3(hin, hin));
    bool resc = false;
    double rmag = 0.0, pr = 1.0;
    if (!(pmag < 0.001)) {
        rmag = sqrt(dot3(r0, r0));
        if (!(pmag > rmag * 100.0)) {
            resc = true;
            pr = sqrt(pmag / rmag);
        }
    }
    d3 ht0 = resc ? scale3(r0, pr) : r0;
    d3 ht1 = add3(ht0, vt);
    double dtmax = P.mu * (2.0 - P.nu) / (2.0 * (1.0 - P.nu)) * dn;
    bool tang = P.mu > 0.0 && dtmax > 0.0;
    bool capped = false, coul = false;
    double htm = 0.0, htm2 = 0.0, xi = 0.0, om = 0.0, so = 0.0, fct = 0.0, A = 0.0, sA = 0.0,
           ct = 0.0, ftm = 0.0, rcap = 1.0, rho = 1.0;
    d3 ht2 = d3{0.0, 0.0, 0.0}, u = d3{0.0, 0.0, 0.0}, ft0 = d3{0.0, 0.0, 0.0};
    if (tang) {
        htm = sqrt(dot3(ht1, ht1));
        if (htm > dtmax) {
            capped = true;
            rho = dtmax / htm;
            ht2 = scale3(ht1, rho);
            htm2 = dtmax;
        } else {
            ht2 = ht1;
            htm2 = htm;
        }
        xi = htm2 / dtmax;
        om = 1.0 - xi;
        so = sqrt(om);
        fct = P.mu * fcn * (1.0 - om * so);
        A = 6.0 * P.mu * ms * fcn * so / dtmax;
        sA = sqrt(A);
        ct = P.gt * sA;
        if (htm2 > 0.0) u = scale3(ht2, 1.0 / htm2);
        double fh = htm2 > 0.0 ? fct / htm2 : 0.0;   // the forward's expression
        ft0 = add3(scale3(ht2, fh), scale3(vt, ct / dt));
        ftm = sqrt(dot3(ft0, ft0));
        double flim = P.mu * fcn;
        if (ftm > flim) {
            coul = true;
            rcap = flim / ftm;
        }
    }
    double wrm = sqrt(dot3(wrel, wrel));
    bool roll = wrm > kXSmall;

    // ---- reverse ----
    d3 ftb = add3(ftb_ext, fib);   // fi = fn + ft
    d3 fnb = fib;
    double fcnb = 0.0;
    d3 vtb = d3{0.0, 0.0, 0.0}, vnb = d3{0.0, 0.0, 0.0};
    if (roll) {   // mr = -ur fcn w_hat, w_hat = wrel / |wrel|
        double wfl = roll_c * P.ur * fcn;
        if (roll_fwd && !(wrm > wfl)) {
            // floored forward: mr = -(ur fcn / wfl) wrel
            acc3s(wrelb, -(P.ur * fcn / wfl), mrb);
        } else {
            d3 wh = scale3(wrel, 1.0 / wrm);
            double mrw = dot3(mrb, wh);
            Pb.ur += -mrw * fcn;
            fcnb += -mrw * P.ur;
            double mc = P.ur * fcn / (wrm > wfl ? wrm : wfl);   // = mcoef unless regularised
            acc3s(wrelb, -mc, mrb);   // d w_hat: (I - w_hat w_hat^T) / |wrel|
            acc3s(wrelb, mc * mrw, wh);
        }
    }
    if (tang) {
        double dtmaxb = 0.0, omb = 0.0, htm2b = 0.0;
        d3 ht2b = houtb;
        d3 ft0b;
        if (coul) {   // ft = ft0 * (flim / |ft0|), flim = mu fcn
            double rb = dot3(ft0, ftb);
            double flimb = rb / ftm;
            double ftmb = -rb * rcap / ftm;
            ft0b = add3(scale3(ftb, rcap), scale3(ft0, ftmb / ftm));
            Pb.mu += flimb * fcn;
            fcnb += flimb * P.mu;
        } else {
            ft0b = ftb;
        }
        // ft0 = fct u + (ct / dt) vt
        double fctb = dot3(ft0b, u);
        d3 ub = scale3(ft0b, fct);
        double ctb = dot3(ft0b, vt) / dt;
        acc3s(vtb, ct / dt, ft0b);
        if (htm2 > 0.0) {   // u = ht2 / htm2
            acc3s(ht2b, 1.0 / htm2, ub);
            htm2b += -dot3(ub, ht2) / (htm2 * htm2);
        }
        if (A > 0.0) {   // ct = gt sqrt(A), A = 6 mu ms fcn sqrt(om) / dtmax
            Pb.gt += ctb * sA;
            double Ab = ctb * P.gt / (2.0 * sA);
            Pb.mu += Ab * A / P.mu;
            fcnb += Ab * A / fcn;
            if (!capped && om > 0.0) omb += Ab * A / (2.0 * om);
            dtmaxb += -Ab * A / dtmax;
        }
        double g15 = 1.0 - om * so;   // fct = mu fcn (1 - om^1.5)
        Pb.mu += fctb * fcn * g15;
        fcnb += fctb * P.mu * g15;
        if (!capped) {
            omb += fctb * P.mu * fcn * (-1.5 * so);
            double xib = -omb;   // om = 1 - xi, xi = htm2 / dtmax
            htm2b += xib / dtmax;
            dtmaxb += -xib * xi / dtmax;
        }
        d3 ht1b;
        double htmb;
        if (capped) {   // ht2 = ht1 dtmax / htm, htm2 = dtmax
            double rhob = dot3(ht1, ht2b);
            ht1b = scale3(ht2b, rho);
            dtmaxb += rhob / htm + htm2b;
            htmb = -rhob * rho / htm;
        } else {
            ht1b = ht2b;
            htmb = htm2b;
        }
        if (htm > 0.0) acc3s(ht1b, htmb / htm, ht1);
        Pb.mu += dtmaxb * dtmax / P.mu;   // dtmax = mu k dn
        dnb += dtmaxb * dtmax / dn;
        acc3(vtb, ht1b);   // ht1 = ht0 + vt
        d3 r0b = ht1b;
        if (resc) {   // ht0 = r0 sqrt(|hin| / |r0|)
            double prb = dot3(r0, ht1b);
            r0b = scale3(ht1b, pr);
            double pmagb = prb / (2.0 * pr * rmag);
            double rmagb = -prb * pr / (2.0 * rmag);
            acc3s(hinb, pmagb / pmag, hin);
            acc3s(r0b, rmagb / rmag, r0);
        }
        // r0 = (-n) x q, q = n x hin   (y = a x b: abar = b x ybar, bbar = ybar x a)
        d3 mnb = cross3(q, r0b);
        d3 qb = cross3(r0b, mn);
        nb.x -= mnb.x; nb.y -= mnb.y; nb.z -= mnb.z;
        acc3(nb, cross3(hin, qb));
        acc3(hinb, cross3(qb, n));
    }
    // fn = -fcn n + (cn / dt) vn
    fcnb += -dot3(fnb, n);
    acc3s(nb, -fcn, fnb);
    double cnb = dot3(fnb, vn) / dt;
    acc3s(vnb, cn / dt, fnb);
    Pb.gn += cnb * sB;   // cn = gn sqrt(B), B = 6 ms E* sqrt(rs dn)
    if (Bq > 0.0) {
        double Bb = cnb * P.gn / (2.0 * sB);
        Pb.estar += Bb * Bq / P.estar;
        dnb += Bb * Bq / (2.0 * dn);
    }
    Pb.estar += fcnb * fcn / P.estar;   // fcn = E* sqrt(rs) dn^1.5
    dnb += fcnb * 1.5 * fcn / dn;
    acc3(vrelb, vtb);   // vt = vrel - vn
    vnb.x -= vtb.x; vnb.y -= vtb.y; vnb.z -= vtb.z;
    double vnsb = dot3(vnb, n);   // vn = vns n
    acc3s(nb, vns, vnb);
    acc3s(vrelb, vnsb, n);   // vns = vrel . n
    acc3s(nb, vnsb, vrel);
}

// Reverse of law_pair.  Output adjoints fab (force on a), tab, tbb (torques), houtb (new
// history).  Accumulates into the input adjoints.  No contact: no contribution.
DEM_HD void law_pair_adj(d3 xa, d3 xb, d3 va, d3 vb, d3 wa, d3 wb, double ra, double rb,
                         double ma, double mb, d3 hin, double dt, const LawParams &P, d3 fab,
                         d3 tab, d3 tbb, d3 houtb, d3 &xab, d3 &xbb, d3 &vab, d3 &vbb, d3 &wab,
                         d3 &wbb, d3 &hinb, LawParamsBar &Pb, double roll_c = 0.0,
                         bool roll_fwd = false)
{
    d3 dx = sub3(xb, xa);
    double dist = sqrt(dot3(dx, dx));
    double dn = (ra + rb) - dist;
    if (!(dn > 0.0)) return;
    double idist = 1.0 / dist;
    d3 n = d3{dx.x * idist, dx.y * idist, dx.z * idist};
    double rs = ra * rb / (ra + rb);
    double ms = ma * mb / (ma + mb);
    d3 s = d3{ra * wa.x + rb * wb.x, ra * wa.y + rb * wb.y, ra * wa.z + rb * wb.z};
    d3 c = cross3(s, n);
    d3 vrel = d3{(vb.x - va.x) - c.x, (vb.y - va.y) - c.y, (vb.z - va.z) - c.z};
    d3 wrel = sub3(wa, wb);
    d3 fa, ft, mr, hout;
    double fout;
    law_core(n, dn, rs, ms, vrel, wrel, hin, dt, P, fa, ft, mr, hout, fout,
             roll_fwd ? roll_c : 0.0);
    // torques: ta = (ra n) x ft + mr, tb = (rb n) x ft - mr
    d3 ua = scale3(n, ra), ub = scale3(n, rb);
    d3 nb = add3(scale3(cross3(ft, tab), ra), scale3(cross3(ft, tbb), rb));
    d3 ftb = add3(cross3(tab, ua), cross3(tbb, ub));
    d3 mrb = sub3(tab, tbb);
    double dnb = 0.0;
    d3 vrelb = d3{0.0, 0.0, 0.0}, wrelb = d3{0.0, 0.0, 0.0};
    law_core_adj(n, dn, rs, ms, vrel, wrel, hin, dt, P, fab, ftb, mrb, houtb, nb, dnb, vrelb,
                 wrelb, hinb, Pb, roll_c, roll_fwd);
    acc3(wab, wrelb);   // wrel = wa - wb
    wbb.x -= wrelb.x; wbb.y -= wrelb.y; wbb.z -= wrelb.z;
    acc3(vbb, vrelb);   // vrel = (vb - va) - c
    vab.x -= vrelb.x; vab.y -= vrelb.y; vab.z -= vrelb.z;
    d3 cb = d3{-vrelb.x, -vrelb.y, -vrelb.z};
    d3 sb = cross3(n, cb);   // c = s x n
    acc3(nb, cross3(cb, s));
    acc3s(wab, ra, sb);   // s = ra wa + rb wb
    acc3s(wbb, rb, sb);
    // n = dx / dist, dn = (ra + rb) - dist, dist = |dx|
    double distb = -dnb - dot3(nb, n) / dist;
    d3 dxb = add3(scale3(nb, 1.0 / dist), scale3(n, distb));
    acc3(xbb, dxb);
    xab.x -= dxb.x; xab.y -= dxb.y; xab.z -= dxb.z;
}

// Reverse of law_walls.  Output adjoints: fwb (wall force on the particle), twb (wall
// torque), houtb (the history slot's new value).  Accumulates into xb, vb, wb, hinb, Pb.
DEM_HD void law_walls_adj(const WallGeom &g, d3 x, d3 v, d3 w, double r, double m, d3 hin,
                          double dt, const LawParams &P, d3 fwb, d3 twb, d3 houtb, d3 &xb,
                          d3 &vb, d3 &wb, d3 &hinb, LawParamsBar &Pb, double roll_c = 0.0,
                          bool roll_fwd = false)
{
    int nw = wall_count(g);
    int prim = -1;
    double best = 0.0;
#pragma unroll 1
    for (int k = 0; k < nw; k++) {
        d3 n;
        double dn = wall_overlap(g, k, x, r, n);
        if (dn > 0.0 && (prim < 0 || dn > best)) { prim = k; best = dn; }
    }
    if (prim < 0) return;
#pragma unroll 1
    for (int k = 0; k < nw; k++) {
        d3 n;
        double dn = wall_overlap(g, k, x, r, n);
        if (!(dn > 0.0)) continue;
        d3 hk = k == prim ? hin : d3{0.0, 0.0, 0.0};
        d3 s = scale3(w, r);
        d3 c = cross3(s, n);
        d3 vrel = d3{(0.0 - v.x) - c.x, (0.0 - v.y) - c.y, (0.0 - v.z) - c.z};
        d3 f, ft, mr, ht;
        double fo;
        law_core(n, dn, r, m, vrel, w, hk, dt, P, f, ft, mr, ht, fo, roll_fwd ? roll_c : 0.0);
        // torque (r n) x ft + mr
        d3 un = scale3(n, r);
        d3 nb = scale3(cross3(ft, twb), r);
        d3 ftb = cross3(twb, un);
        double dnb = 0.0;
        d3 vrelb = d3{0.0, 0.0, 0.0}, wrelb = d3{0.0, 0.0, 0.0}, hkb = d3{0.0, 0.0, 0.0};
        law_core_adj(n, dn, r, m, vrel, w, hk, dt, P, fwb, ftb, twb,
                     k == prim ? houtb : d3{0.0, 0.0, 0.0}, nb, dnb, vrelb, wrelb, hkb, Pb,
                     roll_c, roll_fwd);
        if (k == prim) acc3(hinb, hkb);
        acc3(wb, wrelb);   // wrel = w
        vb.x -= vrelb.x; vb.y -= vrelb.y; vb.z -= vrelb.z;   // vrel = -v - c
        d3 cb = d3{-vrelb.x, -vrelb.y, -vrelb.z};
        d3 sb = cross3(n, cb);   // c = s x n, s = r w
        acc3(nb, cross3(cb, s));
        acc3s(wb, r, sb);
        // geometry: dn(x) and n(x)
        if (k == 0) {
            xb.z -= dnb;   // dn = r - z
        } else if (g.kind == 0) {
            if (k == 1) xb.x -= dnb;        // r - (x + hx)
            else if (k == 2) xb.x += dnb;   // r - (hx - x)
            else if (k == 3) xb.y -= dnb;
            else xb.y += dnb;
        } else {
            double rho = sqrt(x.x * x.x + x.y * x.y);
            if (rho > 0.0) {   // dn = rho + r - R, n = (x, y, 0) / rho
                double rhob = dnb - (nb.x * x.x + nb.y * x.y) / (rho * rho);
                xb.x += nb.x / rho + rhob * x.x / rho;
                xb.y += nb.y / rho + rhob * x.y / rho;
            }
        }
    }
}

}  // namespace dem
