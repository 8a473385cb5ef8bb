// law.hpp -- FP64 contact laws on the device (gfx950).
//
// Pair law: the body of the reference's absent calforceij (call neiglist.cu:208-209),
// following the Zhou/Yu Hertz-Mindlin model named in README.md:11: Hertz normal force
// E* sqrt(R*) dn^1.5, normal dashpot gn*sqrt(6 m* E* sqrt(R* dn)), Mindlin tangential spring
// mu|Fcn|(1-(1-xi)^1.5) with xi = |ht| / dt_max, dt_max = mu(2-nu)/(2(1-nu)) dn, tangential
// dashpot gt*sqrt(6 mu m* |Fcn| sqrt(1-xi) / dt_max), Coulomb cap mu|Fcn|, and rolling
// resistance -ur |Fcn| w_rel/|w_rel|.  The tangential history is carried over by
// vprojec (neiglist.cu:133-146, literal, FP64) and capped by vscale (neiglist.cu:148-153).
// Velocities are per-step displacements (pdis, angv), so dashpot terms divide by dt.
//
// Every expression is written in the same association order as the CPU oracle and the
// library is compiled with -ffp-contract=off, so the two agree to the last bit wherever
// sqrt and division are correctly rounded on both sides.  Divisions shared by the three
// components of a vector are taken once (1/dist, cn/dt, ct/dt, fct/|ht|): FP64 division is
// a multi-instruction sequence on the GPU.  The law is exactly antisymmetric:
// swapping the two bodies negates n, vrel, wrel and the stored history bit-for-bit, so a
// pair's result does not depend on which member is "first".
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

// Host + device: the kernels use the laws, and host entry points evaluate the same code for
// tests of the adjoint.
#define DEM_HD __host__ __device__ __forceinline__

namespace dem {

struct d3 {
    double x, y, z;
};

DEM_HD d3 mk3(double x, double y, double z) { return d3{x, y, z}; }
DEM_HD double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEM_HD d3 cross3(d3 a, d3 b)
{
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DEM_HD d3 scale3(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }

// Material constants of one contact type (particle-particle or particle-wall); the
// reference's Matproperties / Wallproperties members in order.
struct LawParams {
    double estar, mu, nu, gn, gt, ur;
};

// |v| as vprojec takes it: FP64, or with flen the reference's float length() -- sqrtf of
// the FP64 dot product rounded to float, returned as double (cutil_math.h:1086-1089).
DEM_HD double vlen(d3 v, bool flen)
{
    const double d = dot3(v, v);
    return flen ? (double)sqrtf((float)d) : sqrt(d);
}

constexpr double kXSmall = 1e-30;   // neiglist.cu:29

// Common core.  n: unit normal from body i toward the other body; dn > 0 overlap;
// rs, ms effective radius and mass; vrel: displacement increment of the other body's
// contact point relative to i's; wrel = w_i - w_other; hin: incoming tangential history.
// Outputs: force on i, tangential force on i, rolling moment on i, new history, |Ft|.
// roll_c > 0 (the differentiable form's option, off by default) floors the |w_rel| of the
// rolling moment at roll_c mu_r fcn: below it the moment is viscous, -w_rel / roll_c, instead
// of flipping its sign every step about rolling rest (law_core_adj gives the reasoning).
// roll_c = 0 is the reference's moment -mu_r fcn w_rel / |w_rel| bit for bit.  flen: vprojec's
// magnitudes (neiglist.cu:140-142) by the reference's float length() instead of FP64.
DEM_HD void law_core(d3 n, double dn, double rs, double ms, d3 vrel,
                                         d3 wrel, d3 hin, double dt, const LawParams &P,
                                         d3 &fi, d3 &ft, d3 &mr, d3 &hout, double &fout,
                                         double roll_c = 0.0, bool flen = false)
{
    double vns = vrel.x * n.x + vrel.y * n.y + vrel.z * n.z;
    d3 vn = d3{vns * n.x, vns * n.y, vns * n.z};
    d3 vt = d3{vrel.x - vn.x, vrel.y - vn.y, vrel.z - vn.z};
    double sdn = sqrt(dn);
    double fcn = P.estar * sqrt(rs) * dn * sdn;
    double cn = P.gn * sqrt(6.0 * ms * P.estar * sqrt(rs * dn));
    double cnd = cn / dt;   // one division for the three components
    d3 fn = d3{-fcn * n.x + cnd * vn.x, -fcn * n.y + cnd * vn.y, -fcn * n.z + cnd * vn.z};

    // vprojec(hin, n, &ht, 1)
    d3 q = cross3(n, hin);
    d3 ht = cross3(d3{-n.x, -n.y, -n.z}, q);
    double pmag = vlen(hin, flen);
    if (!(pmag < 0.001)) {
        double rmag = vlen(ht, flen);
        if (!(pmag > rmag * 100.0)) {
            double pr = sqrt(pmag / rmag);
            ht = scale3(ht, pr);
        }
    }
    ht = d3{ht.x + vt.x, ht.y + vt.y, ht.z + vt.z};

    double dtmax = P.mu * (2.0 - P.nu) / (2.0 * (1.0 - P.nu)) * dn;
    if (P.mu > 0.0 && dtmax > 0.0) {
        double htm = sqrt(dot3(ht, ht));
        if (htm > dtmax) {
            double ratio = dtmax / htm;   // vscale(ht, dtmax, htm)
            ht = scale3(ht, ratio);
            htm = dtmax;
        }
        double xi = htm / dtmax;
        double om = 1.0 - xi;
        double fct = P.mu * fcn * (1.0 - om * sqrt(om));
        double ct = P.gt * sqrt(6.0 * P.mu * ms * fcn * sqrt(om) / dtmax);
        double fh = htm > 0.0 ? fct / htm : 0.0;   // spring force per unit history
        double ctd = ct / dt;
        ft = d3{fh * ht.x + ctd * vt.x, fh * ht.y + ctd * vt.y, fh * ht.z + ctd * vt.z};
        double ftm = sqrt(dot3(ft, ft));
        double flim = P.mu * fcn;
        if (ftm > flim) {
            double ratio = flim / ftm;   // vscale(ft, flim, ftm)
            ft = scale3(ft, ratio);
        }
    } else {
        ht = d3{0.0, 0.0, 0.0};
        ft = d3{0.0, 0.0, 0.0};
    }
    fout = sqrt(dot3(ft, ft));
    fi = d3{fn.x + ft.x, fn.y + ft.y, fn.z + ft.z};
    hout = ht;

    double wrm = sqrt(dot3(wrel, wrel));
    if (wrm > kXSmall) {
        double wfl = roll_c * P.ur * fcn;
        double m = P.ur * fcn / (wrm > wfl ? wrm : wfl);
        mr = d3{-m * wrel.x, -m * wrel.y, -m * wrel.z};
    } else {
        mr = d3{0.0, 0.0, 0.0};
    }
}

// Contact test of law_pair (the same expressions, so the same decision bit for bit): true
// when spheres a and b overlap.  Needs only the positions and radii.
DEM_HD bool law_touch(d3 xa, d3 xb, double ra, double rb)
{
    d3 dx = d3{xb.x - xa.x, xb.y - xa.y, xb.z - xa.z};
    double dist = sqrt(dot3(dx, dx));
    double dn = (ra + rb) - dist;
    return dn > 0.0;
}

// FP32 copy of a slot's position and radius: the step kernels' contact filter.
DEM_HD float4 pos32(double x, double y, double z, double r)
{
    return make_float4((float)x, (float)y, (float)z, (float)r);
}

// Conservative contact filter on the FP32 copies: false only when the spheres are certainly
// apart, i.e. when law_touch on the FP64 values is certainly false; a true result sends the
// pair to the FP64 law, whose own contact test decides (and gives the exact zero result of a
// pair out of contact).  With u = 2^-24, each FP32 coordinate is within u|x| of the FP64 one
// and the FP32 difference within u|dx| of the exact difference of the copies, so the FP32
// separation vector is within u * sum_k (|xa_k| + |xb_k| + |dx_k|) of the exact one; the
// margin doubles that and adds 8u of the radius sum for the roundings of the sums, products
// and radii.  NaN compares false and keeps the pair.
DEM_HD bool maybe_touch(float4 a, float4 b)
{
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float s = a.w + b.w;
    const float u = 5.9604645e-8f;   // 2^-24
    const float e = fabsf(a.x) + fabsf(b.x) + fabsf(dx) + fabsf(a.y) + fabsf(b.y) + fabsf(dy) +
                    fabsf(a.z) + fabsf(b.z) + fabsf(dz);
    const float cut = s + (2.0f * u) * e + (8.0f * u) * s;
    return !(d2 > cut * cut);
}

// Pair law between bodies a and b.  Returns true on contact; on no contact every output
// is zero (history reset), matching the reference's no-contact branch (neiglist.cu:220-227).
DEM_HD bool law_pair(d3 xa, d3 xb, d3 va, d3 vb, d3 wa, d3 wb, double ra,
                                         double rb, double ma, double mb, d3 hin, double dt,
                                         const LawParams &P, d3 &fa, d3 &ta, d3 &tb,
                                         d3 &hout, double &fout, double roll_c = 0.0,
                                         bool flen = false)
{
    d3 dx = d3{xb.x - xa.x, xb.y - xa.y, xb.z - xa.z};
    double dist = sqrt(dot3(dx, dx));
    double dn = (ra + rb) - dist;
    if (!(dn > 0.0)) {
        fa = ta = tb = hout = d3{0.0, 0.0, 0.0};
        fout = 0.0;
        return false;
    }
    double idist = 1.0 / dist;
    d3 n = d3{dx.x * idist, dx.y * idist, dx.z * idist};
    double rs = ra * rb / (ra + rb);
    double ms = ma * mb / (ma + mb);
    d3 s = d3{ra * wa.x + rb * wb.x, ra * wa.y + rb * wb.y, ra * wa.z + rb * wb.z};
    d3 c = cross3(s, n);
    d3 vrel = d3{(vb.x - va.x) - c.x, (vb.y - va.y) - c.y, (vb.z - va.z) - c.z};
    d3 wrel = d3{wa.x - wb.x, wa.y - wb.y, wa.z - wb.z};
    d3 ft, mr;
    law_core(n, dn, rs, ms, vrel, wrel, hin, dt, P, fa, ft, mr, hout, fout, roll_c, flen);
    d3 ca = cross3(scale3(n, ra), ft);
    ta = d3{ca.x + mr.x, ca.y + mr.y, ca.z + mr.z};
    d3 cb = cross3(scale3(n, rb), ft);
    tb = d3{cb.x - mr.x, cb.y - mr.y, cb.z - mr.z};
    return true;
}

// Wall set: box (floor, x-, x+, y-, y+) or cylinder (floor, side).  Computes the overlap
// and outward unit normal of wall k; returns the number of walls.
struct WallGeom {
    int kind;
    double hx, hy;     // box half sizes
    double radius;     // cylinder radius
};

DEM_HD int wall_count(const WallGeom &g) { return g.kind == 0 ? 5 : 2; }

DEM_HD double wall_overlap(const WallGeom &g, int k, d3 x, double r, d3 &n)
{
    if (k == 0) { n = d3{0.0, 0.0, -1.0}; return r - x.z; }
    if (g.kind == 0) {
        switch (k) {
        case 1: n = d3{-1.0, 0.0, 0.0}; return r - (x.x + g.hx);
        case 2: n = d3{1.0, 0.0, 0.0}; return r - (g.hx - x.x);
        case 3: n = d3{0.0, -1.0, 0.0}; return r - (x.y + g.hy);
        default: n = d3{0.0, 1.0, 0.0}; return r - (g.hy - x.y);
        }
    }
    double rho = sqrt(x.x * x.x + x.y * x.y);
    if (rho > 0.0) n = d3{x.x / rho, x.y / rho, 0.0};
    else n = d3{1.0, 0.0, 0.0};
    return (rho + r) - g.radius;
}

// Wall contact (body of the reference's absent calforceiw, call neiglist.cu:287-288).
// Every overlapped wall contributes; the particle's single history slot belongs to the wall
// with the largest overlap (lowest index on ties), other walls start from zero history.
// Returns the number of walls hit.
DEM_HD int law_walls(const WallGeom &g, d3 x, d3 v, d3 w, double r,
                                         double m, d3 hin, double dt, const LawParams &P,
                                         d3 &fw, d3 &tw, d3 &hout, double &fout,
                                         double roll_c = 0.0,
                                         bool flen = false)
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
    fw = tw = hout = d3{0.0, 0.0, 0.0};
    fout = 0.0;
    if (prim < 0) return 0;
    int hits = 0;
#pragma unroll 1
    for (int k = 0; k < nw; k++) {
        d3 n;
        double dn = wall_overlap(g, k, x, r, n);
        if (!(dn > 0.0)) continue;
        hits++;
        d3 s = d3{r * w.x, r * w.y, r * w.z};
        d3 c = cross3(s, n);
        d3 vrel = d3{(0.0 - v.x) - c.x, (0.0 - v.y) - c.y, (0.0 - v.z) - c.z};
        d3 f, ft, mr, ht;
        double fo;
        law_core(n, dn, r, m, vrel, w, k == prim ? hin : d3{0.0, 0.0, 0.0}, dt, P, f, ft, mr,
                 ht, fo, roll_c, flen);
        d3 cr = cross3(scale3(n, r), ft);
        fw = d3{fw.x + f.x, fw.y + f.y, fw.z + f.z};
        tw = d3{tw.x + (cr.x + mr.x), tw.y + (cr.y + mr.y), tw.z + (cr.z + mr.z)};
        if (k == prim) { hout = ht; fout = fo; }
    }
    return hits;
}

// Grid cell coordinate (neiglist.cu:54-74): floor((p - o) / c) clamped to [0, g-1].
DEM_HD unsigned cell_coord(double p, double o, double c, unsigned g)
{
    double f = floor((p - o) / c);
    if (!(f > 0.0)) return 0u;
    if (f > (double)(g - 1u)) return g - 1u;
    return (unsigned)f;
}

// Candidate test shared by every neighbour kernel: |xb - xa|^2 < (ra + rb + skin)^2.
DEM_HD bool is_candidate(d3 xa, d3 xb, double ra, double rb, double skin)
{
    double dx = xb.x - xa.x;
    double dy = xb.y - xa.y;
    double dz = xb.z - xa.z;
    double d2 = dx * dx + dy * dy + dz * dz;
    double cut = (ra + rb) + skin;
    return d2 < cut * cut;
}

}  // namespace dem
