// adjoint.hpp -- reverse-mode kernels of the DEM step and the smooth packing observables.
//
// One step maps the state (x, v, w and wall history per slot, tangential history per pair)
// through k_pair_force -> k_walls -> k_particle.  Its reverse, given the adjoint of the
// state after the step (SAdj per slot, history adjoint per pair), runs
//   k_adj_particle : integrator reverse (F, T adjoints), wall-law reverse, per slot
//   k_adj_pair     : pair-law reverse, one lane per pair, writing a PAdj record
//   k_adj_gather   : per slot, the PAdj records of its own pairs (CSR) and of the pairs where
//                    it is second (transposed index) -- the forward gather's pattern, so no
//                    atomics and a fixed summation order
// and, for a step that began with a list rebuild, k_adj_unperm / k_adj_unremap move the
// adjoints back to the previous slot order and pair list.  Material adjoints accumulate per
// particle index (each index sits in exactly one slot per step, so no two lanes collide).
// The records, histories and lists a reverse step reads are the forward tapes recorded by
// the recompute of its checkpoint segment.  roll_k (0: exact) regularises the rolling
// moment's linearisation in the chatter regime (law_core_adj).
#pragma once
#include "kernels.hpp"
#include "law_adj.hpp"

namespace dem {

// Adjoint of one slot's state: position, displacement per step, rotation per step, wall
// history (96 bytes).
struct __attribute__((aligned(16))) SAdj {
    double x[3], v[3], w[3], h[3];
};
static_assert(sizeof(SAdj) == 96, "slot adjoint is six 16-byte chunks");

// Per-pair reverse output: adjoints of the first member's x, v, w (9), of the pair-law
// material (E*, mu, gamma_n, gamma_t, mu_r), of the second member's x, v, w (9).
struct __attribute__((aligned(16))) PAdj {
    double a[9];
    double th[5];
    double b[9];
    double pad;
};
static_assert(sizeof(PAdj) == 192, "pair adjoint record is twelve 16-byte chunks");

// Material adjoints per particle index: pair law (5) then wall law (5).
constexpr int kThetaPer = 10;

__device__ __forceinline__ d3 ld3(const double *p) { return d3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void stv3(double *p, d3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// Reverse of k_particle (integrator, gravity, wall add) and of k_walls for one step.
//   forward: v' = v + F dt^2/m, x' = x + v', w' = w + T dt^2/I, wall history' = hits ? h : 0
// lam holds the adjoint of the state after the step and receives the adjoint of the state
// before it, minus the pair contributions (added by k_adj_gather).  ftb[6*a] = (Fbar, Tbar).
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_adj_particle(const PRec *tr, const Hist *twh,
                                                         const unsigned *wflag, SAdj *lam,
                                                         double *ftb, double *thacc, unsigned n,
                                                         const InstDev *inst, InstDev one,
                                                         double roll_k)
{
    unsigned a = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (a >= n) return;
    const PRec R = tr[a];
    const InstDev &I = BATCHED ? inst[R.inst] : one;
    SAdj L = lam[a];
    double dt = I.dt;
    double cm = dt * dt / R.mass, ci = dt * dt / R.inert;
    d3 vt = d3{L.v[0] + L.x[0], L.v[1] + L.x[1], L.v[2] + L.x[2]};
    d3 Fb = scale3(vt, cm);
    d3 Tb = d3{L.w[0] * ci, L.w[1] * ci, L.w[2] * ci};
    d3 xb = ld3(L.x), vb = vt, wb = ld3(L.w), hb = ld3(L.h);
    if (wflag[a]) {
        d3 houtb = hb;
        hb = d3{0.0, 0.0, 0.0};
        LawParamsBar Pb = {0.0, 0.0, 0.0, 0.0, 0.0};
        const Hist H = twh[a];
        // a forward with the rolling floor (I.roll_k > 0) is reversed exactly; otherwise
        // roll_k floors the derivative only
        const bool fwd = I.roll_k > 0.0;
        const double rc = fwd ? roll_floor(I.roll_k, dt, R.inert, 0.0) : roll_k * dt * dt / R.inert;
        law_walls_adj(I.walls, d3{R.px, R.py, R.pz}, d3{R.vx, R.vy, R.vz}, d3{R.wx, R.wy, R.wz},
                      R.rad, R.mass, d3{H.hx, H.hy, H.hz}, dt, I.pw, Fb, Tb, houtb, xb, vb, wb,
                      hb, Pb, rc, fwd);
        double *T = thacc + (size_t)kThetaPer * R.id;
        T[5] += Pb.estar; T[6] += Pb.mu; T[7] += Pb.gn; T[8] += Pb.gt; T[9] += Pb.ur;
    }
    SAdj O;
    stv3(O.x, xb); stv3(O.v, vb); stv3(O.w, wb); stv3(O.h, hb);
    lam[a] = O;
    double *F = ftb + 6 * (size_t)a;
    F[0] = Fb.x; F[1] = Fb.y; F[2] = Fb.z; F[3] = Tb.x; F[4] = Tb.y; F[5] = Tb.z;
}

// Reverse of k_pair_force for one step: pair p = (a, b) adds fa to F_a, -fa to F_b, ta to
// T_a, tb to T_b and writes the new history.  hb (4 doubles per pair, the 4th unused) holds
// the adjoint of the history after the step and receives that of the history before it.
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_adj_pair(const uint2 *ab, unsigned np, const PRec *tr,
                                                     const Hist *th, const double *ftb, double *hb,
                                                     PAdj *pa, const InstDev *inst, LawParams pp,
                                                     double dt, double fwd_k, double roll_k)
{
    const unsigned per = (np + gridDim.x - 1) / gridDim.x;   // contiguous, XCD-aware ranges
    const unsigned p_begin = xcd_block(blockIdx.x, gridDim.x) * per;
    const unsigned p_end = min(p_begin + per, np);
    for (unsigned p = p_begin + threadIdx.x; p < p_end; p += kBlock) {
        uint2 e = ab[p];
        const PRec &A = tr[e.x];
        const PRec &B = tr[e.y];
        LawParams P = pp;
        double tdt = dt, fk = fwd_k;
        if (BATCHED) {
            const InstDev &I = inst[A.inst];
            P = I.pp;
            tdt = I.dt;
            fk = I.roll_k;
        }
        const bool fwd = fk > 0.0;   // the forward's rolling floor, reversed exactly
        const double rc = fwd ? roll_floor(fk, tdt, A.inert, B.inert)
                              : roll_k * tdt * tdt * (1.0 / A.inert + 1.0 / B.inert);
        const double *Fa = ftb + 6 * (size_t)e.x;
        const double *Fb = ftb + 6 * (size_t)e.y;
        d3 fab = d3{Fa[0] - Fb[0], Fa[1] - Fb[1], Fa[2] - Fb[2]};
        d3 tab = d3{Fa[3], Fa[4], Fa[5]};
        d3 tbb = d3{Fb[3], Fb[4], Fb[5]};
        double *HB = hb + 4 * (size_t)p;
        d3 houtb = d3{HB[0], HB[1], HB[2]};
        const Hist H = th[p];
        d3 z = d3{0.0, 0.0, 0.0};
        d3 xab = z, xbb = z, vab = z, vbb = z, wab = z, wbb = z, hinb = z;
        LawParamsBar Pb = {0.0, 0.0, 0.0, 0.0, 0.0};
        law_pair_adj(d3{A.px, A.py, A.pz}, d3{B.px, B.py, B.pz}, d3{A.vx, A.vy, A.vz},
                     d3{B.vx, B.vy, B.vz}, d3{A.wx, A.wy, A.wz}, d3{B.wx, B.wy, B.wz}, A.rad,
                     B.rad, A.mass, B.mass, d3{H.hx, H.hy, H.hz}, tdt, P, fab, tab, tbb, houtb,
                     xab, xbb, vab, vbb, wab, wbb, hinb, Pb, rc, fwd);
        PAdj O;
        stv3(O.a, xab); stv3(O.a + 3, vab); stv3(O.a + 6, wab);
        O.th[0] = Pb.estar; O.th[1] = Pb.mu; O.th[2] = Pb.gn; O.th[3] = Pb.gt; O.th[4] = Pb.ur;
        stv3(O.b, xbb); stv3(O.b + 3, vbb); stv3(O.b + 6, wbb);
        O.pad = 0.0;
        pa[p] = O;
        HB[0] = hinb.x; HB[1] = hinb.y; HB[2] = hinb.z; HB[3] = 0.0;
    }
}

// Per slot: add the pair adjoints of the slot's own pairs (first member) and of the pairs
// where it is second, in the forward gather's order; pair-law material adjoints of the own
// pairs go to the particle index.  Records are read two at a time (both loads in flight
// before either is added; the additions keep the order).
__device__ __forceinline__ void adj_add_own(double *s, double *t, const double2 *x)
{
    s[0] += x[0].x; s[1] += x[0].y; s[2] += x[1].x; s[3] += x[1].y; s[4] += x[2].x;
    s[5] += x[2].y; s[6] += x[3].x; s[7] += x[3].y; s[8] += x[4].x;
    t[0] += x[4].y; t[1] += x[5].x; t[2] += x[5].y; t[3] += x[6].x; t[4] += x[6].y;
}

__device__ __forceinline__ void adj_add_second(double *s, const double2 *x)
{
    s[0] += x[0].x; s[1] += x[0].y; s[2] += x[1].x; s[3] += x[1].y; s[4] += x[2].x;
    s[5] += x[2].y; s[6] += x[3].x; s[7] += x[3].y; s[8] += x[4].x;
}

__global__ __launch_bounds__(kBlock) void k_adj_gather(SAdj *lam, const PAdj *pa, const unsigned *off,
                                                       const unsigned *toff, const unsigned *jt,
                                                       unsigned np, const PRec *tr,
                                                       double *thacc, unsigned n)
{
    static_assert(offsetof(PAdj, th) == 9 * sizeof(double) && offsetof(PAdj, b) == 14 * sizeof(double),
                  "PAdj layout: a[9], th[5], b[9]");
    unsigned a = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (a >= n) return;
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double t[5] = {0, 0, 0, 0, 0};
    unsigned p0 = off[a], p1 = off[a + 1];
    if (p1 > np) p1 = np;
    for (unsigned p = p0; p < p1; p += 2) {
        const bool two = p + 1 < p1;
        const double *A0 = reinterpret_cast<const double *>(pa + p);
        const double *A1 = reinterpret_cast<const double *>(pa + (two ? p + 1 : p));
        double2 x0[7], x1[7];
#pragma unroll
        for (int j = 0; j < 7; j++) x0[j] = ld2(A0 + 2 * j);
#pragma unroll
        for (int j = 0; j < 7; j++) x1[j] = ld2(A1 + 2 * j);
        adj_add_own(s, t, x0);
        if (two) adj_add_own(s, t, x1);
    }
    unsigned q0 = toff[a], q1 = toff[a + 1];
    if (q1 > np) q1 = np;
    for (unsigned k = q0; k < q1; k += 2) {
        const bool two = k + 1 < q1;
        unsigned qa = jt[k], qb = two ? jt[k + 1] : kEmpty;
        const bool ua = qa < np, ub = qb < np;
        const double *B0 = reinterpret_cast<const double *>(pa + (ua ? qa : 0u)) + 14;
        const double *B1 = reinterpret_cast<const double *>(pa + (ub ? qb : 0u)) + 14;
        double2 x0[5], x1[5];
#pragma unroll
        for (int j = 0; j < 5; j++) x0[j] = ld2(B0 + 2 * j);
#pragma unroll
        for (int j = 0; j < 5; j++) x1[j] = ld2(B1 + 2 * j);
        if (ua) adj_add_second(s, x0);
        if (ub) adj_add_second(s, x1);
    }
    SAdj &L = lam[a];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        L.x[k] += s[k];
        L.v[k] += s[3 + k];
        L.w[k] += s[6 + k];
    }
    double *T = thacc + (size_t)kThetaPer * tr[a].id;
#pragma unroll
    for (int k = 0; k < 5; k++) T[k] += t[k];
}

// Reverse of k_reorder: slot s took the record and wall history of old slot perm[s].
__global__ __launch_bounds__(kBlock) void k_adj_unperm(const SAdj *lam_new, SAdj *lam_old,
                                                       const unsigned *perm, unsigned n)
{
    unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    lam_old[perm[s]] = lam_new[s];
}

__global__ __launch_bounds__(kBlock) void k_zero_d(double *a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        a[i] = 0.0;
}

// Reverse of the history remap of k_nbr_fill: new pair p took the history of old pair
// src[p] (kEmpty: started at zero).  Pair identities are unique, so no two new pairs share
// a source.  hb_old must be zero on entry.
__global__ __launch_bounds__(kBlock) void k_adj_unremap(const double *hb_new, double *hb_old,
                                                        const unsigned *src, unsigned np_new,
                                                        unsigned np_old)
{
    unsigned p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np_new) return;
    unsigned q = src[p];
    if (q >= np_old) return;
    const double *S = hb_new + 4 * (size_t)p;
    double *D = hb_old + 4 * (size_t)q;
    D[0] = S[0]; D[1] = S[1]; D[2] = S[2]; D[3] = 0.0;
}

// ----------------------------------------------------------------------------------------
// Smooth packing observables (per instance) and their gradient with respect to positions
// ----------------------------------------------------------------------------------------
//   window  w(x) = prod_k sig(t_k / eps): t_k the signed distances inside the region (side
//           walls moved in by `margin`, zlo < z < zhi), sig the logistic function
//   phi     = sum_i w_i V_i / V_region
//   CN      = sum_i w_i c_i / sum_i w_i,  c_i = sum_j S(delta_ij / eps_c) (soft contacts)
// S is the C1 smoothstep (0 for overlap <= 0, 1 beyond eps_c): only pairs in contact count,
// and every pair in contact is on the current list, so the value does not depend on when
// the list was built.

struct ObsDev {
    double margin, zlo, zhi, eps, eps_c;
};

__device__ __forceinline__ double logistic(double t) { return 1.0 / (1.0 + exp(-t)); }

__device__ __forceinline__ void obs_window(const WallGeom &g, const ObsDev &o, d3 x, double &w,
                                           d3 &gw)
{
    double lw = 1.0;
    d3 acc = d3{0.0, 0.0, 0.0};
    auto fac = [&](double t, d3 dt) {   // d log sig(t/eps) / dt = (1 - sig) / eps
        double s = logistic(t / o.eps);
        lw *= s;
        acc3s(acc, (1.0 - s) / o.eps, dt);
    };
    if (g.kind == 0) {
        fac(x.x + g.hx - o.margin, d3{1.0, 0.0, 0.0});
        fac(g.hx - o.margin - x.x, d3{-1.0, 0.0, 0.0});
        fac(x.y + g.hy - o.margin, d3{0.0, 1.0, 0.0});
        fac(g.hy - o.margin - x.y, d3{0.0, -1.0, 0.0});
    } else {
        double rho = sqrt(x.x * x.x + x.y * x.y);
        d3 dr = rho > 0.0 ? d3{-x.x / rho, -x.y / rho, 0.0} : d3{0.0, 0.0, 0.0};
        fac(g.radius - o.margin - rho, dr);
    }
    fac(x.z - o.zlo, d3{0.0, 0.0, 1.0});
    fac(o.zhi - x.z, d3{0.0, 0.0, -1.0});
    w = lw;
    gw = scale3(acc, lw);
}

// Per pair: soft contact s = S(delta / eps_c) and g = s' (w_a + w_b) n, n from a to b
// (d delta / d x_a = n).
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_obs_pair(const uint2 *ab, unsigned np, const PRec *rec,
                                                     const InstDev *inst, InstDev one, ObsDev o,
                                                     double *pobs)
{
    for (unsigned p = blockIdx.x * kBlock + threadIdx.x; p < np; p += gridDim.x * kBlock) {
        uint2 e = ab[p];
        const PRec &A = rec[e.x];
        const PRec &B = rec[e.y];
        const InstDev &I = BATCHED ? inst[A.inst] : one;
        d3 xa = d3{A.px, A.py, A.pz}, xb = d3{B.px, B.py, B.pz};
        d3 dx = sub3(xb, xa);
        double dist = sqrt(dot3(dx, dx));
        double delta = (A.rad + B.rad) - dist;
        double t = delta / o.eps_c;
        double s = 0.0, sp = 0.0;
        if (t >= 1.0) {
            s = 1.0;
        } else if (t > 0.0) {
            s = t * t * (3.0 - 2.0 * t);
            sp = 6.0 * t * (1.0 - t) / o.eps_c;
        }
        double wa, wb;
        d3 ga, gb;
        obs_window(I.walls, o, xa, wa, ga);
        obs_window(I.walls, o, xb, wb, gb);
        double c = dist > 0.0 ? sp * (wa + wb) / dist : 0.0;
        double *P = pobs + 4 * (size_t)p;
        P[0] = s;
        P[1] = c * dx.x;
        P[2] = c * dx.y;
        P[3] = c * dx.z;
    }
}

// Per slot: soft contact count c, pair-gradient sum G, window w and its gradient; partial
// sums (w V, w, w c) for the per-instance reduction.  sobs: 8 doubles per slot.
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_obs_slot(const PRec *rec, const unsigned *off,
                                                     const unsigned *toff, const unsigned *jt,
                                                     unsigned np, const double *pobs,
                                                     const InstDev *inst, InstDev one, ObsDev o,
                                                     unsigned n, double *sobs, double *part)
{
    unsigned a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    const PRec &R = rec[a];
    const InstDev &I = BATCHED ? inst[R.inst] : one;
    double c = 0.0;
    d3 G = d3{0.0, 0.0, 0.0};
    unsigned p0 = off[a], p1 = off[a + 1];
    if (p1 > np) p1 = np;
    for (unsigned p = p0; p < p1; p++) {
        const double *P = pobs + 4 * (size_t)p;
        c += P[0];
        G.x += P[1]; G.y += P[2]; G.z += P[3];
    }
    unsigned q0 = toff[a], q1 = toff[a + 1];
    if (q1 > np) q1 = np;
    for (unsigned k = q0; k < q1; k++) {
        unsigned q = jt[k];
        if (q >= np) continue;
        const double *P = pobs + 4 * (size_t)q;
        c += P[0];
        G.x -= P[1]; G.y -= P[2]; G.z -= P[3];
    }
    double w;
    d3 gw;
    obs_window(I.walls, o, d3{R.px, R.py, R.pz}, w, gw);
    double V = 4.0 / 3.0 * 3.14159265358979323846 * R.rad * R.rad * R.rad;
    double *S = sobs + 8 * (size_t)a;
    S[0] = c; S[1] = G.x; S[2] = G.y; S[3] = G.z;
    S[4] = V; S[5] = gw.x; S[6] = gw.y; S[7] = gw.z;
    double *Q = part + 3 * (size_t)a;
    Q[0] = w * V;
    Q[1] = w;
    Q[2] = w * c;
}

// One block per instance: fixed-order sums of the three partials over its slot range
// (instances occupy contiguous slot ranges: cell keys are offset per instance).
__global__ __launch_bounds__(kBlock) void k_obs_reduce(const double *part, const unsigned *range,
                                                       double *sums)
{
    __shared__ double sh[3][kBlock];
    unsigned k = blockIdx.x;
    unsigned s0 = range[k], s1 = range[k + 1];
    double acc[3] = {0.0, 0.0, 0.0};
    for (unsigned s = s0 + threadIdx.x; s < s1; s += kBlock) {
        const double *Q = part + 3 * (size_t)s;
        acc[0] += Q[0]; acc[1] += Q[1]; acc[2] += Q[2];
    }
    for (int j = 0; j < 3; j++) sh[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (unsigned h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int j = 0; j < 3; j++) sh[j][threadIdx.x] += sh[j][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 3; j++) sums[3 * k + j] = sh[j][0];
}

// Seed of the reverse sweep: lam.x = gphi/Vreg V grad w + gcn/W (c grad w + G) -
// gcn C/W^2 grad w per slot (coef[3k..3k+2] = gphi/Vreg, gcn/W, gcn C/W^2 of instance k);
// every other adjoint starts at zero.
__global__ __launch_bounds__(kBlock) void k_obs_seed(const PRec *rec, const double *sobs,
                                                     const double *coef, SAdj *lam, unsigned n)
{
    unsigned a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    const double *S = sobs + 8 * (size_t)a;
    const double *C = coef + 3 * (size_t)rec[a].inst;
    double kw = C[0] * S[4] + C[1] * S[0] - C[2];
    SAdj L;
    L.x[0] = kw * S[5] + C[1] * S[1];
    L.x[1] = kw * S[6] + C[1] * S[2];
    L.x[2] = kw * S[7] + C[1] * S[3];
    for (int k = 0; k < 3; k++) { L.v[k] = 0.0; L.w[k] = 0.0; L.h[k] = 0.0; }
    lam[a] = L;
}

}  // namespace dem
