/*
 * dem_oracle.c -- CPU oracle for the DEM packing step (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the checker, never the product.  Only tests/, the smoke() entry and the
 * cpu_baseline leg of bench.py may load it (through ctypes, as liboracle.so).  The product
 * path is the HIP library; it never calls, links or falls back to anything here.
 *
 * It restates, in plain C99 with OpenMP, the reference's contact-candidate-pair-parallel
 * (PCCP) step in the reference's own data layout (particle index order, uint2 hashes,
 * fixed-width neighbour rows Anei/Aref, inclusive scan sjgi, Lp/Ln pair list).  Each
 * function cites the reference file:line it follows.  Routines whose source is absent
 * from the reference (neighborarrayD, pairlistD, calforceij, sumallforceiD, calforceiw,
 * outputporositycnD, the SDK radix sort and scan) follow one written spec, stated at each
 * function below and shared with the HIP kernels: candidates are the 27-cell neighbours
 * within ri + rj + skin, ordered by stencil row and slot; a pair's first member has the lower
 * index; Aref rows hold own pairs then partner pairs; histories follow the (i, j) identity;
 * the law is the Hertz-Mindlin model of the contact-law comment.  The checks that pin each of
 * them are the tests/test_oracle_*.py files.  Parity unpinned: calforceij / calforceiw
 * constants and sign conventions beyond the closed-form cases (reference source absent).
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  Contraction is off so
 * that every expression rounds exactly as written; the HIP kernels are compiled the same
 * way, which makes hashes, sort order and pair sets bit-exact by construction.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned int uint;

#define ORC_EMPTY 0xffffffffu
/* neiglist.cu:27-30 */
#define ORC_CUA 1.0
#define ORC_CUB 1.55
#define ORC_XSMALL 1e-30

/* Material block: reference Matproperties + Wallproperties (neiglist.cu:32-47), in that
 * member order, 12 doubles. */
typedef struct {
    double estarp, ufricp, vpois, dampnp, damptp, upp;
    double estarw, ufricw, vpoisw, dampnw, damptw, upw;
} orc_mat;

/* Wall geometry.  kind 0 = box: side walls x = +-size[0]/2, y = +-size[1]/2, floor z = 0
 * (the reference's worldSize box, neiglist.cu:240 / dempacking.cu:333-339); kind 1 = vertical
 * cylinder of radius `radius` about the z axis with a floor z = 0. */
typedef struct {
    int kind;
    int pad;
    double size[3];
    double radius;
} orc_walls;

/* ------------------------------------------------------------------------------------ */
/* A1: reduced material constants.  dempacking.cu:141-149 and :224-225.                  */
/* ------------------------------------------------------------------------------------ */
void orc_reduced_material(double emod, double vpois, double denp, double diam,
                          double ufricp, double ufricw, double dampnp, double damptp,
                          double dampnw, double damptw, double upp_m, double upw_m,
                          orc_mat *out)
{
    double emodw = emod;                                    /* :141 */
    double vw = vpois;                                      /* :142 */
    double er = 6.0 * emod / (3.14159265 * denp * 9.81 * diam);   /* :146 */
    double erw = 6.0 * emodw / (3.14159265 * denp * 9.81 * diam); /* :147 */
    out->estarp = 1.33333333 * er / (2.0 * (1.0 - vpois * vpois));                  /* :148 */
    out->estarw = 1.33333333 / ((1.0 - vw * vw) / er + (1.0 - vw * vw) / erw);      /* :149 */
    out->ufricp = ufricp;
    out->vpois = vpois;
    out->dampnp = dampnp;
    out->damptp = damptp;
    out->upp = upp_m / diam;                                /* :224 */
    out->ufricw = ufricw;
    out->vpoisw = vw;
    out->dampnw = dampnw;
    out->damptw = damptw;
    out->upw = upw_m / diam;                                /* :225 */
}

/* ------------------------------------------------------------------------------------ */
/* A2: hash.  neiglist.cu:54-97 (calcGridPos, calcGridAddress, calcHashD).               */
/* The floor result is clamped in double before the int conversion; for every finite    */
/* coordinate inside int range this equals the reference's convert-then-clamp.           */
/* ------------------------------------------------------------------------------------ */
static inline uint orc_cell_coord(double p, double o, double c, uint g)
{
    double f = floor((p - o) / c);
    if (!(f > 0.0)) return 0u;          /* also maps NaN to 0 */
    if (f > (double)(g - 1u)) return g - 1u;
    return (uint)f;
}

static inline uint orc_hash_one(const double *p, const uint *g, const double *cell,
                                const double *origin)
{
    uint x = orc_cell_coord(p[0], origin[0], cell[0], g[0]);
    uint y = orc_cell_coord(p[1], origin[1], cell[1], g[1]);
    uint z = orc_cell_coord(p[2], origin[2], cell[2], g[2]);
    return z * g[1] * g[0] + y * g[0] + x;      /* neiglist.cu:73 */
}

void orc_calc_hash(int n, const double *pos, const uint *g, const double *cell,
                   const double *origin, uint *hash /* 2n: (cell, index) */)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        hash[2 * i] = orc_hash_one(pos + 3 * i, g, cell, origin);
        hash[2 * i + 1] = (uint)i;              /* neiglist.cu:95 */
    }
}

/* ------------------------------------------------------------------------------------ */
/* A3: stable sort of (key, value) by key.  The reference calls the CUDA SDK RadixSort    */
/* (dempacking.cu:553-555; source absent).  Pinned by the definition of a stable sort.   */
/* LSD counting sort, 8-bit digits, `bits` key bits (ping-pong in `tmp`).                 */
/* ------------------------------------------------------------------------------------ */
void orc_radix_sort(uint *kv, uint *tmp, int n, int bits)
{
    static const int R = 256;
    uint *src = kv, *dst = tmp;
    int passes = (bits + 7) / 8;
    size_t cnt[256];
    for (int pass = 0; pass < passes; pass++) {
        int sh = 8 * pass;
        memset(cnt, 0, sizeof(cnt));
        for (int i = 0; i < n; i++) cnt[(src[2 * i] >> sh) & 0xffu]++;
        size_t run = 0;
        for (int d = 0; d < R; d++) { size_t c = cnt[d]; cnt[d] = run; run += c; }
        for (int i = 0; i < n; i++) {
            size_t o = cnt[(src[2 * i] >> sh) & 0xffu]++;
            dst[2 * o] = src[2 * i];
            dst[2 * o + 1] = src[2 * i + 1];
        }
        uint *t = src; src = dst; dst = t;
    }
    if (src != kv) memcpy(kv, src, sizeof(uint) * 2 * (size_t)n);
}

/* ------------------------------------------------------------------------------------ */
/* A4: cell start.  particleSystem.cu:58 (0xff fill) + neiglist.cu:102-120.              */
/* ------------------------------------------------------------------------------------ */
void orc_find_cell_start(const uint *sorted, uint *cell_start, int n, uint ncells)
{
    for (uint c = 0; c < ncells; c++) cell_start[c] = ORC_EMPTY;
    for (int i = 0; i < n; i++) {
        uint c = sorted[2 * i];
        if (i == 0 || c != sorted[2 * (i - 1)]) cell_start[c] = (uint)i;
    }
}

/* ------------------------------------------------------------------------------------ */
/* A5: neighbour array (neighborarrayD is absent; call site particleSystem.cu:67-97).      */
/* Spec: candidates of i are j != i in the 27 cells around i's (clamped) cell, visited   */
/* z-, then y-, then x-offset ascending, each cell's particles in sorted order, with     */
/* |xj - xi|^2 < (ri + rj + skin)^2.  With ri = rj = 0.5 and skin = cub - cua this is the */
/* reference's fixed cutoff |xj - xi| < cub = 1.55 (neiglist.cu:27-28).                  */
/* Returns the largest neighbour count seen (caller compares with `width`).              */
/* ------------------------------------------------------------------------------------ */
static inline int orc_is_candidate(const double *pi, const double *pj, double ri, double rj,
                                   double skin)
{
    double dx = pj[0] - pi[0];
    double dy = pj[1] - pi[1];
    double dz = pj[2] - pi[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double cut = (ri + rj) + skin;
    return d2 < cut * cut;
}

/* Takes the unsorted hash array (index order) so the particle's own cell is read
 * directly, like the reference kernel, which receives particleHash. */
int orc_neighbor_array_h(int n, const double *pos, const double *rad, double skin,
                         const uint *hash, const uint *sorted, const uint *cell_start,
                         const uint *g, int width, uint *anei, uint *njgi, uint *njli)
{
    int maxcnt = 0;
    uint gxy = g[0] * g[1];
#pragma omp parallel for schedule(dynamic, 256) reduction(max : maxcnt)
    for (int i = 0; i < n; i++) {
        uint h = hash[2 * i];
        int cz = (int)(h / gxy);
        int cy = (int)((h - (uint)cz * gxy) / g[0]);
        int cx = (int)(h - (uint)cz * gxy - (uint)cy * g[0]);
        int cnt = 0;
        uint ng = 0, nl = 0;
        for (int dz = -1; dz <= 1; dz++) {
            int z = cz + dz;
            if (z < 0 || z >= (int)g[2]) continue;
            for (int dy = -1; dy <= 1; dy++) {
                int y = cy + dy;
                if (y < 0 || y >= (int)g[1]) continue;
                for (int dx = -1; dx <= 1; dx++) {
                    int x = cx + dx;
                    if (x < 0 || x >= (int)g[0]) continue;
                    uint c = (uint)z * gxy + (uint)y * g[0] + (uint)x;
                    uint s = cell_start[c];
                    if (s == ORC_EMPTY) continue;
                    for (; s < (uint)n && sorted[2 * s] == c; s++) {
                        uint j = sorted[2 * s + 1];
                        if (j == (uint)i) continue;
                        if (!orc_is_candidate(pos + 3 * i, pos + 3 * j, rad[i], rad[j], skin))
                            continue;
                        if (cnt < width) anei[(size_t)i * width + cnt] = j;
                        cnt++;
                        if (j > (uint)i) ng++; else nl++;
                    }
                }
            }
        }
        for (int k = cnt; k < width; k++) anei[(size_t)i * width + k] = ORC_EMPTY;
        njgi[i] = ng;
        njli[i] = nl;
        if (cnt > maxcnt) maxcnt = cnt;
    }
    return maxcnt;
}

/* ------------------------------------------------------------------------------------ */
/* A6: prefix sum.  particleSystem.cu:126-140: oldsjgi <- sjgi; sjgi = inclusive scan.    */
/* ------------------------------------------------------------------------------------ */
uint orc_prefix_sum(int n, const uint *njgi, uint *sjgi, uint *oldsjgi)
{
    if (oldsjgi) memcpy(oldsjgi, sjgi, sizeof(uint) * (size_t)n);
    uint run = 0;
    for (int i = 0; i < n; i++) { run += njgi[i]; sjgi[i] = run; }
    return n > 0 ? sjgi[n - 1] : 0u;
}

/* ------------------------------------------------------------------------------------ */
/* A7: pair list + reference table + history remap (pairlistD absent; call site           */
/* particleSystem.cu:143-191, semantics dempacking.cu:640-645).  Pair slots of i are     */
/* [sjgi[i]-njgi[i], sjgi[i]), filled with i's larger-index neighbours in Anei order.    */
/* Aref row i: entries [0, njgi[i]) are i's own slots in that order; entries             */
/* [njgi[i], njgi[i]+njli[i]) are the slots of the pairs (j, i), j < i, in Anei order.   */
/* (sumallforcei receives Aref, njgi and njli but neither Anei nor Lp,                   */
/* particleSystem.cu:243-257, so the side of each entry must follow from its position.)  */
/* History of a new pair (i,j) is copied from the old slot of (i,j) in i's old segment   */
/* [oldsjgi[i-1], oldsjgi[i]) (oldLn == j), else zero; an empty or inverted old segment  */
/* (new particles, dempacking.cu:204,521-522) gives zero.                                */
/* ------------------------------------------------------------------------------------ */
void orc_pair_list(int n, int nold, int width, const uint *anei, const uint *njgi,
                   const uint *njli, const uint *sjgi, const uint *oldsjgi, const uint *oldln,
                   const double *olddispt, const double *oldfricp,
                   uint *aref, uint *lp, uint *ln, double *dispt, double *fricp)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        uint cnt = njgi[i] + njli[i];
        uint start = sjgi[i] - njgi[i];
        uint k = 0, r = 0;
        uint os = 0, oe = 0;
        if (i < nold) { os = i > 0 ? oldsjgi[i - 1] : 0u; oe = oldsjgi[i]; }
        for (uint m = 0; m < (uint)width; m++) aref[(size_t)i * width + m] = ORC_EMPTY;
        for (uint m = 0; m < cnt && m < (uint)width; m++) {
            uint j = anei[(size_t)i * width + m];
            if (j > (uint)i) {
                uint slot = start + k;
                lp[slot] = (uint)i;
                ln[slot] = j;
                aref[(size_t)i * width + k] = slot;
                k++;
                double h0 = 0.0, h1 = 0.0, h2 = 0.0, f = 0.0;
                for (uint q = os; q < oe; q++) {
                    if (oldln[q] == j) {
                        h0 = olddispt[3 * (size_t)q];
                        h1 = olddispt[3 * (size_t)q + 1];
                        h2 = olddispt[3 * (size_t)q + 2];
                        f = oldfricp[q];
                        break;
                    }
                }
                dispt[3 * (size_t)slot] = h0;
                dispt[3 * (size_t)slot + 1] = h1;
                dispt[3 * (size_t)slot + 2] = h2;
                fricp[slot] = f;
            } else {
                /* slot of (j, i): i's rank among j's larger-index neighbours */
                uint cj = njgi[j] + njli[j];
                uint rr = 0, found = ORC_EMPTY;
                for (uint mm = 0; mm < cj && mm < (uint)width; mm++) {
                    uint jj = anei[(size_t)j * width + mm];
                    if (jj == (uint)i) { found = rr; break; }
                    if (jj > j) rr++;
                }
                uint pos = njgi[i] + r;
                if (pos < (uint)width)
                    aref[(size_t)i * width + pos] = found == ORC_EMPTY ? ORC_EMPTY
                                                                       : (sjgi[j] - njgi[j]) + found;
                r++;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* A9: pair contact law (calforceij is absent; signature implied by neiglist.cu:208-209).  */
/* vprojec / vscale follow neiglist.cu:133-153 literally, in FP64.                        */
/* ------------------------------------------------------------------------------------ */
static inline void orc_cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];      /* cutil_math.h:1200-1203 */
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static inline double orc_len(const double *v)
{
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* Float-length compatibility (off by default): the reference's length() is sqrtf of the
 * FP64 dot product rounded to float, returned as double (cutil_math.h:1086-1089).  On, it
 * replaces the FP64 length in vprojec (neiglist.cu:140-142) and the rebuild trigger
 * (neiglist.cu:359,370); the library's DEM_OPT_FLOAT_LENGTH does the same. */
static int orc_float_len = 0;
void orc_set_float_length(int on) { orc_float_len = on != 0; }
static inline double orc_len_ref(const double *v)
{
    double d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    return orc_float_len ? (double)sqrtf((float)d) : sqrt(d);
}

/* neiglist.cu:133-146 */
static inline void orc_vprojec(const double *p, const double *u, double *r, int ityp)
{
    double q[3], mu[3] = {-u[0], -u[1], -u[2]};
    orc_cross(u, p, q);
    orc_cross(mu, q, r);
    if (ityp == 0) return;
    double pmag = orc_len_ref(p);
    if (pmag < 0.001) return;
    double rmag = orc_len_ref(r);
    if (pmag > rmag * 100) return;
    double pr = sqrt(pmag / rmag);
    r[0] = r[0] * pr; r[1] = r[1] * pr; r[2] = r[2] * pr;
}

/* neiglist.cu:148-153 */
static inline void orc_vscale(double *v, double p1, double p2)
{
    double ratio = p1 / p2;
    v[0] = v[0] * ratio; v[1] = v[1] * ratio; v[2] = v[2] * ratio;
}

/* Shared tail of the pair and wall laws: given unit normal n (from particle i toward the
 * other body), overlap dn > 0, effective radius/mass, relative displacement increment vrel
 * of the other body's contact point w.r.t. i's, and the incoming tangential history, build
 * the force on i, the tangential force, the rolling-resistance moment on i and the new
 * history.  Material: E* (estar), sliding friction mu, Poisson nu, damping gn / gt,
 * rolling-friction arm ur. */
static inline void orc_law_core(const double *n, double dn, double rs, double ms,
                                const double *vrel, const double *wrel, const double *hin,
                                double dt, double estar, double mu, double nu, double gn,
                                double gt, double ur, double *fi, double *ft, double *mr,
                                double *hout, double *fout)
{
    double vns = vrel[0] * n[0] + vrel[1] * n[1] + vrel[2] * n[2];
    double vn[3], vt[3];
    for (int k = 0; k < 3; k++) { vn[k] = vns * n[k]; vt[k] = vrel[k] - vn[k]; }
    double sdn = sqrt(dn);
    double fcn = estar * sqrt(rs) * dn * sdn;                 /* Hertz |F_cn| */
    double cn = gn * sqrt(6.0 * ms * estar * sqrt(rs * dn)); /* normal dashpot */
    double cnd = cn / dt;
    double fn[3];
    for (int k = 0; k < 3; k++) fn[k] = -fcn * n[k] + cnd * vn[k];

    double ht[3];
    orc_vprojec(hin, n, ht, 1);
    for (int k = 0; k < 3; k++) ht[k] = ht[k] + vt[k];
    double dtmax = mu * (2.0 - nu) / (2.0 * (1.0 - nu)) * dn;
    if (mu > 0.0 && dtmax > 0.0) {
        double htm = orc_len(ht);
        if (htm > dtmax) { orc_vscale(ht, dtmax, htm); htm = dtmax; }
        double xi = htm / dtmax;
        double om = 1.0 - xi;
        double fct = mu * fcn * (1.0 - om * sqrt(om));
        double ct = gt * sqrt(6.0 * mu * ms * fcn * sqrt(om) / dtmax);
        double fh = htm > 0.0 ? fct / htm : 0.0;
        double ctd = ct / dt;
        for (int k = 0; k < 3; k++) ft[k] = fh * ht[k] + ctd * vt[k];
        double ftm = orc_len(ft);
        double flim = mu * fcn;
        if (ftm > flim) orc_vscale(ft, flim, ftm);
    } else {
        ht[0] = ht[1] = ht[2] = 0.0;
        ft[0] = ft[1] = ft[2] = 0.0;
    }
    *fout = orc_len(ft);
    for (int k = 0; k < 3; k++) { fi[k] = fn[k] + ft[k]; hout[k] = ht[k]; }

    double wrm = orc_len(wrel);
    if (wrm > ORC_XSMALL) {
        double m = ur * fcn / wrm;
        for (int k = 0; k < 3; k++) mr[k] = -m * wrel[k];
    } else {
        mr[0] = mr[1] = mr[2] = 0.0;
    }
}

/* Pair law.  Outputs force on i (fi), torques on i and j, new history.  Returns 1 when
 * the spheres overlap.  On no contact all outputs are zero (history reset). */
int orc_law_pair(const double *xi, const double *xj, const double *vi, const double *vj,
                 const double *wi, const double *wj, double ri, double rj, double mi,
                 double mj, const double *hin, double fin, double dt, const orc_mat *m,
                 double *hout, double *fout, double *fi, double *ti, double *tj)
{
    (void)fin;
    double dx[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    double dist = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    double dn = (ri + rj) - dist;
    if (!(dn > 0.0)) {
        for (int k = 0; k < 3; k++) { hout[k] = 0.0; fi[k] = ti[k] = tj[k] = 0.0; }
        *fout = 0.0;
        return 0;
    }
    double idist = 1.0 / dist;
    double n[3] = {dx[0] * idist, dx[1] * idist, dx[2] * idist};
    double rs = ri * rj / (ri + rj);
    double ms = mi * mj / (mi + mj);
    double s[3], c[3], vrel[3], wrel[3];
    for (int k = 0; k < 3; k++) s[k] = ri * wi[k] + rj * wj[k];
    orc_cross(s, n, c);
    for (int k = 0; k < 3; k++) {
        vrel[k] = (vj[k] - vi[k]) - c[k];
        wrel[k] = wi[k] - wj[k];
    }
    double ft[3], mr[3];
    orc_law_core(n, dn, rs, ms, vrel, wrel, hin, dt, m->estarp, m->ufricp, m->vpois,
                 m->dampnp, m->damptp, m->upp, fi, ft, mr, hout, fout);
    double rn[3], cr[3];
    for (int k = 0; k < 3; k++) rn[k] = ri * n[k];
    orc_cross(rn, ft, cr);
    for (int k = 0; k < 3; k++) ti[k] = cr[k] + mr[k];
    for (int k = 0; k < 3; k++) rn[k] = rj * n[k];
    orc_cross(rn, ft, cr);
    for (int k = 0; k < 3; k++) tj[k] = cr[k] - mr[k];
    return 1;
}

/* ------------------------------------------------------------------------------------ */
/* A8: per-pair force (neiglist.cu:156-233, launcher particleSystem.cu:194-240).         */
/* ------------------------------------------------------------------------------------ */
void orc_force_pairs(uint npairs, const double *pos, const double *pdis, const double *angv,
                     const double *rad, const double *rmass, const uint *lp, const uint *ln,
                     uint *oldln, double *dispt, double *fricp, double *olddispt,
                     double *oldfricp, double dt, const orc_mat *m, double *forcepair,
                     double *torqpairi, double *torqpairj, uint *contactpair)
{
#pragma omp parallel for schedule(static)
    for (long long p = 0; p < (long long)npairs; p++) {
        uint i = lp[p], j = ln[p];
        double hout[3], fout, fi[3], ti[3], tj[3];
        int c = orc_law_pair(pos + 3 * (size_t)i, pos + 3 * (size_t)j, pdis + 3 * (size_t)i,
                             pdis + 3 * (size_t)j, angv + 3 * (size_t)i, angv + 3 * (size_t)j,
                             rad[i], rad[j], rmass[i], rmass[j], dispt + 3 * p, fricp[p], dt, m,
                             hout, &fout, fi, ti, tj);
        for (int k = 0; k < 3; k++) {
            forcepair[3 * p + k] = fi[k];
            torqpairi[3 * p + k] = ti[k];
            torqpairj[3 * p + k] = tj[k];
            dispt[3 * p + k] = hout[k];
            if (olddispt) olddispt[3 * p + k] = hout[k];
        }
        fricp[p] = fout;
        if (oldfricp) oldfricp[p] = fout;
        contactpair[p] = (uint)c;
        if (oldln) oldln[p] = j;
    }
}

/* ------------------------------------------------------------------------------------ */
/* A10: per-particle sum (sumallforceiD absent; launcher particleSystem.cu:243-277).      */
/* Aref row order (see A7): the first njgi entries are i's own pairs (+forcepair,        */
/* torqpairi), the next njli entries pairs where i is second (-forcepair, torqpairj).    */
/* ------------------------------------------------------------------------------------ */
void orc_sum_forces(int n, int width, const uint *njgi, const uint *njli,
                    const double *forcepair, const double *torqpairi, const double *torqpairj,
                    const uint *contactpair, const uint *aref, double *forcei, double *torquei,
                    uint *contacti)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
        uint c = 0;
        uint cnt = njgi[i] + njli[i];
        for (uint m = 0; m < cnt && m < (uint)width; m++) {
            uint p = aref[(size_t)i * width + m];
            if (p == ORC_EMPTY) continue;
            if (m < njgi[i]) {
                for (int k = 0; k < 3; k++) { f[k] += forcepair[3 * (size_t)p + k];
                                              t[k] += torqpairi[3 * (size_t)p + k]; }
            } else {
                for (int k = 0; k < 3; k++) { f[k] -= forcepair[3 * (size_t)p + k];
                                              t[k] += torqpairj[3 * (size_t)p + k]; }
            }
            c += contactpair[p];
        }
        for (int k = 0; k < 3; k++) { forcei[3 * i + k] = f[k]; torquei[3 * i + k] = t[k]; }
        contacti[i] = c;
    }
}

/* ------------------------------------------------------------------------------------ */
/* A11: wall contact + gravity (calculateforceiwD neiglist.cu:238-306; calforceiw absent). */
/* Every wall the sphere overlaps contributes.  The single history slot per particle     */
/* (neiglist.cu:263-264) belongs to the wall with the largest overlap (lowest wall index */
/* on ties); other walls start from zero history.  Wall order: box floor, x-, x+, y-, y+; */
/* cylinder floor, side.                                                                  */
/* ------------------------------------------------------------------------------------ */
static inline int orc_walls_overlaps(const orc_walls *w, const double *x, double r,
                                     double *dn, double (*nrm)[3])
{
    int nw = 0;
    /* floor */
    dn[nw] = r - x[2];
    nrm[nw][0] = 0.0; nrm[nw][1] = 0.0; nrm[nw][2] = -1.0; nw++;
    if (w->kind == 0) {
        double hx = 0.5 * w->size[0], hy = 0.5 * w->size[1];
        dn[nw] = r - (x[0] + hx); nrm[nw][0] = -1.0; nrm[nw][1] = 0.0; nrm[nw][2] = 0.0; nw++;
        dn[nw] = r - (hx - x[0]); nrm[nw][0] = 1.0; nrm[nw][1] = 0.0; nrm[nw][2] = 0.0; nw++;
        dn[nw] = r - (x[1] + hy); nrm[nw][0] = 0.0; nrm[nw][1] = -1.0; nrm[nw][2] = 0.0; nw++;
        dn[nw] = r - (hy - x[1]); nrm[nw][0] = 0.0; nrm[nw][1] = 1.0; nrm[nw][2] = 0.0; nw++;
    } else {
        double rho = sqrt(x[0] * x[0] + x[1] * x[1]);
        dn[nw] = (rho + r) - w->radius;
        if (rho > 0.0) { nrm[nw][0] = x[0] / rho; nrm[nw][1] = x[1] / rho; }
        else { nrm[nw][0] = 1.0; nrm[nw][1] = 0.0; }
        nrm[nw][2] = 0.0;
        nw++;
    }
    return nw;
}

int orc_law_wall(const orc_walls *w, const double *x, const double *v, const double *om,
                 double r, double m, const double *hin, double fin, double dt,
                 const orc_mat *mat, double *hout, double *fout, double *fw, double *tw)
{
    (void)fin;
    double dn[5], nrm[5][3];
    int nw = orc_walls_overlaps(w, x, r, dn, nrm);
    int prim = -1;
    for (int k = 0; k < nw; k++)
        if (dn[k] > 0.0 && (prim < 0 || dn[k] > dn[prim])) prim = k;
    fw[0] = fw[1] = fw[2] = 0.0;
    tw[0] = tw[1] = tw[2] = 0.0;
    hout[0] = hout[1] = hout[2] = 0.0;
    *fout = 0.0;
    int hits = 0;
    static const double zero3[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < nw; k++) {
        if (!(dn[k] > 0.0)) continue;
        hits++;
        const double *n = nrm[k];
        double s[3], c[3], vrel[3];
        for (int q = 0; q < 3; q++) s[q] = r * om[q];
        orc_cross(s, n, c);
        for (int q = 0; q < 3; q++) vrel[q] = (0.0 - v[q]) - c[q];
        double f[3], ft[3], mr[3], ht[3], fo;
        orc_law_core(n, dn[k], r, m, vrel, om, k == prim ? hin : zero3, dt, mat->estarw,
                     mat->ufricw, mat->vpoisw, mat->dampnw, mat->damptw, mat->upw, f, ft, mr,
                     ht, &fo);
        double rn[3], cr[3];
        for (int q = 0; q < 3; q++) rn[q] = r * n[q];
        orc_cross(rn, ft, cr);
        for (int q = 0; q < 3; q++) { fw[q] += f[q]; tw[q] += cr[q] + mr[q]; }
        if (k == prim) { hout[0] = ht[0]; hout[1] = ht[1]; hout[2] = ht[2]; *fout = fo; }
    }
    return hits;
}

void orc_wall_gravity(int n, const orc_walls *w, const double *pos, const double *pdis,
                      const double *angv, const double *rad, const double *rmass, double dt,
                      uint npairs, const orc_mat *mat, double *disptw, double *fricpw,
                      double *forcei, double *torquei)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double f[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
        if (npairs != 0) {                                         /* neiglist.cu:278-282 */
            for (int k = 0; k < 3; k++) { f[k] = forcei[3 * i + k]; t[k] = torquei[3 * i + k]; }
        }
        f[2] = f[2] + (-rmass[i]);                                 /* neiglist.cu:272,284 */
        f[0] = f[0] + 0.0; f[1] = f[1] + 0.0;
        double hout[3], fo, fw[3], tw[3];
        int hit = orc_law_wall(w, pos + 3 * i, pdis + 3 * i, angv + 3 * i, rad[i], rmass[i],
                               disptw + 3 * i, fricpw[i], dt, mat, hout, &fo, fw, tw);
        if (hit) {
            for (int k = 0; k < 3; k++) { f[k] += fw[k]; t[k] += tw[k]; disptw[3 * i + k] = hout[k]; }
            fricpw[i] = fo;
        } else {
            for (int k = 0; k < 3; k++) disptw[3 * i + k] = 0.0;
            fricpw[i] = 0.0;
        }
        for (int k = 0; k < 3; k++) { forcei[3 * i + k] = f[k]; torquei[3 * i + k] = t[k]; }
    }
}

/* ------------------------------------------------------------------------------------ */
/* A12: integrator (neiglist.cu:308-409; launcher + host max particleSystem.cu:315-384).  */
/* Rebuild flag: |q|^2 > ann2 with ann = 0.5*(cub - cua) (neiglist.cu:335,359,370); with  */
/* the float-length compatibility on, (double)sqrtf((float)|q|^2) > ann as the reference.  */
/* maxhz = max(0, max z) (particleSystem.cu:373-379).                                    */
/* ------------------------------------------------------------------------------------ */
void orc_update(int n, double *pos, double *pdis, double *angv, const double *rmass,
                const double *inert, const double *forcei, const double *torquei, double *qi,
                double dt, double ann2, double ann, uint *iall, double *maxhz)
{
    const int flen = orc_float_len;
    uint flag = 0;
    double mz = 0.0;
#pragma omp parallel for schedule(static) reduction(max : flag) reduction(max : mz)
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            pdis[3 * i + k] = pdis[3 * i + k] + forcei[3 * i + k] * dt * dt / rmass[i];
            pos[3 * i + k] = pos[3 * i + k] + pdis[3 * i + k];
            qi[3 * i + k] = qi[3 * i + k] + pdis[3 * i + k];
            angv[3 * i + k] = angv[3 * i + k] + torquei[3 * i + k] * dt * dt / inert[i];
        }
        double q2 = qi[3 * i] * qi[3 * i] + qi[3 * i + 1] * qi[3 * i + 1] + qi[3 * i + 2] * qi[3 * i + 2];
        if (flen ? (double)sqrtf((float)q2) > ann : q2 > ann2) flag = 1;
        if (pos[3 * i + 2] > mz) mz = pos[3 * i + 2];
    }
    *iall = flag;
    *maxhz = mz;
}

/* ------------------------------------------------------------------------------------ */
/* A13: contact total (outputporositycnD absent; particleSystem.cu:386-422).             */
/* ------------------------------------------------------------------------------------ */
uint orc_contact_total(int n, const uint *contacti)
{
    uint s = 0;
    for (int i = 0; i < n; i++) s += contacti[i];
    return s;
}

/* ------------------------------------------------------------------------------------ */
/* A14: one full step in the reference order (dempacking.cu:535-739).                    */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    /* geometry / grid */
    uint g[3];
    uint ncells;
    double cell[3];
    double origin[3];
    double skin, ann2, dt;
    int width;          /* neighbour row width (reference: 30) */
    int pad;
    orc_mat mat;
    orc_walls walls;
} orc_setup;

typedef struct {
    int n, nold;
    uint iall;          /* rebuild at the next step when nonzero */
    uint npairs;
    uint overflow;
    uint pcntot;
    double maxhz;
    /* particle state, index order, AoS double3 */
    double *pos, *pdis, *angv, *rad, *rmass, *inert, *qi;
    double *disptw, *fricpw, *forcei, *torquei;
    uint *contacti;
    /* lists */
    uint *hash, *sorted, *tmp, *cell_start;
    uint *anei, *aref, *njgi, *njli, *sjgi, *oldsjgi;
    uint *lp, *ln, *oldln, *contactpair;
    double *dispt, *fricp, *olddispt, *oldfricp;
    double *forcepair, *torqpairi, *torqpairj;
    uint pair_cap;
} orc_state;

static int orc_key_bits(uint ncells)
{
    int b = 1;
    while (b < 32 && (1u << b) < ncells) b++;
    return b;
}

/* Rebuild the lists (dempacking.cu:539-669).  Returns 0, or 1 on neighbour-row overflow,
 * 2 on pair-capacity overflow. */
int orc_rebuild(const orc_setup *s, orc_state *st)
{
    int n = st->n;
    memset(st->qi, 0, sizeof(double) * 3 * (size_t)n);           /* dempacking.cu:541 */
    orc_calc_hash(n, st->pos, s->g, s->cell, s->origin, st->hash);
    memcpy(st->sorted, st->hash, sizeof(uint) * 2 * (size_t)n);
    orc_radix_sort(st->sorted, st->tmp, n, orc_key_bits(s->ncells));
    orc_find_cell_start(st->sorted, st->cell_start, n, s->ncells);
    int maxcnt = orc_neighbor_array_h(n, st->pos, st->rad, s->skin, st->hash, st->sorted,
                                      st->cell_start, s->g, s->width, st->anei, st->njgi,
                                      st->njli);
    if (maxcnt > s->width) { st->overflow = 1; return 1; }
    uint np = orc_prefix_sum(n, st->njgi, st->sjgi, st->oldsjgi);
    if (np > st->pair_cap) { st->overflow = 2; return 2; }
    orc_pair_list(n, st->nold, s->width, st->anei, st->njgi, st->njli, st->sjgi, st->oldsjgi,
                  st->oldln, st->olddispt, st->oldfricp, st->aref, st->lp, st->ln, st->dispt,
                  st->fricp);
    st->npairs = np;
    st->nold = n;
    st->iall = 0;
    return 0;
}

int orc_step(const orc_setup *s, orc_state *st)
{
    int n = st->n;
    memset(st->forcei, 0, sizeof(double) * 3 * (size_t)n);       /* dempacking.cu:535-537 */
    memset(st->torquei, 0, sizeof(double) * 3 * (size_t)n);
    memset(st->contacti, 0, sizeof(uint) * (size_t)n);
    if (st->iall) {
        int rc = orc_rebuild(s, st);
        if (rc) return rc;
    }
    orc_force_pairs(st->npairs, st->pos, st->pdis, st->angv, st->rad, st->rmass, st->lp,
                    st->ln, st->oldln, st->dispt, st->fricp, st->olddispt, st->oldfricp,
                    s->dt, &s->mat, st->forcepair, st->torqpairi, st->torqpairj,
                    st->contactpair);
    orc_sum_forces(n, s->width, st->njgi, st->njli, st->forcepair, st->torqpairi,
                   st->torqpairj, st->contactpair, st->aref, st->forcei, st->torquei,
                   st->contacti);
    orc_wall_gravity(n, &s->walls, st->pos, st->pdis, st->angv, st->rad, st->rmass, s->dt,
                     st->npairs, &s->mat, st->disptw, st->fricpw, st->forcei, st->torquei);
    orc_update(n, st->pos, st->pdis, st->angv, st->rmass, st->inert, st->forcei, st->torquei,
               st->qi, s->dt, s->ann2, 0.5 * s->skin, &st->iall, &st->maxhz);
    st->pcntot = orc_contact_total(n, st->contacti);
    return 0;
}

/* Allocation helpers so Python can drive orc_step through ctypes without owning every
 * buffer.  Arrays not listed here (pos ... inert) are filled by the caller after create. */
orc_state *orc_state_create(int n, int width, uint pair_cap, uint ncells)
{
    orc_state *st = (orc_state *)calloc(1, sizeof(orc_state));
    size_t N = (size_t)n, W = (size_t)width, P = pair_cap;
    st->n = n; st->nold = 0; st->iall = 1; st->pair_cap = pair_cap;
#define A(f, cnt, T) st->f = (T *)calloc((cnt) > 0 ? (cnt) : 1, sizeof(T))
    A(pos, 3 * N, double); A(pdis, 3 * N, double); A(angv, 3 * N, double);
    A(rad, N, double); A(rmass, N, double); A(inert, N, double); A(qi, 3 * N, double);
    A(disptw, 3 * N, double); A(fricpw, N, double); A(forcei, 3 * N, double);
    A(torquei, 3 * N, double); A(contacti, N, uint);
    A(hash, 2 * N, uint); A(sorted, 2 * N, uint); A(tmp, 2 * N, uint); A(cell_start, ncells, uint);
    A(anei, N * W, uint); A(aref, N * W, uint); A(njgi, N, uint); A(njli, N, uint);
    A(sjgi, N, uint); A(oldsjgi, N, uint);
    A(lp, P, uint); A(ln, P, uint); A(oldln, P, uint); A(contactpair, P, uint);
    A(dispt, 3 * P, double); A(fricp, P, double); A(olddispt, 3 * P, double);
    A(oldfricp, P, double); A(forcepair, 3 * P, double); A(torqpairi, 3 * P, double);
    A(torqpairj, 3 * P, double);
#undef A
    return st;
}

void orc_state_destroy(orc_state *st)
{
    if (!st) return;
    void *ps[] = {st->pos, st->pdis, st->angv, st->rad, st->rmass, st->inert, st->qi,
                  st->disptw, st->fricpw, st->forcei, st->torquei, st->contacti, st->hash,
                  st->sorted, st->tmp, st->cell_start, st->anei, st->aref, st->njgi, st->njli,
                  st->sjgi, st->oldsjgi, st->lp, st->ln, st->oldln, st->contactpair, st->dispt,
                  st->fricp, st->olddispt, st->oldfricp, st->forcepair, st->torqpairi,
                  st->torqpairj};
    for (size_t k = 0; k < sizeof(ps) / sizeof(ps[0]); k++) free(ps[k]);
    free(st);
}

/* Field accessors for ctypes (pointer to the named array and its element count). */
void *orc_state_field(orc_state *st, int which)
{
    switch (which) {
    case 0: return st->pos;      case 1: return st->pdis;     case 2: return st->angv;
    case 3: return st->rad;      case 4: return st->rmass;    case 5: return st->inert;
    case 6: return st->qi;       case 7: return st->disptw;   case 8: return st->fricpw;
    case 9: return st->forcei;   case 10: return st->torquei; case 11: return st->contacti;
    case 12: return st->hash;    case 13: return st->sorted;  case 14: return st->cell_start;
    case 15: return st->anei;    case 16: return st->aref;    case 17: return st->njgi;
    case 18: return st->njli;    case 19: return st->sjgi;    case 20: return st->lp;
    case 21: return st->ln;      case 22: return st->dispt;   case 23: return st->fricp;
    case 24: return st->forcepair; case 25: return st->torqpairi; case 26: return st->torqpairj;
    case 27: return st->contactpair;
    default: return 0;
    }
}

/* scalar view: n, nold, iall, npairs, overflow, pcntot, maxhz */
void orc_state_scalars(orc_state *st, double *out7)
{
    out7[0] = st->n; out7[1] = st->nold; out7[2] = st->iall; out7[3] = st->npairs;
    out7[4] = st->overflow; out7[5] = st->pcntot; out7[6] = st->maxhz;
}

void orc_state_set_iall(orc_state *st, uint v) { st->iall = v; }

/* Grow the active particle count (layer insertion, dempacking.cu:459-461): the arrays were
 * sized for the final count at create; the new particles' state must already be in place. */
void orc_state_set_n(orc_state *st, int n) { st->n = n; st->iall = 1; }

/* Run `steps` steps; returns 0 or the first nonzero orc_step code. */
int orc_run(const orc_setup *s, orc_state *st, int steps)
{
    for (int k = 0; k < steps; k++) {
        int rc = orc_step(s, st);
        if (rc) return rc;
    }
    return 0;
}

/* Brute-force O(n^2) candidate enumeration (test pin for A5/A7): writes (i, j) with i < j
 * in row-major order, returns the count (or -1 if cap exceeded). */
long long orc_brute_pairs(int n, const double *pos, const double *rad, double skin,
                          uint *out, long long cap)
{
    long long c = 0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (orc_is_candidate(pos + 3 * i, pos + 3 * j, rad[i], rad[j], skin)) {
                if (c >= cap) return -1;
                out[2 * c] = (uint)i; out[2 * c + 1] = (uint)j; c++;
            }
    return c;
}
