// host.cpp -- host-side utilities of libdem.so: input deck, reduced units, restart file,
// restitution <-> damping map and seeded workload generators.  Plain C++ (no device code).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dem.h"

namespace {

// Probability that a draw of ln d ~ N(-sigma^2/2, sigma) falls in [lo, hi] (sigma = 0: d = 1
// exactly, so 1 or 0).
double lognormal_window_mass(double sg, double lo, double hi)
{
    if (sg == 0.0) return (lo <= 1.0 && 1.0 <= hi) ? 1.0 : 0.0;
    const double m = -0.5 * sg * sg;
    auto cdf = [&](double d) { return 0.5 * std::erfc(-(std::log(d) - m) / (sg * std::sqrt(2.0))); };
    return cdf(hi) - cdf(lo);
}

// Redraws allowed per sphere of the clipped lognormal before the window is declared empty.
const long kMaxRedraws = 1000000;

// Smallest share of the lognormal a side-car clip window must hold (at most ~1e4 redraws per
// sphere on average).
const double kMinWindowMass = 1e-4;


// splitmix64: small, portable, seedable.
struct Rng {
    unsigned long long s;
    explicit Rng(unsigned long long seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull) {}
    unsigned long long next()
    {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

bool parse_number(const std::string &tok, double &v)
{
    if (tok.empty()) return false;
    std::string t = tok;
    for (char &ch : t)
        if (ch == 'D' || ch == 'd') ch = 'E';   // Fortran exponent
    char *end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0';
}

// Head-on Hertz + dashpot collision in normalised form (the law in law.hpp reduces to it for
// any E*, R*, m* and impact speed): u'' = -u^1.5 - sqrt(6) g u^0.25 u', u(0)=0, u'(0)=1.
// Returns the restitution |u'(t_end)| at the return to u = 0.  Classic RK4 with a fixed
// step; the end of contact is located inside the last step on the cubic Hermite interpolant
// of u (end values and slopes), and the exit speed is that interpolant's slope there.  A
// located crossing (rather than the last step's end, or a linear interpolation) keeps e(g)
// smooth in g when the step count changes, so the inverse map and its central difference
// quotients are well behaved down to relative steps of 1e-6.
double restitution_ode(double g)
{
    const double c = std::sqrt(6.0) * g;
    double u = 0.0, v = 1.0, t = 0.0;
    const double h = 5e-5;
    auto acc = [&](double uu, double vv) {
        double up = uu > 0.0 ? uu : 0.0;
        return -up * std::sqrt(up) - c * std::sqrt(std::sqrt(up)) * vv;
    };
    for (long it = 0; it < 8000000; it++) {
        double k1u = v, k1v = acc(u, v);
        double k2u = v + 0.5 * h * k1v, k2v = acc(u + 0.5 * h * k1u, v + 0.5 * h * k1v);
        double k3u = v + 0.5 * h * k2v, k3v = acc(u + 0.5 * h * k2u, v + 0.5 * h * k2v);
        double k4u = v + h * k3v, k4v = acc(u + h * k3u, v + h * k3v);
        double un = u + h / 6.0 * (k1u + 2 * k2u + 2 * k3u + k4u);
        double vn = v + h / 6.0 * (k1v + 2 * k2v + 2 * k3v + k4v);
        t += h;
        if (un <= 0.0 && t > h) {
            // p(s) = h00 u + h10 h v + h01 un + h11 h vn on s in [0, 1]; p(0) > 0 >= p(1)
            auto p = [&](double s) {
                double s2 = s * s, s3 = s2 * s;
                return (2 * s3 - 3 * s2 + 1) * u + (s3 - 2 * s2 + s) * h * v +
                       (-2 * s3 + 3 * s2) * un + (s3 - s2) * h * vn;
            };
            double lo = 0.0, hi = 1.0;
            for (int b = 0; b < 64; b++) {
                double mid = 0.5 * (lo + hi);
                if (p(mid) > 0.0) lo = mid;
                else hi = mid;
            }
            double s = 0.5 * (lo + hi), s2 = s * s;
            double dp = (6 * s2 - 6 * s) * u + (3 * s2 - 4 * s + 1) * h * v +
                        (-6 * s2 + 6 * s) * un + (3 * s2 - 2 * s) * h * vn;
            return std::fabs(dp / h);
        }
        if (vn <= 0.0 && un <= 0.0) return 0.0;
        u = un;
        v = vn;
        if (v < 0.0 && u < 1e-12) return std::fabs(v);
    }
    return 0.0;   // stuck (overdamped)
}

}  // namespace

extern "C" {

int dem_read_deck(const char *path, dem_deck *d)
{
    if (!path || !d) return DEM_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return DEM_ERR_IO;
    std::string txt;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), fp)) > 0) txt.append(buf, k);
    std::fclose(fp);
    // leading numeric tokens, stopping at the first non-number (the deck's free-text notes)
    std::vector<double> v;
    size_t i = 0;
    while (i < txt.size()) {
        while (i < txt.size() && std::isspace((unsigned char)txt[i])) i++;
        size_t j = i;
        while (j < txt.size() && !std::isspace((unsigned char)txt[j])) j++;
        if (j == i) break;
        double x;
        if (!parse_number(txt.substr(i, j - i), x)) break;
        v.push_back(x);
        i = j;
    }
    // dempacking.cu:112-125 reads 27 values, the first being iforcemode; a deck without
    // that leading line (26 values) is accepted with iforcemode = 0.
    size_t off;
    std::memset(d, 0, sizeof(*d));
    if (v.size() >= 27) { d->iforcemode = (int)v[0]; off = 1; }
    else if (v.size() == 26) { d->iforcemode = 0; off = 0; }
    else return DEM_ERR_IO;
    const double *p = v.data() + off;
    d->emod = p[0]; d->vpois = p[1]; d->denp = p[2];
    d->diam = p[3];
    d->ntot = (int)p[4];
    d->dt = p[5]; d->tstop = (uint)p[6];
    d->dmt1 = p[7]; d->dmt2 = p[8]; d->thick = p[9];
    d->zht1 = p[10]; d->zht2 = p[11];
    d->vel0 = p[12]; d->nrings = (int)p[13];
    d->ufricp = p[14]; d->ufricw = p[15];
    d->dampnp = p[16]; d->damptp = p[17]; d->dampnw = p[18]; d->damptw = p[19]; d->cdamp = p[20];
    d->stiffness = p[21]; d->dampcoeffn = p[22]; d->dampcoefft = p[23];
    d->upp = p[24]; d->upw = p[25];
    if (!(d->diam > 0.0) || !(d->denp > 0.0) || d->ntot < 0) return DEM_ERR_IO;
    return DEM_OK;
}

int dem_read_sidecar(const char *path, dem_sidecar *sc)
{
    if (!path || !sc) return DEM_ERR_INVALID;
    std::memset(sc, 0, sizeof(*sc));
    sc->walls = -1;
    sc->dpar[0] = 1.0;
    sc->seed = 1;
    sc->restitution = -1.0;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return DEM_ERR_IO;
    std::string txt;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) txt.append(buf, got);
    std::fclose(fp);
    std::istringstream lines(txt);
    std::string line;
    while (std::getline(lines, line)) {
        size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::istringstream is(line);
        std::vector<std::string> tok;
        for (std::string t; is >> t;) tok.push_back(t);
        if (tok.empty()) continue;
        const std::string &key = tok[0];
        std::vector<double> v;
        size_t first = 1;
        std::string word;
        if (key == "walls" || key == "diameters" || key == "insertion" || key == "rebuild") {
            if (tok.size() < 2) return DEM_ERR_IO;
            word = tok[1];
            first = 2;
        }
        for (size_t k = first; k < tok.size(); k++) {
            double x;
            if (!parse_number(tok[k], x) || !std::isfinite(x)) return DEM_ERR_IO;
            v.push_back(x);
        }
        auto count = [&](size_t c) { return v.size() == c; };
        auto whole = [](double x, double lo, double hi) {
            return x >= lo && x <= hi && x == std::floor(x);
        };
        if (key == "walls") {
            if (word == "box" && count(3) && v[0] > 0.0 && v[1] > 0.0 && v[2] > 0.0) {
                sc->walls = DEM_WALLS_BOX;
                for (int k = 0; k < 3; k++) sc->box[k] = v[k];
            } else if (word == "cylinder" && count(1) && v[0] > 0.0) {
                sc->walls = DEM_WALLS_CYLINDER;
                sc->radius = v[0];
            } else {
                return DEM_ERR_IO;
            }
        } else if (key == "height") {
            if (!count(1) || !(v[0] > 0.0)) return DEM_ERR_IO;
            sc->height = v[0];
        } else if (key == "particles") {
            if (!count(1) || !whole(v[0], 1.0, 2147483647.0)) return DEM_ERR_IO;
            sc->particles = (int)v[0];
        } else if (key == "diameters") {
            if (word == "mono" && count(1) && v[0] > 0.0) {
                sc->diameters = 0;
            } else if (word == "lognormal" && count(3) && v[0] >= 0.0 && v[1] > 0.0 && v[2] >= v[1] &&
                       lognormal_window_mass(v[0], v[1], v[2]) >= kMinWindowMass) {
                // the clip window must hold a drawable share of the distribution
                sc->diameters = 1;
            } else if (word == "bimodal" && count(3) && v[0] > 0.0 && v[1] > 0.0 && v[2] >= 0.0 &&
                       v[2] <= 1.0) {
                sc->diameters = 2;
            } else {
                return DEM_ERR_IO;
            }
            for (size_t k = 0; k < 3; k++) sc->dpar[k] = k < v.size() ? v[k] : 0.0;
        } else if (key == "insertion") {
            if (word == "layers" && count(0)) {
                sc->insertion = 0;
            } else if (word == "lattice" && count(3) && v[0] > 0.0 && v[1] >= 0.0) {
                sc->insertion = 1;
            } else if (word == "random" && count(1) && v[0] > 0.0 && v[0] < 0.6) {
                sc->insertion = 2;
            } else {
                return DEM_ERR_IO;
            }
            for (size_t k = 0; k < 3; k++) sc->ipar[k] = k < v.size() ? v[k] : 0.0;
        } else if (key == "seed") {
            if (!count(1) || !whole(v[0], 0.0, 9007199254740992.0)) return DEM_ERR_IO;
            sc->seed = (unsigned long long)v[0];
        } else if (key == "restitution") {
            if (!count(1) || !(v[0] > 0.0 && v[0] <= 1.0)) return DEM_ERR_IO;
            sc->restitution = v[0];
        } else if (key == "rebuild") {
            if (word == "trigger" && count(0)) sc->rebuild_interval = 0;
            else if (word == "interval" && count(1) && whole(v[0], 1.0, 1e9)) sc->rebuild_interval = (int)v[0];
            else return DEM_ERR_IO;
        } else {
            return DEM_ERR_IO;
        }
    }
    return DEM_OK;
}

int dem_deck_material(const dem_deck *d, dem_material *m)
{
    if (!d || !m) return DEM_ERR_INVALID;
    double emod = d->emod, emodw = d->emod;                                  // :141
    double vw = d->vpois;                                                    // :142
    emod = 6.0 * emod / (3.14159265 * d->denp * 9.81 * d->diam);             // :146
    emodw = 6.0 * emodw / (3.14159265 * d->denp * 9.81 * d->diam);           // :147
    m->p.estarp = 1.33333333 * emod / (2.0 * (1.0 - d->vpois * d->vpois));   // :148
    m->w.estarw = 1.33333333 / ((1.0 - vw * vw) / emod + (1.0 - vw * vw) / emodw);   // :149
    m->p.ufricp = d->ufricp;
    m->p.vpois = d->vpois;
    m->p.dampnp = d->dampnp;
    m->p.damptp = d->damptp;
    m->p.upp = d->upp / d->diam;                                             // :224
    m->w.ufricw = d->ufricw;
    m->w.vpoisw = vw;
    m->w.dampnw = d->dampnw;
    m->w.damptw = d->damptw;
    m->w.upw = d->upw / d->diam;                                             // :225
    return DEM_OK;
}

int dem_write_restart(const char *path)
{
    if (!path) return DEM_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "wb");   // dempacking.cu:908-913
    if (!fp) return DEM_ERR_IO;
    std::fprintf(fp, "program comes to end!\n");
    std::fprintf(fp, "%8d %8d\n", 1, 1);
    std::fclose(fp);
    return DEM_OK;
}

int dem_read_restart(const char *path, int *nrestart, int *npregas, int *nparticledat)
{
    if (!path || !nrestart || !npregas || !nparticledat) return DEM_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "rb");   // dempacking.cu:361-366
    if (!fp) return DEM_ERR_IO;
    int a = 0, b = 0, c = 0;
    int got = std::fscanf(fp, "%d %d %d", &a, &b, &c);
    std::fclose(fp);
    if (got != 3) return DEM_ERR_INVALID;   // e.g. the end-of-run text
    *nrestart = a;
    *npregas = b;
    *nparticledat = c;
    return DEM_OK;
}

int dem_write_restart_resume(const char *path, int nrestart, int npregas, int nparticledat)
{
    if (!path) return DEM_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "wb");   // the write left commented at dempacking.cu:829
    if (!fp) return DEM_ERR_IO;
    std::fprintf(fp, "%8d %8d %8d", nrestart, npregas, nparticledat);
    return std::fclose(fp) == 0 ? DEM_OK : DEM_ERR_IO;
}

double dem_restitution_for_damping(double g)
{
    if (!(g >= 0.0)) return -1.0;
    return restitution_ode(g);
}

double dem_damping_for_restitution(double e)
{
    if (!(e > 0.0) || !(e <= 1.0)) return -1.0;
    if (e >= 1.0) return 0.0;
    // bracket, then the Illinois variant of regula falsi on f(g) = e(g) - e (e decreases
    // with g), to the last bits of g
    double lo = 0.0, hi = 0.1;
    double flo = 1.0 - e, fhi = restitution_ode(hi) - e;
    while (fhi > 0.0 && hi < 100.0) {
        lo = hi;
        flo = fhi;
        hi *= 2.0;
        fhi = restitution_ode(hi) - e;
    }
    if (fhi > 0.0) return hi;
    int side = 0;
    for (int it = 0; it < 200 && hi - lo > 1e-13 * hi; it++) {
        double g = (lo * fhi - hi * flo) / (fhi - flo);
        if (!(g > lo && g < hi)) g = 0.5 * (lo + hi);
        double fg = restitution_ode(g) - e;
        if (std::fabs(fg) < 1e-15) return g;
        if (fg > 0.0) {
            lo = g;
            flo = fg;
            if (side == -1) fhi *= 0.5;
            side = -1;
        } else {
            hi = g;
            fhi = fg;
            if (side == 1) flo *= 0.5;
            side = 1;
        }
    }
    return 0.5 * (lo + hi);
}

int dem_gen_diameters(int kind, uint n, double p0, double p1, double p2, unsigned long long seed,
                      double *diam)
{
    if (!diam) return DEM_ERR_INVALID;
    Rng r(seed);
    if (kind == 0) {
        if (!(p0 > 0.0)) return DEM_ERR_INVALID;
        for (uint i = 0; i < n; i++) diam[i] = p0;
        return DEM_OK;
    }
    if (kind == 1) {
        // ln d ~ N(-sigma^2/2, sigma): mean diameter 1, rejection-clipped to [p1, p2]
        double sg = p0, lo = p1, hi = p2;
        if (!(sg >= 0.0) || !(lo > 0.0) || !(hi >= lo)) return DEM_ERR_INVALID;
        if (lognormal_window_mass(sg, lo, hi) <= 0.0) return DEM_ERR_INVALID;
        for (uint i = 0; i < n; i++) {
            double d;
            long draws = 0;
            do {
                if (++draws > kMaxRedraws) return DEM_ERR_INVALID;   // a window too far out
                d = std::exp(-0.5 * sg * sg + sg * r.normal());
            } while (d < lo || d > hi);
            diam[i] = d;
        }
        return DEM_OK;
    }
    if (kind == 2) {
        // bimodal: big p0, small p1, big volume fraction p2; exact counts, seeded shuffle
        double db = p0, ds = p1, fv = p2;
        if (!(db > 0.0) || !(ds > 0.0) || !(fv >= 0.0) || !(fv <= 1.0)) return DEM_ERR_INVALID;
        double nb = fv / (db * db * db), ns = (1.0 - fv) / (ds * ds * ds);
        uint nbig = (uint)std::llround((double)n * nb / (nb + ns));
        for (uint i = 0; i < n; i++) diam[i] = i < nbig ? db : ds;
        for (uint i = n; i > 1; i--) {
            uint j = (uint)(r.next() % i);
            std::swap(diam[i - 1], diam[j]);
        }
        return DEM_OK;
    }
    return DEM_ERR_INVALID;
}

int dem_gen_insert(const dem_walls *w, double z0, double z1, uint n, const double *diam,
                   unsigned long long seed, double *pos)
{
    if (!w || !diam || !pos || !(z1 > z0)) return DEM_ERR_INVALID;
    if (n == 0) return DEM_OK;
    double dmax = 0.0;
    for (uint i = 0; i < n; i++) dmax = std::max(dmax, diam[i]);
    double hx = w->kind == DEM_WALLS_BOX ? 0.5 * w->size[0] : w->radius;
    double hy = w->kind == DEM_WALLS_BOX ? 0.5 * w->size[1] : w->radius;
    double c = dmax;
    int gx = (int)(2.0 * hx / c) + 1, gy = (int)(2.0 * hy / c) + 1, gz = (int)((z1 - z0) / c) + 1;
    size_t ncell = (size_t)gx * gy * gz;
    if (ncell > (size_t)1 << 31) return DEM_ERR_INVALID;
    std::vector<int> head(ncell, -1), next(n, -1);
    auto cell_of = [&](double x, double y, double z, int &ix, int &iy, int &iz) {
        ix = std::min(gx - 1, std::max(0, (int)std::floor((x + hx) / c)));
        iy = std::min(gy - 1, std::max(0, (int)std::floor((y + hy) / c)));
        iz = std::min(gz - 1, std::max(0, (int)std::floor((z - z0) / c)));
    };
    // insert larger spheres first (easier packing); positions keep the caller's indices
    std::vector<uint> order(n);
    for (uint i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint a, uint b) { return diam[a] > diam[b]; });
    Rng r(seed);
    const int kAttempts = 20000;
    for (uint oi = 0; oi < n; oi++) {
        uint i = order[oi];
        double ri = 0.5 * diam[i];
        bool placed = false;
        for (int att = 0; att < kAttempts && !placed; att++) {
            double x, y;
            if (w->kind == DEM_WALLS_BOX) {
                x = -hx + ri + (2.0 * (hx - ri)) * r.uniform();
                y = -hy + ri + (2.0 * (hy - ri)) * r.uniform();
            } else {
                double rr = (w->radius - ri) * std::sqrt(r.uniform());
                double th = 6.283185307179586 * r.uniform();
                x = rr * std::cos(th);
                y = rr * std::sin(th);
            }
            double z = z0 + ri + (z1 - z0 - 2.0 * ri) * r.uniform();
            int ix, iy, iz;
            cell_of(x, y, z, ix, iy, iz);
            bool ok = true;
            for (int dz = -1; dz <= 1 && ok; dz++)
                for (int dy = -1; dy <= 1 && ok; dy++)
                    for (int dx = -1; dx <= 1 && ok; dx++) {
                        int X = ix + dx, Y = iy + dy, Z = iz + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= gx || Y >= gy || Z >= gz) continue;
                        for (int j = head[((size_t)Z * gy + Y) * gx + X]; j >= 0; j = next[j]) {
                            double ex = pos[3 * (size_t)j] - x, ey = pos[3 * (size_t)j + 1] - y,
                                   ez = pos[3 * (size_t)j + 2] - z;
                            double rs = ri + 0.5 * diam[j];
                            if (ex * ex + ey * ey + ez * ez < rs * rs) { ok = false; break; }
                        }
                    }
            if (!ok) continue;
            pos[3 * (size_t)i] = x;
            pos[3 * (size_t)i + 1] = y;
            pos[3 * (size_t)i + 2] = z;
            size_t cidx = ((size_t)iz * gy + iy) * gx + ix;
            next[i] = head[cidx];
            head[cidx] = (int)i;
            placed = true;
        }
        if (!placed) return DEM_ERR_CAPACITY;
    }
    return DEM_OK;
}

int dem_gen_lattice(const dem_walls *w, double z0, double a, double jitter, uint n,
                    unsigned long long seed, double *pos)
{
    if (!w || !pos || !(a > 0.0)) return DEM_ERR_INVALID;
    double hx = w->kind == DEM_WALLS_BOX ? 0.5 * w->size[0] : w->radius;
    double hy = w->kind == DEM_WALLS_BOX ? 0.5 * w->size[1] : w->radius;
    double m = 0.5 + jitter;   // keep jittered unit spheres inside the walls
    int nx = (int)std::floor((2.0 * (hx - m)) / a) + 1;
    int ny = (int)std::floor((2.0 * (hy - m)) / a) + 1;
    if (nx < 1 || ny < 1) return DEM_ERR_INVALID;
    std::vector<double> layer;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            double x = -hx + m + ix * a, y = -hy + m + iy * a;
            if (w->kind == DEM_WALLS_CYLINDER && std::sqrt(x * x + y * y) > w->radius - m) continue;
            layer.push_back(x);
            layer.push_back(y);
        }
    size_t per = layer.size() / 2;
    if (per == 0) return DEM_ERR_INVALID;
    Rng r(seed);
    for (uint i = 0; i < n; i++) {
        size_t L = i / per, k = i % per;
        pos[3 * (size_t)i] = layer[2 * k] + jitter * (2.0 * r.uniform() - 1.0);
        pos[3 * (size_t)i + 1] = layer[2 * k + 1] + jitter * (2.0 * r.uniform() - 1.0);
        pos[3 * (size_t)i + 2] = z0 + (double)L * a + jitter * (2.0 * r.uniform() - 1.0);
    }
    return DEM_OK;
}

int dem_particle_props(uint n, const double *diam, double *rad, double *rmass, double *inert)
{
    if (!diam || !rad || !rmass || !inert) return DEM_ERR_INVALID;
    for (uint i = 0; i < n; i++) {
        rad[i] = 0.5 * diam[i];
        rmass[i] = 8 * rad[i] * rad[i] * rad[i];             // dempacking.cu:311
        inert[i] = 0.4 * rmass[i] * (rad[i] * rad[i]);       // dempacking.cu:312
    }
    return DEM_OK;
}

int dem_make_grid(const dem_walls *w, double h, double c, uint gsize[3], double origin[3])
{
    if (!w || !gsize || !origin || !(c > 0.0) || !(h > 0.0)) return DEM_ERR_INVALID;
    double wx = w->kind == DEM_WALLS_BOX ? w->size[0] : 2.0 * w->radius;
    double wy = w->kind == DEM_WALLS_BOX ? w->size[1] : 2.0 * w->radius;
    origin[0] = -0.5 * wx;                                   // dempacking.cu:333-335
    origin[1] = -0.5 * wy;
    origin[2] = 0.0;
    gsize[0] = (uint)(int)(wx / c) + 1;                      // dempacking.cu:340-342
    gsize[1] = (uint)(int)(wy / c) + 1;
    gsize[2] = (uint)(int)(h / c) + 1;
    return DEM_OK;
}

}  // extern "C"
