// gpusphere.cpp -- DEM packing driver: the reference's main() flow (dempacking.cu:28-918) on
// top of libdem.so's context API.
//
//   gpusphere [hopp3d.inp] [--side FILE] [--tstop N] [--quiet] [--checkpoint-every N]
//             [--max-cpu-seconds S]
//
// Reads the positional deck, converts to reduced units, builds the layer lattice with the
// reference's glibc rand() velocities, inserts one layer every tnewpar = 0.5 time units,
// steps with the displacement-triggered list rebuild, and writes view.dat, ppor.dat (every
// 1000 steps), timesteps.dat, particle.dat (step 1 and every 50000), jobfinished and
// restart.dat in the reference's formats.  State stays in HBM; no per-layer reallocation.
//
// Checkpoint and resume (the reference's preflow.dat / restart.dat scheme, dempacking.cu:
// 358-367, 794-875, finished): a checkpoint writes preflow.dat in the reference's text format
// and preflow.bin, the library's lossless full-state checkpoint (pair list and contact
// histories included), then restart.dat as "0 0 nparticledat".  It is taken every N steps
// with --checkpoint-every N, at the reference's CPU-time marks (3600 s, then every 7200 s,
// checked every 1000 steps), and before the reference's CPU-time stop (41400 s, or
// --max-cpu-seconds), after which the driver exits with status 3.  A run that finds
// restart.dat starting with nrestart = 0 loads preflow.bin and continues from the step after
// it; the steps it runs are the uninterrupted run's steps, bit for bit.
//
// --side FILE reads a deck side-car (dem_read_sidecar): container, particle count, diameter
// distribution, insertion (the reference's layers, a jittered lattice or random insertion),
// seed, restitution and rebuild policy -- the knobs the 12-line deck has no line for, so the
// deck keeps the reference's layout.  Without it the run is the reference's.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/dem.h"

namespace {

int fail(const char *what, int rc)
{
    std::fprintf(stderr, "gpusphere: %s failed: %s\n", what, dem_status_string(rc));
    return 1;
}

void write_particles(const char *path, int zone, const std::vector<double> &pos,
                     const std::vector<double> &rad, unsigned n, double diam)
{
    std::FILE *fp = std::fopen(path, "ab+");
    if (!fp) {
        std::printf("Cannot open file particle01.dat! %d\n", zone);
        return;
    }
    std::fprintf(fp, "ZONE T=\"ZONE%03d\"\n", zone);   // dempacking.cu:786-790
    for (unsigned i = 0; i < n; i++)
        std::fprintf(fp, "%15.6g%15.6g%15.6g%15.6g%8d\n", pos[3 * i] * diam, pos[3 * i + 1] * diam,
                     pos[3 * i + 2] * diam, rad[i] * diam, (int)i);
    std::fclose(fp);
}

// preflow.dat (dempacking.cu:812-820), preflow.bin and restart.dat for a resume after step it.
int write_checkpoint(dem_ctx *ctx, unsigned long long it, unsigned nparticles, int nparticledat,
                     std::vector<double> &pos, std::vector<double> &pdis, std::vector<double> &angv)
{
    int rc = dem_get_state(ctx, pos.data(), pdis.data(), angv.data(), nullptr, 0);
    if (rc) return fail("dem_get_state", rc);
    if (std::FILE *fp = std::fopen("preflow.dat", "wb+")) {
        std::fprintf(fp, "%8d,%8d,%8d\n", (int)it, (int)nparticles, 0);
        for (unsigned i = 0; i < nparticles; i++)
            std::fprintf(fp, "%15.6f %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n",
                         pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], pdis[3 * i], pdis[3 * i + 1],
                         pdis[3 * i + 2], angv[3 * i], angv[3 * i + 1], angv[3 * i + 2]);
        std::fclose(fp);
    } else {
        std::printf("Cannot open file preflow.dat!");
    }
    if ((rc = dem_save_checkpoint(ctx, "preflow.bin"))) return fail("dem_save_checkpoint", rc);
    if ((rc = dem_write_restart_resume("restart.dat", 0, 0, nparticledat)))
        return fail("dem_write_restart_resume", rc);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *deck_path = "hopp3d.inp";
    const char *side_path = nullptr;
    long long tstop_override = -1, ckpt_every = 0;
    double max_cpu = 41400.0;   // dempacking.cu:841
    bool quiet = false;
    for (int a = 1; a < argc; a++) {
        if (!std::strcmp(argv[a], "--tstop") && a + 1 < argc) tstop_override = std::atoll(argv[++a]);
        else if (!std::strcmp(argv[a], "--checkpoint-every") && a + 1 < argc) ckpt_every = std::atoll(argv[++a]);
        else if (!std::strcmp(argv[a], "--max-cpu-seconds") && a + 1 < argc) max_cpu = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--side") && a + 1 < argc) side_path = argv[++a];
        else if (!std::strcmp(argv[a], "--quiet")) quiet = true;
        else deck_path = argv[a];
    }
    dem_deck d;
    int rc = dem_read_deck(deck_path, &d);
    if (rc) {
        std::printf("Cannot open file hopp3d.inp !");
        return 1;
    }
    // echo (dempacking.cu:128-139)
    std::printf("%d\n", d.iforcemode);
    std::printf("%lf %lf %lf\n", d.emod, d.vpois, d.denp);
    std::printf("%lf\n", d.diam);
    std::printf("%d\n", d.ntot);
    std::printf("%lf %u\n", d.dt, d.tstop);
    std::printf("%lf %lf %lf\n", d.dmt1, d.dmt2, d.thick);
    std::printf("%lf %lf\n", d.zht1, d.zht2);
    std::printf("%lf %d\n", d.vel0, d.nrings);
    std::printf("%lf %lf\n", d.ufricp, d.ufricw);
    std::printf("%lf %lf %lf %lf %lf\n", d.dampnp, d.damptp, d.dampnw, d.damptw, d.cdamp);
    std::printf("%lf %lf %lf\n", d.stiffness, d.dampcoeffn, d.dampcoefft);
    std::printf("%lf %lf\n", d.upp, d.upw);
    unsigned long long tstop = tstop_override >= 0 ? (unsigned long long)tstop_override : d.tstop;
    dem_sidecar side;
    std::memset(&side, 0, sizeof(side));
    side.walls = -1;
    side.dpar[0] = 1.0;
    side.seed = 1;
    side.restitution = -1.0;
    if (side_path && (rc = dem_read_sidecar(side_path, &side))) return fail("dem_read_sidecar", rc);
    if (side.particles > 0) d.ntot = side.particles;

    dem_material mat;
    dem_deck_material(&d, &mat);
    if (side.restitution > 0.0) {   // both laws' dashpots from the restitution coefficient
        double g = dem_damping_for_restitution(side.restitution);
        if (g < 0.0) return fail("dem_damping_for_restitution", DEM_ERR_INVALID);
        mat.p.dampnp = mat.p.damptp = mat.w.dampnw = mat.w.damptw = g;
    }
    const double diam = d.diam;
    const double pi = std::atan(1.0) * 4.0;   // dempacking.cu:160
    const double dt = d.dt;
    // geometry in units of d (dempacking.cu:213-223)
    const double thick = d.thick / diam;
    const double dmt2 = d.dmt2 / diam;
    const double dmt1 = d.dmt1 / diam;
    const double hgt1 = d.zht1 / diam, hgt2 = hgt1 + d.zht2 / diam, hopzt = hgt2;
    (void)dmt1;

    if (std::FILE *fp6 = std::fopen("view.dat", "wb+")) {   // dempacking.cu:227-234
        std::fprintf(fp6, "number of particles=%4d\n", d.ntot);
        std::fprintf(fp6, "diameters =%6.3f\n", diam);
        std::fprintf(fp6, "time step=%f bed width=%6.3f thickniss=%6.3f\n", dt, d.dmt1 / diam, dmt2);
        std::fprintf(fp6, "glabal damping=%10.4f bed hbot=%10.4f htop=%10.4f\n", d.cdamp, hgt1, hgt2);
        std::fclose(fp6);
    }

    // container (dempacking.cu:213-223, or the side-car's) and diameters
    const unsigned ntot = (unsigned)d.ntot;
    if (ntot == 0) return fail("deck", DEM_ERR_INVALID);
    double world[3] = {dmt2, thick, hopzt};
    dem_walls walls;
    std::memset(&walls, 0, sizeof(walls));
    walls.kind = DEM_WALLS_BOX;
    if (side.walls == DEM_WALLS_BOX) {
        for (int k = 0; k < 3; k++) world[k] = side.box[k];
    } else if (side.walls == DEM_WALLS_CYLINDER) {
        walls.kind = DEM_WALLS_CYLINDER;
        walls.radius = side.radius;
        world[0] = world[1] = 2.0 * side.radius;
    }
    if (side.height > 0.0) world[2] = side.height;
    std::vector<double> diamv(ntot, side.dpar[0]);
    if (side.diameters != 0 &&
        (rc = dem_gen_diameters(side.diameters, ntot, side.dpar[0], side.dpar[1], side.dpar[2],
                                side.seed, diamv.data())))
        return fail("dem_gen_diameters", rc);
    double dmax = 0.0, vol = 0.0;
    std::vector<double> cumvol(ntot + 1, 0.0);   // volume of particles [0, i)
    bool unit_spheres = true;
    for (unsigned i = 0; i < ntot; i++) {
        const double x = diamv[i];
        dmax = std::max(dmax, x);
        vol += pi / 6.0 * x * x * x;
        cumvol[i + 1] = vol;
        unit_spheres = unit_spheres && x == 1.0;
    }
    const double area = walls.kind == DEM_WALLS_BOX ? world[0] * world[1] : pi * side.radius * side.radius;
    double fill_height = 0.0;   // random insertion: the height that gives the asked fill
    if (side.insertion == 2) {
        fill_height = vol / side.ipar[0] / area;
        if (side.height > 0.0) fill_height = std::min(fill_height, side.height);
        else world[2] = fill_height;
    }
    walls.size[0] = world[0];
    walls.size[1] = world[1];
    walls.size[2] = world[2];

    std::vector<double> pos(3 * (size_t)ntot, 0.0), pdis(3 * (size_t)ntot, 0.0), angv(3 * (size_t)ntot, 0.0);
    std::vector<double> rad(ntot), rmass(ntot), inert(ntot);
    if ((rc = dem_particle_props(ntot, diamv.data(), rad.data(), rmass.data(), inert.data())))
        return fail("dem_particle_props", rc);
    long nplayer = 1;
    if (side.insertion == 0) {
        // layer lattice and initial velocities (dempacking.cu:242-316)
        if (walls.kind != DEM_WALLS_BOX) return fail("insertion layers needs a box", DEM_ERR_INVALID);
        const double wx = world[0], wy = world[1], wx5 = wx * 0.5;
        // the reference's ring counts (unsigned)(w) - 3 need both sides >= 4 d once the
        // lattice is used (wx > 3.98): narrower boxes would wrap or divide by zero
        if (wx5 > 1.99 && (wx < 4.0 || wy < 4.0))
            return fail("insertion layers needs a box at least 4 d wide on both sides", DEM_ERR_INVALID);
        std::vector<double> xp, yp;
        double nrings = (unsigned)(wx) - 3;
        double nringsy = (unsigned)(wy) - 3;
        if (wx5 > 1.99) {
            double dringsx = wx / (float)(nrings);
            std::vector<double> rring((size_t)nrings), rringy((size_t)nringsy);
            for (long ir = 0; ir < (long)nrings; ir++) rring[ir] = ((float)(ir) + 0.5) * dringsx - 0.5 * wx;
            double dringsy = wy / (float)(nringsy);
            for (long ir = 0; ir < (long)nringsy; ir++) rringy[ir] = ((float)(ir) + 0.5) * dringsy - 0.5 * wy;
            for (long ir = 0; ir < (long)nrings; ir++)
                for (long ij = 0; ij < (long)nringsy; ij++) {
                    xp.push_back(rring[ir]);
                    yp.push_back(rringy[ij]);
                }
            nplayer = (long)xp.size();
        } else {
            xp.push_back(0.0);
            yp.push_back(0.0);
            nplayer = 1;
        }
        long nlayers = (long)(ntot / nplayer) + 1;
        double zp0 = world[2] - 5.0;
        unsigned i = 0;
        for (long ilay = 0; ilay < nlayers; ilay++)
            for (long ip = 0; ip < nplayer; ip++) {
                if (i == ntot) break;
                pos[3 * i] = xp[ip];
                pos[3 * i + 1] = yp[ip];
                pos[3 * i + 2] = zp0;
                double rno = std::rand() / (double)(RAND_MAX);
                double ang = pi * (rno * .33333 + .33333);
                pdis[3 * i + 2] = -d.vel0 * std::sin(ang) * dt;
                double dxy = -d.vel0 * std::cos(ang) * dt;
                rno = std::rand() / (double)(RAND_MAX);
                ang = 2.0 * pi * rno;
                pdis[3 * i] = dxy * std::cos(ang);
                pdis[3 * i + 1] = dxy * std::sin(ang);
                i++;
            }
    } else if (side.insertion == 1) {
        if ((rc = dem_gen_lattice(&walls, side.ipar[2], side.ipar[0], side.ipar[1], ntot, side.seed,
                                  pos.data())))
            return fail("dem_gen_lattice", rc);
    } else {
        if ((rc = dem_gen_insert(&walls, 0.0, fill_height, ntot, diamv.data(), side.seed + 1,
                                 pos.data())))
            return fail("dem_gen_insert", rc);
    }

    // grid (dempacking.cu:329-344): cell = largest diameter + skin, 1.55 for the deck's spheres
    dem_instance inst;
    std::memset(&inst, 0, sizeof(inst));
    inst.n = ntot;
    const double skin = 1.55 - 1.0;   // cub - cua (neiglist.cu:27-28)
    const double cell = dmax + skin;
    double origin[3];
    if ((rc = dem_make_grid(&walls, world[2], cell, inst.gsize, origin))) return fail("dem_make_grid", rc);
    for (int k = 0; k < 3; k++) {
        inst.cell[k] = cell;
        inst.origin[k] = origin[k];
    }
    inst.walls = walls;
    inst.mat = mat;
    inst.dt = dt;
    inst.skin = skin;

    dem_ctx *ctx = nullptr;
    if ((rc = dem_create(&ctx, 1, &inst, 0))) return fail("dem_create", rc);
    if ((rc = dem_set_state(ctx, pos.data(), pdis.data(), angv.data(), rad.data(), rmass.data(),
                            inert.data(), 0)))
        return fail("dem_set_state", rc);
    if (side.rebuild_interval > 0 &&
        (rc = dem_set_option(ctx, DEM_OPT_REBUILD_INTERVAL, side.rebuild_interval)))
        return fail("dem_set_option", rc);

    // the reference's layers enter one by one; lattice and random insertion start complete
    unsigned nparticles = side.insertion == 0 ? 0u : ntot;
    int nparticledat = 0;
    unsigned long long it0 = 1;
    {
        int nrestart = 1, npregas = 0, ndat = 0;
        if (dem_read_restart("restart.dat", &nrestart, &npregas, &ndat) == DEM_OK && nrestart == 0) {
            if ((rc = dem_load_checkpoint(ctx, "preflow.bin"))) return fail("dem_load_checkpoint", rc);
            dem_stats st;
            if ((rc = dem_get_stats(ctx, &st))) return fail("dem_get_stats", rc);
            it0 = st.steps + 1;
            nparticles = dem_active_count(ctx);
            nparticledat = ndat;
            std::printf("resuming at step %llu with %u particles\n", it0, nparticles);
        }
    }
    const double tnewpar = 0.50;
    double realtime = 3600.0;   // dempacking.cu:161
    std::clock_t start = std::clock();
    for (unsigned long long it = it0; it <= tstop; it++) {
        double time = dt * (double)it;
        if (nparticles < ntot && time > (float)nparticles * tnewpar / float(nplayer)) {   // :378-381
            unsigned nn = nparticles + (unsigned)nplayer;
            if (nn > ntot) nn = ntot;
            if ((rc = dem_set_active(ctx, nn))) return fail("dem_set_active", rc);
            nparticles = nn;
        }
        if ((rc = dem_step(ctx, 1))) return fail("dem_step", rc);
        if (it % 1000 == 0) {   // dempacking.cu:758-777
            dem_stats st;
            if ((rc = dem_get_stats(ctx, &st))) return fail("dem_get_stats", rc);
            double avercn = st.pcntot / (nparticles * 1.0 + 1.0e-30);
            double xvolumespace = area * st.maxhz;
            // the reference counts unit spheres (N pi / 6, dempacking.cu:769); other diameters
            // count their own volume
            double pvol = unit_spheres ? nparticles * pi / 6 : cumvol[nparticles];
            double porosity5 = 1.0 - pvol / xvolumespace;
            double bedaver = 1.0 - porosity5;
            if (std::FILE *fp = std::fopen("ppor.dat", "ab+")) {
                std::fprintf(fp, "%8d %8d %8d %15.6f %15.6f %15.6f %15.6f\n", (int)it, (int)nparticles,
                             (int)(st.pcntot / 2), porosity5, bedaver, avercn, st.maxhz);
                std::fclose(fp);
            }
            if (!quiet)
                std::printf("%8d %8d %8d %15.6f %15.6f %15.6f %15.6f\n", (int)it, (int)nparticles,
                            (int)(st.pcntot / 2), porosity5, bedaver, avercn, st.maxhz);
            double tret = (double)(std::clock() - start) / CLOCKS_PER_SEC;   // :796-804
            if (std::FILE *fp = std::fopen("timesteps.dat", "ab+")) {
                std::fprintf(fp, "%15.6f %15.6f\n", tret, it * 1.0);
                std::fclose(fp);
            }
        }
        if (it % 50000 == 0 || it == 1) {   // dempacking.cu:779-792
            if ((rc = dem_get_state(ctx, pos.data(), nullptr, nullptr, nullptr, 0)))
                return fail("dem_get_state", rc);
            nparticledat++;
            write_particles("particle.dat", nparticledat, pos, rad, nparticles, diam);
        }
        bool ckpt = ckpt_every > 0 && it % (unsigned long long)ckpt_every == 0 && it < tstop;
        bool stop = false;
        if (it % 1000 == 0) {   // dempacking.cu:805-823, 837-841
            double tret = (double)(std::clock() - start) / CLOCKS_PER_SEC;
            if (tret >= realtime) {
                ckpt = true;
                realtime += 7200.0;
            }
            if (tret > max_cpu && it < tstop) ckpt = stop = true;
        }
        if (ckpt && write_checkpoint(ctx, it, nparticles, nparticledat, pos, pdis, angv)) return 1;
        if (stop) {
            std::printf("CPU-time limit reached after step %llu: checkpoint written\n", it);
            dem_destroy(ctx);
            return 3;
        }
    }
    if (std::FILE *fp = std::fopen("jobfinished", "wb")) {   // dempacking.cu:904-907
        std::fprintf(fp, "job finished\n");
        std::fclose(fp);
    }
    if ((rc = dem_write_restart("restart.dat"))) return fail("dem_write_restart", rc);
    std::printf("program comes to end!\n");
    dem_destroy(ctx);
    return 0;
}
