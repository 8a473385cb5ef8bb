// Host readers of libdem under AddressSanitizer / UndefinedBehaviorSanitizer (built and run by
// tests/test_sanitize.py with g++ -fsanitize=address,undefined, no device code): the input
// deck, the deck side-car and the restart.dat readers on the shipped files, every truncation
// of them and malformed lines; the seeded generators at their limits; and the checkpoint
// reader (csrc/ckpt.hpp, the checks dem_load_checkpoint runs before anything reaches the
// device) on a well-formed file in the library's layout and on damaged copies of it --
// every header field, sizes, checksum, out-of-range indices, truncations.
//
// argv: golden-dir decks-dir tmp-dir  followed by the 14 layout numbers of
// dem_checkpoint_layout.  Exit status 0 when every expectation holds.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../discrete-element-method-for-packing-cuda-c-code_amd/csrc/ckpt.hpp"
#include "../../include/dem.h"

namespace {

int g_fail = 0;
std::string g_tmp;

void expect(bool ok, const char *what, long detail = 0)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s (%ld)\n", what, detail);
        g_fail++;
    }
}

std::vector<unsigned char> slurp(const std::string &path)
{
    std::vector<unsigned char> b;
    std::FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) return b;
    unsigned char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), fp)) > 0) b.insert(b.end(), buf, buf + k);
    std::fclose(fp);
    return b;
}

std::string spill(const std::string &name, const void *p, size_t n)
{
    std::string path = g_tmp + "/" + name;
    std::FILE *fp = std::fopen(path.c_str(), "wb");
    if (fp) {
        if (n) std::fwrite(p, 1, n, fp);
        std::fclose(fp);
    }
    return path;
}

std::string spill(const std::string &name, const std::string &s)
{
    return spill(name, s.data(), s.size());
}

// ---- deck, side-car, restart -----------------------------------------------------------

void decks(const std::string &golden, const std::string &decks_dir)
{
    dem_deck d;
    const std::string deck = golden + "/hopp3d.inp";
    expect(dem_read_deck(deck.c_str(), &d) == DEM_OK, "shipped deck");
    std::vector<unsigned char> b = slurp(deck);
    expect(!b.empty(), "deck bytes");
    int ok = 0, err = 0;
    for (size_t cut = 0; cut < b.size(); cut++) {   // every truncation
        std::string p = spill("cut.inp", b.data(), cut);
        int rc = dem_read_deck(p.c_str(), &d);
        expect(rc == DEM_OK || rc == DEM_ERR_IO, "truncated deck status", (long)cut);
        (rc == DEM_OK ? ok : err)++;
    }
    expect(err > 0, "some truncations refused");
    const char *bad[] = {"", "nan nan nan\n", "1e999\n2\n", "-1 -1 -1 -1\n", "\x01\x02\xff\n",
                         "1\n1e7 0.29 2500\n0.01\n2000\n0.001 1e30\n"};
    for (const char *t : bad) {
        std::string p = spill("bad.inp", std::string(t));
        int rc = dem_read_deck(p.c_str(), &d);
        expect(rc == DEM_OK || rc == DEM_ERR_IO, "malformed deck status", rc);
    }
    std::string huge(1 << 20, '7');
    std::string p = spill("long.inp", huge + "\n");
    int rc = dem_read_deck(p.c_str(), &d);
    expect(rc == DEM_OK || rc == DEM_ERR_IO, "long-line deck status", rc);
    expect(dem_read_deck((g_tmp + "/missing.inp").c_str(), &d) == DEM_ERR_IO, "missing deck");

    dem_sidecar sc;
    for (int k = 0; k < 3; k++) {
        std::string side = decks_dir + "/config" + std::to_string(k) + ".side";
        expect(dem_read_sidecar(side.c_str(), &sc) == DEM_OK, "shipped side-car", k);
        std::vector<unsigned char> s = slurp(side);
        for (size_t cut = 0; cut < s.size(); cut++) {
            std::string q = spill("cut.side", s.data(), cut);
            int r = dem_read_sidecar(q.c_str(), &sc);
            expect(r == DEM_OK || r == DEM_ERR_IO, "truncated side-car status", (long)cut);
        }
    }
    const char *lines[] = {"walls", "walls box", "walls box 1 2 3 4", "walls cylinder nan",
                           "diameters lognormal 1e308 1e-308 1e308", "diameters lognormal 0 1 1",
                           "diameters bimodal 1 0.5 0.3 9", "insertion random 1e-320",
                           "insertion lattice 1e308 1e308 1e308", "seed 1e300", "seed 18446744073709551615",
                           "particles 4294967296", "rebuild interval 1e10", "height -0",
                           "restitution inf", "\t\t  ", "# only a comment", "walls box 1e999 1 1"};
    for (const char *l : lines) {
        std::string q = spill("line.side", std::string(l) + "\n");
        int r = dem_read_sidecar(q.c_str(), &sc);
        expect(r == DEM_OK || r == DEM_ERR_IO, "malformed side-car status", r);
    }

    int a = -1, bb = -1, c = -1;
    std::string rs = golden + "/restart.dat";
    // the shipped file is the end-of-run text ("program comes to end!"): no integers, a fresh
    // start (DEM_ERR_INVALID)
    int r = dem_read_restart(rs.c_str(), &a, &bb, &c);
    expect(r == DEM_ERR_INVALID, "shipped restart.dat status", r);
    std::vector<unsigned char> rb = slurp(rs);
    for (size_t cut = 0; cut < rb.size(); cut++) {
        std::string q = spill("cut.dat", rb.data(), cut);
        int r2 = dem_read_restart(q.c_str(), &a, &bb, &c);
        expect(r2 == DEM_OK || r2 == DEM_ERR_INVALID, "truncated restart status", (long)cut);
    }
    std::string q = spill("w.dat", std::string("1 2"));
    expect(dem_write_restart_resume(q.c_str(), 0, 3, 5) == DEM_OK, "restart resume write");
    expect(dem_read_restart(q.c_str(), &a, &bb, &c) == DEM_OK && a == 0 && bb == 3 && c == 5,
           "restart resume round trip");
}

// ---- generators ------------------------------------------------------------------------

void generators()
{
    std::vector<double> d(1000), pos(3000);
    expect(dem_gen_diameters(1, 1000, 0.1, 0.7, 1.3, 3, d.data()) == DEM_OK, "lognormal");
    expect(dem_gen_diameters(1, 10, 0.0, 1.1, 1.3, 3, d.data()) != DEM_OK, "empty window");
    expect(dem_gen_diameters(1, 10, 0.1, 1.2, 1.2, 3, d.data()) != DEM_OK, "zero-width window");
    expect(dem_gen_diameters(2, 1000, 1.0, 0.5, 0.3, 3, d.data()) == DEM_OK, "bimodal");
    expect(dem_gen_diameters(2, 1000, 1.0, 0.5, 2.0, 3, d.data()) != DEM_OK, "bimodal fraction");
    expect(dem_gen_diameters(7, 10, 1.0, 0.0, 0.0, 3, d.data()) != DEM_OK, "unknown kind");
    dem_walls w;
    std::memset(&w, 0, sizeof(w));
    w.kind = DEM_WALLS_BOX;
    w.size[0] = w.size[1] = 3.0;
    w.size[2] = 10.0;
    for (int i = 0; i < 1000; i++) d[i] = 1.0;
    // 1000 unit spheres cannot fit into 3 x 3 x 2: the insertion gives up with a status
    int rc = dem_gen_insert(&w, 0.0, 2.0, 1000, d.data(), 5, pos.data());
    expect(rc != DEM_OK, "overfull insertion refused", rc);
    expect(dem_gen_insert(&w, 0.0, 10.0, 20, d.data(), 5, pos.data()) == DEM_OK, "insertion");
    expect(dem_gen_lattice(&w, 1.0, 1.2, 0.05, 20, 1, pos.data()) == DEM_OK, "lattice");
    unsigned g[3];
    double o[3];
    expect(dem_make_grid(&w, 10.0, 1.55, g, o) == DEM_OK, "grid");
    expect(dem_make_grid(&w, 10.0, 0.0, g, o) != DEM_OK, "zero cell refused");
}

// ---- checkpoint reader -------------------------------------------------------------------

struct Synth {
    dem_ckpt::Layout L;
    dem_ckpt::Expect E;
    std::vector<unsigned char> body;   // everything before the checksum
    size_t off_inst, off_rec, off_inv, off_ctl, off_off, off_toff, off_wlist, off_ab, off_jt;
    unsigned n, np;
};

void put(std::vector<unsigned char> &b, const void *p, size_t k)
{
    const unsigned char *c = static_cast<const unsigned char *>(p);
    b.insert(b.end(), c, c + k);
}

void set_word(std::vector<unsigned char> &b, size_t at, unsigned v)
{
    std::memcpy(b.data() + at, &v, 4);
}

Synth make_synth(const dem_ckpt::Layout &L)
{
    Synth S;
    S.L = L;
    const unsigned ninst = 2, n0 = 20, n1 = 30, n = n0 + n1, np = 120;
    S.n = n;
    S.np = np;
    S.E.ninst = ninst;
    S.E.cap_n = n;
    S.E.p_off = {0, n0};
    S.E.n = {n0, n1};
    dem_ckpt::Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, dem_ckpt::kMagic, 8);
    h.version = dem_ckpt::kVersion;
    h.endian = dem_ckpt::kEndian;
    h.sz_rec = L.sz_rec;
    h.sz_hist = L.sz_hist;
    h.sz_inst = L.sz_inst;
    h.sz_ctl = L.sz_ctl;
    h.ninst = ninst;
    h.cap_n = n;
    h.n = n;
    h.npairs = np;
    h.have_list = 1;
    h.n_list = n;
    h.interval = 0;
    h.steps = 123;
    std::vector<unsigned char> &b = S.body;
    put(b, &h, sizeof(h));
    S.off_inst = b.size();
    std::vector<unsigned char> inst(L.sz_inst * ninst, 0);
    for (unsigned i = 0; i < ninst; i++) {
        unsigned char *a = inst.data() + L.sz_inst * i;
        unsigned po = i ? n0 : 0, nn = i ? n1 : n0, co = i ? 64 : 0, gs[3] = {4, 4, 4};
        std::memcpy(a + L.inst_p_off, &po, 4);
        std::memcpy(a + L.inst_n, &nn, 4);
        std::memcpy(a + L.inst_cell_off, &co, 4);
        std::memcpy(a + L.inst_gsize, gs, 12);
    }
    put(b, inst.data(), inst.size());
    S.off_rec = b.size();
    std::vector<unsigned char> rec(L.sz_rec * n, 0);
    for (unsigned s = 0; s < n; s++) {
        unsigned id = (s * 7) % n, in = id < n0 ? 0u : 1u;
        std::memcpy(rec.data() + L.sz_rec * s + L.rec_id, &id, 4);
        std::memcpy(rec.data() + L.sz_rec * s + L.rec_inst, &in, 4);
    }
    put(b, rec.data(), rec.size());
    std::vector<unsigned char> wh(L.sz_hist * n, 0);
    put(b, wh.data(), wh.size());
    S.off_inv = b.size();
    std::vector<unsigned> inv(n);
    for (unsigned i = 0; i < n; i++) inv[(i * 7) % n] = i;
    put(b, inv.data(), 4 * n);
    S.off_ctl = b.size();
    std::vector<unsigned char> ctl(L.sz_ctl, 0);
    unsigned nwall = 5;
    std::memcpy(ctl.data() + L.ctl_npairs, &np, 4);
    std::memcpy(ctl.data() + L.ctl_nwall, &nwall, 4);
    put(b, ctl.data(), ctl.size());
    std::vector<unsigned> off(n + 1), toff(n + 1), wflag(n, 0), wlist(n, 0), ab(2 * np), jt(np);
    for (unsigned s = 0; s <= n; s++) off[s] = toff[s] = (unsigned)((unsigned long long)np * s / n);
    for (unsigned k = 0; k < np; k++) {
        ab[2 * k] = k * n / np;
        ab[2 * k + 1] = (k * 13) % n;
        jt[k] = k % 3 ? (k * 11) % np : dem_ckpt::kNone;
    }
    for (unsigned k = 0; k < nwall; k++) wlist[k] = k * 3;
    S.off_off = b.size();
    put(b, off.data(), 4 * (n + 1));
    S.off_toff = b.size();
    put(b, toff.data(), 4 * (n + 1));
    put(b, wflag.data(), 4 * n);
    S.off_wlist = b.size();
    put(b, wlist.data(), 4 * n);
    S.off_ab = b.size();
    put(b, ab.data(), 8 * np);
    std::vector<unsigned char> hist(L.sz_hist * np, 0);
    put(b, hist.data(), hist.size());
    S.off_jt = b.size();
    put(b, jt.data(), 4 * np);
    return S;
}

std::string write_ckpt(const std::vector<unsigned char> &body, bool fix_sum, bool flip_sum = false)
{
    dem_ckpt::Fnv64 f;
    f.add(body.data(), body.size());
    unsigned long long sum = f.h;
    if (!fix_sum || flip_sum) sum ^= 1ull;
    std::vector<unsigned char> all(body);
    put(all, &sum, 8);
    return spill("ck.bin", all.data(), all.size());
}

int read_ckpt(const Synth &S, const std::string &path)
{
    dem_ckpt::Data d;
    return dem_ckpt::read(path.c_str(), S.L, S.E, d);
}

void checkpoints(const dem_ckpt::Layout &L)
{
    Synth S = make_synth(L);
    {
        dem_ckpt::Data d;
        std::string p = write_ckpt(S.body, true);
        int rc = dem_ckpt::read(p.c_str(), L, S.E, d);
        expect(rc == DEM_OK, "well-formed checkpoint", rc);
        expect(rc == DEM_OK && d.h.steps == 123 && d.ab.size() == 2 * S.np && d.cells == 128,
               "checkpoint contents");
    }
    expect(read_ckpt(S, write_ckpt(S.body, true, true)) == DEM_ERR_IO, "checksum flip");
    // every header word, changed, with and without a matching checksum
    for (size_t w = 2; w < sizeof(dem_ckpt::Header) / 4; w++) {
        for (unsigned v : {0u, 1u, 3u, 0x7fffffffu, 0xffffffffu}) {
            std::vector<unsigned char> b(S.body);
            unsigned old;
            std::memcpy(&old, b.data() + 4 * w, 4);
            if (old == v) continue;
            set_word(b, 4 * w, v);
            bool benign = w == 14 || w == 15 || w == 18 || w == 19 || w >= 20;   // since,
            // last_flag, pads, steps/rebuilds: any value is sound
            if (w == 16 && v <= S.n) benign = true;                  // n_list within cap_n
            if (w == 17 && v <= 0x7fffffffu) benign = true;          // interval
            if (w == 13 && v <= 1u) benign = true;                   // need_rebuild
            int rc = read_ckpt(S, write_ckpt(b, true));
            if (benign) expect(rc == DEM_OK, "benign header change refused", (long)(w * 100 + v % 97));
            else expect(rc != DEM_OK, "header change accepted", (long)(w * 100 + v % 97));
            expect(read_ckpt(S, write_ckpt(b, false)) != DEM_OK, "stale checksum accepted", (long)w);
        }
    }
    std::vector<unsigned char> magic(S.body);
    magic[0] = 'X';
    expect(read_ckpt(S, write_ckpt(magic, true)) == DEM_ERR_IO, "magic");
    // semantic damage with a matching checksum: each must be refused
    struct Edit {
        size_t at;
        unsigned v;
        const char *what;
    } edits[] = {
        {S.off_rec + L.rec_id, S.n, "record index"},
        {S.off_rec + L.sz_rec * 7 + L.rec_inst, 2, "record instance"},
        {S.off_inv + 4 * 3, S.n, "inverse map"},
        {S.off_ctl + L.ctl_npairs, S.np - 1, "control pair count"},
        {S.off_ctl + L.ctl_overflow, 1, "control overflow"},
        {S.off_ctl + L.ctl_zero, 1, "control zero word"},
        {S.off_ctl + L.ctl_nwall, S.n + 1, "wall count"},
        {S.off_off, 1, "first offset"},
        {S.off_off + 4 * 10, 0, "offsets not monotone"},
        {S.off_off + 4 * S.n, S.np + 1, "last offset"},
        {S.off_toff + 4 * 5, 0xffffffffu, "transposed offsets"},
        {S.off_ab + 8 * 17, S.n, "pair first member"},
        {S.off_ab + 8 * 17 + 4, S.n + 100, "pair second member"},
        {S.off_jt + 4 * 4, S.np, "transposed entry"},
        {S.off_wlist + 4 * 2, S.n, "wall list"},
        {S.off_inst + L.inst_p_off, 1, "instance offset"},
        {S.off_inst + L.sz_inst + L.inst_n, 31, "instance count"},
        {S.off_inst + L.sz_inst + L.inst_cell_off, 63, "instance cell offset"},
        {S.off_inst + L.inst_gsize, 0xffffffffu, "grid size"},
    };
    for (const Edit &e : edits) {
        std::vector<unsigned char> b(S.body);
        set_word(b, e.at, e.v);
        expect(read_ckpt(S, write_ckpt(b, true)) != DEM_OK, e.what);
    }
    // every truncation and an appended tail
    std::vector<unsigned char> all = slurp(write_ckpt(S.body, true));
    for (size_t cut = 0; cut < all.size(); cut++) {
        std::string p = spill("cut.bin", all.data(), cut);
        expect(read_ckpt(S, p) == DEM_ERR_IO, "truncated checkpoint accepted", (long)cut);
    }
    std::vector<unsigned char> longer(all);
    longer.push_back(0);
    expect(read_ckpt(S, spill("long.bin", longer.data(), longer.size())) == DEM_ERR_IO, "tail");
    // a file of another context: other instance count / capacity
    dem_ckpt::Expect E2 = S.E;
    E2.cap_n = S.n + 1;
    dem_ckpt::Data d;
    std::string p = write_ckpt(S.body, true);
    expect(dem_ckpt::read(p.c_str(), L, E2, d) == DEM_ERR_INVALID, "other capacity");
    E2 = S.E;
    E2.n = {21, 29};
    expect(dem_ckpt::read(p.c_str(), L, E2, d) == DEM_ERR_INVALID, "other instance sizes");
    expect(dem_ckpt::read((g_tmp + "/none.bin").c_str(), L, S.E, d) == DEM_ERR_IO, "missing");
    // random bytes of every length up to a few headers
    unsigned long long x = 88172645463325252ull;
    for (size_t len = 0; len < 400; len += 7) {
        std::vector<unsigned char> junk(len);
        for (auto &c : junk) {
            x ^= x << 13;
            x ^= x >> 7;
            x ^= x << 17;
            c = (unsigned char)x;
        }
        if (len >= 8) std::memcpy(junk.data(), dem_ckpt::kMagic, 8);
        expect(read_ckpt(S, spill("junk.bin", junk.data(), junk.size())) != DEM_OK, "junk", (long)len);
    }
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 4 + 14) {
        std::fprintf(stderr, "usage: %s golden-dir decks-dir tmp-dir <14 layout numbers>\n", argv[0]);
        return 2;
    }
    g_tmp = argv[3];
    unsigned v[14];
    for (int k = 0; k < 14; k++) v[k] = (unsigned)std::strtoul(argv[4 + k], nullptr, 10);
    dem_ckpt::Layout L{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], v[11],
                       v[12], v[13]};
    decks(argv[1], argv[2]);
    generators();
    checkpoints(L);
    std::printf("host readers: %d failures\n", g_fail);
    return g_fail ? 1 : 0;
}
