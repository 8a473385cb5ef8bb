// ckpt.hpp -- the full-state checkpoint file of dem_save_checkpoint / dem_load_checkpoint:
// header, FNV-1a checksum, writer, and the host-side reader with every consistency check a
// damaged or foreign file must fail before any of it reaches the device.  Plain C++ (no
// device types): the sizes and field offsets of the particle record, instance entry and
// control block come in as a Layout, so the reader also builds on its own (host sanitizer
// runs of the checks).
//
// File: Header (96 bytes) | instance table | records | wall histories | inverse slot map |
// control block | [list: offsets (n+1) | transposed offsets (n+1) | wall flags (n) | wall
// list (n) | pairs (2 words each) | histories | transposed index] | FNV-1a 64 of all of it.
#pragma once
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/dem.h"

namespace dem_ckpt {

constexpr char kMagic[8] = {'D', 'E', 'M', 'C', 'K', 'P', 'T', '1'};
constexpr unsigned kVersion = 1;
constexpr unsigned kEndian = 0x01020304u;
constexpr unsigned kNone = 0xffffffffu;

struct Header {
    char magic[8];
    unsigned version, endian;
    unsigned sz_rec, sz_hist, sz_inst, sz_ctl;
    unsigned ninst, cap_n, n, npairs;
    unsigned have_list, need_rebuild, since, last_flag;
    unsigned n_list, interval, pad0, pad1;
    unsigned long long steps, rebuilds;
};
static_assert(sizeof(Header) == 96, "checkpoint header layout");

// Sizes and byte offsets of the fields the reader checks.
struct Layout {
    unsigned sz_rec, sz_hist, sz_inst, sz_ctl;
    unsigned rec_inst, rec_id;                                   // record: instance, index
    unsigned inst_p_off, inst_n, inst_cell_off, inst_gsize;     // instance entry
    unsigned ctl_npairs, ctl_overflow, ctl_nwall, ctl_zero;     // control block
};

// What the file must hold to fit a context: its instances' particle ranges.
struct Expect {
    unsigned ninst = 0, cap_n = 0;
    std::vector<unsigned> p_off, n;
};

struct Data {
    Header h;
    std::vector<unsigned char> inst, rec, wh, ctl, hist;
    std::vector<unsigned> inv, off, toff, wflag, wlist, ab, jt;   // ab: 2 words per pair
    unsigned long long cells = 0;                                 // total grid cells
};

struct Fnv64 {
    unsigned long long h = 1469598103934665603ull;
    void add(const void *p, size_t bytes)
    {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t k = 0; k < bytes; k++) {
            h ^= b[k];
            h *= 1099511628211ull;
        }
    }
};

struct Writer {
    std::FILE *fp;
    Fnv64 sum;
    bool ok = true;
    void put(const void *p, size_t bytes)
    {
        if (!bytes) return;
        sum.add(p, bytes);
        ok = ok && std::fwrite(p, 1, bytes, fp) == bytes;
    }
};

struct Reader {
    std::FILE *fp;
    Fnv64 sum;
    bool ok = true;
    void get(void *p, size_t bytes)
    {
        if (!bytes || !ok) return;
        ok = std::fread(p, 1, bytes, fp) == bytes;
        if (ok) sum.add(p, bytes);
    }
};

inline unsigned word_at(const unsigned char *base, size_t off)
{
    unsigned v;
    std::memcpy(&v, base + off, sizeof(v));
    return v;
}

// Bytes the file must have for this header (saturates instead of wrapping).
inline unsigned long long expected_bytes(const Header &h)
{
    const unsigned long long N = h.cap_n, n = h.n, np = h.npairs;
    unsigned long long b = sizeof(Header) + (unsigned long long)h.sz_inst * h.ninst +
                           N * (h.sz_rec + h.sz_hist + 4ull) + h.sz_ctl + 8ull;
    if (h.have_list) b += (n + 1) * 8ull + n * 8ull + np * (8ull + h.sz_hist + 4ull);
    return b;
}

// The list read from a file is consistent, so a damaged file can never make a kernel index
// out of bounds: CSR offsets monotone and ending at npairs, slots < n, transposed entries
// pairs or empty, wall candidates < n.
inline bool list_ok(unsigned n, unsigned np, const Data &d, unsigned nwall)
{
    if (d.off[0] != 0 || d.toff[0] != 0 || d.off[n] != np || d.toff[n] != np || nwall > n)
        return false;
    for (unsigned s = 0; s < n; s++)
        if (d.off[s] > d.off[s + 1] || d.toff[s] > d.toff[s + 1]) return false;
    for (unsigned k = 0; k < np; k++) {
        if (d.ab[2 * (size_t)k] >= n || d.ab[2 * (size_t)k + 1] >= n) return false;
        if (d.jt[k] >= np && d.jt[k] != kNone) return false;
    }
    for (unsigned k = 0; k < nwall; k++)
        if (d.wlist[k] >= n) return false;
    return true;
}

// Reads and checks the file at `path`: DEM_OK with `d` filled, DEM_ERR_IO for a damaged,
// truncated or foreign file (magic, version, byte order, layout, size, checksum or an index
// out of range), DEM_ERR_INVALID for a sound file that does not fit `E` (other instances or
// particle counts).  Nothing is allocated before the header's sizes have been checked
// against `E` and against the file's length.
inline int read(const char *path, const Layout &L, const Expect &E, Data &d)
{
    if (!path) return DEM_ERR_INVALID;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return DEM_ERR_IO;
    struct Closer {
        std::FILE *fp;
        ~Closer() { std::fclose(fp); }
    } closer{fp};
    if (std::fseek(fp, 0, SEEK_END) != 0) return DEM_ERR_IO;
    const long flen = std::ftell(fp);
    if (flen < 0 || std::fseek(fp, 0, SEEK_SET) != 0) return DEM_ERR_IO;
    Reader r{fp};
    Header &h = d.h;
    r.get(&h, sizeof(h));
    if (!r.ok || std::memcmp(h.magic, kMagic, sizeof(h.magic)) != 0 || h.version != kVersion ||
        h.endian != kEndian || h.sz_rec != L.sz_rec || h.sz_hist != L.sz_hist ||
        h.sz_inst != L.sz_inst || h.sz_ctl != L.sz_ctl)
        return DEM_ERR_IO;
    // the context must have been created for the same instances (particle counts)
    if (h.ninst != E.ninst || h.cap_n != E.cap_n || h.n == 0 || h.n > h.cap_n ||
        (E.ninst > 1 && h.n != h.cap_n) || h.npairs > 0xfffffff0u || (!h.have_list && h.npairs) ||
        h.have_list > 1u || h.need_rebuild > 1u || h.n_list > h.cap_n ||
        h.interval > 0x7fffffffu)
        return DEM_ERR_INVALID;
    if (expected_bytes(h) != (unsigned long long)flen) return DEM_ERR_IO;
    d.inst.resize((size_t)L.sz_inst * h.ninst);
    r.get(d.inst.data(), d.inst.size());
    if (!r.ok) return DEM_ERR_IO;
    d.cells = 0;
    for (unsigned i = 0; i < h.ninst; i++) {
        const unsigned char *a = d.inst.data() + (size_t)L.sz_inst * i;
        if (word_at(a, L.inst_p_off) != E.p_off[i] || word_at(a, L.inst_n) != E.n[i] ||
            word_at(a, L.inst_cell_off) != d.cells)
            return DEM_ERR_INVALID;
        const unsigned long long gxy =
            (unsigned long long)word_at(a, L.inst_gsize) * word_at(a, L.inst_gsize + 4);
        if (gxy >= 0xffffffffull) return DEM_ERR_INVALID;
        d.cells += gxy * word_at(a, L.inst_gsize + 8);   // < 2^64: no wrap
        if (d.cells >= 0xffffffffull) return DEM_ERR_INVALID;
    }
    const size_t N = h.cap_n, n = h.n, np = h.npairs;
    d.rec.resize(L.sz_rec * N);
    d.wh.resize(L.sz_hist * N);
    d.inv.resize(N);
    d.ctl.resize(L.sz_ctl);
    r.get(d.rec.data(), d.rec.size());
    r.get(d.wh.data(), d.wh.size());
    r.get(d.inv.data(), sizeof(unsigned) * N);
    r.get(d.ctl.data(), d.ctl.size());
    if (h.have_list) {
        d.off.resize(n + 1);
        d.toff.resize(n + 1);
        d.wflag.resize(n);
        d.wlist.resize(n);
        d.ab.resize(2 * np);
        d.hist.resize(L.sz_hist * np);
        d.jt.resize(np);
        r.get(d.off.data(), sizeof(unsigned) * (n + 1));
        r.get(d.toff.data(), sizeof(unsigned) * (n + 1));
        r.get(d.wflag.data(), sizeof(unsigned) * n);
        r.get(d.wlist.data(), sizeof(unsigned) * n);
        r.get(d.ab.data(), sizeof(unsigned) * 2 * np);
        r.get(d.hist.data(), d.hist.size());
        r.get(d.jt.data(), sizeof(unsigned) * np);
    }
    const unsigned long long want = r.sum.h;
    unsigned long long sum = 0;
    if (!r.ok || std::fread(&sum, sizeof(sum), 1, fp) != 1 || sum != want) return DEM_ERR_IO;
    for (size_t s = 0; s < N; s++) {
        const unsigned char *R = d.rec.data() + L.sz_rec * s;
        if (word_at(R, L.rec_id) >= N || d.inv[s] >= N || word_at(R, L.rec_inst) >= h.ninst)
            return DEM_ERR_IO;
    }
    // the kernels bound their pair loops by the control block's count and gate on its zero word
    const unsigned char *C = d.ctl.data();
    if (word_at(C, L.ctl_npairs) != h.npairs || word_at(C, L.ctl_overflow) != 0u ||
        word_at(C, L.ctl_zero) != 0u)
        return DEM_ERR_IO;
    if (h.have_list && !list_ok((unsigned)n, (unsigned)np, d, word_at(C, L.ctl_nwall)))
        return DEM_ERR_IO;
    return DEM_OK;
}

}  // namespace dem_ckpt
