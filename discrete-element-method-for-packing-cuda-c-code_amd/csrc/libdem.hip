// libdem.hip -- host side of libdem.so: context API and the reference-named drop-in entry
// points.  Single translation unit: the kernels live in kernels.hpp / sort.hpp / law.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/dem.h"
#include "kernels.hpp"
#include "adjoint.hpp"
#include "law_adj.hpp"
#include "sort.hpp"
#include "ckpt.hpp"

using namespace dem;

namespace {

int g_last_error = 0;

#define DEM_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "GPUassert: %s %s %d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                        \
            g_last_error = DEM_ERR_HIP;                                                    \
            return DEM_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

#define DEM_HIP_V(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "GPUassert: %s %s %d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                        \
            g_last_error = DEM_ERR_HIP;                                                    \
            return;                                                                        \
        }                                                                                  \
    } while (0)

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

LawParams law_from(const Matproperties &m)
{
    return LawParams{m.estarp, m.ufricp, m.vpois, m.dampnp, m.damptp, m.upp};
}
LawParams law_from(const Wallproperties &w)
{
    return LawParams{w.estarw, w.ufricw, w.vpoisw, w.dampnw, w.damptw, w.upw};
}

// Scratch for scans and sorts, grown on demand (never per step once sized).
struct Scratch {
    unsigned *tiles = nullptr;
    size_t tiles_cap = 0;
    unsigned *counts = nullptr;   // radix digit counts [256 * ntiles]
    unsigned *offs = nullptr;     // scanned counts (+1)
    size_t counts_cap = 0;
    unsigned *total = nullptr;    // 4 device words
    bool fixed = false;           // a view of preallocated memory: never (re)allocates
    int ensure(size_t ntiles_scan, size_t ncounts)
    {
        if (fixed)
            return (total && ntiles_scan <= tiles_cap && ncounts <= counts_cap) ? DEM_OK
                                                                               : DEM_ERR_CAPACITY;
        if (!total) DEM_HIP(hipMalloc(&total, 4 * sizeof(unsigned)));
        if (ntiles_scan > tiles_cap) {
            if (tiles) DEM_HIP(hipFree(tiles));
            tiles_cap = std::max(ntiles_scan, (size_t)64);
            DEM_HIP(hipMalloc(&tiles, tiles_cap * sizeof(unsigned)));
        }
        if (ncounts > counts_cap) {
            if (counts) DEM_HIP(hipFree(counts));
            if (offs) DEM_HIP(hipFree(offs));
            counts_cap = ncounts;
            DEM_HIP(hipMalloc(&counts, counts_cap * sizeof(unsigned)));
            DEM_HIP(hipMalloc(&offs, (counts_cap + 1) * sizeof(unsigned)));
        }
        return DEM_OK;
    }
    void release()
    {
        if (tiles) (void)hipFree(tiles);
        if (counts) (void)hipFree(counts);
        if (offs) (void)hipFree(offs);
        if (total) (void)hipFree(total);
        tiles = counts = offs = total = nullptr;
        tiles_cap = counts_cap = 0;
    }
};

// Exclusive scan of n words: out[i] = sum(in[0..i)); if `total` is non-null it receives the
// total and out[n] is written too.
int scan_excl(Scratch &S, const unsigned *in, unsigned n, unsigned *out, unsigned *total,
              hipStream_t st, const unsigned *active)
{
    unsigned nt = std::max(1u, cdiv(n, kScanTile));
    if (S.ensure(nt, 0)) return DEM_ERR_HIP;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, st, in, n, S.tiles, active);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, st, S.tiles, nt, total, active);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kScanThreads), 0, st, in, n, S.tiles, out,
                       (const unsigned *)total, active);
    return hipGetLastError() == hipSuccess ? DEM_OK : DEM_ERR_HIP;
}

// Exclusive scans of two n-word arrays at once: out_k[i] = sum(in_k[0..i)), out_k[n] and
// total[k] = the totals (total: 2 device words).
int scan2_excl(Scratch &S, const unsigned *in0, const unsigned *in1, unsigned n, unsigned *out0,
               unsigned *out1, unsigned *total, hipStream_t st, const unsigned *active)
{
    unsigned nt = std::max(1u, cdiv(n, kScanTile));
    if (S.ensure(2 * (size_t)nt, 0)) return DEM_ERR_HIP;
    hipLaunchKernelGGL(k_scan2_tiles, dim3(nt), dim3(kScanThreads), 0, st, in0, in1, n, S.tiles, active);
    hipLaunchKernelGGL(k_scan2_top, dim3(1), dim3(kScanThreads), 0, st, S.tiles, nt, total, active);
    hipLaunchKernelGGL(k_scan2_apply, dim3(nt), dim3(kScanThreads), 0, st, in0, in1, n,
                       (const unsigned *)S.tiles, out0, out1, (const unsigned *)total, active);
    return hipGetLastError() == hipSuccess ? DEM_OK : DEM_ERR_HIP;
}

// Stable LSD radix sort of n (key, value) pairs over the low `bits` key bits.  Reads `in`
// (left untouched unless it is one of the ping-pong buffers), ping-pongs a <-> b, returns
// the buffer holding the result.
uint2 *radix_sort(Scratch &S, const uint2 *in, uint2 *a, uint2 *b, unsigned n, int bits,
                  hipStream_t st, const unsigned *active, int *rc)
{
    *rc = DEM_OK;
    int passes = (bits + 7) / 8;
    unsigned nt = std::max(1u, cdiv(n, kSortTile));
    size_t ncounts = (size_t)256 * nt;
    if (S.ensure(std::max(1u, cdiv(ncounts, kScanTile)), ncounts)) { *rc = DEM_ERR_HIP; return nullptr; }
    const uint2 *src = in;
    uint2 *dst = a;
    if (passes == 0) return const_cast<uint2 *>(in);
    for (int p = 0; p < passes; p++) {
        if (src == dst) dst = (dst == a) ? b : a;
        int sh = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(kSortThreads), 0, st, src, n, sh, nt,
                           S.counts, active);
        if (scan_excl(S, S.counts, (unsigned)ncounts, S.offs, nullptr, st, active)) {
            *rc = DEM_ERR_HIP;
            return nullptr;
        }
        hipLaunchKernelGGL(k_radix_scatter, dim3(nt), dim3(kSortThreads), 0, st, src, n, sh, nt,
                           (const unsigned *)S.offs, dst, active);
        src = dst;
        dst = (dst == a) ? b : a;
    }
    if (hipGetLastError() != hipSuccess) *rc = DEM_ERR_HIP;
    return const_cast<uint2 *>(src);
}

int current_device_cus()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0)
            cus = p.multiProcessorCount;
    }
    return cus;
}

}  // namespace

// =========================================================================================
// Context
// =========================================================================================

struct dem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ninst = 0;
    bool batched = false;
    std::vector<InstDev> hinst;
    InstDev *dinst = nullptr;
    unsigned n = 0, ncells = 0;   // n = active particles (<= cap_n, single instance may grow)
    unsigned cap_n = 0;           // allocated particle slots
    int cus = 256;

    PRec *rec[2] = {nullptr, nullptr};
    Hist *whist[2] = {nullptr, nullptr};
    unsigned *inv[2] = {nullptr, nullptr};
    int rc = 0;   // index of the current record buffer (flips at rebuilds and fused steps)
    int pc = 0;   // index of the current wall histories and inverse slot map (flips at rebuilds)
    unsigned *contacti = nullptr;
    unsigned *wflag = nullptr, *wlist = nullptr;   // wall candidates of the current rebuild
    // Position + radius per slot (32 bytes: four slots per cache line): written by the reorder,
    // read by the candidate searches of the same rebuild.
    double4 *cpos[2] = {nullptr, nullptr};
    // Its FP32 copy (16 bytes), cp32[rc] mirroring rec[rc]: written by the reorder and by both
    // step paths, read by the step kernels' contact filter.
    float4 *cp32[2] = {nullptr, nullptr};
    unsigned *cid = nullptr;   // particle index per slot, written by the reorder
    double *wforce = nullptr;                      // k_walls output, 8 doubles per slot
    double *fdbg = nullptr;
    bool record_forces = false;

    // rebuild sort: kv_hash (cell, index) by slot, kv_a the scatter, kv_sorted (= kv_b) the
    // sorted pairs; cstart[c] = first slot of cell c (ncells + 1 entries, cstart[ncells] = n);
    // ccount the per-cell counts
    uint2 *kv_hash = nullptr, *kv_a = nullptr, *kv_b = nullptr, *kv_sorted = nullptr;
    unsigned *cstart = nullptr, *ccount = nullptr;
    unsigned *ng = nullptr, *nl = nullptr, *toff = nullptr;
    unsigned long long *rmask = nullptr;   // the count's row masks, 2 kRows words per slot
    unsigned *rs0 = nullptr;               // the count's row starts, kRows words per slot
    unsigned *oslot = nullptr;             // per new slot, the particle's old-list slot
    unsigned *off[2] = {nullptr, nullptr};
    uint2 *ab[2] = {nullptr, nullptr};
    unsigned cap[2] = {0, 0};
    int lc = 0;   // index of the current pair list
    // Tangential histories of the current list: hist[hc] holds them, hist[1-hc] is the
    // other buffer (the fused step writes it and flips hc; a rebuild fills it for the new
    // list and flips hc).  Both always hold at least the current list.
    Hist *hist[2] = {nullptr, nullptr};
    unsigned hcap[2] = {0, 0};
    int hc = 0;
    bool fused = true;        // step path in use: fused per-particle (k_step) or split
    int fused_mode = 2;       // DEM_OPT_FUSED_STEP: 0 split, 1 fused, 2 auto (timed probes)
    unsigned fused_fast = kFast;   // DEM_OPT_FUSED_FAST: one-pass limit of the fused step
    unsigned xcd_chunk = 8;        // DEM_OPT_XCD_CHUNK: blocks per XCD chunk of the step kernels
    unsigned tpar = 0;             // parity of the next queued fused step (DevCtl::trig)
    int probe_left = 0;       // auto: steps left in the current probe (fused, then split;
                              // the first probe runs the pair twice and keeps the second)
    unsigned reb_since_probe = 0, probe_every = 8;
    bool probed = false;
    float t_probe[2] = {0.f, 0.f};
    unsigned long long probe_delay = 0;   // wall-clock ticks of the busy period before a probe
    hipEvent_t evp0 = nullptr, evp1 = nullptr;
    // kernel timing of dem_step (DEM_OPT_TIME_STEPS): a ring of event sets, read back lazily
    struct TimedStep {
        hipEvent_t e[3] = {nullptr, nullptr, nullptr};   // before, mid, after the step kernels
        hipEvent_t r[2] = {nullptr, nullptr};            // around a rebuild
        bool used = false, reb = false, fused = false;
        unsigned np = 0;
    };
    static constexpr int kRing = 64;
    TimedStep ring[kRing];
    int ring_pos = 0;
    bool timing = false;
    // queued steps under the displacement trigger: up to `pipeline` steps in flight, each
    // reporting its outcome in a host-mapped slot (2 = not run)
    static constexpr int kQueue = 16;
    int pipeline = 16;
    int pipeline_batch = 8;   // DEM_OPT_PIPELINE_BATCH: queued steps per recorded event
    hipEvent_t evq[kQueue] = {};
    unsigned *qflag_h = nullptr, *qflag_d = nullptr;
    double t_acc[2][2] = {{0, 0}, {0, 0}};   // [path][first kernel group, per-particle kernel]
    double np_acc[2] = {0, 0}, n_acc[2] = {0, 0}, t_reb = 0, n_reb = 0;
    int last_path = 0;        // path of the last step: 0 none, 1 split, 2 fused
    bool out_valid = false;   // fused: rec[1-rc] / hist[1-hc] hold the last step's inputs
    POut *out = nullptr;
    unsigned *jt = nullptr;
    unsigned *jf = nullptr;   // jt's companion: the slot of the pair's first member
    unsigned *bo = nullptr;   // the second member's slot of each pair of the current list
    // Per pair of the current list, bit k: hist[k] holds an exact zero for the pair.  Valid
    // (zs_ok) only across consecutive fused steps: a rebuild, a split step or a restore of the
    // histories leaves it stale, and the next fused step then rewrites it from scratch.
    unsigned char *zs = nullptr;
    bool zs_ok = false;
    unsigned out_cap = 0;
    unsigned npairs = 0;
    unsigned np_peak = 0;     // largest list built by rebuild_host since the last reset
    bool have_list = false;
    unsigned n_list = 0;

    Scratch S;
    DevCtl *ctl = nullptr;
    size_t ctl_bytes = 0;
    unsigned char *mirror_h = nullptr, *mirror_d = nullptr;

    int interval = 0;
    unsigned pair_block = 256, pair_grid_per_cu = 16, particle_block = 128;   // launch shapes
    unsigned long long steps = 0, rebuilds = 0;
    unsigned since = 0;
    bool need_rebuild = true;
    unsigned last_flag = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evs = nullptr;
    float last_ms = 0.f;

    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    int graph_k = 0;
    unsigned graph_np_grid = 0;
    bool graph_fused = false;
    int gpar[4] = {0, 0, 0, 0};   // (rc, lc, hc, pc) after one captured group
    // adjoint recording (null when not recording): pre-step records, incoming histories,
    // rebuild permutation and history-source map
    PRec *tape_rec = nullptr;
    Hist *tape_hist = nullptr, *tape_whist = nullptr;
    unsigned *tape_perm = nullptr, *tape_src = nullptr;
    struct AdjData *adj = nullptr;
    double adj_roll_k = 1.0;   // rolling-moment regularisation of the reverse pass (0: exact)
    unsigned long long buf_gen = 0;     // bumped whenever a pair buffer is reallocated
    unsigned long long graph_gen = ~0ull;
};

namespace {

int alloc_list(dem_ctx *c, int side, unsigned cap)
{
    if (c->ab[side]) DEM_HIP(hipFree(c->ab[side]));
    c->ab[side] = nullptr;
    DEM_HIP(hipMalloc(&c->ab[side], (size_t)std::max(cap, 1u) * sizeof(uint2)));
    c->cap[side] = cap;
    c->buf_gen++;
    return DEM_OK;
}

int alloc_hist(dem_ctx *c, int side, unsigned cap)
{
    if (c->hist[side]) DEM_HIP(hipFree(c->hist[side]));
    c->hist[side] = nullptr;
    DEM_HIP(hipMalloc(&c->hist[side], (size_t)std::max(cap, 1u) * sizeof(Hist)));
    c->hcap[side] = cap;
    c->buf_gen++;
    return DEM_OK;
}

// zs is read by k_step in 4-byte words: whole words cover every pair.
size_t zs_bytes(unsigned cap) { return ((size_t)cap + 3) & ~(size_t)3; }

int alloc_out(dem_ctx *c, unsigned cap)
{
    if (c->out) DEM_HIP(hipFree(c->out));
    if (c->jt) DEM_HIP(hipFree(c->jt));
    if (c->jf) DEM_HIP(hipFree(c->jf));
    if (c->bo) DEM_HIP(hipFree(c->bo));
    if (c->zs) DEM_HIP(hipFree(c->zs));
    c->out = nullptr;
    c->jt = nullptr;
    c->jf = nullptr;
    c->bo = nullptr;
    c->zs = nullptr;
    c->zs_ok = false;
    DEM_HIP(hipMalloc(&c->out, (size_t)std::max(cap, 1u) * sizeof(POut)));
    DEM_HIP(hipMalloc(&c->jt, (size_t)std::max(cap, 1u) * sizeof(unsigned)));
    DEM_HIP(hipMalloc(&c->jf, (size_t)std::max(cap, 1u) * sizeof(unsigned)));
    DEM_HIP(hipMalloc(&c->bo, (size_t)std::max(cap, 1u) * sizeof(unsigned)));
    DEM_HIP(hipMalloc(&c->zs, zs_bytes(std::max(cap, 1u))));
    c->out_cap = cap;
    c->buf_gen++;
    return DEM_OK;
}

unsigned wall_grid(const dem_ctx *c)
{
    // grid-stride over the device-resident candidate count; at most one lane per particle
    unsigned g = cdiv(std::max(c->n, 1u), kBlock);
    return std::max(1u, std::min(g, (unsigned)c->cus * 4u));
}

unsigned pair_grid(const dem_ctx *c, unsigned np)
{
    unsigned g = cdiv(std::max(np, 1u), c->pair_block);
    unsigned gmax = (unsigned)c->cus * c->pair_grid_per_cu;
    return std::max(1u, std::min(g, gmax));
}

// Stage events of a timed rebuild (dem_time_rebuild): recorded on the context stream at the
// stage boundaries when ev is non-null.
enum { kEvStart, kEvHash, kEvSort, kEvCells, kEvReorder, kEvCount, kEvScan, kEvFillStart,
       kEvFill, kEvRemap, kEvTranspose, kEvCountAll };
inline void stage_event(hipEvent_t *ev, int k, hipStream_t st)
{
    if (ev && ev[k]) (void)hipEventRecord(ev[k], st);
}

// Phase A+B of a rebuild: hash, sort, cell ranges, reorder, neighbour counts and scans.
int rebuild_count(dem_ctx *c, const unsigned *active, hipEvent_t *ev = nullptr)
{
    hipStream_t st = c->stream;
    unsigned n = c->n;
    unsigned gb = cdiv(n, kBlock);
    InstDev one = c->hinst[0];
    // counting sort by cell (kernels.hpp): counts, their scan (the cells' first slots), the
    // scatter by arrival rank; k_reorder then restores the index order inside each cell.
    // The ranks live in ng and the scattered pairs' old slots in nl until the neighbour count
    // writes them.
    hipLaunchKernelGGL(k_fill, dim3(std::min(cdiv(c->ncells, kBlock), 4096u)), dim3(kBlock), 0, st,
                       c->ccount, c->ncells, 0u, active);
    if (c->batched)
        hipLaunchKernelGGL(k_hash<true>, dim3(gb), dim3(kBlock), 0, st, c->rec[c->rc], n, c->dinst,
                           one, c->kv_hash, c->ccount, c->ng, c->ctl, active);
    else
        hipLaunchKernelGGL(k_hash<false>, dim3(gb), dim3(kBlock), 0, st, c->rec[c->rc], n,
                           c->dinst, one, c->kv_hash, c->ccount, c->ng, c->ctl, active);
    stage_event(ev, kEvHash, st);
    if (scan_excl(c->S, c->ccount, c->ncells, c->cstart, c->S.total + 2, st, active))
        return DEM_ERR_HIP;
    hipLaunchKernelGGL(k_sort_scatter, dim3(gb), dim3(kBlock), 0, st, (const uint2 *)c->kv_hash,
                       (const unsigned *)c->ng, (const unsigned *)c->cstart, n, c->kv_a, c->nl,
                       active);
    stage_event(ev, kEvSort, st);   // the cell ranges are the scan's output
    c->kv_sorted = c->kv_b;
    int nr = 1 - c->rc;
    const unsigned n_old = c->have_list ? c->n_list : 0u;   // particles the old list holds
    if (c->batched)
        hipLaunchKernelGGL(k_reorder<true>, dim3(gb), dim3(kBlock), 0, st, (const uint2 *)c->kv_a,
                           (const unsigned *)c->nl, (const unsigned *)c->cstart, c->kv_sorted,
                           n, c->rec[c->rc], c->rec[nr], c->whist[c->pc], c->whist[1 - c->pc],
                           c->inv[1 - c->pc], (const InstDev *)c->dinst, one, c->wflag,
                           c->wlist, c->ctl, c->tape_perm, c->cpos[nr], c->cp32[nr], c->cid,
                           c->oslot, n_old, active);
    else
        hipLaunchKernelGGL(k_reorder<false>, dim3(gb), dim3(kBlock), 0, st,
                           (const uint2 *)c->kv_a, (const unsigned *)c->nl,
                           (const unsigned *)c->cstart, c->kv_sorted, n,
                           c->rec[c->rc], c->rec[nr],
                           c->whist[c->pc], c->whist[1 - c->pc], c->inv[1 - c->pc],
                           (const InstDev *)c->dinst, one, c->wflag, c->wlist, c->ctl,
                           c->tape_perm, c->cpos[nr], c->cp32[nr], c->cid, c->oslot, n_old,
                           active);
    c->rc = nr;
    c->pc = 1 - c->pc;
    stage_event(ev, kEvReorder, st);
    if (c->batched)
        hipLaunchKernelGGL(k_nbr_count<true>, dim3(gb), dim3(kBlock), 0, st, c->rec[c->rc],
                           (const double4 *)c->cpos[c->rc], (const unsigned *)c->cid,
                           (const uint2 *)c->kv_sorted, n, c->dinst, one, c->cstart,
                           c->ng, c->nl, c->rmask, c->rs0, active);
    else
        hipLaunchKernelGGL(k_nbr_count<false>, dim3(gb), dim3(kBlock), 0, st, c->rec[c->rc],
                           (const double4 *)c->cpos[c->rc], (const unsigned *)c->cid,
                           (const uint2 *)c->kv_sorted, n, c->dinst, one, c->cstart,
                           c->ng, c->nl, c->rmask, c->rs0, active);
    stage_event(ev, kEvCount, st);
    int nl = 1 - c->lc;
    if (scan2_excl(c->S, c->ng, c->nl, n, c->off[nl], c->toff, c->S.total, st, active))
        return DEM_ERR_HIP;
    hipLaunchKernelGGL(k_check_cap, dim3(1), dim3(64), 0, st, c->ctl, (const unsigned *)c->S.total,
                       (const unsigned *)(c->S.total + 1),
                       std::min(std::min(c->cap[nl], c->out_cap), c->hcap[1 - c->hc]), active);
    stage_event(ev, kEvScan, st);
    DEM_HIP(hipGetLastError());
    return DEM_OK;
}

// Phase C: fill the pair list with the remapped history and build the transposed index.
int rebuild_fill(dem_ctx *c, const unsigned *active, hipEvent_t *ev = nullptr)
{
    hipStream_t st = c->stream;
    stage_event(ev, kEvFillStart, st);
    unsigned n = c->n;
    unsigned gb = cdiv(n, kBlock);
    InstDev one = c->hinst[0];
    int nl = 1 - c->lc;
    int ol = c->lc;
    int nh = 1 - c->hc, oh = c->hc;
    unsigned capn = std::min(std::min(c->cap[nl], c->out_cap), c->hcap[nh]);
    unsigned capo = std::min(std::min(c->cap[ol], c->out_cap), c->hcap[oh]);
    if (c->batched)
        hipLaunchKernelGGL(k_nbr_fill<true>, dim3(gb), dim3(kBlock), 0, st, c->rec[c->rc],
                           (const double4 *)c->cpos[c->rc], (const unsigned *)c->cid,
                           (const uint2 *)c->kv_sorted, n, c->dinst, one, c->cstart,
                           c->off[nl], c->toff, capn, c->ab[nl], c->bo, c->jt, c->jf, c->ctl,
                           (const unsigned long long *)c->rmask, (const unsigned *)c->rs0, active);
    else
        hipLaunchKernelGGL(k_nbr_fill<false>, dim3(gb), dim3(kBlock), 0, st, c->rec[c->rc],
                           (const double4 *)c->cpos[c->rc], (const unsigned *)c->cid,
                           (const uint2 *)c->kv_sorted, n, c->dinst, one, c->cstart,
                           c->off[nl], c->toff, capn, c->ab[nl], c->bo, c->jt, c->jf, c->ctl,
                           (const unsigned long long *)c->rmask, (const unsigned *)c->rs0, active);
    stage_event(ev, kEvFill, st);
    // histories by pair identity, one lane per pair over the device pair count
    const unsigned rg = std::max(1u, std::min(cdiv(std::max(capn, 1u), kBlock), (unsigned)c->cus * 16u));
    hipLaunchKernelGGL(k_remap, dim3(rg), dim3(kBlock), 0, st, (const uint2 *)c->ab[nl],
                       (const unsigned *)c->off[nl], (const DevCtl *)c->ctl, capn, (const unsigned *)c->oslot, c->hist[nh],
                       c->have_list, (const uint2 *)c->ab[ol], (const unsigned *)c->off[ol],
                       (const Hist *)c->hist[oh], c->n_list, capo, c->tape_src, active);
    stage_event(ev, kEvRemap, st);
    hipLaunchKernelGGL(k_transpose, dim3(rg), dim3(kBlock), 0, st, (const unsigned *)c->bo,
                       (const unsigned *)c->off[nl], (const unsigned *)c->toff, n, capn, c->jt,
                       (const unsigned *)c->jf, active);
    stage_event(ev, kEvTranspose, st);
    DEM_HIP(hipGetLastError());
    c->lc = nl;
    c->hc = nh;
    c->have_list = true;
    c->n_list = n;
    c->out_valid = false;
    c->zs_ok = false;
    return DEM_OK;
}

// Capacity hook of a recorded rebuild: called with the new pair total before the fill, it
// sizes the history-source tape and points c->tape_src at it.
typedef int (*pair_hook_fn)(dem_ctx *c, void *arg, unsigned np);

// Host-driven rebuild: count, read the pair total, grow capacities if needed, fill.
int rebuild_host(dem_ctx *c, pair_hook_fn hook = nullptr, void *hook_arg = nullptr,
                 hipEvent_t *ev = nullptr)
{
    stage_event(ev, kEvStart, c->stream);
    int rc = rebuild_count(c, nullptr, ev);
    if (rc) return rc;
    unsigned tot[2];
    DEM_HIP(hipMemcpyAsync(tot, c->S.total, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    DEM_HIP(hipStreamSynchronize(c->stream));
    unsigned np = tot[0];
    int nl = 1 - c->lc;
    unsigned grow = (unsigned)std::min<unsigned long long>(0xfffffff0ull, (unsigned long long)np * 3 / 2 + 1024);
    if (np > c->cap[nl] && (rc = alloc_list(c, nl, grow))) return rc;
    if (np > c->hcap[1 - c->hc] && (rc = alloc_hist(c, 1 - c->hc, grow))) return rc;
    if (np > c->out_cap && (rc = alloc_out(c, grow))) return rc;
    // k_check_cap clamped npairs to the capacity seen before growing; store the real total
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c->stream, &c->ctl->npairs, np);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, c->stream, &c->ctl->overflow, 0u);
    if (hook && (rc = hook(c, hook_arg, np))) return rc;
    if ((rc = rebuild_fill(c, nullptr, ev))) return rc;
    if (np > c->hcap[1 - c->hc]) {
        // the buffer the fused step writes next must hold the new list too (its contents
        // were the old list's histories, read by the fill just launched)
        DEM_HIP(hipStreamSynchronize(c->stream));
        if ((rc = alloc_hist(c, 1 - c->hc, grow))) return rc;
    }
    c->npairs = np;
    c->np_peak = std::max(c->np_peak, (unsigned)np);
    c->rebuilds++;
    c->since = 0;
    c->need_rebuild = false;
    return DEM_OK;
}

// One step.  Split path: k_pair_force (one lane per pair, history in place, 80-byte output
// per pair) -> k_walls -> k_particle (gather, integrate in place).  Fused path: k_step alone
// (one lane per particle, wall contacts included, both sides of each pair evaluated, records
// and histories from the current buffers into the other ones; the buffer indices flip).  `mid` is
// recorded before the per-particle kernel, `end` after it.
template <bool FLEN>
void launch_force_step_t(dem_ctx *c, unsigned np_grid, hipEvent_t mid, hipEvent_t end,
                         const unsigned *gate, unsigned *slot)
{
    hipStream_t st = c->stream;
    unsigned pb = c->particle_block;
    unsigned gb = cdiv(c->n, pb);
    InstDev one = c->hinst[0];
    const unsigned *np = &c->ctl->npairs;
    const PRec *rin = c->rec[c->rc];
    const InstDev *dinst = (const InstDev *)c->dinst;
    double *fdbg = c->record_forces ? c->fdbg : nullptr;
    if (!c->fused) {
        Hist *h = c->hist[c->hc];
        if (c->batched)
            hipLaunchKernelGGL((k_pair_force<true, FLEN>), dim3(np_grid), dim3(c->pair_block), 0, st,
                               (const uint2 *)c->ab[c->lc], np, rin, (const Hist *)h, h, c->out,
                               dinst, one.pp, one.dt, one.roll_k, c->tape_hist, gate, c->xcd_chunk);
        else
            hipLaunchKernelGGL((k_pair_force<false, FLEN>), dim3(np_grid), dim3(c->pair_block), 0, st,
                               (const uint2 *)c->ab[c->lc], np, rin, (const Hist *)h, h, c->out,
                               dinst, one.pp, one.dt, one.roll_k, c->tape_hist, gate, c->xcd_chunk);
    }
    if (!c->fused) {   // the fused step evaluates the wall contacts itself
        if (c->batched)
            hipLaunchKernelGGL((k_walls<true, FLEN>), dim3(wall_grid(c)), dim3(kBlock), 0, st, rin,
                               c->whist[c->pc], (const unsigned *)c->wlist, (const DevCtl *)c->ctl,
                               dinst, one, c->wforce, c->tape_whist, gate);
        else
            hipLaunchKernelGGL((k_walls<false, FLEN>), dim3(wall_grid(c)), dim3(kBlock), 0, st, rin,
                               c->whist[c->pc], (const unsigned *)c->wlist, (const DevCtl *)c->ctl,
                               dinst, one, c->wforce, c->tape_whist, gate);
    }
    if (mid) (void)hipEventRecord(mid, st);
    if (c->fused) {
        int nr = 1 - c->rc, nh = 1 - c->hc;
        // zero-history bookkeeping: bit nh of zs describes the buffer written now; with
        // valid bookkeeping (lazy) pairs already known to be zero there are not rewritten
        const unsigned zmode = (1u << nh) | (c->zs_ok ? 4u : 0u);
        // queued (slot given): gated by the previous step's trigger word, raises its own and
        // reports a trigger in the slot, with no publish kernel; otherwise ungated, raising
        // need_rebuild, which k_publish copies to the mirror
        const bool queued = slot != nullptr;
        const unsigned par = c->tpar;
        const unsigned *kgate = queued ? &c->ctl->trig[1 - par] : (gate ? gate : &c->ctl->zero);
        unsigned *ktrig = queued ? &c->ctl->trig[par] : &c->ctl->need_rebuild;
        if (queued) c->tpar = 1 - par;
        c->zs_ok = true;
        // the other record buffer takes the whole updated record; wall histories are per
        // slot and updated in place by the step's wall contact, so they move with the records
        if (c->batched)
            hipLaunchKernelGGL((k_step<true, FLEN>), dim3(gb), dim3(pb), k_step_lds(pb), st, rin, c->rec[nr],
                               (const unsigned *)c->bo, (const unsigned *)c->off[c->lc],
                               (const unsigned *)c->toff, (const unsigned *)c->jt, (const unsigned *)c->jf,
                               (const Hist *)c->hist[c->hc], c->hist[nh], (const unsigned *)c->wflag,
                               c->whist[c->pc], (const float4 *)c->cp32[c->rc], c->cp32[nr],
                               c->zs, zmode, c->n, dinst, one, c->contacti, fdbg, c->ctl,
                               c->tape_rec, c->tape_hist, c->tape_whist, kgate, c->fused_fast,
                               c->xcd_chunk, ktrig, queued ? slot : nullptr);
        else
            hipLaunchKernelGGL((k_step<false, FLEN>), dim3(gb), dim3(pb), k_step_lds(pb), st, rin, c->rec[nr],
                               (const unsigned *)c->bo, (const unsigned *)c->off[c->lc],
                               (const unsigned *)c->toff, (const unsigned *)c->jt, (const unsigned *)c->jf,
                               (const Hist *)c->hist[c->hc], c->hist[nh], (const unsigned *)c->wflag,
                               c->whist[c->pc], (const float4 *)c->cp32[c->rc], c->cp32[nr],
                               c->zs, zmode, c->n, dinst, one, c->contacti, fdbg, c->ctl,
                               c->tape_rec, c->tape_hist, c->tape_whist, kgate, c->fused_fast,
                               c->xcd_chunk, ktrig, queued ? slot : nullptr);
        c->rc = nr;
        c->hc = nh;
        c->out_valid = true;
        c->last_path = 2;
    } else if (c->batched) {
        hipLaunchKernelGGL((k_particle<true, FLEN>), dim3(gb), dim3(pb), 0, st, c->rec[c->rc],
                           (const unsigned *)c->wflag, (const double *)c->wforce,
                           (const unsigned *)c->off[c->lc], (const unsigned *)c->toff,
                           (const unsigned *)c->jt, (const POut *)c->out, c->cp32[c->rc], c->n, dinst,
                           one, c->contacti, fdbg, c->ctl, c->tape_rec, gate, c->xcd_chunk);
    } else {
        hipLaunchKernelGGL((k_particle<false, FLEN>), dim3(gb), dim3(pb), 0, st, c->rec[c->rc],
                           (const unsigned *)c->wflag, (const double *)c->wforce,
                           (const unsigned *)c->off[c->lc], (const unsigned *)c->toff,
                           (const unsigned *)c->jt, (const POut *)c->out, c->cp32[c->rc], c->n, dinst,
                           one, c->contacti, fdbg, c->ctl, c->tape_rec, gate, c->xcd_chunk);
    }
    if (!c->fused) {
        c->last_path = 1;
        c->zs_ok = false;   // the split step rewrites hist[hc] in place
    }
    if (end) (void)hipEventRecord(end, st);
    if (!(c->fused && slot))   // a queued fused step needs no publish kernel
        hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, c->ctl, c->mirror_d, gate, slot);
}

// The float-length option selects the kernels' FLEN instantiation (the default path carries
// no test of it).
void launch_force_step(dem_ctx *c, unsigned np_grid, hipEvent_t mid = nullptr,
                       hipEvent_t end = nullptr, const unsigned *gate = nullptr,
                       unsigned *slot = nullptr)
{
    if (c->hinst[0].flen)
        launch_force_step_t<true>(c, np_grid, mid, end, gate, slot);
    else
        launch_force_step_t<false>(c, np_grid, mid, end, gate, slot);
}

InstDev make_inst(const dem_instance &d, unsigned cell_off, unsigned p_off)
{
    InstDev I;
    std::memset(&I, 0, sizeof(I));
    I.pp = law_from(d.mat.p);
    I.pw = law_from(d.mat.w);
    I.dt = d.dt;
    I.skin = d.skin;
    double ann = 0.5 * d.skin;
    I.ann2 = ann * ann;
    for (int k = 0; k < 3; k++) {
        I.origin[k] = d.origin[k];
        I.cell[k] = d.cell[k];
        I.gsize[k] = d.gsize[k];
    }
    I.walls.kind = d.walls.kind;
    I.walls.hx = 0.5 * d.walls.size[0];
    I.walls.hy = 0.5 * d.walls.size[1];
    I.walls.radius = d.walls.radius;
    I.cell_off = cell_off;
    I.p_off = p_off;
    I.n = d.n;
    return I;
}

// Change the list skin of every instance (the cell edge follows: cell = (cell - skin) + new
// skin, i.e. the 2 r_max part is kept) and rebuild the grids over the same extents.  The
// physics does not depend on the skin (pairs outside contact contribute nothing); the
// rebuild trigger |q| > skin/2 does.
int install_grid(dem_ctx *c, unsigned cell_off);
int regrid(dem_ctx *c, double skin)
{
    if (!(skin > 0.0)) return DEM_ERR_INVALID;
    unsigned cell_off = 0;
    for (int i = 0; i < c->ninst; i++) {
        InstDev &I = c->hinst[i];
        unsigned long long cells = 1;
        for (int k = 0; k < 3; k++) {
            double extent = (double)I.gsize[k] * I.cell[k];
            double cell = (I.cell[k] - I.skin) + skin;
            I.cell[k] = cell;
            I.gsize[k] = (unsigned)(int)(extent / cell) + 1u;
            cells *= I.gsize[k];
        }
        if (cells + cell_off >= 0xffffffffull) return DEM_ERR_INVALID;
        I.skin = skin;
        double ann = 0.5 * skin;
        I.ann2 = ann * ann;
        I.cell_off = cell_off;
        cell_off += (unsigned)cells;
    }
    return install_grid(c, cell_off);
}

// Upload the host instance table and size the cell arrays for `cell_off` cells in total.
int install_grid(dem_ctx *c, unsigned cell_off)
{
    DEM_HIP(hipStreamSynchronize(c->stream));
    DEM_HIP(hipMemcpy(c->dinst, c->hinst.data(), sizeof(InstDev) * c->ninst, hipMemcpyHostToDevice));
    if (cell_off != c->ncells) {
        DEM_HIP(hipFree(c->cstart));
        DEM_HIP(hipFree(c->ccount));
        c->cstart = c->ccount = nullptr;
        DEM_HIP(hipMalloc(&c->cstart, sizeof(unsigned) * ((size_t)cell_off + 1)));
        DEM_HIP(hipMalloc(&c->ccount, sizeof(unsigned) * (size_t)cell_off));
        c->ncells = cell_off;
    }
    c->need_rebuild = true;
    c->buf_gen++;
    return DEM_OK;
}

void adj_free(dem_ctx *c);

void free_ctx(dem_ctx *c)
{
    if (!c) return;
    adj_free(c);
    void *bufs[] = {c->dinst, c->rec[0], c->rec[1], c->whist[0], c->whist[1], c->inv[0], c->inv[1],
                    c->contacti, c->wflag, c->wlist, c->wforce, c->fdbg, c->cpos[0], c->cpos[1], c->cp32[0], c->cp32[1], c->cid, c->kv_hash, c->kv_a, c->kv_b, c->cstart, c->ccount, c->ng,
                    c->nl, c->rmask, c->rs0, c->oslot, c->toff, c->off[0], c->off[1], c->ab[0], c->ab[1], c->hist[0],
                    c->hist[1], c->out, c->jt, c->jf, c->bo, c->zs, c->ctl};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    c->S.release();
    if (c->mirror_h) (void)hipHostFree(c->mirror_h);
    for (int k = 0; k < 2; k++) {
        if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
        if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
    }
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evs) (void)hipEventDestroy(c->evs);
    for (auto &T : c->ring) {
        for (auto &e : T.e) if (e) (void)hipEventDestroy(e);
        for (auto &e : T.r) if (e) (void)hipEventDestroy(e);
    }
    for (auto &e : c->evq) if (e) (void)hipEventDestroy(e);
    if (c->qflag_h) (void)hipHostFree(c->qflag_h);
    if (c->evp0) (void)hipEventDestroy(c->evp0);
    if (c->evp1) (void)hipEventDestroy(c->evp1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// A path probe times one step between two events.  Launched first, a short busy period
// keeps the GPU occupied while the host queues the step's kernels, so the probe measures the
// kernels back to back (as queued steps run them), not the host's launch gaps.
void launch_probe_delay(dem_ctx *c)
{
    if (c->probe_delay)
        hipLaunchKernelGGL(k_delay, dim3(1), dim3(64), 0, c->stream, c->probe_delay);
}

int sync_mirror(dem_ctx *c)
{
    DEM_HIP(hipEventRecord(c->evs, c->stream));
    DEM_HIP(hipEventSynchronize(c->evs));
    const volatile DevCtl *m = reinterpret_cast<const volatile DevCtl *>(c->mirror_h);
    c->last_flag = m->need_rebuild;
    return DEM_OK;
}

// Read back one timed step of the ring (waits for its last event).
int ring_flush_slot(dem_ctx *c, dem_ctx::TimedStep &T)
{
    if (!T.used) return DEM_OK;
    DEM_HIP(hipEventSynchronize(T.e[2]));
    float a = 0.f, b = 0.f;
    DEM_HIP(hipEventElapsedTime(&a, T.e[0], T.e[1]));
    DEM_HIP(hipEventElapsedTime(&b, T.e[1], T.e[2]));
    int p = T.fused ? 1 : 0;
    c->t_acc[p][0] += a;
    c->t_acc[p][1] += b;
    c->np_acc[p] += T.np;
    c->n_acc[p] += 1;
    if (T.reb) {
        float r = 0.f;
        DEM_HIP(hipEventElapsedTime(&r, T.r[0], T.r[1]));
        c->t_reb += r;
        c->n_reb += 1;
    }
    T.used = false;
    return DEM_OK;
}

int ring_flush_all(dem_ctx *c)
{
    for (int k = 0; k < dem_ctx::kRing; k++) {
        int rc = ring_flush_slot(c, c->ring[(c->ring_pos + k) % dem_ctx::kRing]);
        if (rc) return rc;
    }
    return DEM_OK;
}

void drop_graph(dem_ctx *c)
{
    for (int k = 0; k < 2; k++) {
        if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
        if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
        c->gexec[k] = nullptr;
        c->graph[k] = nullptr;
    }
    c->graph_k = 0;
}

// Grow a list / history buffer keeping its contents (the live list and history).
int grow_list_keep(dem_ctx *c, int side, unsigned cap)
{
    if (c->cap[side] >= cap) return DEM_OK;
    uint2 *nab = nullptr;
    DEM_HIP(hipMalloc(&nab, (size_t)cap * sizeof(uint2)));
    DEM_HIP(hipMemcpyAsync(nab, c->ab[side], (size_t)c->cap[side] * sizeof(uint2), hipMemcpyDeviceToDevice, c->stream));
    DEM_HIP(hipStreamSynchronize(c->stream));
    DEM_HIP(hipFree(c->ab[side]));
    c->ab[side] = nab;
    c->cap[side] = cap;
    c->buf_gen++;
    return DEM_OK;
}

int grow_hist_keep(dem_ctx *c, int side, unsigned cap)
{
    if (c->hcap[side] >= cap) return DEM_OK;
    Hist *nh = nullptr;
    DEM_HIP(hipMalloc(&nh, (size_t)cap * sizeof(Hist)));
    DEM_HIP(hipMemcpyAsync(nh, c->hist[side], (size_t)c->hcap[side] * sizeof(Hist), hipMemcpyDeviceToDevice, c->stream));
    DEM_HIP(hipStreamSynchronize(c->stream));
    DEM_HIP(hipFree(c->hist[side]));
    c->hist[side] = nh;
    c->hcap[side] = cap;
    c->buf_gen++;
    return DEM_OK;
}

}  // namespace

extern "C" {

int dem_last_error(void)
{
    int e = g_last_error;
    g_last_error = 0;
    return e;
}

const char *dem_status_string(int s)
{
    switch (s) {
    case DEM_OK: return "ok";
    case DEM_ERR_INVALID: return "invalid argument";
    case DEM_ERR_HIP: return "HIP runtime error";
    case DEM_ERR_CAPACITY: return "capacity exceeded";
    case DEM_ERR_NO_DEVICE: return "no usable device";
    case DEM_ERR_IO: return "file error";
    default: return "unknown status";
    }
}

void *dem_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

int dem_dev_free(void *p)
{
    if (p) DEM_HIP(hipFree(p));
    return DEM_OK;
}

int dem_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
    static const hipMemcpyKind kinds[4] = {hipMemcpyHostToHost, hipMemcpyHostToDevice,
                                           hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if (kind < 0 || kind > 3) return DEM_ERR_INVALID;
    DEM_HIP(hipMemcpy(dst, src, bytes, kinds[kind]));
    return DEM_OK;
}

int dem_memzero(void *dst, size_t bytes)
{
    DEM_HIP(hipMemset(dst, 0, bytes));
    DEM_HIP(hipDeviceSynchronize());
    return DEM_OK;
}

int dem_device_sync(void)
{
    DEM_HIP(hipDeviceSynchronize());
    return DEM_OK;
}

int dem_device_info(int *cu_count, char *name, int name_len)
{
    int dev = 0;
    DEM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    DEM_HIP(hipGetDeviceProperties(&p, dev));
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", p.gcnArchName);
    }
    return DEM_OK;
}

int dem_create(dem_ctx **out, int ninst, const dem_instance *inst, uint pair_capacity)
{
    if (!out || ninst < 1 || !inst) return DEM_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return DEM_ERR_NO_DEVICE;
    dem_ctx *c = new dem_ctx();
    DEM_HIP(hipGetDevice(&c->device));
    c->cus = current_device_cus();
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess &&
            khz > 0 && khz <= 1000000)
            c->probe_delay = (unsigned long long)khz * 60ull / 1000ull;   // 60 us
    }
    c->ninst = ninst;
    c->batched = ninst > 1;
    unsigned cell_off = 0, p_off = 0;
    for (int i = 0; i < ninst; i++) {
        const dem_instance &d = inst[i];
        if (d.n == 0 || d.gsize[0] == 0 || d.gsize[1] == 0 || d.gsize[2] == 0 || !(d.dt > 0.0) ||
            !(d.cell[0] > 0.0) || !(d.cell[1] > 0.0) || !(d.cell[2] > 0.0) || d.skin < 0.0) {
            free_ctx(c);
            return DEM_ERR_INVALID;
        }
        unsigned long long cells = (unsigned long long)d.gsize[0] * d.gsize[1] * d.gsize[2];
        if (cells + cell_off >= 0xffffffffull || (unsigned long long)p_off + d.n >= 0xffffffffull) {
            free_ctx(c);
            return DEM_ERR_INVALID;
        }
        c->hinst.push_back(make_inst(d, cell_off, p_off));
        cell_off += (unsigned)cells;
        p_off += d.n;
    }
    c->n = p_off;
    c->cap_n = p_off;
    c->ncells = cell_off;
    unsigned n = c->cap_n;
    unsigned pc = pair_capacity ? pair_capacity : (unsigned)std::min<unsigned long long>(0xfffffff0ull, 12ull * n + 1024);

    DEM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    DEM_HIP(hipEventCreate(&c->ev0));
    DEM_HIP(hipEventCreate(&c->ev1));
    DEM_HIP(hipEventCreateWithFlags(&c->evs, hipEventDisableTiming));
    DEM_HIP(hipEventCreate(&c->evp0));
    DEM_HIP(hipEventCreate(&c->evp1));
    for (auto &e : c->evq) DEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    DEM_HIP(hipHostMalloc((void **)&c->qflag_h, sizeof(unsigned) * dem_ctx::kQueue,
                          hipHostMallocMapped | hipHostMallocCoherent));
    DEM_HIP(hipHostGetDevicePointer((void **)&c->qflag_d, c->qflag_h, 0));
    DEM_HIP(hipMalloc(&c->dinst, sizeof(InstDev) * ninst));
    DEM_HIP(hipMemcpy(c->dinst, c->hinst.data(), sizeof(InstDev) * ninst, hipMemcpyHostToDevice));
    for (int k = 0; k < 2; k++) {
        DEM_HIP(hipMalloc(&c->rec[k], sizeof(PRec) * (size_t)n));
        DEM_HIP(hipMalloc(&c->whist[k], sizeof(Hist) * (size_t)n));
        DEM_HIP(hipMalloc(&c->inv[k], sizeof(unsigned) * (size_t)n));
        DEM_HIP(hipMalloc(&c->off[k], sizeof(unsigned) * ((size_t)n + 1)));
        if (alloc_list(c, k, pc) || alloc_hist(c, k, pc)) { free_ctx(c); return DEM_ERR_HIP; }
    }
    if (alloc_out(c, pc)) { free_ctx(c); return DEM_ERR_HIP; }
    DEM_HIP(hipMalloc(&c->contacti, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->wflag, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->wlist, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->wforce, sizeof(double) * 8 * (size_t)n));
    DEM_HIP(hipMalloc(&c->cpos[0], sizeof(double4) * (size_t)n));
    DEM_HIP(hipMalloc(&c->cpos[1], sizeof(double4) * (size_t)n));
    DEM_HIP(hipMalloc(&c->cp32[0], sizeof(float4) * (size_t)n));
    DEM_HIP(hipMalloc(&c->cp32[1], sizeof(float4) * (size_t)n));
    DEM_HIP(hipMalloc(&c->cid, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->kv_hash, sizeof(uint2) * (size_t)n));
    DEM_HIP(hipMalloc(&c->kv_a, sizeof(uint2) * (size_t)n));
    DEM_HIP(hipMalloc(&c->kv_b, sizeof(uint2) * (size_t)n));
    DEM_HIP(hipMalloc(&c->cstart, sizeof(unsigned) * ((size_t)c->ncells + 1)));
    DEM_HIP(hipMalloc(&c->ccount, sizeof(unsigned) * (size_t)c->ncells));
    DEM_HIP(hipMalloc(&c->ng, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->nl, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->rmask, sizeof(unsigned long long) * 2 * kRows * (size_t)n));
    DEM_HIP(hipMalloc(&c->rs0, sizeof(unsigned) * kRows * (size_t)n));
    DEM_HIP(hipMalloc(&c->oslot, sizeof(unsigned) * (size_t)n));
    DEM_HIP(hipMalloc(&c->toff, sizeof(unsigned) * ((size_t)n + 1)));
    c->ctl_bytes = sizeof(DevCtl) + (size_t)ninst * (sizeof(unsigned long long) + sizeof(unsigned));
    DEM_HIP(hipMalloc(&c->ctl, c->ctl_bytes));
    DEM_HIP(hipMemset(c->ctl, 0, c->ctl_bytes));
    DEM_HIP(hipHostMalloc((void **)&c->mirror_h, c->ctl_bytes, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(c->mirror_h, 0, c->ctl_bytes);
    DEM_HIP(hipHostGetDevicePointer((void **)&c->mirror_d, c->mirror_h, 0));
    if (c->S.ensure(64, 256 * 64)) { free_ctx(c); return DEM_ERR_HIP; }
    *out = c;
    return DEM_OK;
}

int dem_destroy(dem_ctx *c)
{
    if (!c) return DEM_ERR_INVALID;
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return DEM_OK;
}

int dem_set_option(dem_ctx *c, int option, double value)
{
    if (!c) return DEM_ERR_INVALID;
    switch (option) {
    case DEM_OPT_RECORD_FORCES:
        c->record_forces = value != 0.0;
        if (c->record_forces && !c->fdbg) DEM_HIP(hipMalloc(&c->fdbg, sizeof(double) * 6 * (size_t)c->cap_n));
        drop_graph(c);
        return DEM_OK;
    case DEM_OPT_REBUILD_INTERVAL:
        // a finite whole number in [0, INT_MAX] (NaN fails the first test)
        if (!(value >= 0.0 && value <= 2147483647.0) || (double)(long long)value != value)
            return DEM_ERR_INVALID;
        c->interval = (int)value;
        drop_graph(c);
        return DEM_OK;
    case DEM_OPT_SKIN: {
        drop_graph(c);
        return regrid(c, value);
    }
    case DEM_OPT_PAIR_BLOCK:
    case DEM_OPT_PARTICLE_BLOCK: {
        unsigned b = (unsigned)value;
        if (b != 64 && b != 128 && b != 256) return DEM_ERR_INVALID;
        if (option == DEM_OPT_PAIR_BLOCK) c->pair_block = b;
        else c->particle_block = b;
        drop_graph(c);
        return DEM_OK;
    }
    case DEM_OPT_FUSED_STEP: {
        int m = (int)value;
        if (m < 0 || m > 2 || (double)m != value) return DEM_ERR_INVALID;
        c->fused_mode = m;
        if (m < 2) c->fused = m == 1;
        c->probed = false;
        c->probe_left = 0;
        drop_graph(c);
        return DEM_OK;
    }
    case DEM_OPT_FUSED_FAST: {
        if (!(value >= 0.0 && value <= (double)kFast) || (double)(unsigned)value != value)
            return DEM_ERR_INVALID;
        c->fused_fast = (unsigned)value;
        drop_graph(c);
        return DEM_OK;
    }
    case DEM_OPT_XCD_CHUNK: {
        if (!(value >= 0.0 && value <= 65536.0) || (double)(unsigned)value != value)
            return DEM_ERR_INVALID;
        c->xcd_chunk = (unsigned)value;
        drop_graph(c);
        return DEM_OK;
    }
    case DEM_OPT_PIPELINE: {
        int d = (int)value;
        if (d < 1 || d > dem_ctx::kQueue || (double)d != value) return DEM_ERR_INVALID;
        c->pipeline = d;
        return DEM_OK;
    }
    case DEM_OPT_PIPELINE_BATCH: {
        int b = (int)value;
        if (b < 1 || b > dem_ctx::kQueue || (double)b != value) return DEM_ERR_INVALID;
        c->pipeline_batch = b;
        return DEM_OK;
    }
    case DEM_OPT_TIME_STEPS:
        if (value != 0.0 && !c->ring[0].e[0]) {
            for (auto &T : c->ring) {
                for (auto &e : T.e) DEM_HIP(hipEventCreate(&e));
                for (auto &e : T.r) DEM_HIP(hipEventCreate(&e));
            }
        }
        if (c->timing && value == 0.0 && ring_flush_all(c)) return DEM_ERR_HIP;
        c->timing = value != 0.0;
        return DEM_OK;
    case DEM_OPT_ADJ_ROLL_REG:
        if (!(value >= 0.0)) return DEM_ERR_INVALID;
        c->adj_roll_k = value;
        return DEM_OK;
    case DEM_OPT_ROLL_FLOOR: {
        if (!(value >= 0.0 && value <= 1e6)) return DEM_ERR_INVALID;
        DEM_HIP(hipStreamSynchronize(c->stream));
        for (InstDev &I : c->hinst) I.roll_k = value;
        DEM_HIP(hipMemcpy(c->dinst, c->hinst.data(), sizeof(InstDev) * c->ninst, hipMemcpyHostToDevice));
        drop_graph(c);
        return DEM_OK;
    }
    case DEM_OPT_FLOAT_LENGTH: {
        if (!(value == 0.0 || value == 1.0)) return DEM_ERR_INVALID;
        DEM_HIP(hipStreamSynchronize(c->stream));
        for (InstDev &I : c->hinst) I.flen = (unsigned)value;
        DEM_HIP(hipMemcpy(c->dinst, c->hinst.data(), sizeof(InstDev) * c->ninst, hipMemcpyHostToDevice));
        drop_graph(c);
        return DEM_OK;
    }
    case DEM_OPT_PAIR_GRID: {
        if (!(value >= 1.0 && value <= 256.0)) return DEM_ERR_INVALID;
        c->pair_grid_per_cu = (unsigned)value;
        drop_graph(c);
        return DEM_OK;
    }
    default: return DEM_ERR_INVALID;
    }
}

int dem_set_state(dem_ctx *c, const double *pos, const double *pdis, const double *angv,
                  const double *rad, const double *rmass, const double *inert, int on_device)
{
    if (!c || !pos || !pdis || !angv || !rad || !rmass || !inert) return DEM_ERR_INVALID;
    size_t n = c->cap_n;
    std::vector<double> hp, hv, hw, hr, hm, hi;
    if (on_device) {
        hp.resize(3 * n); hv.resize(3 * n); hw.resize(3 * n); hr.resize(n); hm.resize(n); hi.resize(n);
        DEM_HIP(hipMemcpy(hp.data(), pos, 24 * n, hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(hv.data(), pdis, 24 * n, hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(hw.data(), angv, 24 * n, hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(hr.data(), rad, 8 * n, hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(hm.data(), rmass, 8 * n, hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(hi.data(), inert, 8 * n, hipMemcpyDeviceToHost));
        pos = hp.data(); pdis = hv.data(); angv = hw.data(); rad = hr.data(); rmass = hm.data(); inert = hi.data();
    }
    std::vector<PRec> recs(n);
    std::vector<unsigned> inv(n);
    for (int k = 0; k < c->ninst; k++) {
        const InstDev &I = c->hinst[k];
        for (unsigned j = 0; j < I.n; j++) {
            size_t i = (size_t)I.p_off + j;
            PRec &r = recs[i];
            r.px = pos[3 * i]; r.py = pos[3 * i + 1]; r.pz = pos[3 * i + 2]; r.rad = rad[i];
            r.vx = pdis[3 * i]; r.vy = pdis[3 * i + 1]; r.vz = pdis[3 * i + 2]; r.mass = rmass[i];
            r.wx = angv[3 * i]; r.wy = angv[3 * i + 1]; r.wz = angv[3 * i + 2]; r.inert = inert[i];
            r.qx = r.qy = r.qz = 0.0;
            r.inst = (unsigned)k;
            r.id = (unsigned)i;
            inv[i] = (unsigned)i;
        }
    }
    DEM_HIP(hipStreamSynchronize(c->stream));
    // both buffers: particles not yet active (dem_set_active) keep slot = index in either
    for (int k = 0; k < 2; k++) {
        DEM_HIP(hipMemcpy(c->rec[k], recs.data(), sizeof(PRec) * n, hipMemcpyHostToDevice));
        DEM_HIP(hipMemset(c->whist[k], 0, sizeof(Hist) * n));
        DEM_HIP(hipMemcpy(c->inv[k], inv.data(), sizeof(unsigned) * n, hipMemcpyHostToDevice));
    }
    DEM_HIP(hipMemset(c->contacti, 0, sizeof(unsigned) * n));
    c->have_list = false;
    c->out_valid = false;
    c->zs_ok = false;
    c->last_path = 0;
    c->need_rebuild = true;
    c->npairs = 0;
    DEM_HIP(hipMemset(c->ctl, 0, c->ctl_bytes));
    return DEM_OK;
}

static int fetch_records(dem_ctx *c, std::vector<PRec> &recs)
{
    recs.resize(c->cap_n);
    DEM_HIP(hipStreamSynchronize(c->stream));
    DEM_HIP(hipMemcpy(recs.data(), c->rec[c->rc], sizeof(PRec) * (size_t)c->cap_n, hipMemcpyDeviceToHost));
    return DEM_OK;
}

// Activate particles [n, n_new) of a single-instance context (layer insertion of the pour,
// dempacking.cu:378-462): they join at the next step, which rebuilds the lists.
int dem_set_active(dem_ctx *c, uint n_new)
{
    // before the first list build any prefix may be chosen; afterwards the count only grows
    // (active particles have been moved to sorted slots)
    if (!c || c->ninst != 1 || n_new > c->cap_n || n_new == 0) return DEM_ERR_INVALID;
    if (c->have_list && n_new < c->n) return DEM_ERR_INVALID;
    if (n_new == c->n) return DEM_OK;
    DEM_HIP(hipStreamSynchronize(c->stream));
    c->n = n_new;
    c->need_rebuild = true;
    drop_graph(c);
    return DEM_OK;
}

uint dem_active_count(dem_ctx *c) { return c ? c->n : 0u; }

int dem_get_state(dem_ctx *c, double *pos, double *pdis, double *angv, double *qi, int on_device)
{
    if (!c) return DEM_ERR_INVALID;
    std::vector<PRec> recs;
    int rc = fetch_records(c, recs);
    if (rc) return rc;
    size_t n = c->cap_n;
    std::vector<double> p(3 * n), v(3 * n), w(3 * n), q(3 * n);
    for (size_t s = 0; s < n; s++) {
        const PRec &r = recs[s];
        size_t i = r.id;
        p[3 * i] = r.px; p[3 * i + 1] = r.py; p[3 * i + 2] = r.pz;
        v[3 * i] = r.vx; v[3 * i + 1] = r.vy; v[3 * i + 2] = r.vz;
        w[3 * i] = r.wx; w[3 * i + 1] = r.wy; w[3 * i + 2] = r.wz;
        q[3 * i] = r.qx; q[3 * i + 1] = r.qy; q[3 * i + 2] = r.qz;
    }
    hipMemcpyKind k = on_device ? hipMemcpyHostToDevice : hipMemcpyHostToHost;
    if (pos) DEM_HIP(hipMemcpy(pos, p.data(), 24 * n, k));
    if (pdis) DEM_HIP(hipMemcpy(pdis, v.data(), 24 * n, k));
    if (angv) DEM_HIP(hipMemcpy(angv, w.data(), 24 * n, k));
    if (qi) DEM_HIP(hipMemcpy(qi, q.data(), 24 * n, k));
    return DEM_OK;
}

int dem_get_wall_history(dem_ctx *c, double *disptw, double *fricpw)
{
    if (!c) return DEM_ERR_INVALID;
    std::vector<PRec> recs;
    int rc = fetch_records(c, recs);
    if (rc) return rc;
    size_t n = c->cap_n;
    std::vector<Hist> wh(n);
    DEM_HIP(hipMemcpy(wh.data(), c->whist[c->pc], sizeof(Hist) * n, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < n; s++) {
        size_t i = recs[s].id;
        if (disptw) { disptw[3 * i] = wh[s].hx; disptw[3 * i + 1] = wh[s].hy; disptw[3 * i + 2] = wh[s].hz; }
        if (fricpw) fricpw[i] = wh[s].fr;
    }
    return DEM_OK;
}

int dem_request_rebuild(dem_ctx *c)
{
    if (!c) return DEM_ERR_INVALID;
    c->need_rebuild = true;
    return DEM_OK;
}

int dem_rebuild(dem_ctx *c)
{
    if (!c) return DEM_ERR_INVALID;
    int rc = rebuild_host(c);
    if (rc) return rc;
    return DEM_OK;
}

// =========================================================================================
// Full dynamic state: restore from a snapshot (device) or a checkpoint file (host)
// =========================================================================================

namespace {

// The dynamic state of a context between steps: records, wall histories, inverse slot map
// and control block, and the live pair list with its transposed index, wall candidates and
// tangential histories.  Restoring it (and nothing else) makes the following steps repeat
// the original run bit for bit: the slot order, the list and the summation order are the
// original's.
struct StateSrc {
    const PRec *rec = nullptr;
    const Hist *whist = nullptr, *hist = nullptr;
    const unsigned *inv = nullptr, *off = nullptr, *toff = nullptr, *jt = nullptr,
                   *wflag = nullptr, *wlist = nullptr;
    const uint2 *ab = nullptr;
    const DevCtl *ctl = nullptr;
    size_t nrec = 0;          // records, wall histories and inverse-map entries to copy
    bool both = false;        // copy them into both buffers (slots of inactive particles)
    unsigned npairs = 0, n_list = 0, since = 0, last_flag = 0;
    bool have_list = false, need_rebuild = false;
    unsigned long long steps = 0, rebuilds = 0;
};

int restore_state(dem_ctx *c, const StateSrc &S, hipMemcpyKind kind)
{
    hipStream_t st = c->stream;
    size_t n = c->n, np = S.have_list ? S.npairs : 0;
    int rc;
    DEM_HIP(hipStreamSynchronize(st));
    unsigned grow = (unsigned)(np * 3 / 2 + 1024);
    if (np > c->cap[c->lc] && (rc = alloc_list(c, c->lc, grow))) return rc;
    for (int k = 0; k < 2; k++)
        if (np > c->hcap[k] && (rc = alloc_hist(c, k, grow))) return rc;
    if (np > c->out_cap && (rc = alloc_out(c, (unsigned)(np * 3 / 2 + 1024)))) return rc;
    for (int k = 0; k < (S.both ? 2 : 1); k++) {
        int r = S.both ? k : c->rc, p = S.both ? k : c->pc;
        DEM_HIP(hipMemcpyAsync(c->rec[r], S.rec, sizeof(PRec) * S.nrec, kind, st));
        DEM_HIP(hipMemcpyAsync(c->whist[p], S.whist, sizeof(Hist) * S.nrec, kind, st));
        DEM_HIP(hipMemcpyAsync(c->inv[p], S.inv, sizeof(unsigned) * S.nrec, kind, st));
    }
    c->zs_ok = false;
    c->out_valid = false;
    hipLaunchKernelGGL(k_cpos_from_rec, dim3(cdiv(S.nrec, kBlock)), dim3(kBlock), 0, st,
                       (const PRec *)c->rec[c->rc], (unsigned)S.nrec, c->cpos[c->rc], c->cp32[c->rc]);
    DEM_HIP(hipMemcpyAsync(c->ctl, S.ctl, sizeof(DevCtl), kind, st));
    if (S.have_list) {
        DEM_HIP(hipMemcpyAsync(c->off[c->lc], S.off, sizeof(unsigned) * (n + 1), kind, st));
        DEM_HIP(hipMemcpyAsync(c->toff, S.toff, sizeof(unsigned) * (n + 1), kind, st));
        DEM_HIP(hipMemcpyAsync(c->wflag, S.wflag, sizeof(unsigned) * n, kind, st));
        DEM_HIP(hipMemcpyAsync(c->wlist, S.wlist, sizeof(unsigned) * n, kind, st));
        if (np) {
            DEM_HIP(hipMemcpyAsync(c->ab[c->lc], S.ab, sizeof(uint2) * np, kind, st));
            DEM_HIP(hipMemcpyAsync(c->hist[c->hc], S.hist, sizeof(Hist) * np, kind, st));
            DEM_HIP(hipMemcpyAsync(c->jt, S.jt, sizeof(unsigned) * np, kind, st));
            hipLaunchKernelGGL(k_owner, dim3(cdiv(np, kBlock)), dim3(kBlock), 0, st,
                               (const uint2 *)c->ab[c->lc], (const unsigned *)c->jt, (unsigned)np,
                               c->jf, c->bo);
        }
    }
    DEM_HIP(hipGetLastError());
    // host sources must outlive the copies
    if (kind == hipMemcpyHostToDevice) DEM_HIP(hipStreamSynchronize(st));
    c->npairs = S.npairs;
    c->n_list = S.n_list;
    c->since = S.since;
    c->last_flag = S.last_flag;
    c->have_list = S.have_list;
    c->need_rebuild = S.need_rebuild;
    c->steps = S.steps;
    c->rebuilds = S.rebuilds;
    drop_graph(c);
    return DEM_OK;
}

// Checkpoint file: a fixed header, the instance table, then the state arrays in the
// library's slot layout, then an FNV-1a 64-bit checksum of everything before it.
using dem_ckpt::Header;

// Layout of the structures the checkpoint reader checks (ckpt.hpp works on raw bytes).
dem_ckpt::Layout ckpt_layout()
{
    dem_ckpt::Layout L;
    L.sz_rec = sizeof(PRec);
    L.sz_hist = sizeof(Hist);
    L.sz_inst = sizeof(InstDev);
    L.sz_ctl = sizeof(DevCtl);
    L.rec_inst = offsetof(PRec, inst);
    L.rec_id = offsetof(PRec, id);
    L.inst_p_off = offsetof(InstDev, p_off);
    L.inst_n = offsetof(InstDev, n);
    L.inst_cell_off = offsetof(InstDev, cell_off);
    L.inst_gsize = offsetof(InstDev, gsize);
    L.ctl_npairs = offsetof(DevCtl, npairs);
    L.ctl_overflow = offsetof(DevCtl, overflow);
    L.ctl_nwall = offsetof(DevCtl, nwall);
    L.ctl_zero = offsetof(DevCtl, zero);
    return L;
}

}  // namespace

int dem_save_checkpoint(dem_ctx *c, const char *path)
{
    if (!c || !path || !*path) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    const size_t N = c->cap_n, n = c->n;
    const size_t np = c->have_list ? c->npairs : 0;
    std::vector<PRec> rec(N);
    std::vector<Hist> wh(N), hist(np);
    std::vector<unsigned> inv(N), off, toff, wflag, wlist, jt(np);
    std::vector<uint2> ab(np);
    DevCtl ctl;
    DEM_HIP(hipMemcpy(rec.data(), c->rec[c->rc], sizeof(PRec) * N, hipMemcpyDeviceToHost));
    DEM_HIP(hipMemcpy(wh.data(), c->whist[c->pc], sizeof(Hist) * N, hipMemcpyDeviceToHost));
    DEM_HIP(hipMemcpy(inv.data(), c->inv[c->pc], sizeof(unsigned) * N, hipMemcpyDeviceToHost));
    DEM_HIP(hipMemcpy(&ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost));
    if (c->have_list) {
        off.resize(n + 1);
        toff.resize(n + 1);
        wflag.resize(n);
        wlist.resize(n);
        DEM_HIP(hipMemcpy(off.data(), c->off[c->lc], sizeof(unsigned) * (n + 1), hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(toff.data(), c->toff, sizeof(unsigned) * (n + 1), hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(wflag.data(), c->wflag, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
        DEM_HIP(hipMemcpy(wlist.data(), c->wlist, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
        if (np) {
            DEM_HIP(hipMemcpy(ab.data(), c->ab[c->lc], sizeof(uint2) * np, hipMemcpyDeviceToHost));
            DEM_HIP(hipMemcpy(hist.data(), c->hist[c->hc], sizeof(Hist) * np, hipMemcpyDeviceToHost));
            DEM_HIP(hipMemcpy(jt.data(), c->jt, sizeof(unsigned) * np, hipMemcpyDeviceToHost));
        }
    }
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, dem_ckpt::kMagic, sizeof(h.magic));
    h.version = dem_ckpt::kVersion;
    h.endian = dem_ckpt::kEndian;
    h.sz_rec = sizeof(PRec);
    h.sz_hist = sizeof(Hist);
    h.sz_inst = sizeof(InstDev);
    h.sz_ctl = sizeof(DevCtl);
    h.ninst = (unsigned)c->ninst;
    h.cap_n = (unsigned)N;
    h.n = (unsigned)n;
    h.npairs = (unsigned)np;
    h.have_list = c->have_list;
    h.need_rebuild = c->need_rebuild;
    h.since = c->since;
    h.last_flag = c->last_flag;
    h.n_list = c->n_list;
    h.interval = (unsigned)c->interval;
    h.steps = c->steps;
    h.rebuilds = c->rebuilds;
    // write to a side file and rename it over the target: an interrupted save leaves the
    // previous checkpoint intact
    std::vector<char> tmp(std::strlen(path) + 8);
    std::snprintf(tmp.data(), tmp.size(), "%s.part", path);
    std::FILE *fp = std::fopen(tmp.data(), "wb");
    if (!fp) return DEM_ERR_IO;
    dem_ckpt::Writer w{fp};
    w.put(&h, sizeof(h));
    w.put(c->hinst.data(), sizeof(InstDev) * c->hinst.size());
    w.put(rec.data(), sizeof(PRec) * N);
    w.put(wh.data(), sizeof(Hist) * N);
    w.put(inv.data(), sizeof(unsigned) * N);
    w.put(&ctl, sizeof(ctl));
    if (c->have_list) {
        w.put(off.data(), sizeof(unsigned) * (n + 1));
        w.put(toff.data(), sizeof(unsigned) * (n + 1));
        w.put(wflag.data(), sizeof(unsigned) * n);
        w.put(wlist.data(), sizeof(unsigned) * n);
        w.put(ab.data(), sizeof(uint2) * np);
        w.put(hist.data(), sizeof(Hist) * np);
        w.put(jt.data(), sizeof(unsigned) * np);
    }
    unsigned long long sum = w.sum.h;
    bool ok = w.ok && std::fwrite(&sum, sizeof(sum), 1, fp) == 1;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.data(), path) != 0) {
        std::remove(tmp.data());
        return DEM_ERR_IO;
    }
    return DEM_OK;
}

int dem_checkpoint_layout(uint *out)
{
    if (!out) return DEM_ERR_INVALID;
    const dem_ckpt::Layout L = ckpt_layout();
    const unsigned v[14] = {L.sz_rec, L.sz_hist, L.sz_inst, L.sz_ctl, L.rec_inst, L.rec_id,
                            L.inst_p_off, L.inst_n, L.inst_cell_off, L.inst_gsize,
                            L.ctl_npairs, L.ctl_overflow, L.ctl_nwall, L.ctl_zero};
    std::memcpy(out, v, sizeof(v));
    return DEM_OK;
}

int dem_load_checkpoint(dem_ctx *c, const char *path)
{
    if (!c || !path) return DEM_ERR_INVALID;
    dem_ckpt::Expect E;
    E.ninst = (unsigned)c->ninst;
    E.cap_n = c->cap_n;
    for (const InstDev &I : c->hinst) {
        E.p_off.push_back(I.p_off);
        E.n.push_back(I.n);
    }
    dem_ckpt::Data d;
    int rc0 = dem_ckpt::read(path, ckpt_layout(), E, d);
    if (rc0) return rc0;
    const Header &h = d.h;
    const unsigned long long cells = d.cells;
    const size_t N = h.cap_n, np = h.npairs;
    std::vector<InstDev> inst(h.ninst);
    std::memcpy(inst.data(), d.inst.data(), sizeof(InstDev) * h.ninst);
    std::vector<PRec> rec(N);
    std::vector<Hist> wh(N), hist(np);
    std::vector<uint2> ab(np);
    DevCtl ctl;
    std::memcpy(rec.data(), d.rec.data(), sizeof(PRec) * N);
    std::memcpy(wh.data(), d.wh.data(), sizeof(Hist) * N);
    std::memcpy(&ctl, d.ctl.data(), sizeof(DevCtl));
    if (h.have_list && np) {
        std::memcpy(ab.data(), d.ab.data(), sizeof(uint2) * np);
        std::memcpy(hist.data(), d.hist.data(), sizeof(Hist) * np);
    }
    std::vector<unsigned> &inv = d.inv, &off = d.off, &toff = d.toff, &wflag = d.wflag,
                          &wlist = d.wlist, &jt = d.jt;

    adj_free(c);   // tapes and checkpoints of another trajectory
    c->hinst = inst;
    int rc = install_grid(c, (unsigned)cells);
    if (rc) return rc;
    c->n = h.n;
    c->interval = (int)h.interval;
    c->last_path = 0;
    StateSrc S;
    S.rec = rec.data();
    S.whist = wh.data();
    S.inv = inv.data();
    S.ctl = &ctl;
    S.nrec = N;
    S.both = true;
    if (h.have_list) {
        S.off = off.data();
        S.toff = toff.data();
        S.wflag = wflag.data();
        S.wlist = wlist.data();
        S.ab = ab.data();
        S.hist = hist.data();
        S.jt = jt.data();
    }
    S.npairs = h.npairs;
    S.n_list = h.n_list;
    S.since = h.since;
    S.last_flag = h.last_flag;
    S.have_list = h.have_list != 0;
    S.need_rebuild = h.need_rebuild != 0;
    S.steps = h.steps;
    S.rebuilds = h.rebuilds;
    return restore_state(c, S, hipMemcpyHostToDevice);
}

namespace {

// dem_step under the displacement trigger with steps queued ahead of the host.  A queued
// step that starts after its predecessor raised the trigger does nothing (the fused step
// reads its predecessor's trigger word, the split path's kernels the gate word k_publish
// sets); a step that raises the trigger writes 1 into its host-mapped slot.  When the host
// sees the trigger it drains the queue, undoes the host-side buffer flips of the steps that
// did not run, rebuilds and issues them again.  The steps that run, and their results, are
// the synchronous loop's.  Steps are issued in batches of `pipeline_batch` with one event
// after each batch (every packet between two dependent kernels lengthens the gap between
// them); the queue is refilled when a whole batch fits.
int step_queued(dem_ctx *c, int nsteps)
{
    struct Issued {
        int rc, hc, last_path;
        bool out_valid;
        int ev;   // queue slot of the event recorded after this step's batch
    };
    Issued q[dem_ctx::kQueue];
    const int depth = c->pipeline;
    const int batch = std::min(c->pipeline_batch, depth);
    const unsigned *gate = &c->ctl->skip;
    volatile unsigned *flag = c->qflag_h;
    int done = 0, head = 0, inflight = 0;
    // queued fused steps gate on each other's trigger words: start from both clear
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, c->stream, c->ctl->trig, 2u, 0u,
                       (const unsigned *)nullptr);
    c->tpar = 0;
    while (done < nsteps) {
        if (inflight == 0) {
            if (c->need_rebuild) {
                int rc = rebuild_host(c);
                if (rc) return rc;
                if (c->fused_mode == 2 && (!c->probed || ++c->reb_since_probe >= c->probe_every)) {
                    c->probe_left = c->probed ? 2 : 4;
                    c->reb_since_probe = 0;
                }
            }
            if (c->probe_left > 0) {   // a timed probe step runs on its own
                c->fused = c->probe_left % 2 == 0;
                launch_probe_delay(c);
                DEM_HIP(hipEventRecord(c->evp0, c->stream));
                launch_force_step(c, pair_grid(c, c->npairs));
                DEM_HIP(hipEventRecord(c->evp1, c->stream));
                DEM_HIP(hipGetLastError());
                DEM_HIP(hipEventSynchronize(c->evp1));
                DEM_HIP(hipEventElapsedTime(&c->t_probe[c->fused ? 1 : 0], c->evp0, c->evp1));
                if (--c->probe_left == 0) {
                    c->fused = c->t_probe[1] <= c->t_probe[0];
                    c->probed = true;
                }
                c->steps++;
                c->since++;
                int rc = sync_mirror(c);
                if (rc) return rc;
                c->need_rebuild = c->last_flag != 0u;
                done++;
                continue;
            }
        }
        while (inflight + batch <= depth && done + inflight < nsteps) {
            const int nb = std::min(batch, nsteps - done - inflight);
            const int kl = (head + inflight + nb - 1) % dem_ctx::kQueue;   // the batch's last
            for (int i = 0; i < nb; i++) {
                int k = (head + inflight) % dem_ctx::kQueue;
                q[k] = Issued{c->rc, c->hc, c->last_path, c->out_valid, kl};
                flag[k] = 0u;   // set to 1 by the step if it raises the trigger
                launch_force_step(c, pair_grid(c, c->npairs), nullptr, nullptr, gate, c->qflag_d + k);
                inflight++;
            }
            DEM_HIP(hipEventRecord(c->evq[kl], c->stream));
        }
        DEM_HIP(hipGetLastError());
        DEM_HIP(hipEventSynchronize(c->evq[q[head].ev]));
        unsigned f = flag[head];   // the oldest queued step always runs
        head = (head + 1) % dem_ctx::kQueue;
        inflight--;
        done++;
        c->steps++;
        c->since++;
        if (f == 1u) {
            // the trigger: the queued steps after this one did nothing -- drain them, undo
            // their host-side buffer flips, and rebuild before issuing them again
            if (inflight > 0) {
                int last = (head + inflight - 1) % dem_ctx::kQueue;   // ends a batch
                DEM_HIP(hipEventSynchronize(c->evq[q[last].ev]));
                const Issued &u = q[head];
                c->rc = u.rc;
                c->hc = u.hc;
                c->last_path = u.last_path;
                c->out_valid = u.out_valid;
                inflight = 0;
            }
            c->need_rebuild = true;
        }
    }
    c->last_flag = c->need_rebuild ? 1u : 0u;
    return DEM_OK;
}

}  // namespace

int dem_step(dem_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return DEM_ERR_INVALID;
    if (c->interval == 0 && !c->timing && c->pipeline > 1) {
        DEM_HIP(hipEventRecord(c->ev0, c->stream));
        int rc = step_queued(c, nsteps);
        if (rc) return rc;
        DEM_HIP(hipEventRecord(c->ev1, c->stream));
        DEM_HIP(hipEventSynchronize(c->ev1));
        DEM_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
        return DEM_OK;
    }
    DEM_HIP(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; s++) {
        dem_ctx::TimedStep *T = nullptr;
        if (c->timing) {
            T = &c->ring[c->ring_pos];
            c->ring_pos = (c->ring_pos + 1) % dem_ctx::kRing;
            int rc = ring_flush_slot(c, *T);   // the slot's previous step, long finished
            if (rc) return rc;
            T->reb = false;
        }
        if (c->need_rebuild || (c->interval > 0 && c->since >= (unsigned)c->interval)) {
            if (T) {
                T->reb = true;
                DEM_HIP(hipEventRecord(T->r[0], c->stream));
            }
            int rc = rebuild_host(c);
            if (rc) return rc;
            if (T) DEM_HIP(hipEventRecord(T->r[1], c->stream));
            // auto path: after the first rebuild and every probe_every-th one, time one step
            // of each path (same results bit for bit) and keep the faster
            if (c->fused_mode == 2 && (!c->probed || ++c->reb_since_probe >= c->probe_every)) {
                c->probe_left = c->probed ? 2 : 4;
                c->reb_since_probe = 0;
            }
        }
        bool probe = c->probe_left > 0;
        if (probe) {
            c->fused = c->probe_left % 2 == 0;
            launch_probe_delay(c);
            DEM_HIP(hipEventRecord(c->evp0, c->stream));
        }
        if (T) {
            T->fused = c->fused;
            T->np = c->npairs;
            T->used = true;
            DEM_HIP(hipEventRecord(T->e[0], c->stream));
            launch_force_step(c, pair_grid(c, c->npairs), T->e[1], T->e[2]);
        } else {
            launch_force_step(c, pair_grid(c, c->npairs));
        }
        DEM_HIP(hipGetLastError());
        if (probe) {
            DEM_HIP(hipEventRecord(c->evp1, c->stream));
            DEM_HIP(hipEventSynchronize(c->evp1));
            DEM_HIP(hipEventElapsedTime(&c->t_probe[c->fused ? 1 : 0], c->evp0, c->evp1));
            if (--c->probe_left == 0) {
                c->fused = c->t_probe[1] <= c->t_probe[0];
                c->probed = true;
            }
        }
        c->steps++;
        c->since++;
        if (c->interval == 0) {
            int rc = sync_mirror(c);
            if (rc) return rc;
            c->need_rebuild = c->last_flag != 0u;
        }
    }
    DEM_HIP(hipEventRecord(c->ev1, c->stream));
    DEM_HIP(hipEventSynchronize(c->ev1));
    DEM_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    if (c->interval > 0) {
        const volatile DevCtl *m = reinterpret_cast<const volatile DevCtl *>(c->mirror_h);
        c->last_flag = m->need_rebuild;
    }
    return DEM_OK;
}

// Graph form: groups of K force steps followed by a list rebuild, captured once per buffer
// parity (each rebuild swaps the particle and pair buffers, so two graphs alternate) and
// replayed without host synchronisation.  The schedule equals dem_step with
// DEM_OPT_REBUILD_INTERVAL = K.  A rebuild inside a graph cannot grow buffers: capacities
// get 50% headroom over the current list, and an overflow is reported as DEM_ERR_CAPACITY
// after the replay (the state is then invalid).
int dem_step_graph(dem_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return DEM_ERR_INVALID;
    int K = c->interval;
    if (K <= 0 || nsteps % K != 0) return DEM_ERR_INVALID;
    int rc;
    if (c->need_rebuild || !c->have_list || c->since != 0) {
        if ((rc = rebuild_host(c))) return rc;
    }
    // one capacity for every pair buffer (lists, histories, output and transposed index), so
    // the kernels captured for either parity agree on it
    unsigned want = (unsigned)std::min<unsigned long long>(0xfffffff0ull, (unsigned long long)c->npairs * 3 / 2 + 4096);
    want = std::max(want, std::max(std::max(c->cap[0], c->cap[1]), std::max(c->hcap[0], c->hcap[1])));
    want = std::max(want, c->out_cap);
    if (c->cap[1 - c->lc] < want && (rc = alloc_list(c, 1 - c->lc, want))) return rc;
    if ((rc = grow_list_keep(c, c->lc, want))) return rc;
    if (c->hcap[1 - c->hc] < want && (rc = alloc_hist(c, 1 - c->hc, want))) return rc;
    if ((rc = grow_hist_keep(c, c->hc, want))) return rc;
    if (c->out_cap < want) {
        // out / jt / jf / bo hold the live transposed index and partner slots: keep their contents
        POut *no = nullptr;
        unsigned *nj = nullptr, *nf = nullptr, *nb = nullptr;
        DEM_HIP(hipMalloc(&no, (size_t)want * sizeof(POut)));
        DEM_HIP(hipMalloc(&nj, (size_t)want * sizeof(unsigned)));
        DEM_HIP(hipMalloc(&nf, (size_t)want * sizeof(unsigned)));
        DEM_HIP(hipMalloc(&nb, (size_t)want * sizeof(unsigned)));
        DEM_HIP(hipMemcpyAsync(nj, c->jt, (size_t)c->out_cap * sizeof(unsigned), hipMemcpyDeviceToDevice, c->stream));
        DEM_HIP(hipMemcpyAsync(nf, c->jf, (size_t)c->out_cap * sizeof(unsigned), hipMemcpyDeviceToDevice, c->stream));
        DEM_HIP(hipMemcpyAsync(nb, c->bo, (size_t)c->out_cap * sizeof(unsigned), hipMemcpyDeviceToDevice, c->stream));
        DEM_HIP(hipStreamSynchronize(c->stream));
        DEM_HIP(hipFree(c->out));
        DEM_HIP(hipFree(c->jt));
        DEM_HIP(hipFree(c->jf));
        DEM_HIP(hipFree(c->bo));
        DEM_HIP(hipFree(c->zs));
        c->out = no;
        c->jt = nj;
        c->jf = nf;
        c->bo = nb;
        c->zs = nullptr;
        c->zs_ok = false;
        DEM_HIP(hipMalloc(&c->zs, zs_bytes(want)));
        c->out_cap = want;
        c->buf_gen++;
    }
    unsigned np_grid = pair_grid(c, want);
    if (!c->gexec[0] || c->graph_k != K || c->graph_np_grid != np_grid || c->graph_gen != c->buf_gen ||
        c->graph_fused != c->fused) {
        drop_graph(c);
        hipStream_t st = c->stream;
        for (int g = 0; g < 2; g++) {
            DEM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; k++) launch_force_step(c, np_grid);
            rc = rebuild_count(c, nullptr);
            if (!rc) rc = rebuild_fill(c, nullptr);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(st, &gr);
            if (rc || e != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                drop_graph(c);
                return rc ? rc : DEM_ERR_HIP;
            }
            c->graph[g] = gr;
            DEM_HIP(hipGraphInstantiate(&c->gexec[g], gr, nullptr, nullptr, 0));
            // buffer parities after one group: the second graph starts from them
            if (g == 0) {
                c->gpar[0] = c->rc; c->gpar[1] = c->lc; c->gpar[2] = c->hc; c->gpar[3] = c->pc;
            }
        }
        // a group flips each parity the same way every time, so two groups restore them:
        // the host indices are back at the live state
        c->graph_k = K;
        c->graph_np_grid = np_grid;
        c->graph_gen = c->buf_gen;
        c->graph_fused = c->fused;
    }
    int groups = nsteps / K;
    DEM_HIP(hipEventRecord(c->ev0, c->stream));
    for (int g = 0; g < groups; g++) DEM_HIP(hipGraphLaunch(c->gexec[g & 1], c->stream));
    DEM_HIP(hipEventRecord(c->ev1, c->stream));
    DEM_HIP(hipEventSynchronize(c->ev1));
    DEM_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    if (groups & 1) {
        // odd number of groups: the live buffers are those after one group, and the graphs
        // must start from the other one next time
        int r0 = c->rc, l0 = c->lc, h0 = c->hc, p0 = c->pc;
        c->rc = c->gpar[0]; c->lc = c->gpar[1]; c->hc = c->gpar[2]; c->pc = c->gpar[3];
        c->gpar[0] = r0; c->gpar[1] = l0; c->gpar[2] = h0; c->gpar[3] = p0;
        std::swap(c->gexec[0], c->gexec[1]);
        std::swap(c->graph[0], c->graph[1]);
    }
    c->out_valid = false;
    unsigned w[2];
    DEM_HIP(hipMemcpy(w, &c->ctl->npairs, 2 * sizeof(unsigned), hipMemcpyDeviceToHost));
    c->npairs = w[0];
    c->steps += (unsigned long long)nsteps;
    c->rebuilds += (unsigned long long)groups;
    c->since = 0;
    c->need_rebuild = false;
    const volatile DevCtl *m = reinterpret_cast<const volatile DevCtl *>(c->mirror_h);
    c->last_flag = m->need_rebuild;
    if (w[1]) return DEM_ERR_CAPACITY;
    return DEM_OK;
}

// Same loop as dem_step, with HIP events on the context stream around the rebuild, the pair
// kernel and the per-particle kernel of every step.  out_ms[0] = mean pair-kernel time,
// out_ms[1] = mean per-particle-kernel time, out_ms[2] = mean rebuild time (0 if none),
// out_ms[3] = total device time; *nrebuild = rebuilds run.
int dem_time_kernels(dem_ctx *c, int nsteps, double *out_ms, int *nrebuild)
{
    if (!c || nsteps < 1 || !out_ms) return DEM_ERR_INVALID;
    hipEvent_t e0, e1, e2, e3, r0, r1;
    DEM_HIP(hipEventCreate(&e0));
    DEM_HIP(hipEventCreate(&e1));
    DEM_HIP(hipEventCreate(&e2));
    DEM_HIP(hipEventCreate(&e3));
    DEM_HIP(hipEventCreate(&r0));
    DEM_HIP(hipEventCreate(&r1));
    double t_pair = 0.0, t_part = 0.0, t_reb = 0.0, t_all = 0.0;
    int nreb = 0;
    for (int s = 0; s < nsteps; s++) {
        if (c->need_rebuild || (c->interval > 0 && c->since >= (unsigned)c->interval)) {
            DEM_HIP(hipEventRecord(r0, c->stream));
            int rc = rebuild_host(c);
            if (rc) return rc;
            DEM_HIP(hipEventRecord(r1, c->stream));
            DEM_HIP(hipEventSynchronize(r1));
            float ms;
            DEM_HIP(hipEventElapsedTime(&ms, r0, r1));
            t_reb += ms;
            nreb++;
        }
        DEM_HIP(hipEventRecord(e0, c->stream));
        launch_force_step(c, pair_grid(c, c->npairs), e1, e2);
        DEM_HIP(hipGetLastError());
        c->steps++;
        c->since++;
        int rc = sync_mirror(c);
        if (rc) return rc;
        if (c->interval == 0) c->need_rebuild = c->last_flag != 0u;
        float a, b;
        DEM_HIP(hipEventElapsedTime(&a, e0, e1));
        DEM_HIP(hipEventElapsedTime(&b, e1, e2));
        t_pair += a;
        t_part += b;
        t_all += a + b;
    }
    (void)e3;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2); (void)hipEventDestroy(e3);
    (void)hipEventDestroy(r0); (void)hipEventDestroy(r1);
    out_ms[0] = t_pair / nsteps;
    out_ms[1] = t_part / nsteps;
    out_ms[2] = nreb ? t_reb / nreb : 0.0;
    out_ms[3] = t_all + t_reb;
    if (nrebuild) *nrebuild = nreb;
    return DEM_OK;
}

// One host-driven rebuild now, timed stage by stage with HIP events on the context stream:
// out[0..8] = ms of hash (cell counts), sort (count scan and scatter), cell ranges (0: the
// scan gives them), reorder, neighbour count, the two scans (+ capacity check), pair fill,
// history remap, transpose; out[9] = radix passes (0: counting sort); out[10] = pairs of the
// new list; out[11] = pairs of the old list; out[12] = cells; out[13] = particles; out[14] =
// ms of a second rebuild timed with events at its two phases' ends only (count phase from
// the hash to the scans, fill phase from the fill to the transpose; the host's read of the
// pair total between them excluded), free of the per-stage events' gaps.  The trajectory is
// unchanged (a rebuild at any step gives the same forces: lists hold every contact, sums
// follow particle order).
int dem_time_rebuild(dem_ctx *c, double *out)
{
    if (!c || !out) return DEM_ERR_INVALID;
    int rc = ring_flush_all(c);
    if (rc) return rc;
    hipEvent_t ev[kEvCountAll];
    for (int k = 0; k < kEvCountAll; k++) DEM_HIP(hipEventCreate(&ev[k]));
    const unsigned np_old = c->have_list ? c->npairs : 0u;
    rc = rebuild_host(c, nullptr, nullptr, ev);
    if (!rc) rc = hipEventSynchronize(ev[kEvTranspose]) == hipSuccess ? DEM_OK : DEM_ERR_HIP;
    if (!rc) {
        static const int a[9] = {kEvStart, kEvHash, kEvSort, kEvSort, kEvReorder, kEvCount,
                                 kEvFillStart, kEvFill, kEvRemap};
        static const int b[9] = {kEvHash, kEvSort, kEvSort, kEvReorder, kEvCount, kEvScan,
                                 kEvFill, kEvRemap, kEvTranspose};
        for (int k = 0; k < 9; k++) {
            float ms = 0.0f;
            if (a[k] != b[k] && hipEventElapsedTime(&ms, ev[a[k]], ev[b[k]]) != hipSuccess)
                rc = DEM_ERR_HIP;
            out[k] = ms;
        }
        out[9] = 0;
        out[10] = c->npairs;
        out[11] = np_old;
        out[12] = c->ncells;
        out[13] = c->n;
    }
    if (!rc) {
        hipEvent_t ph[kEvCountAll] = {};
        for (int k : {kEvStart, kEvScan, kEvFillStart, kEvTranspose}) ph[k] = ev[k];
        rc = rebuild_host(c, nullptr, nullptr, ph);
        if (!rc) rc = hipEventSynchronize(ev[kEvTranspose]) == hipSuccess ? DEM_OK : DEM_ERR_HIP;
        float m0 = 0.0f, m1 = 0.0f;
        if (!rc && (hipEventElapsedTime(&m0, ev[kEvStart], ev[kEvScan]) != hipSuccess ||
                    hipEventElapsedTime(&m1, ev[kEvFillStart], ev[kEvTranspose]) != hipSuccess))
            rc = DEM_ERR_HIP;
        out[14] = (double)m0 + (double)m1;
    }
    for (int k = 0; k < kEvCountAll; k++) (void)hipEventDestroy(ev[k]);
    return rc;
}

// Accumulated kernel times of dem_step under DEM_OPT_TIME_STEPS (ms), then reset:
// out[0..3] = split path: first group (pair kernel + walls) total, per-particle kernel total,
// steps, sum of pair counts; out[4..7] = fused path: walls total, step kernel total, steps,
// sum of pair counts; out[8] = rebuild total, out[9] = rebuilds.
int dem_get_step_timing(dem_ctx *c, double *out)
{
    if (!c || !out) return DEM_ERR_INVALID;
    int rc = ring_flush_all(c);
    if (rc) return rc;
    for (int p = 0; p < 2; p++) {
        out[4 * p] = c->t_acc[p][0];
        out[4 * p + 1] = c->t_acc[p][1];
        out[4 * p + 2] = c->n_acc[p];
        out[4 * p + 3] = c->np_acc[p];
        c->t_acc[p][0] = c->t_acc[p][1] = c->n_acc[p] = c->np_acc[p] = 0.0;
    }
    out[8] = c->t_reb;
    out[9] = c->n_reb;
    c->t_reb = c->n_reb = 0.0;
    return DEM_OK;
}

static int reduce_instances(dem_ctx *c)
{
    // per-instance max z / contact totals, reduced into the device control block's tail (the
    // per-wave atomics stay in HBM) and copied to the mirror's tail
    size_t tail = c->ctl_bytes - sizeof(DevCtl);
    unsigned char *dt = reinterpret_cast<unsigned char *>(c->ctl) + sizeof(DevCtl);
    unsigned long long *dz = reinterpret_cast<unsigned long long *>(dt);
    unsigned *dp = reinterpret_cast<unsigned *>(dz + c->ninst);
    DEM_HIP(hipMemsetAsync(dt, 0, tail, c->stream));
    unsigned g = std::max(1u, std::min(cdiv(c->n, kBlock), (unsigned)c->cus * 4u));
    hipLaunchKernelGGL(k_inst_reduce, dim3(g), dim3(kBlock), 0, c->stream,
                       (const PRec *)c->rec[c->rc], (const unsigned *)c->contacti, c->n, dz, dp,
                       c->ninst == 1 ? 1u : 0u);
    DEM_HIP(hipGetLastError());
    DEM_HIP(hipMemcpyAsync(c->mirror_h + sizeof(DevCtl), dt, tail, hipMemcpyDeviceToHost, c->stream));
    DEM_HIP(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

// One trial of `steps` steps with the reference trigger, timed with HIP events: mean force-
// step time (pair + walls + particle kernels), mean rebuild time, and the smallest gap (in
// steps) between trigger-raised rebuilds (-1 if none fired).
static int tune_trial(dem_ctx *c, int steps, double *force_ms, double *rebuild_ms, int *min_gap,
                      hipEvent_t e0, hipEvent_t e1)
{
    double tf = 0.0, tr = 0.0;
    int nf = 0, nr = 0, last = -1, gap = -1;
    for (int s = 0; s < steps; s++) {
        if (c->need_rebuild) {
            bool triggered = c->last_flag != 0u;
            DEM_HIP(hipEventRecord(e0, c->stream));
            int rc = rebuild_host(c);
            if (rc) return rc;
            DEM_HIP(hipEventRecord(e1, c->stream));
            DEM_HIP(hipEventSynchronize(e1));
            float ms;
            DEM_HIP(hipEventElapsedTime(&ms, e0, e1));
            tr += ms;
            nr++;
            if (triggered) {
                if (last >= 0 && (gap < 0 || s - last < gap)) gap = s - last;
                last = s;
            } else {
                last = s;
            }
        }
        DEM_HIP(hipEventRecord(e0, c->stream));
        launch_force_step(c, pair_grid(c, c->npairs));
        DEM_HIP(hipEventRecord(e1, c->stream));
        c->steps++;
        c->since++;
        int rc = sync_mirror(c);
        if (rc) return rc;
        c->need_rebuild = c->last_flag != 0u;
        float ms;
        DEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        tf += ms;
        nf++;
    }
    *force_ms = nf ? tf / nf : 0.0;
    *rebuild_ms = nr ? tr / nr : 0.0;
    *min_gap = gap;
    return DEM_OK;
}

int dem_autotune(dem_ctx *c, int steps, dem_tune_result *out)
{
    if (!c || !out || steps < 8) return DEM_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    hipEvent_t e0, e1;
    DEM_HIP(hipEventCreate(&e0));
    DEM_HIP(hipEventCreate(&e1));
    int saved_interval = c->interval;
    c->interval = 0;   // trials use the reference trigger, so every trial step is exact
    static const double skins[] = {0.3, 0.45, 0.55, 0.7, 0.9};
    double best_score = 1e300, best_skin = c->hinst[0].skin;
    int best_k = 1;
    int trials = 0;
    for (double sk : skins) {
        int rc = regrid(c, sk);
        if (rc) return rc;
        double tf, tr;
        int gap;
        if ((rc = tune_trial(c, steps, &tf, &tr, &gap, e0, e1))) return rc;
        trials++;
        // safe fixed interval: 3/4 of the shortest observed trigger gap (the trial length if
        // the trigger never fired)
        int k = gap > 0 ? std::max(1, (int)(0.75 * gap)) : steps;
        double score = tf + tr / k;
        if (score < best_score) { best_score = score; best_skin = sk; best_k = k; }
    }
    int rc = regrid(c, best_skin);
    if (rc) return rc;
    // launch shapes at the chosen skin (force steps only): the split path's pair block, pair
    // grid and particle block; the fused path's block; and, in auto mode, the faster path
    static const unsigned blocks[] = {64, 128, 256};
    static const unsigned grids[] = {4, 8, 16, 32};
    bool try_split = c->fused_mode != 1, try_fused = c->fused_mode != 0;
    double best_t = 1e300, best_fused_t = 1e300;
    unsigned bp = c->pair_block, bg = c->pair_grid_per_cu, bq = c->particle_block, fq = c->particle_block;
    int sub = std::max(4, steps / 4);
    if (try_split) {
        c->fused = false;
        for (unsigned pb : blocks)
            for (unsigned pg : grids)
                for (unsigned qb : blocks) {
                    if (qb != 256 && !(pb == 256 && pg == 16)) continue;   // particle block swept once
                    c->pair_block = pb;
                    c->pair_grid_per_cu = pg;
                    c->particle_block = qb;
                    double tf, tr;
                    int gap;
                    if ((rc = tune_trial(c, sub, &tf, &tr, &gap, e0, e1))) return rc;
                    trials++;
                    if (tf < best_t) { best_t = tf; bp = pb; bg = pg; bq = qb; }
                }
    }
    if (try_fused) {
        c->fused = true;
        for (unsigned qb : blocks) {
            c->particle_block = qb;
            double tf, tr;
            int gap;
            if ((rc = tune_trial(c, sub, &tf, &tr, &gap, e0, e1))) return rc;
            trials++;
            if (tf < best_fused_t) { best_fused_t = tf; fq = qb; }
        }
    }
    bool use_fused = try_fused && (!try_split || best_fused_t < best_t);
    c->fused = use_fused;
    c->pair_block = bp;
    c->pair_grid_per_cu = bg;
    c->particle_block = use_fused ? fq : bq;
    if (use_fused) best_t = best_fused_t;
    if (c->fused_mode == 2) {
        c->probed = true;
        c->reb_since_probe = 0;
        c->probe_left = 0;
    }
    c->interval = saved_interval;
    drop_graph(c);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out->skin = best_skin;
    out->cell = c->hinst[0].cell[0];
    out->interval = best_k;
    out->pair_block = bp;
    out->pair_grid_per_cu = bg;
    out->particle_block = c->particle_block;
    out->step_path = use_fused ? 1 : 0;
    out->ms_per_step = best_t;
    out->score_ms = best_score;
    out->trials = trials;
    return DEM_OK;
}

int dem_get_stats(dem_ctx *c, dem_stats *o)
{
    if (!c || !o) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    if (c->steps && reduce_instances(c)) return DEM_ERR_HIP;
    const volatile DevCtl *m = reinterpret_cast<const volatile DevCtl *>(c->mirror_h);
    std::memset(o, 0, sizeof(*o));
    o->steps = c->steps;
    o->rebuilds = c->rebuilds;
    o->n = c->n;
    o->npairs = c->npairs;
    o->pair_capacity = std::min(c->cap[c->lc], c->out_cap);
    o->ninstances = (unsigned)c->ninst;
    o->last_rebuild_flag = c->last_flag;
    o->overflow = m->overflow;
    const volatile unsigned long long *mz = reinterpret_cast<const volatile unsigned long long *>(c->mirror_h + sizeof(DevCtl));
    const volatile unsigned *mp = reinterpret_cast<const volatile unsigned *>(mz + c->ninst);
    unsigned long long zb = mz[0];
    double z;
    std::memcpy(&z, &zb, 8);
    o->maxhz = z;
    o->pcntot = mp[0];
    o->step_path = c->fused ? 1u : 0u;
    return DEM_OK;
}

int dem_get_instance_stats(dem_ctx *c, double *maxhz, uint *pcntot)
{
    if (!c) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    if (c->steps && reduce_instances(c)) return DEM_ERR_HIP;
    const volatile unsigned long long *mz = reinterpret_cast<const volatile unsigned long long *>(c->mirror_h + sizeof(DevCtl));
    const volatile unsigned *mp = reinterpret_cast<const volatile unsigned *>(mz + c->ninst);
    for (int k = 0; k < c->ninst; k++) {
        if (maxhz) {
            unsigned long long zb = mz[k];
            std::memcpy(&maxhz[k], &zb, 8);
        }
        if (pcntot) pcntot[k] = mp[k];
    }
    return DEM_OK;
}

int dem_sync(dem_ctx *c)
{
    if (!c) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

double dem_last_ms(dem_ctx *c) { return c ? (double)c->last_ms : -1.0; }

int dem_get_hash(dem_ctx *c, uint *kv)
{
    if (!c || !kv || !c->kv_sorted) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    // the device holds the pairs by slot; the reference's array is in particle-index order
    std::vector<uint2> h(c->n);
    DEM_HIP(hipMemcpy(h.data(), c->kv_hash, sizeof(uint2) * (size_t)c->n, hipMemcpyDeviceToHost));
    for (const uint2 &e : h) {
        if (e.y >= c->n) return DEM_ERR_INVALID;
        kv[2 * (size_t)e.y] = e.x;
        kv[2 * (size_t)e.y + 1] = e.y;
    }
    return DEM_OK;
}

int dem_get_sorted(dem_ctx *c, uint *kv)
{
    if (!c || !kv || !c->kv_sorted) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    DEM_HIP(hipMemcpy(kv, c->kv_sorted, sizeof(uint2) * (size_t)c->n, hipMemcpyDeviceToHost));
    return DEM_OK;
}

int dem_get_cell_start(dem_ctx *c, uint *cs, uint ncells)
{
    if (!c || !cs || ncells < c->ncells || !c->kv_sorted) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    // the device holds each cell's first slot for every cell (ncells + 1 entries); the
    // reference's cellStart marks empty cells 0xffffffff
    std::vector<unsigned> pre((size_t)c->ncells + 1);
    DEM_HIP(hipMemcpy(pre.data(), c->cstart, sizeof(unsigned) * pre.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < c->ncells; k++) cs[k] = pre[k] < pre[k + 1] ? pre[k] : kEmpty;
    return DEM_OK;
}

int dem_get_pairs(dem_ctx *c, uint *ids, double *hist, double *out, uint cap)
{
    if (!c) return DEM_ERR_INVALID;
    if (cap < c->npairs) return DEM_ERR_CAPACITY;
    size_t np = c->npairs;
    if (!np) return DEM_OK;
    std::vector<PRec> recs;
    int rc = fetch_records(c, recs);
    if (rc) return rc;
    if (ids) {
        std::vector<uint2> ab(np);
        DEM_HIP(hipMemcpy(ab.data(), c->ab[c->lc], sizeof(uint2) * np, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < np; p++) {
            ids[2 * p] = recs[ab[p].x].id;
            ids[2 * p + 1] = recs[ab[p].y].id;
        }
    }
    if (hist) DEM_HIP(hipMemcpy(hist, c->hist[c->hc], sizeof(Hist) * np, hipMemcpyDeviceToHost));
    if (out) {
        if (c->last_path == 2) {
            // the fused step keeps no per-pair output: evaluate the last step's pair laws
            // again from its inputs, which stay in the other buffers until the next step
            // (zeros when no step has run on the current list)
            if (!c->out_valid) {
                std::memset(out, 0, sizeof(POut) * np);
                return DEM_OK;
            }
            InstDev one = c->hinst[0];
            unsigned g = pair_grid(c, (unsigned)np);
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(g), dim3(c->pair_block), 0, c->stream,
                                   (const uint2 *)c->ab[c->lc], (const unsigned *)&c->ctl->npairs,
                                   (const PRec *)c->rec[1 - c->rc], (const Hist *)c->hist[1 - c->hc],
                                   (Hist *)nullptr, c->out, (const InstDev *)c->dinst, one.pp, one.dt,
                                   one.roll_k, (Hist *)nullptr, (const unsigned *)nullptr, c->xcd_chunk);
            };
            if (c->batched)
                one.flen ? go(k_pair_force<true, true>) : go(k_pair_force<true, false>);
            else
                one.flen ? go(k_pair_force<false, true>) : go(k_pair_force<false, false>);
            DEM_HIP(hipGetLastError());
            DEM_HIP(hipStreamSynchronize(c->stream));
        }
        DEM_HIP(hipMemcpy(out, c->out, sizeof(POut) * np, hipMemcpyDeviceToHost));
    }
    return DEM_OK;
}

int dem_get_contacts(dem_ctx *c, uint *contacti)
{
    if (!c || !contacti) return DEM_ERR_INVALID;
    std::vector<PRec> recs;
    int rc = fetch_records(c, recs);
    if (rc) return rc;
    std::vector<unsigned> ct(c->cap_n);
    DEM_HIP(hipMemcpy(ct.data(), c->contacti, sizeof(unsigned) * (size_t)c->cap_n, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < c->cap_n; s++) contacti[recs[s].id] = s < c->n ? ct[s] : 0u;
    return DEM_OK;
}

int dem_get_forces(dem_ctx *c, double *force, double *torque, uint *contacti)
{
    if (!c || !c->record_forces || !c->fdbg) return DEM_ERR_INVALID;
    std::vector<PRec> recs;
    int rc = fetch_records(c, recs);
    if (rc) return rc;
    size_t n = c->cap_n;
    std::vector<double> f(6 * n);
    std::vector<unsigned> ct(n);
    DEM_HIP(hipMemcpy(f.data(), c->fdbg, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
    DEM_HIP(hipMemcpy(ct.data(), c->contacti, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < n; s++) {
        size_t i = recs[s].id;
        bool act = s < c->n;
        for (int k = 0; k < 3; k++) {
            if (force) force[3 * i + k] = act ? f[6 * s + k] : 0.0;
            if (torque) torque[3 * i + k] = act ? f[6 * s + 3 + k] : 0.0;
        }
        if (contacti) contacti[i] = act ? ct[s] : 0u;
    }
    return DEM_OK;
}

// =========================================================================================
// Host evaluation of the contact laws and their vector-Jacobian products (the same inline
// code the kernels run), for tests and tools.
// =========================================================================================

namespace {
LawParams law6(const double *m) { return LawParams{m[0], m[1], m[2], m[3], m[4], m[5]}; }
d3 v3(const double *p) { return d3{p[0], p[1], p[2]}; }
void put3(double *p, d3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
WallGeom wall_geom(const dem_walls *w)
{
    WallGeom g;
    g.kind = w->kind;
    g.hx = 0.5 * w->size[0];
    g.hy = 0.5 * w->size[1];
    g.radius = w->radius;
    return g;
}
}  // namespace

int dem_law_pair_eval_floor(const double *in, const double *mat, double roll_c, double *out)
{
    if (!in || !mat || !out || !(roll_c >= 0.0)) return -1;
    d3 fa, ta, tb, ho;
    double fo;
    bool c = law_pair(v3(in), v3(in + 3), v3(in + 6), v3(in + 9), v3(in + 12), v3(in + 15), in[18],
                      in[19], in[20], in[21], v3(in + 22), in[25], law6(mat), fa, ta, tb, ho, fo,
                      roll_c);
    put3(out, fa); put3(out + 3, ta); put3(out + 6, tb); put3(out + 9, ho);
    out[12] = fo;
    return c ? 1 : 0;
}

int dem_law_pair_eval(const double *in, const double *mat, double *out)
{
    return dem_law_pair_eval_floor(in, mat, 0.0, out);
}

int dem_contact_filter(const double *in, int count, int *keep)
{
    if (!in || !keep || count < 0) return DEM_ERR_INVALID;
    for (int i = 0; i < count; i++) {
        const double *q = in + 8 * (size_t)i;
        keep[i] = maybe_touch(pos32(q[0], q[1], q[2], q[3]), pos32(q[4], q[5], q[6], q[7])) ? 1 : 0;
    }
    return DEM_OK;
}

int dem_law_pair_vjp_floor(const double *in, const double *mat, double roll_c,
                           const double *outbar, double *inbar, double *matbar)
{
    if (!in || !mat || !outbar || !inbar || !matbar || !(roll_c >= 0.0)) return -1;
    d3 g[7];
    for (int k = 0; k < 7; k++) g[k] = d3{0.0, 0.0, 0.0};
    LawParamsBar pb = {0.0, 0.0, 0.0, 0.0, 0.0};
    law_pair_adj(v3(in), v3(in + 3), v3(in + 6), v3(in + 9), v3(in + 12), v3(in + 15), in[18],
                 in[19], in[20], in[21], v3(in + 22), in[25], law6(mat), v3(outbar),
                 v3(outbar + 3), v3(outbar + 6), v3(outbar + 9), g[0], g[1], g[2], g[3], g[4],
                 g[5], g[6], pb, roll_c, roll_c > 0.0);
    for (int k = 0; k < 7; k++) put3(inbar + 3 * k, g[k]);
    matbar[0] = pb.estar; matbar[1] = pb.mu; matbar[2] = pb.gn; matbar[3] = pb.gt; matbar[4] = pb.ur;
    return DEM_OK;
}

int dem_law_pair_vjp(const double *in, const double *mat, const double *outbar, double *inbar,
                     double *matbar)
{
    return dem_law_pair_vjp_floor(in, mat, 0.0, outbar, inbar, matbar);
}

int dem_law_walls_eval_floor(const dem_walls *walls, const double *in, const double *mat,
                             double roll_c, double *out)
{
    if (!walls || !in || !mat || !out || !(roll_c >= 0.0)) return -1;
    d3 fw, tw, hw;
    double fo;
    int hits = law_walls(wall_geom(walls), v3(in), v3(in + 3), v3(in + 6), in[9], in[10],
                         v3(in + 11), in[14], law6(mat), fw, tw, hw, fo, roll_c);
    put3(out, fw); put3(out + 3, tw); put3(out + 6, hw);
    out[9] = fo;
    return hits;
}

int dem_law_walls_eval(const dem_walls *walls, const double *in, const double *mat, double *out)
{
    return dem_law_walls_eval_floor(walls, in, mat, 0.0, out);
}

int dem_law_walls_vjp_floor(const dem_walls *walls, const double *in, const double *mat,
                            double roll_c, const double *outbar, double *inbar, double *matbar)
{
    if (!walls || !in || !mat || !outbar || !inbar || !matbar || !(roll_c >= 0.0)) return -1;
    d3 g[4];
    for (int k = 0; k < 4; k++) g[k] = d3{0.0, 0.0, 0.0};
    LawParamsBar pb = {0.0, 0.0, 0.0, 0.0, 0.0};
    law_walls_adj(wall_geom(walls), v3(in), v3(in + 3), v3(in + 6), in[9], in[10], v3(in + 11),
                  in[14], law6(mat), v3(outbar), v3(outbar + 3), v3(outbar + 6), g[0], g[1], g[2],
                  g[3], pb, roll_c, roll_c > 0.0);
    for (int k = 0; k < 4; k++) put3(inbar + 3 * k, g[k]);
    matbar[0] = pb.estar; matbar[1] = pb.mu; matbar[2] = pb.gn; matbar[3] = pb.gt; matbar[4] = pb.ur;
    return DEM_OK;
}

int dem_law_walls_vjp(const dem_walls *walls, const double *in, const double *mat,
                      const double *outbar, double *inbar, double *matbar)
{
    return dem_law_walls_vjp_floor(walls, in, mat, 0.0, outbar, inbar, matbar);
}

// =========================================================================================
// Drop-in: reference-named entry points (particleSystem.cu:13-424)
// =========================================================================================

namespace {
Matproperties g_par = {0, 0, 0, 0, 0, 0};
Wallproperties g_parw = {0, 0, 0, 0, 0, 0};
bool g_walls_set = false;
dem_walls g_walls = {DEM_WALLS_BOX, 0, {0, 0, 0}, 0};
Scratch g_S;
unsigned *g_words = nullptr;                 // [0] iall, [1] overflow, [2] pcntot, [3] scan total, [4..5] maxz

constexpr double kCua = 1.0;                 // neiglist.cu:27
constexpr double kCub = 1.55;                // neiglist.cu:28

int compat_words()
{
    if (!g_words) DEM_HIP(hipMalloc(&g_words, 8 * sizeof(unsigned)));
    return DEM_OK;
}

// neighborarray's sorted copies and cell prefix, grown on demand.
struct NbrScratch {
    double4 *spos = nullptr;
    unsigned *sid = nullptr, *pre = nullptr, *cnt = nullptr;
    size_t n_cap = 0, c_cap = 0;
    int ensure(size_t n, size_t cells)
    {
        if (n > n_cap) {
            if (spos) DEM_HIP(hipFree(spos));
            if (sid) DEM_HIP(hipFree(sid));
            spos = nullptr;
            sid = nullptr;
            n_cap = 0;
            DEM_HIP(hipMalloc(&spos, n * sizeof(double4)));
            DEM_HIP(hipMalloc(&sid, n * sizeof(unsigned)));
            n_cap = n;
        }
        if (cells + 1 > c_cap) {
            if (pre) DEM_HIP(hipFree(pre));
            if (cnt) DEM_HIP(hipFree(cnt));
            pre = cnt = nullptr;
            c_cap = 0;
            DEM_HIP(hipMalloc(&pre, (cells + 1) * sizeof(unsigned)));
            DEM_HIP(hipMalloc(&cnt, (cells + 1) * sizeof(unsigned)));
            c_cap = cells + 1;
        }
        return DEM_OK;
    }
} g_nbr;

CompatGrid make_grid(const uint g[3], const double cell[3], const double origin[3])
{
    CompatGrid G;
    for (int k = 0; k < 3; k++) { G.g[k] = g[k]; G.cell[k] = cell[k]; G.origin[k] = origin[k]; }
    return G;
}

void check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::fprintf(stderr, "GPUassert: %s (%s)\n", hipGetErrorString(e), what);
        g_last_error = DEM_ERR_HIP;
    }
}
}  // namespace

void setMaterialProperties(const Matproperties *par, const Wallproperties *parw)
{
    if (par) g_par = *par;
    if (parw) g_parw = *parw;
}

void setWallGeometry(const dem_walls *walls)
{
    if (walls) { g_walls = *walls; g_walls_set = true; }
    else g_walls_set = false;
}

void calcHash(double *Pos, uint *particleHash, uint gridSize[3], double cellSize[3],
              double worldOrigin[3], int numBodies)
{
    if (numBodies <= 0) return;
    hipLaunchKernelGGL(c_calc_hash, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0,
                       (const double *)Pos, (unsigned)numBodies,
                       make_grid(gridSize, cellSize, worldOrigin), (uint2 *)particleHash);
    check_launch("calcHash");
}

void RadixSort(KeyValuePair *pData0, KeyValuePair *pData1, uint elements, uint bits)
{
    if (elements == 0) return;
    int rc;
    uint2 *a = reinterpret_cast<uint2 *>(pData0), *b = reinterpret_cast<uint2 *>(pData1);
    int nb = (int)std::min(bits, 32u);
    // ping-pong starting from pData0: pass k reads one buffer and writes the other
    uint2 *res = radix_sort(g_S, a, b, a, elements, nb, 0, nullptr, &rc);
    if (rc) { g_last_error = rc; return; }
    if (res != a) DEM_HIP_V(hipMemcpyAsync(a, res, sizeof(uint2) * elements, hipMemcpyDeviceToDevice, 0));
    check_launch("RadixSort");
}

// ---- RadixSort without allocation --------------------------------------------------------
// The sort's scratch (digit counts, their scan, the scan's tile sums) lives in one file-scope
// device array, sized for kNoAllocMaxElements keys; each device holds its own copy, whose
// address is resolved once per device (hipGetSymbolAddress) and kept in g_noalloc_ptr.
// Every word the sort reads is written earlier in the same call (the histogram pass writes
// every count of every pass, the scan every tile sum and offset), so nothing needs clearing.
constexpr unsigned kNoAllocMaxElements = 1u << 23;                                // 8,388,608
constexpr size_t kNaSortTiles = kNoAllocMaxElements / kSortTile;                  // 4,096
constexpr size_t kNaCounts = 256 * kNaSortTiles;                                 // 1,048,576
constexpr size_t kNaScanTiles = (kNaCounts + kScanTile - 1) / kScanTile;         // 512
constexpr size_t kNaOffs = kNaCounts;                                            // word offsets
constexpr size_t kNaTiles = kNaOffs + kNaCounts + 64;
constexpr size_t kNaTotal = kNaTiles + kNaScanTiles;
constexpr size_t kNaWords = kNaTotal + 64;                                       // 8.0 MiB + 2.5 KiB
static_assert(kNoAllocMaxElements % kSortTile == 0, "whole sort tiles");

static __device__ __align__(256) unsigned scratch[kNaWords];

namespace {
constexpr int kNaMaxDevices = 64;
unsigned *g_noalloc_ptr[kNaMaxDevices] = {};
std::mutex g_noalloc_mu;

// Device address of this device's copy of `scratch` (resolved on the first call per device).
int noalloc_scratch(unsigned **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return DEM_ERR_NO_DEVICE;
    if (dev < 0 || dev >= kNaMaxDevices) return DEM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        unsigned *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            std::fprintf(stderr, "RadixSortNoAlloc: hipGetSymbolAddress failed (%s)\n",
                         hipGetErrorString(e));
            return DEM_ERR_HIP;
        }
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return DEM_OK;
}
}  // namespace

int RadixSortNoAlloc(KeyValuePair *pData0, KeyValuePair *pData1, uint elements, uint bits)
{
    if (elements == 0) return DEM_OK;
    if (!pData0 || !pData1 || pData0 == pData1) {
        g_last_error = DEM_ERR_INVALID;
        return DEM_ERR_INVALID;
    }
    if (elements > kNoAllocMaxElements) {
        std::fprintf(stderr, "RadixSortNoAlloc: %u elements exceed its fixed scratch (at most %u)\n",
                     elements, kNoAllocMaxElements);
        g_last_error = DEM_ERR_CAPACITY;
        return DEM_ERR_CAPACITY;
    }
    unsigned *base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) {
        g_last_error = rc;
        return rc;
    }
    Scratch S;   // a view of the device array; ensure() checks sizes and allocates nothing
    S.fixed = true;
    S.counts = base;
    S.offs = base + kNaOffs;
    S.counts_cap = kNaCounts;
    S.tiles = base + kNaTiles;
    S.tiles_cap = kNaScanTiles;
    S.total = base + kNaTotal;
    uint2 *a = reinterpret_cast<uint2 *>(pData0), *b = reinterpret_cast<uint2 *>(pData1);
    uint2 *res = radix_sort(S, a, b, a, elements, (int)std::min(bits, 32u), 0, nullptr, &rc);
    if (rc) {
        g_last_error = rc;
        return rc;
    }
    if (res != a &&
        hipMemcpyAsync(a, res, sizeof(uint2) * elements, hipMemcpyDeviceToDevice, 0) != hipSuccess) {
        g_last_error = DEM_ERR_HIP;
        return DEM_ERR_HIP;
    }
    if (hipGetLastError() != hipSuccess) {
        g_last_error = DEM_ERR_HIP;
        return DEM_ERR_HIP;
    }
    return DEM_OK;
}

void findCellStart(uint *particleHash, uint *cellStart, int numBodies, uint numGridCells)
{
    hipLaunchKernelGGL(k_fill, dim3(std::min(cdiv(numGridCells, kBlock), 4096u)), dim3(kBlock), 0, 0,
                       cellStart, numGridCells, kEmpty, (const unsigned *)nullptr);
    if (numBodies > 0)
        hipLaunchKernelGGL(k_cell_range, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0,
                           (const uint2 *)particleHash, (unsigned)numBodies, cellStart,
                           (unsigned *)nullptr, (const unsigned *)nullptr);
    check_launch("findCellStart");
}

void neighborarray(double *pos, uint *particleHash, uint *cellStart, uint gridSize[3],
                   double cellSize[3], double worldOrigin[3], int maxParticlesPerCell,
                   int numBodies, uint *Anei, uint *njgi, uint *njli)
{
    if (numBodies <= 0 || maxParticlesPerCell <= 0) return;
    (void)cellStart;   // the cell ranges follow from the sorted hash itself (c_cell_count)
    const unsigned long long cells = (unsigned long long)gridSize[0] * gridSize[1] * gridSize[2];
    if (cells == 0 || cells >= 0xffffffffull) { g_last_error = DEM_ERR_INVALID; return; }
    if (compat_words() || g_nbr.ensure((size_t)numBodies, (size_t)cells)) {
        g_last_error = DEM_ERR_HIP;
        return;
    }
    const unsigned n = (unsigned)numBodies;
    InstDev I{};
    for (int k = 0; k < 3; k++) I.gsize[k] = gridSize[k];
    I.cell_off = 0;
    I.skin = kCub - kCua;
    DEM_HIP_V(hipMemsetAsync(g_words + 1, 0, sizeof(unsigned), 0));
    hipLaunchKernelGGL(c_sorted_copy, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, 0, (const double *)pos,
                       (const uint2 *)particleHash, n, g_nbr.spos, g_nbr.sid);
    hipLaunchKernelGGL(k_fill, dim3(std::min(cdiv(cells, kBlock), 4096u)), dim3(kBlock), 0, 0,
                       g_nbr.cnt, (unsigned)cells, 0u, (const unsigned *)nullptr);
    hipLaunchKernelGGL(c_cell_count, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, 0,
                       (const uint2 *)particleHash, n, (unsigned)cells, g_nbr.cnt);
    if (scan_excl(g_S, g_nbr.cnt, (unsigned)cells, g_nbr.pre, g_words + 3, 0, nullptr)) {
        g_last_error = DEM_ERR_HIP;
        return;
    }
    hipLaunchKernelGGL(c_neighbor_search, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, 0,
                       (const double4 *)g_nbr.spos, (const unsigned *)g_nbr.sid,
                       (const uint2 *)particleHash, (const unsigned *)g_nbr.pre, n, I,
                       maxParticlesPerCell, Anei, njgi, njli, g_words + 1);
    check_launch("neighborarray");
    unsigned ov = 0;
    DEM_HIP_V(hipMemcpy(&ov, g_words + 1, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (ov) {
        std::fprintf(stderr, "neighborarray: a particle has more than %d candidates\n", maxParticlesPerCell);
        g_last_error = DEM_ERR_CAPACITY;
    }
}

void prefixsum(uint numBodies, uint *njgi, uint *sjgi, uint *oldsjgi, uint *totalcontacts)
{
    if (numBodies == 0) { if (totalcontacts) *totalcontacts = 0; return; }
    DEM_HIP_V(hipMemcpyAsync(oldsjgi, sjgi, numBodies * sizeof(uint), hipMemcpyDeviceToDevice, 0));
    if (scan_excl(g_S, njgi, numBodies, sjgi, nullptr, 0, nullptr)) { g_last_error = DEM_ERR_HIP; return; }
    hipLaunchKernelGGL(c_add_arrays, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0, sjgi,
                       (const unsigned *)njgi, numBodies);
    check_launch("prefixsum");
    DEM_HIP_V(hipMemcpy(totalcontacts, sjgi + numBodies - 1, sizeof(uint), hipMemcpyDeviceToHost));
}

void pairlist(uint *sjgi, uint *oldsjgi, uint *oldLn, uint *Anei, uint *njgi, uint *njli,
              double *olddispt, double *oldfricp, int maxParticlesPerCell, int maxCnPerParticle,
              uint numBodies, uint *Aref, uint *Lp, uint *Ln, double *dispt, double *fricp)
{
    if (numBodies == 0) return;
    size_t slots = (size_t)numBodies * (size_t)maxCnPerParticle;
    // sentinel fills of particleSystem.cu:166-170 (Lp/Ln 0xff; dispt/fricp zero)
    hipLaunchKernelGGL(k_fill, dim3(std::min(cdiv(slots, kBlock), 4096u)), dim3(kBlock), 0, 0, Lp,
                       (unsigned)slots, kEmpty, (const unsigned *)nullptr);
    hipLaunchKernelGGL(k_fill, dim3(std::min(cdiv(slots, kBlock), 4096u)), dim3(kBlock), 0, 0, Ln,
                       (unsigned)slots, kEmpty, (const unsigned *)nullptr);
    DEM_HIP_V(hipMemsetAsync(dispt, 0, slots * 3 * sizeof(double), 0));
    DEM_HIP_V(hipMemsetAsync(fricp, 0, slots * sizeof(double), 0));
    hipLaunchKernelGGL(c_pair_list, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0,
                       (const unsigned *)sjgi, (const unsigned *)oldsjgi, (const unsigned *)oldLn,
                       (const unsigned *)Anei, (const unsigned *)njgi, (const unsigned *)njli,
                       (const double *)olddispt, (const double *)oldfricp, maxParticlesPerCell,
                       numBodies, Aref, Lp, Ln, dispt, fricp);
    hipLaunchKernelGGL(c_pair_list_partner, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0,
                       (const unsigned *)sjgi, (const unsigned *)njgi, (const unsigned *)njli,
                       maxParticlesPerCell, numBodies, (const unsigned *)Ln, Aref);
    check_launch("pairlist");
}

void calculateforcepair(double *pos, double *pdis, double *angv, double *rad, double *rmass,
                        uint *Lp, uint *Ln, uint *oldLn, double *dispt, double *fricp,
                        double *olddispt, double *oldfricp, double dt, double *forcepair,
                        double *torqpairi, double *torqpairj, uint *contactpair, uint totalcontacts)
{
    if (totalcontacts == 0) return;   // the reference launches zero threads (no-op)
    hipLaunchKernelGGL(c_force_pair, dim3(cdiv(totalcontacts, kBlock)), dim3(kBlock), 0, 0,
                       totalcontacts, (const double *)pos, (const double *)pdis,
                       (const double *)angv, (const double *)rad, (const double *)rmass,
                       (const unsigned *)Lp, (const unsigned *)Ln, oldLn, dispt, fricp, olddispt,
                       oldfricp, dt, law_from(g_par), forcepair, torqpairi, torqpairj, contactpair);
    check_launch("calculateforcepair");
}

void sumallforcei(int maxParticlesPerCell, int numBodies, uint *njgi, uint *njli, double *forcepair,
                  double *torqpairi, double *torqpairj, uint *contactpair, uint *Aref,
                  double *forcei, double *torquei, uint *contacti)
{
    if (numBodies <= 0) return;
    hipLaunchKernelGGL(c_sum_forces, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0, numBodies,
                       maxParticlesPerCell, (const unsigned *)njgi, (const unsigned *)njli,
                       (const double *)forcepair, (const double *)torqpairi,
                       (const double *)torqpairj, (const unsigned *)contactpair,
                       (const unsigned *)Aref, forcei, torquei, contacti);
    check_launch("sumallforcei");
}

void calculateforceiw(double worldSize[3], double *pos, double *pdis, double *angv, double *rad,
                      double *rmass, double dt, uint totalcontacts, int numBodies, double *disptw,
                      double *fricpw, double *forcei, double *torquei)
{
    if (numBodies <= 0) return;
    WallGeom G;
    if (g_walls_set) {
        G.kind = g_walls.kind;
        G.hx = 0.5 * g_walls.size[0];
        G.hy = 0.5 * g_walls.size[1];
        G.radius = g_walls.radius;
    } else {
        G.kind = DEM_WALLS_BOX;
        G.hx = 0.5 * worldSize[0];
        G.hy = 0.5 * worldSize[1];
        G.radius = 0.0;
    }
    hipLaunchKernelGGL(c_wall_gravity, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0, numBodies,
                       G, (const double *)pos, (const double *)pdis, (const double *)angv,
                       (const double *)rad, (const double *)rmass, dt, totalcontacts,
                       law_from(g_parw), disptw, fricpw, forcei, torquei);
    check_launch("calculateforceiw");
}

void updateposition(double worldSize[3], int numBodies, double *pos, double *pdis, double *angv,
                    double *rmass, double *inert, double *forcei, double *torquei, double *qi,
                    double dt, double cdamp, uint *iall, double *maxhz)
{
    (void)worldSize;
    (void)cdamp;   // the reference reads cdamp but integrates with adamp = bdamp = 1 (neiglist.cu:354-355)
    if (numBodies <= 0) { if (iall) *iall = 0; if (maxhz) *maxhz = 0.0; return; }
    if (compat_words()) return;
    DEM_HIP_V(hipMemsetAsync(g_words, 0, 8 * sizeof(unsigned), 0));
    double ann = 0.5 * (kCub - kCua);
    hipLaunchKernelGGL(c_update, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0, numBodies, pos,
                       pdis, angv, (const double *)rmass, (const double *)inert,
                       (const double *)forcei, (const double *)torquei, qi, dt, ann * ann,
                       g_words, reinterpret_cast<unsigned long long *>(g_words + 4));
    check_launch("updateposition");
    unsigned w[8];
    DEM_HIP_V(hipMemcpy(w, g_words, sizeof(w), hipMemcpyDeviceToHost));
    if (iall) *iall = w[0];
    if (maxhz) std::memcpy(maxhz, &w[4], sizeof(double));
}

void outputporositycn(uint numBodies, uint *contacti, double maxhz, double worldSize[3], uint *pcntot)
{
    (void)maxhz;
    (void)worldSize;
    if (numBodies == 0) { if (pcntot) *pcntot = 0; return; }
    if (compat_words()) return;
    DEM_HIP_V(hipMemsetAsync(g_words + 2, 0, sizeof(unsigned), 0));
    hipLaunchKernelGGL(c_contact_total, dim3(cdiv(numBodies, kBlock)), dim3(kBlock), 0, 0, numBodies,
                       (const unsigned *)contacti, g_words + 2);
    check_launch("outputporositycn");
    DEM_HIP_V(hipMemcpy(pcntot, g_words + 2, sizeof(uint), hipMemcpyDeviceToHost));
}

}  // extern "C"

// =========================================================================================
// Adjoint form: checkpointed forward, segment recompute with tapes, reverse sweep
// =========================================================================================

namespace {

#define ADJ_FREE(p)                     \
    do {                                \
        if (p) (void)hipFree((void *)p); \
        p = nullptr;                    \
    } while (0)

// Full dynamic state of a context at a segment boundary (the list included, so the
// recompute replays the forward bit for bit).
struct AdjSnap {
    PRec *rec = nullptr;
    Hist *whist = nullptr, *hist = nullptr;
    unsigned *inv = nullptr, *off = nullptr, *toff = nullptr, *jt = nullptr, *wflag = nullptr,
             *wlist = nullptr;
    uint2 *ab = nullptr;
    DevCtl *ctl = nullptr;
    size_t pcap = 0;
    unsigned npairs = 0, n_list = 0, since = 0, last_flag = 0;
    bool have_list = false, need_rebuild = false, valid = false;
    unsigned long long steps = 0, rebuilds = 0;
};

// A pair-list version made by a rebuild inside a recomputed segment, with that rebuild's
// slot permutation and history-source map.
struct AdjList {
    uint2 *ab = nullptr;
    unsigned *off = nullptr, *toff = nullptr, *jt = nullptr, *wflag = nullptr, *perm = nullptr,
             *src = nullptr;
    size_t pcap = 0;
    unsigned np = 0;
};

struct ListView {
    const uint2 *ab;
    const unsigned *off, *toff, *jt, *wflag, *perm, *src;
    unsigned np;
};

}  // namespace

struct AdjData {
    int nsteps = 0, K = 0, nseg = 0;
    unsigned n = 0, max_np = 0;
    std::vector<AdjSnap> ck;          // nseg + 1 boundaries
    // Tapes and list versions of a segment, two sets: the recompute of segment s writes set
    // s mod 2 while the reverse of segment s + 1 reads the other set on the reverse stream.
    PRec *t_rec[2] = {nullptr, nullptr};     // K records per slot
    Hist *t_hist[2] = {nullptr, nullptr};    // K x max_np incoming pair histories
    Hist *t_whist[2] = {nullptr, nullptr};   // K incoming wall histories per slot
    std::vector<AdjList> pool[2];            // list versions built during the segment
    int tape_sets = 0;
    hipStream_t rs = nullptr;                // reverse stream (default priority)
    hipEvent_t rec_done[2] = {nullptr, nullptr}, rev_done[2] = {nullptr, nullptr};
    bool rev_pending[2] = {false, false};
    SAdj *lam[2] = {nullptr, nullptr};
    double *hb[2] = {nullptr, nullptr};   // 4 doubles per pair
    double *ftb = nullptr, *thacc = nullptr;
    PAdj *pa = nullptr;
    double *pobs = nullptr, *sobs = nullptr, *part = nullptr, *sums = nullptr, *coef = nullptr;
    size_t pobs_cap = 0;
    unsigned *range = nullptr;
    ObsDev obs{};
    std::vector<double> vreg, wsum, csum;
    int cur = 0, tape_K = 0;
    bool have_grad = false;
    float fwd_ms = 0.f, bwd_ms = 0.f;
};

namespace {

void snap_free(AdjSnap &S)
{
    ADJ_FREE(S.rec); ADJ_FREE(S.whist); ADJ_FREE(S.hist); ADJ_FREE(S.inv); ADJ_FREE(S.off);
    ADJ_FREE(S.toff); ADJ_FREE(S.jt); ADJ_FREE(S.wflag); ADJ_FREE(S.wlist); ADJ_FREE(S.ab);
    ADJ_FREE(S.ctl);
    S.pcap = 0;
    S.valid = false;
}

void list_free(AdjList &L)
{
    ADJ_FREE(L.ab); ADJ_FREE(L.off); ADJ_FREE(L.toff); ADJ_FREE(L.jt); ADJ_FREE(L.wflag);
    ADJ_FREE(L.perm); ADJ_FREE(L.src);
    L.pcap = 0;
}

void adj_release(AdjData *A)
{
    if (!A) return;
    for (AdjSnap &S : A->ck) snap_free(S);
    if (A->rs) (void)hipStreamSynchronize(A->rs);
    for (int k = 0; k < 2; k++) {
        for (AdjList &L : A->pool[k]) list_free(L);
        ADJ_FREE(A->t_rec[k]); ADJ_FREE(A->t_hist[k]); ADJ_FREE(A->t_whist[k]);
        if (A->rec_done[k]) (void)hipEventDestroy(A->rec_done[k]);
        if (A->rev_done[k]) (void)hipEventDestroy(A->rev_done[k]);
    }
    if (A->rs) (void)hipStreamDestroy(A->rs);
    ADJ_FREE(A->lam[0]); ADJ_FREE(A->lam[1]); ADJ_FREE(A->hb[0]); ADJ_FREE(A->hb[1]);
    ADJ_FREE(A->ftb); ADJ_FREE(A->thacc); ADJ_FREE(A->pa);
    ADJ_FREE(A->pobs); ADJ_FREE(A->sobs); ADJ_FREE(A->part); ADJ_FREE(A->sums); ADJ_FREE(A->coef);
    ADJ_FREE(A->range);
    delete A;
}

void adj_free(dem_ctx *c)
{
    if (c && c->adj) {
        (void)hipStreamSynchronize(c->stream);
        adj_release(c->adj);
        c->adj = nullptr;
    }
}

int snap_save(dem_ctx *c, AdjSnap &S)
{
    hipStream_t st = c->stream;
    size_t n = c->n, np = c->have_list ? c->npairs : 0;
    if (!S.rec) {
        DEM_HIP(hipMalloc(&S.rec, sizeof(PRec) * n));
        DEM_HIP(hipMalloc(&S.whist, sizeof(Hist) * n));
        DEM_HIP(hipMalloc(&S.inv, sizeof(unsigned) * n));
        DEM_HIP(hipMalloc(&S.off, sizeof(unsigned) * (n + 1)));
        DEM_HIP(hipMalloc(&S.toff, sizeof(unsigned) * (n + 1)));
        DEM_HIP(hipMalloc(&S.wflag, sizeof(unsigned) * n));
        DEM_HIP(hipMalloc(&S.wlist, sizeof(unsigned) * n));
        DEM_HIP(hipMalloc(&S.ctl, sizeof(DevCtl)));
    }
    if (S.pcap < std::max<size_t>(np, 1)) {
        ADJ_FREE(S.ab); ADJ_FREE(S.hist); ADJ_FREE(S.jt);
        S.pcap = std::max<size_t>(np, 1);
        DEM_HIP(hipMalloc(&S.ab, sizeof(uint2) * S.pcap));
        DEM_HIP(hipMalloc(&S.hist, sizeof(Hist) * S.pcap));
        DEM_HIP(hipMalloc(&S.jt, sizeof(unsigned) * S.pcap));
    }
    DEM_HIP(hipMemcpyAsync(S.rec, c->rec[c->rc], sizeof(PRec) * n, hipMemcpyDeviceToDevice, st));
    DEM_HIP(hipMemcpyAsync(S.whist, c->whist[c->pc], sizeof(Hist) * n, hipMemcpyDeviceToDevice, st));
    DEM_HIP(hipMemcpyAsync(S.inv, c->inv[c->pc], sizeof(unsigned) * n, hipMemcpyDeviceToDevice, st));
    DEM_HIP(hipMemcpyAsync(S.ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToDevice, st));
    if (c->have_list) {
        DEM_HIP(hipMemcpyAsync(S.off, c->off[c->lc], sizeof(unsigned) * (n + 1), hipMemcpyDeviceToDevice, st));
        DEM_HIP(hipMemcpyAsync(S.toff, c->toff, sizeof(unsigned) * (n + 1), hipMemcpyDeviceToDevice, st));
        DEM_HIP(hipMemcpyAsync(S.wflag, c->wflag, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, st));
        DEM_HIP(hipMemcpyAsync(S.wlist, c->wlist, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, st));
        if (np) {
            DEM_HIP(hipMemcpyAsync(S.ab, c->ab[c->lc], sizeof(uint2) * np, hipMemcpyDeviceToDevice, st));
            DEM_HIP(hipMemcpyAsync(S.hist, c->hist[c->hc], sizeof(Hist) * np, hipMemcpyDeviceToDevice, st));
            DEM_HIP(hipMemcpyAsync(S.jt, c->jt, sizeof(unsigned) * np, hipMemcpyDeviceToDevice, st));
        }
    }
    S.npairs = (unsigned)np;
    S.n_list = c->n_list;
    S.since = c->since;
    S.last_flag = c->last_flag;
    S.have_list = c->have_list;
    S.need_rebuild = c->need_rebuild;
    S.steps = c->steps;
    S.rebuilds = c->rebuilds;
    S.valid = true;
    return DEM_OK;
}

int snap_restore(dem_ctx *c, const AdjSnap &S)
{
    StateSrc V;
    V.rec = S.rec;
    V.whist = S.whist;
    V.inv = S.inv;
    V.ctl = S.ctl;
    V.nrec = c->n;
    V.off = S.off;
    V.toff = S.toff;
    V.wflag = S.wflag;
    V.wlist = S.wlist;
    V.ab = S.ab;
    V.hist = S.hist;
    V.jt = S.jt;
    V.npairs = S.npairs;
    V.n_list = S.n_list;
    V.since = S.since;
    V.last_flag = S.last_flag;
    V.have_list = S.have_list;
    V.need_rebuild = S.need_rebuild;
    V.steps = S.steps;
    V.rebuilds = S.rebuilds;
    return restore_state(c, V, hipMemcpyDeviceToDevice);
}

int list_reserve(AdjList &L, size_t n, size_t np)
{
    if (!L.off) {
        DEM_HIP(hipMalloc(&L.off, sizeof(unsigned) * (n + 1)));
        DEM_HIP(hipMalloc(&L.toff, sizeof(unsigned) * (n + 1)));
        DEM_HIP(hipMalloc(&L.wflag, sizeof(unsigned) * n));
        DEM_HIP(hipMalloc(&L.perm, sizeof(unsigned) * n));
    }
    if (L.pcap < std::max<size_t>(np, 1)) {
        ADJ_FREE(L.ab); ADJ_FREE(L.jt); ADJ_FREE(L.src);
        L.pcap = std::max<size_t>(np, 1);
        DEM_HIP(hipMalloc(&L.ab, sizeof(uint2) * L.pcap));
        DEM_HIP(hipMalloc(&L.jt, sizeof(unsigned) * L.pcap));
        DEM_HIP(hipMalloc(&L.src, sizeof(unsigned) * L.pcap));
    }
    return DEM_OK;
}

struct HookArg {
    AdjList *L;
    size_t n;
};

int record_hook(dem_ctx *c, void *arg, unsigned np)
{
    HookArg *h = static_cast<HookArg *>(arg);
    int rc = list_reserve(*h->L, h->n, np);
    if (rc) return rc;
    c->tape_src = h->L->src;
    return DEM_OK;
}

// One step of dem_step's schedule.  With `L` non-null a rebuild records its permutation,
// history sources and resulting list into L (*rebuilt set); with `tape_step` >= 0 the force
// step writes the tapes of that step.
int adj_step(dem_ctx *c, AdjData *A, int set, int tape_step, AdjList *L, bool *rebuilt)
{
    int rc;
    if (rebuilt) *rebuilt = false;
    if (c->need_rebuild || (c->interval > 0 && c->since >= (unsigned)c->interval)) {
        if (L) {
            if ((rc = list_reserve(*L, c->n, 1))) return rc;
            c->tape_perm = L->perm;
            HookArg h{L, c->n};
            rc = rebuild_host(c, record_hook, &h);
            c->tape_perm = nullptr;
            c->tape_src = nullptr;
            if (rc) return rc;
            hipStream_t st = c->stream;
            size_t n = c->n, np = c->npairs;
            DEM_HIP(hipMemcpyAsync(L->off, c->off[c->lc], sizeof(unsigned) * (n + 1), hipMemcpyDeviceToDevice, st));
            DEM_HIP(hipMemcpyAsync(L->toff, c->toff, sizeof(unsigned) * (n + 1), hipMemcpyDeviceToDevice, st));
            DEM_HIP(hipMemcpyAsync(L->wflag, c->wflag, sizeof(unsigned) * n, hipMemcpyDeviceToDevice, st));
            if (np) {
                DEM_HIP(hipMemcpyAsync(L->ab, c->ab[c->lc], sizeof(uint2) * np, hipMemcpyDeviceToDevice, st));
                DEM_HIP(hipMemcpyAsync(L->jt, c->jt, sizeof(unsigned) * np, hipMemcpyDeviceToDevice, st));
            }
            L->np = (unsigned)np;
            if (rebuilt) *rebuilt = true;
        } else if ((rc = rebuild_host(c))) {
            return rc;
        }
    }
    if (A && c->npairs > A->max_np && tape_step >= 0) return DEM_ERR_CAPACITY;
    if (tape_step >= 0) {
        c->tape_rec = A->t_rec[set] + (size_t)tape_step * c->n;
        c->tape_hist = A->t_hist[set] + (size_t)tape_step * A->max_np;
        c->tape_whist = A->t_whist[set] + (size_t)tape_step * c->n;
    }
    launch_force_step(c, pair_grid(c, c->npairs));
    c->tape_rec = nullptr;
    c->tape_hist = nullptr;
    c->tape_whist = nullptr;
    DEM_HIP(hipGetLastError());
    c->steps++;
    c->since++;
    if (c->interval == 0) {
        if ((rc = sync_mirror(c))) return rc;
        c->need_rebuild = c->last_flag != 0u;
    }
    return DEM_OK;
}

int obs_alloc(dem_ctx *c, AdjData *A)
{
    size_t n = c->n;
    if (!A->sobs) {
        DEM_HIP(hipMalloc(&A->sobs, sizeof(double) * 8 * n));
        DEM_HIP(hipMalloc(&A->part, sizeof(double) * 3 * n));
        DEM_HIP(hipMalloc(&A->sums, sizeof(double) * 3 * c->ninst));
        DEM_HIP(hipMalloc(&A->coef, sizeof(double) * 3 * c->ninst));
        DEM_HIP(hipMalloc(&A->range, sizeof(unsigned) * (c->ninst + 1)));
        std::vector<unsigned> r(c->ninst + 1);
        for (int k = 0; k < c->ninst; k++) r[k] = c->hinst[k].p_off;
        r[c->ninst] = (unsigned)n;
        DEM_HIP(hipMemcpy(A->range, r.data(), sizeof(unsigned) * r.size(), hipMemcpyHostToDevice));
    }
    size_t np = std::max<size_t>(c->npairs, 1);
    if (A->pobs_cap < np) {
        ADJ_FREE(A->pobs);
        A->pobs_cap = np * 3 / 2 + 1024;
        DEM_HIP(hipMalloc(&A->pobs, sizeof(double) * 4 * A->pobs_cap));
    }
    return DEM_OK;
}

double region_volume(const InstDev &I, const ObsDev &o)
{
    double h = std::max(o.zhi - o.zlo, 0.0);
    if (I.walls.kind == 0)
        return std::max(2.0 * (I.walls.hx - o.margin), 0.0) * std::max(2.0 * (I.walls.hy - o.margin), 0.0) * h;
    double r = std::max(I.walls.radius - o.margin, 0.0);
    return 3.14159265358979323846 * r * r * h;
}

// Smooth observables of the current state (per instance): fills A's per-slot terms and the
// sums; phi/cn may be null.
int observe(dem_ctx *c, AdjData *A, const ObsDev &o, double *phi, double *cn)
{
    int rc;
    if (!c->have_list && (rc = rebuild_host(c))) return rc;
    if ((rc = obs_alloc(c, A))) return rc;
    hipStream_t st = c->stream;
    unsigned n = c->n, np = c->npairs;
    InstDev one = c->hinst[0];
    const PRec *rec = c->rec[c->rc];
    unsigned pg = std::max(1u, std::min(cdiv(np, kBlock), (unsigned)c->cus * 16u));
    if (c->batched)
        hipLaunchKernelGGL(k_obs_pair<true>, dim3(pg), dim3(kBlock), 0, st, (const uint2 *)c->ab[c->lc],
                           np, rec, (const InstDev *)c->dinst, one, o, A->pobs);
    else
        hipLaunchKernelGGL(k_obs_pair<false>, dim3(pg), dim3(kBlock), 0, st, (const uint2 *)c->ab[c->lc],
                           np, rec, (const InstDev *)c->dinst, one, o, A->pobs);
    if (c->batched)
        hipLaunchKernelGGL(k_obs_slot<true>, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, rec,
                           (const unsigned *)c->off[c->lc], (const unsigned *)c->toff,
                           (const unsigned *)c->jt, np, (const double *)A->pobs,
                           (const InstDev *)c->dinst, one, o, n, A->sobs, A->part);
    else
        hipLaunchKernelGGL(k_obs_slot<false>, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, st, rec,
                           (const unsigned *)c->off[c->lc], (const unsigned *)c->toff,
                           (const unsigned *)c->jt, np, (const double *)A->pobs,
                           (const InstDev *)c->dinst, one, o, n, A->sobs, A->part);
    hipLaunchKernelGGL(k_obs_reduce, dim3(c->ninst), dim3(kBlock), 0, st, (const double *)A->part,
                       (const unsigned *)A->range, A->sums);
    DEM_HIP(hipGetLastError());
    std::vector<double> sums(3 * (size_t)c->ninst);
    DEM_HIP(hipMemcpyAsync(sums.data(), A->sums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, st));
    DEM_HIP(hipStreamSynchronize(st));
    A->obs = o;
    A->vreg.assign(c->ninst, 0.0);
    A->wsum.assign(c->ninst, 0.0);
    A->csum.assign(c->ninst, 0.0);
    for (int k = 0; k < c->ninst; k++) {
        double v = region_volume(c->hinst[k], o);
        A->vreg[k] = v;
        A->wsum[k] = sums[3 * k + 1];
        A->csum[k] = sums[3 * k + 2];
        if (phi) phi[k] = v > 0.0 ? sums[3 * k] / v : 0.0;
        if (cn) cn[k] = sums[3 * k + 1] > 0.0 ? sums[3 * k + 2] / sums[3 * k + 1] : 0.0;
    }
    return DEM_OK;
}

bool obs_valid(const dem_obs *o)
{
    return o && o->eps > 0.0 && o->eps_c > 0.0 && o->zhi > o->zlo && o->margin >= 0.0;
}

ObsDev obs_dev(const dem_obs *o)
{
    return ObsDev{o->margin, o->zlo, o->zhi, o->eps, o->eps_c};
}

AdjData *adj_get(dem_ctx *c)
{
    if (!c->adj) c->adj = new AdjData();
    return c->adj;
}

// Reverse sweep over the recomputed segment [ck[seg], ck[seg+1]): `ks` steps.
int reverse_segment(dem_ctx *c, AdjData *A, int set, int ks, const std::vector<ListView> &views,
                    const std::vector<int> &step_ver, const std::vector<char> &step_reb, int &hc)
{
    hipStream_t st = A->rs;
    unsigned n = c->n;
    InstDev one = c->hinst[0];
    unsigned gn = cdiv(n, kBlock);
    for (int j = ks - 1; j >= 0; j--) {
        const ListView &V = views[step_ver[j]];
        const PRec *tr = A->t_rec[set] + (size_t)j * n;
        const Hist *th = A->t_hist[set] + (size_t)j * A->max_np;
        const Hist *twh = A->t_whist[set] + (size_t)j * n;
        if (c->batched)
            hipLaunchKernelGGL(k_adj_particle<true>, dim3(gn), dim3(kBlock), 0, st, tr, twh, V.wflag,
                               A->lam[A->cur], A->ftb, A->thacc, n, (const InstDev *)c->dinst, one,
                               c->adj_roll_k);
        else
            hipLaunchKernelGGL(k_adj_particle<false>, dim3(gn), dim3(kBlock), 0, st, tr, twh, V.wflag,
                               A->lam[A->cur], A->ftb, A->thacc, n, (const InstDev *)c->dinst, one,
                               c->adj_roll_k);
        unsigned pg = std::max(1u, std::min(cdiv(V.np, kBlock), (unsigned)c->cus * 16u));
        if (c->batched)
            hipLaunchKernelGGL(k_adj_pair<true>, dim3(pg), dim3(kBlock), 0, st, V.ab, V.np, tr, th,
                               (const double *)A->ftb, A->hb[hc], A->pa, (const InstDev *)c->dinst,
                               one.pp, one.dt, one.roll_k, c->adj_roll_k);
        else
            hipLaunchKernelGGL(k_adj_pair<false>, dim3(pg), dim3(kBlock), 0, st, V.ab, V.np, tr, th,
                               (const double *)A->ftb, A->hb[hc], A->pa, (const InstDev *)c->dinst,
                               one.pp, one.dt, one.roll_k, c->adj_roll_k);
        hipLaunchKernelGGL(k_adj_gather, dim3(gn), dim3(kBlock), 0, st, A->lam[A->cur],
                           (const PAdj *)A->pa, V.off, V.toff, V.jt, V.np, tr, A->thacc, n);
        if (step_reb[j]) {
            const ListView &P = views[step_ver[j] - 1];
            hipLaunchKernelGGL(k_adj_unperm, dim3(gn), dim3(kBlock), 0, st,
                               (const SAdj *)A->lam[A->cur], A->lam[1 - A->cur], V.perm, n);
            A->cur = 1 - A->cur;
            size_t zn = 4 * (size_t)std::max(P.np, 1u);
            hipLaunchKernelGGL(k_zero_d, dim3(std::max(1u, std::min(cdiv(zn, kBlock), 4096u))), dim3(kBlock),
                               0, st, A->hb[1 - hc], zn);
            hipLaunchKernelGGL(k_adj_unremap, dim3(std::max(1u, cdiv(V.np, kBlock))), dim3(kBlock), 0, st,
                               (const double *)A->hb[hc], A->hb[1 - hc], V.src, V.np, P.np);
            hc = 1 - hc;
        }
        DEM_HIP(hipGetLastError());
    }
    return DEM_OK;
}

}  // namespace

extern "C" {

int dem_set_material(dem_ctx *c, int instance, const dem_material *mat)
{
    if (!c || !mat || instance < 0 || instance >= c->ninst) return DEM_ERR_INVALID;
    DEM_HIP(hipStreamSynchronize(c->stream));
    InstDev &I = c->hinst[instance];
    I.pp = law_from(mat->p);
    I.pw = law_from(mat->w);
    DEM_HIP(hipMemcpy(c->dinst + instance, &I, sizeof(InstDev), hipMemcpyHostToDevice));
    drop_graph(c);
    return DEM_OK;
}

int dem_observe(dem_ctx *c, const dem_obs *obs, double *phi, double *cn)
{
    if (!c || !obs_valid(obs) || c->n != c->cap_n) return DEM_ERR_INVALID;
    return observe(c, adj_get(c), obs_dev(obs), phi, cn);
}

int dem_adj_forward(dem_ctx *c, int nsteps, int every, const dem_obs *obs, double *phi, double *cn)
{
    if (!c || nsteps < 1 || every < 1 || !obs_valid(obs) || c->n != c->cap_n) return DEM_ERR_INVALID;
    // the reverse sweep differentiates the FP64 law; the float-length option is not supported
    for (const InstDev &I : c->hinst)
        if (I.flen) return DEM_ERR_INVALID;
    int rc;
    DEM_HIP(hipStreamSynchronize(c->stream));
    drop_graph(c);
    AdjData *A = adj_get(c);
    for (AdjSnap &S : A->ck) snap_free(S);
    A->ck.clear();
    A->have_grad = false;
    A->nsteps = nsteps;
    A->K = every;
    A->nseg = (nsteps + every - 1) / every;
    A->n = c->n;
    A->ck.resize(A->nseg + 1);
    c->np_peak = c->have_list ? c->npairs : 0;
    // under the displacement trigger the segments run as queued steps (dem_step's schedule
    // and results, without a host wait per step)
    bool queued = c->interval == 0 && !c->timing && c->pipeline > 1;
    DEM_HIP(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < A->nseg; s++) {
        if ((rc = snap_save(c, A->ck[s]))) return rc;
        int ks = std::min(every, nsteps - s * every);
        if (queued) {
            if ((rc = step_queued(c, ks))) return rc;
        } else {
            for (int j = 0; j < ks; j++)
                if ((rc = adj_step(c, nullptr, 0, -1, nullptr, nullptr))) return rc;
        }
    }
    if ((rc = snap_save(c, A->ck[A->nseg]))) return rc;
    A->max_np = c->np_peak;
    DEM_HIP(hipEventRecord(c->ev1, c->stream));
    DEM_HIP(hipEventSynchronize(c->ev1));
    DEM_HIP(hipEventElapsedTime(&A->fwd_ms, c->ev0, c->ev1));
    c->last_ms = A->fwd_ms;
    return observe(c, A, obs_dev(obs), phi, cn);
}

int dem_adj_backward(dem_ctx *c, const double *gphi, const double *gcn, double *theta_bar)
{
    if (!c || !c->adj || c->adj->ck.empty() || !gphi || !gcn) return DEM_ERR_INVALID;
    AdjData *A = c->adj;
    if (c->n != A->n) return DEM_ERR_INVALID;
    int rc;
    hipStream_t st = c->stream;
    size_t n = c->n, K = (size_t)A->K, mnp = std::max<size_t>(A->max_np, 1);
    int sets = A->nseg > 1 ? 2 : 1;
    if (A->tape_K != A->K || A->tape_sets < sets) {
        for (int k = 0; k < 2; k++) {
            ADJ_FREE(A->t_rec[k]); ADJ_FREE(A->t_whist[k]);
        }
        for (int k = 0; k < sets; k++) {
            DEM_HIP(hipMalloc(&A->t_rec[k], sizeof(PRec) * K * n));
            DEM_HIP(hipMalloc(&A->t_whist[k], sizeof(Hist) * K * n));
        }
        A->tape_K = A->K;
        A->tape_sets = sets;
    }
    if (!A->rs) {
        DEM_HIP(hipStreamCreateWithFlags(&A->rs, hipStreamNonBlocking));
        for (int k = 0; k < 2; k++) {
            DEM_HIP(hipEventCreateWithFlags(&A->rec_done[k], hipEventDisableTiming));
            DEM_HIP(hipEventCreateWithFlags(&A->rev_done[k], hipEventDisableTiming));
        }
    }
    if (!A->lam[0]) {
        DEM_HIP(hipMalloc(&A->lam[0], sizeof(SAdj) * n));
        DEM_HIP(hipMalloc(&A->lam[1], sizeof(SAdj) * n));
        DEM_HIP(hipMalloc(&A->ftb, sizeof(double) * 6 * n));
        DEM_HIP(hipMalloc(&A->thacc, sizeof(double) * kThetaPer * n));
    }
    // pair-sized buffers follow the largest list of the forward
    ADJ_FREE(A->t_hist[0]); ADJ_FREE(A->t_hist[1]);
    ADJ_FREE(A->hb[0]); ADJ_FREE(A->hb[1]); ADJ_FREE(A->pa);
    for (int k = 0; k < sets; k++) DEM_HIP(hipMalloc(&A->t_hist[k], sizeof(Hist) * K * mnp));
    DEM_HIP(hipMalloc(&A->hb[0], sizeof(double) * 4 * mnp));
    DEM_HIP(hipMalloc(&A->hb[1], sizeof(double) * 4 * mnp));
    DEM_HIP(hipMalloc(&A->pa, sizeof(PAdj) * mnp));

    DEM_HIP(hipEventRecord(c->ev0, st));
    // seed at the final state
    if ((rc = snap_restore(c, A->ck[A->nseg]))) return rc;
    if ((rc = observe(c, A, A->obs, nullptr, nullptr))) return rc;
    std::vector<double> coef(3 * (size_t)c->ninst);
    for (int k = 0; k < c->ninst; k++) {
        double W = A->wsum[k];
        coef[3 * k] = A->vreg[k] > 0.0 ? gphi[k] / A->vreg[k] : 0.0;
        coef[3 * k + 1] = W > 0.0 ? gcn[k] / W : 0.0;
        coef[3 * k + 2] = W > 0.0 ? gcn[k] * A->csum[k] / (W * W) : 0.0;
    }
    DEM_HIP(hipMemcpyAsync(A->coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, st));
    unsigned gn = cdiv(n, kBlock);
    A->cur = 0;
    hipLaunchKernelGGL(k_obs_seed, dim3(gn), dim3(kBlock), 0, st, (const PRec *)c->rec[c->rc],
                       (const double *)A->sobs, (const double *)A->coef, A->lam[0], (unsigned)n);
    int hc = 0;
    size_t z = 4 * mnp;
    hipLaunchKernelGGL(k_zero_d, dim3(std::max(1u, std::min(cdiv(z, kBlock), 4096u))), dim3(kBlock), 0, st,
                       A->hb[0], z);
    z = (size_t)kThetaPer * n;
    hipLaunchKernelGGL(k_zero_d, dim3(std::max(1u, std::min(cdiv(z, kBlock), 4096u))), dim3(kBlock), 0, st,
                       A->thacc, z);
    DEM_HIP(hipGetLastError());

    // Segments from the last: restore the checkpoint and recompute the segment with tapes on
    // the context stream (the host follows the rebuild trigger step by step), then queue its
    // reverse sweep on the reverse stream.  The reverse of segment s overlaps the recompute
    // of segment s - 1, which writes the other tape set; before a set is rewritten the host
    // waits for the reverse that last read it (segment s + 2).
    A->rev_pending[0] = A->rev_pending[1] = false;
    for (int s = A->nseg - 1; s >= 0; s--) {
        const int set = s % sets;
        if (A->rev_pending[set]) {
            DEM_HIP(hipEventSynchronize(A->rev_done[set]));
            A->rev_pending[set] = false;
        }
        const AdjSnap &S = A->ck[s];
        if ((rc = snap_restore(c, S))) return rc;
        int ks = std::min(A->K, A->nsteps - s * A->K);
        std::vector<ListView> views;
        views.push_back(ListView{S.ab, S.off, S.toff, S.jt, S.wflag, nullptr, nullptr, S.npairs});
        std::vector<int> step_ver(ks);
        std::vector<char> step_reb(ks, 0);
        int used = 0;
        std::vector<AdjList> &pool = A->pool[set];
        for (int j = 0; j < ks; j++) {
            if ((size_t)used >= pool.size()) pool.emplace_back();
            bool reb = false;
            AdjList &L = pool[used];
            if ((rc = adj_step(c, A, set, j, &L, &reb))) return rc;
            if (reb) {
                views.push_back(ListView{L.ab, L.off, L.toff, L.jt, L.wflag, L.perm, L.src, L.np});
                used++;
                step_reb[j] = 1;
            }
            step_ver[j] = (int)views.size() - 1;
        }
        DEM_HIP(hipEventRecord(A->rec_done[set], st));
        DEM_HIP(hipStreamWaitEvent(A->rs, A->rec_done[set], 0));
        if ((rc = reverse_segment(c, A, set, ks, views, step_ver, step_reb, hc))) return rc;
        DEM_HIP(hipEventRecord(A->rev_done[set], A->rs));
        A->rev_pending[set] = true;
    }
    // the last reverse (segment 0) is the latest work on the reverse stream
    DEM_HIP(hipStreamWaitEvent(st, A->rev_done[0], 0));
    A->rev_pending[0] = A->rev_pending[1] = false;
    DEM_HIP(hipEventRecord(c->ev1, st));
    DEM_HIP(hipEventSynchronize(c->ev1));
    DEM_HIP(hipEventElapsedTime(&A->bwd_ms, c->ev0, c->ev1));
    c->last_ms = A->bwd_ms;
    A->have_grad = true;
    if (theta_bar) {
        std::vector<double> th((size_t)kThetaPer * n);
        DEM_HIP(hipMemcpy(th.data(), A->thacc, sizeof(double) * th.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < c->ninst; k++) {
            const InstDev &I = c->hinst[k];
            double acc[kThetaPer] = {0.0};
            for (unsigned i = I.p_off; i < I.p_off + I.n; i++)
                for (int j = 0; j < kThetaPer; j++) acc[j] += th[(size_t)kThetaPer * i + j];
            for (int j = 0; j < kThetaPer; j++) theta_bar[(size_t)kThetaPer * k + j] = acc[j];
        }
    }
    // leave the context at the end of the forward
    return snap_restore(c, A->ck[A->nseg]);
}

int dem_adj_state_grad(dem_ctx *c, double *xbar, double *vbar, double *wbar)
{
    if (!c || !c->adj || !c->adj->have_grad) return DEM_ERR_INVALID;
    AdjData *A = c->adj;
    size_t n = c->n;
    std::vector<SAdj> lam(n);
    std::vector<unsigned> inv(n);
    DEM_HIP(hipStreamSynchronize(c->stream));
    DEM_HIP(hipMemcpy(lam.data(), A->lam[A->cur], sizeof(SAdj) * n, hipMemcpyDeviceToHost));
    DEM_HIP(hipMemcpy(inv.data(), A->ck[0].inv, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        const SAdj &L = lam[inv[i]];
        for (int k = 0; k < 3; k++) {
            if (xbar) xbar[3 * i + k] = L.x[k];
            if (vbar) vbar[3 * i + k] = L.v[k];
            if (wbar) wbar[3 * i + k] = L.w[k];
        }
    }
    return DEM_OK;
}

int dem_adj_info(dem_ctx *c, double *out)
{
    if (!c || !c->adj || !out) return DEM_ERR_INVALID;
    AdjData *A = c->adj;
    out[0] = A->fwd_ms;
    out[1] = A->bwd_ms;
    out[2] = A->nseg;
    out[3] = A->max_np;
    // device bytes held by the checkpoints and by the tapes of one segment
    size_t snaps = 0;
    for (const AdjSnap &S : A->ck)
        if (S.valid)
            snaps += (sizeof(PRec) + sizeof(Hist) + 6 * sizeof(unsigned)) * (size_t)A->n +
                     S.pcap * (sizeof(uint2) + sizeof(Hist) + sizeof(unsigned));
    size_t tapes = (size_t)A->tape_sets * A->tape_K *
                   (A->n * (sizeof(PRec) + sizeof(Hist)) + (size_t)A->max_np * sizeof(Hist));
    out[4] = (double)snaps;
    out[5] = (double)tapes;
    return DEM_OK;
}

int dem_adj_release(dem_ctx *c)
{
    if (!c) return DEM_ERR_INVALID;
    adj_free(c);
    return DEM_OK;
}

}  // extern "C"
