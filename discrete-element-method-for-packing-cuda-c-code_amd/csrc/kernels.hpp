// kernels.hpp -- gfx950 kernels of the DEM step.
//
// Context path (cell-sorted slots, CSR pair lists):
//   rebuild : k_hash (cell counts) -> scan -> k_sort_scatter -> k_reorder -> k_nbr_count
//             (candidate search, row records) -> scans -> k_nbr_fill (pair list) -> k_remap
//             (tangential histories by pair identity) -> k_transpose
//   step    : fused (default) k_step -- one lane per particle: wall contact, contact filter,
//             the laws of its contacts balanced over the wave, the sums in the reference
//             order, gravity, integration, trigger; or split: k_pair_force (one lane per
//             candidate pair), k_walls (one lane per wall candidate) -> k_particle (gather
//             of the pair forces in the reference order, then as k_step)
// Reference-layout path (drop-in entry points, particle-index order, AoS double3):
//   c_* kernels with the reference's buffers and semantics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "law.hpp"

namespace dem {

constexpr unsigned kEmpty = 0xffffffffu;
constexpr int kBlock = 256;

// 128-byte particle record in slot order.  pos/rad, pdis/mass and angv/inert pack the data a
// pair gathers into the first 96 bytes; q (displacement since the last rebuild), the
// instance and the particle index fill the rest of the cache line.
struct __attribute__((aligned(16))) PRec {
    double px, py, pz, rad;
    double vx, vy, vz, mass;
    double wx, wy, wz, inert;
    double qx, qy, qz;
    unsigned inst, id;
};
static_assert(sizeof(PRec) == 128, "record must be one 128-byte line");

// Per-pair tangential history: dispt (second relative to first) and fricp.
struct __attribute__((aligned(16))) Hist {
    double hx, hy, hz, fr;
};

// Per-pair force output of the last step (80 bytes): force on first, torque on first,
// torque on second, contact flag (1.0 / 0.0).
struct __attribute__((aligned(16))) POut {
    double fx, fy, fz, flag;
    double tax, tay, taz, tbx;
    double tby, tbz;
};
static_assert(sizeof(POut) == 80, "pair output is five 16-byte chunks");

// Per-instance parameters (device table for the batched form; passed by value otherwise).
struct InstDev {
    LawParams pp;          // particle-particle
    LawParams pw;          // particle-wall
    double dt, skin, ann2;
    double roll_k;         // rolling-moment floor of the differentiable form, k (0: reference)
    double origin[3], cell[3];
    WallGeom walls;
    unsigned gsize[3], cell_off;
    unsigned p_off, n;
    unsigned flen;         // 1: the reference's float length() in vprojec and the trigger
                           // (context-wide: kernels take it from the launch's copy `one`)
    unsigned pad2;
};

// Device control block, copied to a host-mapped mirror by k_publish after each step.  Its
// allocation continues with per-instance max-z bit patterns (u64[ninst]) and contact totals
// (u32[ninst]), reduced there by k_inst_reduce on demand and copied to the mirror's tail.
struct DevCtl {
    unsigned need_rebuild;   // trigger raised by the last step (atomicOr)
    unsigned steps;          // steps completed (incremented by k_publish)
    unsigned npairs;         // written by the pair scan
    unsigned overflow;       // pair capacity exceeded at the last rebuild
    unsigned nwall;          // wall-candidate count of the last rebuild
    unsigned skip;           // copy of need_rebuild made by k_publish: the gate of queued steps
    unsigned zero;           // always 0: k_step's gate when the step is not queued
    unsigned pad;
    // Triggers of queued fused steps, by step parity: a queued k_step raises trig[s & 1] and
    // is gated by trig[(s - 1) & 1], so no publish kernel is needed between queued steps (a
    // gated step sets its own word, so the gate carries on to the steps after it).  Cleared
    // by every rebuild and before every queued run.
    unsigned trig[2];
    unsigned pad2[2];
};
static_assert(sizeof(DevCtl) == 48, "DevCtl is 12 words (the per-instance tail is 8-aligned)");

__device__ __forceinline__ double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }
__device__ __forceinline__ void st2(double *p, double a, double b)
{
    *reinterpret_cast<double2 *>(p) = make_double2(a, b);
}

__device__ __forceinline__ bool is_active(const unsigned *active) { return !active || *active != 0u; }

// roll_c of the law for bodies of inertias ia, ib (ib = 0: a wall) under floor factor k:
// k dt^2 (1/I_a + 1/I_b), the change one step of a rolling moment of unit size makes to the
// relative spin per unit moment; 0 when k = 0 (the reference's moment).
__device__ __forceinline__ double roll_floor(double k, double dt, double ia, double ib)
{
    if (!(k > 0.0)) return 0.0;
    return k * dt * dt * (ib > 0.0 ? 1.0 / ia + 1.0 / ib : 1.0 / ia);
}



// Gate of a queued step: a step launched before the host saw the previous step's rebuild
// trigger does nothing (every kernel of it returns at once) when that trigger was raised.
__device__ __forceinline__ bool gated_off(const unsigned *gate) { return gate && *gate != 0u; }

// XCD-aware block order.  Workgroups are dispatched to the 8 XCDs round-robin (block b to
// XCD b mod 8), each XCD with its own L2.  Renumbering so that XCD x runs the contiguous
// range of logical blocks [base_x, base_x + count_x), in order, keeps a particle's
// neighbours (nearby slots: slots are cell-sorted) and the pair data its partners read on
// the same L2.  A bijection on [0, nb) for any nb.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    unsigned q = nb / kXcds, r = nb % kXcds;
    unsigned x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

// XCD-aware order in chunks: XCD x runs chunks x, x + 8, x + 16, ... of `chunk` consecutive
// logical blocks, so work concentrated in one part of the slot range (the contacts of a bed
// forming at the floor, in the lowest slots) is shared by all XCDs while each chunk keeps its
// neighbours on one L2.  Blocks past the last whole round of 8 chunks take xcd_block's order.
// chunk == 0: one contiguous range per XCD (xcd_block).  A bijection on [0, nb).
__device__ __forceinline__ unsigned xcd_chunk_block(unsigned b, unsigned nb, unsigned chunk)
{
    if (chunk == 0u) return xcd_block(b, nb);
    const unsigned round = kXcds * chunk;
    const unsigned full = nb / round;
    const unsigned x = b % kXcds, i = b / kXcds;
    if (i / chunk < full) return (i / chunk) * round + x * chunk + i % chunk;
    const unsigned base = full * round;
    return base + xcd_block(b - base, nb - base);
}

// ----------------------------------------------------------------------------------------
// Rebuild
// ----------------------------------------------------------------------------------------

// The list rebuild sorts the (cell, index) pairs of calcHashD (neiglist.cu:77-97) by cell,
// equal cells in ascending particle index -- the result of the reference's stable radix sort
// of the index-ordered hash array (dempacking.cu:553-555) -- as a counting sort over the
// cells: k_hash counts each cell's particles (one atomic per particle, which also gives the
// particle an arrival rank in its cell), an exclusive scan of the counts gives each cell's
// first slot (cell c holds slots [cstart[c], cstart[c + 1]); cstart has ncells + 1 entries),
// k_sort_scatter places each pair at its cell's first slot plus its rank, and k_reorder
// restores the index order inside each cell while it moves the records.  Cells are offset per
// instance so one sort serves every instance.

// Adds one to cnt[h] for every calling lane and returns the lane's arrival rank there.  The
// calling lanes must be a prefix of the wave (lanes 0 .. k-1).  Lanes that ask for the same
// cell as their neighbour form runs, and the first lane of each run takes the run's ranks
// with one atomic and hands them on (a cell split over several runs takes one atomic per run),
// so inputs in cell order cost one atomic per run, not per lane (atomics on one address
// serialize).
__device__ __forceinline__ unsigned run_count(unsigned *cnt, unsigned h)
{
    const unsigned lane = __lane_id();
    const unsigned hp = (unsigned)__shfl_up((int)h, 1, 64);
    const unsigned long long heads = __ballot(lane == 0u || hp != h);
    const unsigned long long live = __ballot(1);
    const unsigned long long le = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);   // lanes <= lane
    const unsigned hl = 63u - (unsigned)__builtin_clzll(heads & le);   // this lane's run head
    const unsigned long long after = heads & ~le;                      // heads past this lane
    const unsigned end = after ? (unsigned)__builtin_ctzll(after)
                               : 64u - (unsigned)__builtin_clzll(live);   // one past the run
    unsigned base = 0;
    if (lane == hl) base = atomicAdd(&cnt[h], end - hl);
    base = (unsigned)__shfl((int)base, (int)hl, 64);
    return base + (lane - hl);
}

// calcHashD on slot-ordered records: kv[s] = (cell, index) of the particle in slot s, its
// cell's count incremented (cnt zeroed beforehand) and its arrival rank in rank[s].
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_hash(const PRec *rec, unsigned n, const InstDev *inst,
                                                 InstDev one, uint2 *kv, unsigned *cnt,
                                                 unsigned *rank, DevCtl *ctl,
                                                 const unsigned *active)
{
    if (!is_active(active)) return;
    unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s == 0 && ctl) {
        ctl->need_rebuild = 0u;
        ctl->skip = 0u;
        ctl->nwall = 0u;
        ctl->trig[0] = 0u;
        ctl->trig[1] = 0u;
    }
    if (s >= n) return;
    const PRec &r = rec[s];
    double px = r.px, py = r.py, pz = r.pz;
    unsigned id = r.id;
    const InstDev &I = BATCHED ? inst[r.inst] : one;
    unsigned x = cell_coord(px, I.origin[0], I.cell[0], I.gsize[0]);
    unsigned y = cell_coord(py, I.origin[1], I.cell[1], I.gsize[1]);
    unsigned z = cell_coord(pz, I.origin[2], I.cell[2], I.gsize[2]);
    unsigned h = z * I.gsize[1] * I.gsize[0] + y * I.gsize[0] + x + I.cell_off;
    kv[s] = make_uint2(h, id);
    // slots are in the last rebuild's cell order and few particles change cell
    rank[s] = run_count(cnt, h);
}

// Counting-sort scatter: the pair of slot s goes to its cell's first slot plus its arrival
// rank (a permutation of [0, n): ranks are 0 .. count - 1 within each cell), and the slot s
// itself -- the particle's slot in the records being sorted -- goes to the same place of
// oslot_out, so the reorder needs no index-to-slot lookup.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(const uint2 *kv, const unsigned *rank,
                                                         const unsigned *cstart, unsigned n,
                                                         uint2 *out, unsigned *oslot_out,
                                                         const unsigned *active)
{
    if (!is_active(active)) return;
    const unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint2 e = kv[s];
    const unsigned d = cstart[e.x] + rank[s];
    out[d] = e;
    oslot_out[d] = s;
}

__global__ __launch_bounds__(kBlock) void k_fill(unsigned *a, unsigned n, unsigned v,
                                                 const unsigned *active)
{
    if (!is_active(active)) return;
    for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) a[i] = v;
}

// findCellStartD (neiglist.cu:102-120) plus the matching cell end.
__global__ __launch_bounds__(kBlock) void k_cell_range(const uint2 *sorted, unsigned n,
                                                       unsigned *cstart, unsigned *cend,
                                                       const unsigned *active)
{
    if (!is_active(active)) return;
    unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    unsigned c = sorted[s].x;
    if (s == 0 || sorted[s - 1].x != c) cstart[c] = s;
    if (cend && (s == n - 1 || sorted[s + 1].x != c)) cend[c] = s + 1;
}

// Largest overlap the particle could reach with any wall after moving by `slack`: a
// particle with wall_reach(...) <= 0 cannot touch a wall before the next rebuild, because no
// particle moves more than half the skin between rebuilds (the rebuild trigger).
__device__ __forceinline__ double wall_reach(const WallGeom &g, double x, double y, double z,
                                             double r, double slack)
{
    double d = z;   // floor
    if (g.kind == 0) {
        d = fmin(d, fmin(x + g.hx, g.hx - x));
        d = fmin(d, fmin(y + g.hy, g.hy - y));
    } else {
        d = fmin(d, g.radius - sqrt(x * x + y * y));
    }
    return (r + slack) - d;
}

// Move every particle into its sorted slot; q restarts at zero (dempacking.cu:541).  The
// pairs come from k_sort_scatter in cell order, in arrival order inside a cell: the lane of
// position s moves the particle it finds there to slot s' = its cell's first slot plus the
// number of particles of that cell with a smaller index, and writes the pair to sorted[s'].
// Also lists the particles that may touch a wall before the next rebuild (for k_walls).
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_reorder(const uint2 *scat, const unsigned *sslot,
                                                    const unsigned *cstart,
                                                    uint2 *sorted, unsigned n,
                                                    const PRec *rec_old, PRec *rec_new,
                                                    const Hist *wh_old, Hist *wh_new,
                                                    unsigned *inv_new,
                                                    const InstDev *inst, InstDev one,
                                                    unsigned *wflag, unsigned *wlist,
                                                    DevCtl *ctl, unsigned *perm,
                                                    double4 *cpos, float4 *cp32, unsigned *cid,
                                                    unsigned *oslot, unsigned n_old,
                                                    const unsigned *active)
{
    if (!is_active(active)) return;
    unsigned s = blockIdx.x * kBlock + threadIdx.x;
    bool cand = false;
    unsigned id = 0, o = 0;
    uint2 e = make_uint2(0u, 0u);
    double2 v[8];   // the record, eight 16-byte chunks
    Hist w = Hist{0.0, 0.0, 0.0, 0.0};
    if (s < n) {
        e = scat[s];
        id = e.y;
        o = sslot[s];   // the particle's slot in rec_old (k_sort_scatter)
        const unsigned c0 = cstart[e.x], c1 = cstart[e.x + 1];
        // every load of the record in flight before any store (the stores cannot be moved
        // above loads of a buffer the compiler cannot tell apart from theirs)
        const double *src = reinterpret_cast<const double *>(rec_old + o);
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = ld2(src + 2 * k);
        w = wh_old[o];
        // stable position inside the cell while the record is in flight (a cell holds few
        // particles: its edge is the largest diameter plus the skin)
        unsigned r = 0;
        for (unsigned k = c0; k < c1; k++) r += scat[k].y < id ? 1u : 0u;
        s = c0 + r;
        sorted[s] = e;
        const unsigned ri = (unsigned)(unsigned long long)__double_as_longlong(v[7].y);   // inst
        const InstDev &I = BATCHED ? inst[ri] : one;
        cand = wall_reach(I.walls, v[0].x, v[0].y, v[1].x, v[1].y, 0.5 * I.skin) > 0.0;
        wflag[s] = cand ? 1u : 0u;
    }
    // wave-aggregated append to the wall-candidate list
    unsigned long long m = __ballot(cand);
    if (m) {
        unsigned base = 0;
        unsigned lane = __lane_id();
        unsigned leader = (unsigned)__ffsll((long long)m) - 1u;
        if (lane == leader) base = atomicAdd(&ctl->nwall, (unsigned)__popcll(m));
        base = __shfl(base, (int)leader, 64);
        if (cand) wlist[base + __popcll(m & ((1ull << lane) - 1ull))] = s;
    }
    if (s >= n) return;
    if (perm) perm[s] = o;   // recorded for the adjoint of the reorder
    // the particle's slot in the old list (kEmpty for particles the old list did not hold),
    // for the history remap of this rebuild
    oslot[s] = id < n_old ? o : kEmpty;
    double *dst = reinterpret_cast<double *>(rec_new + s);
#pragma unroll
    for (int k = 0; k < 6; k++) st2(dst + 2 * k, v[k].x, v[k].y);
    st2(dst + 12, 0.0, 0.0);     // q restarts at zero
    st2(dst + 14, 0.0, v[7].y);  // (instance, index) copied as raw bits
    wh_new[s] = w;
    inv_new[id] = s;
    // compact copies for the candidate searches of this rebuild (a 32-byte position and
    // radius, a 4-byte index: four candidates per cache line instead of one)
    cpos[s] = make_double4(v[0].x, v[0].y, v[1].x, v[1].y);
    cp32[s] = pos32(v[0].x, v[0].y, v[1].x, v[1].y);
    cid[s] = (unsigned)((unsigned long long)__double_as_longlong(v[7].y) >> 32);
}

// Slot range [s0, s1) of the consecutive cells c0 .. c1 of one stencil row.  Cells of a row
// are adjacent in key order and slots are sorted by key, so their particles are the slots
// [cstart[c0], cstart[c1 + 1]).  Returns false if the row is empty.
__device__ __forceinline__ bool row_range(const unsigned *cstart, unsigned c0, unsigned c1,
                                          unsigned &s0, unsigned &s1)
{
    s0 = cstart[c0];
    s1 = cstart[c1 + 1u];
    return s0 < s1;
}

// Block-wide minimum of lo and maximum of hi over the lanes (every lane of the block calls
// it; red holds two words per wave).
__device__ __forceinline__ void block_minmax(unsigned lo, unsigned hi, unsigned *red,
                                             unsigned &blo, unsigned &bhi)
{
#pragma unroll
    for (int d = 32; d; d >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, d, 64));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, d, 64));
    }
    const unsigned w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    __syncthreads();   // the previous call's readers are done
    if ((threadIdx.x & 63u) == 0) {
        red[2 * w] = lo;
        red[2 * w + 1] = hi;
    }
    __syncthreads();
    blo = red[0];
    bhi = red[1];
    for (unsigned k = 1; k < nw; k++) {
        blo = min(blo, red[2 * k]);
        bhi = max(bhi, red[2 * k + 1]);
    }
}

// Positions staged per chunk of the cooperative search, and the longest union of a stencil
// row's ranges it stages (a block spanning a cell layer's end reaches a whole layer: its
// lanes read their ranges from cpos directly instead).
constexpr unsigned kTile = 512;
constexpr unsigned kTileMax = 4 * kTile;
// Candidates a lane tests per window of the staged search (tile arrays hold kTile + kWin
// entries: a window may read past the staged ones and discards those bits).
constexpr unsigned kWin = 8;

// Candidate search (neighborarrayD's role) for the consecutive slots of a block, every lane
// of which calls it (live: the lane has a slot a < n): every b in the 27 cells around a's
// cell with is_candidate true (a itself included: the caller drops it), in stencil-row order (z-, y-offset ascending) and
// ascending slot order within a row -- the list order.  For each row the lanes' slot ranges
// -- contiguous, and overlapping for nearby cells -- are united, the union's positions are
// staged in LDS (tile: kTile entries) by coalesced loads, and each lane tests its own range
// from there, so each position is read from HBM once per block instead of once per lane.
// f(b, s0, s1, id) gets each candidate b with the lane's range [s0, s1) of its row and b's
// particle index (cid, staged with the positions); row_done(r, s0, s1) follows each row
// r = 0..8 on every lane (s0 = kEmpty, s1 = 0: no range).
template <bool BATCHED, class F, class R>
__device__ __forceinline__ void for_candidates_block(const double4 *cpos, const unsigned *cid,
                                                     unsigned a, bool live, unsigned key,
                                                     const InstDev &I, const unsigned *cstart,
                                                     double4 *tile,
                                                     unsigned *tcid, unsigned *red, F &&f,
                                                     R &&row_done)
{
    double4 ra = make_double4(0.0, 0.0, 0.0, 0.0);
    int cz = 0, cy = 0, cx = 0, x0 = 0, x1 = 0;
    const unsigned gx = I.gsize[0], gy = I.gsize[1], gz = I.gsize[2], gxy = gx * gy;
    if (live) {
        ra = cpos[a];
        const unsigned local = key - I.cell_off;
        cz = (int)(local / gxy);
        cy = (int)((local - (unsigned)cz * gxy) / gx);
        cx = (int)(local - (unsigned)cz * gxy - (unsigned)cy * gx);
        x0 = cx > 0 ? cx - 1 : 0;
        x1 = cx + 1 < (int)gx ? cx + 1 : (int)gx - 1;
    }
    const d3 xa = d3{ra.x, ra.y, ra.z};
    int r = -1;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            r++;
            unsigned s0 = kEmpty, s1 = 0;
            const int z = cz + dz, y = cy + dy;
            if (live && z >= 0 && z < (int)gz && y >= 0 && y < (int)gy) {
                const unsigned rowc = I.cell_off + (unsigned)z * gxy + (unsigned)y * gx;
                if (!row_range(cstart, rowc + (unsigned)x0, rowc + (unsigned)x1, s0, s1)) {
                    s0 = kEmpty;
                    s1 = 0;
                }
            }
            unsigned lo, hi;
            block_minmax(s0, s1, red, lo, hi);
            if (lo < hi && hi - lo > kTileMax) {   // uniform: each lane reads its own range
                for (unsigned b = s0; b < s1; b++) {
                    const double4 rb = cpos[b];
                    if (is_candidate(xa, d3{rb.x, rb.y, rb.z}, ra.w, rb.w, I.skin))
                        f(b, s0, s1, cid[b]);
                }
            } else {
                for (unsigned c0 = lo; c0 < hi; c0 += kTile) {   // (none when lo >= hi)
                    const unsigned c1 = min(c0 + kTile, hi);
                    __syncthreads();   // the previous chunk's readers are done
                    for (unsigned k = threadIdx.x; c0 + k < c1; k += blockDim.x) {
                        tile[k] = cpos[c0 + k];
                        tcid[k] = cid[c0 + k];
                    }
                    __syncthreads();
                    // The lane's slots [b0, b0 + len) of this chunk, tested kWin at a time
                    // with no branch per candidate: window w of every lane starts at the same
                    // offset from the lane's own b0, so the hit bits land at fixed positions
                    // of a 64-slot mask, and f runs afterwards for the set bits in ascending
                    // slot order (a candidate test hits ~4% of the time, but some lane of the
                    // wave almost always does, so a branch per candidate ran f's body for
                    // nearly every candidate).  A lane whose range ends inside a window reads
                    // up to kWin - 1 entries past it (tile holds kTile + kWin entries) and
                    // clears those bits; the lane's own slot passes the test and f drops it.
                    const unsigned b0 = max(s0, c0), b1 = min(s1, c1);
                    const unsigned len = b1 > b0 ? b1 - b0 : 0u;
                    for (unsigned w0 = 0; __ballot(w0 < len) != 0ull; w0 += 64u) {
                        unsigned mlo = 0u, mhi = 0u;
                        for (unsigned w = 0; w < 64u; w += kWin) {
                            const bool act = w0 + w < len;
                            if (__ballot(act) == 0ull) break;
                            const double4 *tw = tile + (act ? b0 - c0 + w0 + w : 0u);
                            unsigned m = 0u;
#pragma unroll
                            for (unsigned t = 0; t < kWin; t++) {
                                const double4 rb = tw[t];
                                m |= (unsigned)is_candidate(xa, d3{rb.x, rb.y, rb.z}, ra.w, rb.w,
                                                            I.skin) << t;
                            }
                            if (w < 32u) mlo |= m << w;
                            else mhi |= m << (w - 32u);
                        }
                        const unsigned rem = len > w0 ? len - w0 : 0u;   // slots left in range
                        if (rem < 32u) {
                            mlo &= (1u << rem) - 1u;
                            mhi = 0u;
                        } else if (rem < 64u) {
                            mhi &= (1u << (rem - 32u)) - 1u;
                        }
                        while (mlo) {
                            const unsigned b = b0 + w0 + (unsigned)__builtin_ctz(mlo);
                            mlo &= mlo - 1u;
                            f(b, s0, s1, tcid[b - c0]);
                        }
                        while (mhi) {
                            const unsigned b = b0 + w0 + 32u + (unsigned)__builtin_ctz(mhi);
                            mhi &= mhi - 1u;
                            f(b, s0, s1, tcid[b - c0]);
                        }
                    }
                }
            }
            row_done(r, s0, s1);
        }
}

// The count's record of each stencil row of a lane for the fill: the row's first slot s0 and
// two bit masks over its slots (bit k for slot s0 + k), the candidates with a larger particle
// index (the lane's own pairs) and those with a smaller one (its transposed entries).  Rows
// longer than kRowBits slots are marked kLongRow (the fill tests them again); rows without
// slots kEmpty.
constexpr unsigned kRowBits = 64;
constexpr unsigned kRows = 9;
constexpr unsigned kLongRow = kEmpty - 1u;

// Slot range [s0, s1) of stencil row r (z offset r / 3 - 1, y offset r % 3 - 1) of the cell
// with key `key`, as for_candidates_block forms it; false if the row is empty or off the grid.
__device__ __forceinline__ bool stencil_row(unsigned key, const InstDev &I, const unsigned *cstart,
                                            int r, unsigned &s0,
                                            unsigned &s1)
{
    const unsigned gx = I.gsize[0], gy = I.gsize[1], gz = I.gsize[2], gxy = gx * gy;
    const unsigned local = key - I.cell_off;
    const int cz = (int)(local / gxy);
    const int cy = (int)((local - (unsigned)cz * gxy) / gx);
    const int cx = (int)(local - (unsigned)cz * gxy - (unsigned)cy * gx);
    const int x0 = cx > 0 ? cx - 1 : 0;
    const int x1 = cx + 1 < (int)gx ? cx + 1 : (int)gx - 1;
    const int z = cz + r / 3 - 1, y = cy + r % 3 - 1;
    if (z < 0 || z >= (int)gz || y < 0 || y >= (int)gy) return false;
    const unsigned rowc = I.cell_off + (unsigned)z * gxy + (unsigned)y * gx;
    return row_range(cstart, rowc + (unsigned)x0, rowc + (unsigned)x1, s0, s1);
}

template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_nbr_count(const PRec *rec, const double4 *cpos,
                                                      const unsigned *cid, const uint2 *sorted,
                                                      unsigned n, const InstDev *inst,
                                                      InstDev one, const unsigned *cstart,
                                                      unsigned *ng,
                                                      unsigned *nl, unsigned long long *rmask,
                                                      unsigned *rs0, const unsigned *active)
{
    if (!is_active(active)) return;   // uniform over the grid
    __shared__ double4 tile[kTile + kWin];
    __shared__ unsigned tcid[kTile];
    __shared__ unsigned red[2 * (kBlock / 64)];
    const unsigned a = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool live = a < n;
    const InstDev &I = BATCHED ? inst[rec[live ? a : 0u].inst] : one;
    unsigned cg = 0, cl = 0;
    unsigned long long mg = 0ull, ml = 0ull;
    const unsigned ida = live ? cid[a] : 0u;
    for_candidates_block<BATCHED>(
        cpos, cid, a, live, live ? sorted[a].x : 0u, I, cstart, tile, tcid, red,
        [&](unsigned b, unsigned r0, unsigned r1, unsigned idb) {
            if (b == a) return;   // the lane's own slot
            const unsigned long long bit = r1 - r0 <= kRowBits ? 1ull << (b - r0) : 0ull;
            if (idb > ida) {
                cg++;
                mg |= bit;
            } else {
                cl++;
                ml |= bit;
            }
        },
        [&](int r, unsigned r0, unsigned r1) {
            if (live) {
                const size_t k = (size_t)r * n + a;
                rmask[2 * k] = mg;
                rmask[2 * k + 1] = ml;
                rs0[k] = r0 == kEmpty ? kEmpty : (r1 - r0 <= kRowBits ? r0 : kLongRow);
            }
            mg = ml = 0ull;
        });
    if (live) {
        ng[a] = cg;
        nl[a] = cl;
    }
}

// Binary search of `target` in the ascending values get(lo) .. get(hi - 1); kEmpty if absent.
template <class Get>
__device__ __forceinline__ unsigned find_sorted(Get get, unsigned lo, unsigned hi, unsigned target)
{
    while (lo < hi) {
        const unsigned mid = (lo + hi) >> 1;
        const unsigned v = get(mid);
        if (v == target) return mid;
        if (v < target) lo = mid + 1;
        else hi = mid;
    }
    return kEmpty;
}

// `target` among the ascending second members of ab[lo, hi); kEmpty if absent.
__device__ __forceinline__ unsigned find_second(const uint2 *ab, unsigned lo, unsigned hi,
                                                unsigned target)
{
    return find_sorted([&](unsigned i) { return ab[i].y; }, lo, hi, target);
}

// Entries of one block's own-pair and transposed ranges the fill stages in LDS and writes
// out together (consecutive lanes on consecutive entries) instead of as per-lane appends.
constexpr unsigned kFillStage = 2048;

// Pair list of slot a: its partners with a larger particle index, in ascending slot order,
// at [off[a], off[a+1]) -- each pair is stored once, first member = smaller particle index,
// exactly the reference's Lp < Ln orientation (dempacking.cu:640-644), while segments follow
// the cell-sorted slot order.  Partners with a smaller index go to the transposed list
// [toff[a], toff[a+1]): jf = the partner (the pair's first member), jt = a for now; the
// transpose turns jt into the pair's index and the remap fills the tangential histories
// (k_transpose, k_remap).  Both
// ranges of a block are contiguous; entries are staged in LDS and written out by consecutive
// lanes (a range longer than kFillStage is written lane by lane).
template <bool BATCHED>
__global__ __launch_bounds__(kBlock) void k_nbr_fill(
    const PRec *rec, const double4 *cpos, const unsigned *cid, const uint2 *sorted, unsigned n,
    const InstDev *inst, InstDev one,
    const unsigned *cstart, const unsigned *off, const unsigned *toff,
    unsigned cap, uint2 *ab, unsigned *bo, unsigned *jt, unsigned *jf, DevCtl *ctl,
    const unsigned long long *rmask, const unsigned *rs0, const unsigned *active)
{
    if (!is_active(active)) return;   // uniform over the grid
    __shared__ uint2 s_ab[kFillStage];
    __shared__ unsigned s_bo[kFillStage], s_jf[kFillStage], s_jt[kFillStage];
    const unsigned s0 = xcd_block(blockIdx.x, gridDim.x) * kBlock;
    const unsigned a = s0 + threadIdx.x;
    const unsigned se = min(s0 + kBlock, n);
    // the block's contiguous own-pair and transposed ranges
    const unsigned G0 = off[s0], G1 = off[se], T0 = toff[s0], T1 = toff[se];
    const bool stage_g = G1 - G0 <= kFillStage, stage_t = T1 - T0 <= kFillStage;
    bool over = false;
    if (a < n) {
        const InstDev &I = BATCHED ? inst[rec[a].inst] : one;
        unsigned pg = off[a];
        unsigned pl = toff[a];
        const unsigned ida = cid[a];
        auto own = [&](unsigned b) {
            if (pg >= cap) {
                over = true;
                return;
            }
            if (stage_g) {
                s_ab[pg - G0] = make_uint2(a, b);
                s_bo[pg - G0] = b;
            } else {
                ab[pg] = make_uint2(a, b);
                bo[pg] = b;
            }
            pg++;
        };
        auto transposed = [&](unsigned b) {
            if (pl < cap) {
                if (stage_t) {
                    s_jf[pl - T0] = b;
                    s_jt[pl - T0] = a;
                } else {
                    jf[pl] = b;
                    jt[pl] = a;
                }
            }
            pl++;
        };
        // the count's record of each row, in the same row and slot order: the larger-index
        // candidates go to the own list, the others to the transposed list (each list keeps
        // candidate order); rows longer than kRowBits are tested again
        for (int r = 0; r < (int)kRows; r++) {
            const size_t k = (size_t)r * n + a;
            const unsigned r0 = rs0[k];
            if (r0 == kEmpty) continue;
            if (r0 != kLongRow) {
                unsigned long long g = rmask[2 * k], l = rmask[2 * k + 1];
                while (g) {
                    own(r0 + (unsigned)__builtin_ctzll(g));
                    g &= g - 1ull;
                }
                while (l) {
                    transposed(r0 + (unsigned)__builtin_ctzll(l));
                    l &= l - 1ull;
                }
            } else {
                unsigned b0, b1;
                stencil_row(sorted[a].x, I, cstart, r, b0, b1);
                const double4 ra = cpos[a];
                const d3 xa = d3{ra.x, ra.y, ra.z};
                for (unsigned b = b0; b < b1; b++) {
                    if (b == a) continue;
                    const double4 rb = cpos[b];
                    if (!is_candidate(xa, d3{rb.x, rb.y, rb.z}, ra.w, rb.w, I.skin)) continue;
                    if (cid[b] > ida) own(b);
                    else transposed(b);
                }
            }
        }
    }
    __syncthreads();
    if (stage_g)
        for (unsigned k = threadIdx.x; G0 + k < min(G1, cap); k += kBlock) {
            ab[G0 + k] = s_ab[k];
            bo[G0 + k] = s_bo[k];
        }
    if (stage_t)
        for (unsigned k = threadIdx.x; T0 + k < min(T1, cap); k += kBlock) {
            jf[T0 + k] = s_jf[k];
            jt[T0 + k] = s_jt[k];
        }
    if (over && ctl) atomicOr(&ctl->overflow, 1u);
}

// Tangential history of each pair of a new list, one lane per pair: the history of the same
// particle pair in the old list, searched under the members' old slots (oslot, which the
// reorder of this rebuild records), else zero -- the
// reference's pairlistD remap (particleSystem.cu:143-191; new particles have empty old
// segments, dempacking.cu:204,521-522).  Consecutive lanes take consecutive pairs, mostly
// of one first member and so of one old segment, and write whole history lines.  A rebuild
// changes few partners, so the pair is first looked for in the old segment at the place it
// has in the new one (one load, kept when the second member matches: second members are
// distinct in a segment), and searched for only when that misses.  Block blk of
// nblk takes a contiguous, XCD-aware range of the device pair count (clamped to the capacity),
// so a fixed grid serves a captured rebuild.  src (optional) records each pair's old index for
// the adjoint of the remap.
__device__ __forceinline__ void remap_blocks(unsigned blk, unsigned nblk, const uint2 *ab,
                                             const unsigned *off, const DevCtl *ctl, unsigned cap,
                                             const unsigned *oslot, Hist *hist, bool have_old,
                                             const uint2 *ab_old, const unsigned *off_old,
                                             const Hist *hist_old, unsigned n_old,
                                             unsigned cap_old, unsigned *src)
{
    const unsigned np = min(ctl->npairs, cap);
    const unsigned per = (np + nblk - 1) / nblk;
    const unsigned p_begin = xcd_block(blk, nblk) * per;
    const unsigned p_end = min(p_begin + per, np);
    for (unsigned p = p_begin + threadIdx.x; p < p_end; p += blockDim.x) {
        const uint2 e = ab[p];
        Hist h = Hist{0.0, 0.0, 0.0, 0.0};
        unsigned q = kEmpty;
        if (have_old) {
            const unsigned oa = oslot[e.x], ob = oslot[e.y];
            const unsigned j = p - off[e.x];   // the pair's place in its new segment
            if (oa < n_old && ob != kEmpty) {
                const unsigned lo = min(off_old[oa], cap_old), hi = min(off_old[oa + 1], cap_old);
                const unsigned qp = lo + j;
                q = (qp < hi && ab_old[qp].y == ob) ? qp : find_second(ab_old, lo, hi, ob);
                if (q != kEmpty) h = hist_old[q];
            }
        }
        hist[p] = h;
        if (src) src[p] = q;
    }
}

// `target` among the ascending second slots bo[lo, hi) of a segment; kEmpty if absent.
__device__ __forceinline__ unsigned find_slot(const unsigned *bo, unsigned lo, unsigned hi,
                                              unsigned target)
{
    return find_sorted([&](unsigned i) { return bo[i]; }, lo, hi, target);
}

// Resolve each transposed entry k (first member jf[k], second member jt[k], as the fill left
// them) to the index of that pair in the first member's segment, whose second slots ascend --
// one lane per entry; block blk of nblk takes a contiguous, XCD-aware range of the n
// particles' toff[n] entries (clamped to the capacity), so a fixed grid serves a captured
// rebuild.
__device__ __forceinline__ void transpose_blocks(unsigned blk, unsigned nblk, const unsigned *bo,
                                                 const unsigned *off, const unsigned *toff,
                                                 unsigned n, unsigned cap, unsigned *jt,
                                                 const unsigned *jf)
{
    const unsigned nt = min(toff[n], cap);
    const unsigned per = (nt + nblk - 1) / nblk;
    const unsigned k_begin = xcd_block(blk, nblk) * per;
    const unsigned k_end = min(k_begin + per, nt);
    for (unsigned k = k_begin + threadIdx.x; k < k_end; k += blockDim.x) {
        const unsigned b = jf[k], a = jt[k];
        unsigned lo = off[b], hi = off[b + 1];
        if (hi > cap) hi = cap;
        jt[k] = lo < hi ? find_slot(bo, lo, hi, a) : kEmpty;
    }
}

__global__ __launch_bounds__(kBlock) void k_remap(const uint2 *ab, const unsigned *off, const DevCtl *ctl, unsigned cap,
                                                  const unsigned *oslot, Hist *hist,
                                                  bool have_old, const uint2 *ab_old,
                                                  const unsigned *off_old, const Hist *hist_old,
                                                  unsigned n_old, unsigned cap_old, unsigned *src,
                                                  const unsigned *active)
{
    if (!is_active(active)) return;
    remap_blocks(blockIdx.x, gridDim.x, ab, off, ctl, cap, oslot, hist, have_old, ab_old, off_old,
                 hist_old, n_old, cap_old, src);
}

__global__ __launch_bounds__(kBlock) void k_transpose(const unsigned *bo, const unsigned *off,
                                                      const unsigned *toff, unsigned n,
                                                      unsigned cap, unsigned *jt,
                                                      const unsigned *jf, const unsigned *active)
{
    if (!is_active(active)) return;
    transpose_blocks(blockIdx.x, gridDim.x, bo, off, toff, n, cap, jt, jf);
}

// Compact positions of restored records: cpos[s] = (px, py, pz, rad) of rec[s], and its FP32
// copy.
__global__ __launch_bounds__(kBlock) void k_cpos_from_rec(const PRec *rec, unsigned n,
                                                          double4 *cpos, float4 *cp32)
{
    unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const double *R = reinterpret_cast<const double *>(rec + s);
    double2 p01 = ld2(R), p23 = ld2(R + 2);
    double *C = reinterpret_cast<double *>(cpos + s);
    st2(C, p01.x, p01.y);
    st2(C + 2, p23.x, p23.y);
    cp32[s] = pos32(p01.x, p01.y, p23.x, p23.y);
}

// jf from a restored list: jf[k] = first member of pair jt[k].
__global__ __launch_bounds__(kBlock) void k_owner(const uint2 *ab, const unsigned *jt,
                                                  unsigned np, unsigned *jf, unsigned *bo)
{
    unsigned k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    unsigned q = jt[k];
    jf[k] = q < np ? ab[q].x : kEmpty;
    bo[k] = ab[k].y;
}

// Pair-capacity check after the scan: npairs = min(total, cap); overflow when total > cap.
// Every later kernel bounds its pair indices by npairs, so an overflowed list is truncated,
// never read or written past its allocation.
__global__ void k_check_cap(DevCtl *ctl, const unsigned *total_g, const unsigned *total_l,
                            unsigned cap, const unsigned *active)
{
    if (!is_active(active)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned p = *total_g;
        bool over = p > cap || *total_l > cap;
        ctl->npairs = over ? cap : p;
        if (over) ctl->overflow = 1u;
    }
}

__global__ void k_set_u32(unsigned *p, unsigned v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ----------------------------------------------------------------------------------------
// Step
// ----------------------------------------------------------------------------------------

// One lane per candidate pair (the reference's calculateforcepairD, neiglist.cu:156-233).
// Grid-stride over npairs read from the device, so a fixed grid serves any list size.
template <bool BATCHED, bool FLEN = false>
__global__ __launch_bounds__(kBlock) void k_pair_force(const uint2 *ab, const unsigned *npairs,
                                                       const PRec *rec, const Hist *hist,
                                                       Hist *hist_out, POut *out,
                                                       const InstDev *inst, LawParams pp,
                                                       double dt, double roll_k, Hist *tape_hist,
                                                       const unsigned *gate, unsigned xchunk)
{
    if (gated_off(gate)) return;
    const unsigned np = *npairs;
    // each block takes a contiguous range of pairs, XCD-aware (pairs are ordered by their
    // first member's slot, so a range touches nearby records)
    const unsigned bs = blockDim.x;
    const unsigned per = (np + gridDim.x - 1) / gridDim.x;
    const unsigned p_begin = xcd_chunk_block(blockIdx.x, gridDim.x, xchunk) * per;
    const unsigned p_end = min(p_begin + per, np);
    for (unsigned p = p_begin + threadIdx.x; p < p_end; p += bs) {
        uint2 e = ab[p];
        const double *A = reinterpret_cast<const double *>(rec + e.x);
        const double *B = reinterpret_cast<const double *>(rec + e.y);
        double2 a0 = ld2(A), a1 = ld2(A + 2), a2 = ld2(A + 4), a3 = ld2(A + 6), a4 = ld2(A + 8),
                a5 = ld2(A + 10);
        double2 b0 = ld2(B), b1 = ld2(B + 2), b2 = ld2(B + 4), b3 = ld2(B + 6), b4 = ld2(B + 8),
                b5 = ld2(B + 10);
        const double *H = reinterpret_cast<const double *>(hist + p);
        double2 h0 = ld2(H), h1 = ld2(H + 2);
        if (tape_hist) {   // incoming history, for the adjoint
            double *T = reinterpret_cast<double *>(tape_hist + p);
            st2(T, h0.x, h0.y);
            st2(T + 2, h1.x, h1.y);
        }
        LawParams P = pp;
        double tdt = dt, rk = roll_k;
        if (BATCHED) {
            const InstDev &I = inst[rec[e.x].inst];
            P = I.pp;
            tdt = I.dt;
            rk = I.roll_k;
        }
        d3 fa, ta, tb, ho;
        double fo;
        bool c = law_pair(d3{a0.x, a0.y, a1.x}, d3{b0.x, b0.y, b1.x}, d3{a2.x, a2.y, a3.x},
                          d3{b2.x, b2.y, b3.x}, d3{a4.x, a4.y, a5.x}, d3{b4.x, b4.y, b5.x},
                          a1.y, b1.y, a3.y, b3.y, d3{h0.x, h0.y, h1.x}, tdt, P, fa, ta, tb, ho,
                          fo, roll_floor(rk, tdt, a5.y, b5.y), FLEN);
        double *O = reinterpret_cast<double *>(out + p);
        st2(O, fa.x, fa.y);
        st2(O + 2, fa.z, c ? 1.0 : 0.0);
        st2(O + 4, ta.x, ta.y);
        st2(O + 6, ta.z, tb.x);
        st2(O + 8, tb.y, tb.z);
        if (hist_out) {   // in place (hist_out == hist) on the split path
            double *HW = reinterpret_cast<double *>(hist_out + p);
            st2(HW, ho.x, ho.y);
            st2(HW + 2, ho.z, fo);
        }
    }
}

// Wall contact of slot a (calculateforceiwD's wall part, neiglist.cu:286-301) from its record
// R (pos/rad, pdis/mass, angv/inert): reads the slot's wall history, writes it back (zeroed
// when no wall is touched) and returns the number of walls hit with the wall force and
// torque.  tape_whist receives the incoming history, for the adjoint.
template <bool BATCHED, bool FLEN = false>
__device__ __forceinline__ int wall_slot(const double *R, const InstDev &I, Hist *whist,
                                         unsigned a, Hist *tape_whist, d3 &fw, d3 &tw)
{
    double2 r0 = ld2(R), r1 = ld2(R + 2), r2 = ld2(R + 4), r3 = ld2(R + 6), r4 = ld2(R + 8),
            r5 = ld2(R + 10);
    double *WH = reinterpret_cast<double *>(whist + a);
    double2 w0 = ld2(WH), w1 = ld2(WH + 2);
    if (tape_whist) {
        double *T = reinterpret_cast<double *>(tape_whist + a);
        st2(T, w0.x, w0.y);
        st2(T + 2, w1.x, w1.y);
    }
    d3 hw;
    double fow;
    int hits = law_walls(I.walls, d3{r0.x, r0.y, r1.x}, d3{r2.x, r2.y, r3.x},
                         d3{r4.x, r4.y, r5.x}, r1.y, r3.y, d3{w0.x, w0.y, w1.x}, I.dt, I.pw,
                         fw, tw, hw, fow, roll_floor(I.roll_k, I.dt, r5.y, 0.0), FLEN);
    if (hits) {
        st2(WH, hw.x, hw.y);
        st2(WH + 2, hw.z, fow);
    } else {
        st2(WH, 0.0, 0.0);
        st2(WH + 2, 0.0, 0.0);
    }
    return hits;
}

// Wall contacts of the rebuild's wall candidates for the split path: wforce[slot] =
// (F xyz, hits, T xyz, 0).  Grid-stride over the device-resident candidate count.
template <bool BATCHED, bool FLEN = false>
__global__ __launch_bounds__(kBlock) void k_walls(const PRec *rec, Hist *whist,
                                                  const unsigned *wlist, const DevCtl *ctl,
                                                  const InstDev *inst, InstDev one,
                                                  double *wforce, Hist *tape_whist,
                                                  const unsigned *gate)
{
    if (gated_off(gate)) return;
    const unsigned nw = ctl->nwall;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
        unsigned a = wlist[i];
        const InstDev &I = BATCHED ? inst[rec[a].inst] : one;
        d3 fw, tw;
        int hits = wall_slot<BATCHED, FLEN>(reinterpret_cast<const double *>(rec + a), I, whist,
                                            a, tape_whist, fw, tw);
        double *W = wforce + 8 * (size_t)a;
        st2(W, fw.x, fw.y);
        st2(W + 2, fw.z, (double)hits);
        st2(W + 4, tw.x, tw.y);
        st2(W + 6, tw.z, 0.0);
    }
}

// Gravity, wall force, force record, integrator and rebuild trigger of slot a, after its
// pair forces have been summed (calculateforceiwD, neiglist.cu:238-306, and
// updatepositionD, neiglist.cu:308-409).  R is the record before the step, W receives the
// record after it (the same record on the split path, the other buffer on the fused one).
// Returns 1 if the particle moved more than half the skin since the last rebuild.
template <bool BATCHED, bool FLEN = false>
__device__ __forceinline__ unsigned finish_particle(const double *R, double *W, unsigned a,
                                                    double fx, double fy, double fz, double tx,
                                                    double ty, double tz, unsigned c, bool wall,
                                                    d3 fw, d3 tw, const InstDev *inst,
                                                    const InstDev &one,
                                                    unsigned *contacti, double *fdbg,
                                                    PRec *tape_rec, float4 *cp32_out)
{
    double2 r0 = ld2(R), r1 = ld2(R + 2), r2 = ld2(R + 4), r3 = ld2(R + 6), r4 = ld2(R + 8),
            r5 = ld2(R + 10), r6 = ld2(R + 12);
    double qz = R[14];
    uint2 tail = *reinterpret_cast<const uint2 *>(R + 15);
    if (tape_rec) {   // the record before the update, for the adjoint
        double *T = reinterpret_cast<double *>(tape_rec + a);
        st2(T, r0.x, r0.y); st2(T + 2, r1.x, r1.y); st2(T + 4, r2.x, r2.y);
        st2(T + 6, r3.x, r3.y); st2(T + 8, r4.x, r4.y); st2(T + 10, r5.x, r5.y);
        st2(T + 12, r6.x, r6.y);
        T[14] = qz;
        *reinterpret_cast<uint2 *>(T + 15) = tail;
    }
    const InstDev &I = BATCHED ? inst[tail.x] : one;
    d3 x = d3{r0.x, r0.y, r1.x};
    d3 v = d3{r2.x, r2.y, r3.x};
    double m = r3.y;
    d3 w = d3{r4.x, r4.y, r5.x};
    double inert = r5.y;
    // gravity (neiglist.cu:272, 284)
    fx = fx + 0.0;
    fy = fy + 0.0;
    fz = fz + (-m);
    // walls (wall candidates of this rebuild that touch a wall)
    if (wall) {
        fx += fw.x; fy += fw.y; fz += fw.z;
        tx += tw.x; ty += tw.y; tz += tw.z;
    }
    if (fdbg) {
        double *D = fdbg + 6 * (size_t)a;
        D[0] = fx; D[1] = fy; D[2] = fz; D[3] = tx; D[4] = ty; D[5] = tz;
    }
    // integrator
    double dt = I.dt;
    v.x = v.x + fx * dt * dt / m;
    v.y = v.y + fy * dt * dt / m;
    v.z = v.z + fz * dt * dt / m;
    x.x = x.x + v.x; x.y = x.y + v.y; x.z = x.z + v.z;
    double qx = r6.x + v.x, qy = r6.y + v.y, qzz = qz + v.z;
    w.x = w.x + tx * dt * dt / inert;
    w.y = w.y + ty * dt * dt / inert;
    w.z = w.z + tz * dt * dt / inert;
    st2(W, x.x, x.y);
    st2(W + 2, x.z, r1.y);
    st2(W + 4, v.x, v.y);
    st2(W + 6, v.z, m);
    st2(W + 8, w.x, w.y);
    st2(W + 10, w.z, inert);
    st2(W + 12, qx, qy);
    W[14] = qzz;
    *reinterpret_cast<uint2 *>(W + 15) = tail;
    cp32_out[a] = pos32(x.x, x.y, x.z, r1.y);
    contacti[a] = c;
    double q2 = qx * qx + qy * qy + qzz * qzz;
    // the reference's rmov = length(qqi) > ann (neiglist.cu:335,359,370) under the
    // float-length option; |q|^2 > ann^2 in FP64 otherwise
    if (FLEN) return (double)sqrtf((float)q2) > 0.5 * I.skin ? 1u : 0u;
    return q2 > I.ann2 ? 1u : 0u;
}

// Per-particle kernel of the split path: gather of the particle's pair forces (own pairs in
// list order, then pairs where it is second, via the transposed index) from k_pair_force's
// output, then finish_particle.  Max z and contact totals are not reduced here (they are
// diagnostics): k_inst_reduce computes them on demand.
template <bool BATCHED, bool FLEN = false>
__global__ __launch_bounds__(kBlock) void k_particle(
    PRec *rec, const unsigned *wflag, const double *wforce, const unsigned *off,
    const unsigned *toff, const unsigned *jt, const POut *out, float4 *cp32, unsigned n,
    const InstDev *inst, InstDev one, unsigned *contacti, double *fdbg, DevCtl *ctl,
    PRec *tape_rec, const unsigned *gate, unsigned xchunk)
{
    if (gated_off(gate)) return;
    const unsigned np = ctl->npairs;
    unsigned a = xcd_chunk_block(blockIdx.x, gridDim.x, xchunk) * blockDim.x + threadIdx.x;
    bool live = a < n;
    unsigned flag = 0;
    if (live) {
        double fx = 0.0, fy = 0.0, fz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
        unsigned c = 0;
        // Sum in the reference's order (sumallforceiD over an Aref row): the particle's own
        // pairs (partners with a larger index) in ascending slot order, then the pairs where
        // it is the second member, also in ascending slot order.  The same terms in the same
        // order as the oracle, so the totals agree bit for bit.
        unsigned p0 = off[a], p1 = off[a + 1];
        if (p1 > np) p1 = np;
        for (unsigned p = p0; p < p1; p++) {
            const double *O = reinterpret_cast<const double *>(out + p);
            double2 o0 = ld2(O), o1 = ld2(O + 2), o2 = ld2(O + 4), o3 = ld2(O + 6);
            fx += o0.x; fy += o0.y; fz += o1.x;
            tx += o2.x; ty += o2.y; tz += o3.x;
            c += o1.y != 0.0 ? 1u : 0u;
        }
        unsigned q0 = toff[a], q1 = toff[a + 1];
        if (q1 > np) q1 = np;
        for (unsigned k = q0; k < q1; k++) {
            unsigned q = jt[k];
            if (q >= np) continue;
            const double *O = reinterpret_cast<const double *>(out + q);
            double2 o0 = ld2(O), o1 = ld2(O + 2), o3 = ld2(O + 6), o4 = ld2(O + 8);
            fx -= o0.x; fy -= o0.y; fz -= o1.x;
            tx += o3.y; ty += o4.x; tz += o4.y;
            c += o1.y != 0.0 ? 1u : 0u;
        }
        bool wall = false;
        d3 fw = d3{0.0, 0.0, 0.0}, tw = d3{0.0, 0.0, 0.0};
        if (wflag[a]) {   // k_walls' result for this rebuild's wall candidates
            const double *WF = wforce + 8 * (size_t)a;
            double2 w0 = ld2(WF), w1 = ld2(WF + 2), w2 = ld2(WF + 4), w3 = ld2(WF + 6);
            if (w1.y != 0.0) {
                wall = true;
                fw = d3{w0.x, w0.y, w1.x};
                tw = d3{w2.x, w2.y, w3.x};
            }
        }
        double *R = reinterpret_cast<double *>(rec + a);
        flag = finish_particle<BATCHED, FLEN>(R, R, a, fx, fy, fz, tx, ty, tz, c, wall, fw, tw, inst,
                                        one, contacti, fdbg, tape_rec, cp32);
    }
    // rebuild trigger: one atomic per wave that saw a particle past the threshold (rare)
    if (__any(flag) && __lane_id() == 0) atomicOr(&ctl->need_rebuild, 1u);
}

// Law of pair (A first, B second) from the records' first 96 bytes.
template <bool BATCHED, bool FLEN = false>
__device__ __forceinline__ bool pair_from_records(const double *A, const double *B, d3 h,
                                                  const InstDev *inst, const LawParams &pp,
                                                  double dt, double roll_k, unsigned inst_a,
                                                  d3 &fa, d3 &ta, d3 &tb, d3 &ho, double &fo)
{
    double2 a0 = ld2(A), a1 = ld2(A + 2), a2 = ld2(A + 4), a3 = ld2(A + 6), a4 = ld2(A + 8),
            a5 = ld2(A + 10);
    double2 b0 = ld2(B), b1 = ld2(B + 2), b2 = ld2(B + 4), b3 = ld2(B + 6), b4 = ld2(B + 8),
            b5 = ld2(B + 10);
    LawParams P = pp;
    double tdt = dt, rk = roll_k;
    if (BATCHED) {
        const InstDev &I = inst[inst_a];
        P = I.pp;
        tdt = I.dt;
        rk = I.roll_k;
    }
    return law_pair(d3{a0.x, a0.y, a1.x}, d3{b0.x, b0.y, b1.x}, d3{a2.x, a2.y, a3.x},
                    d3{b2.x, b2.y, b3.x}, d3{a4.x, a4.y, a5.x}, d3{b4.x, b4.y, b5.x}, a1.y, b1.y,
                    a3.y, b3.y, h, tdt, P, fa, ta, tb, ho, fo, roll_floor(rk, tdt, a5.y, b5.y),
                    FLEN);
}

// Fused step (one lane per particle): the lane evaluates the law of each of its own pairs
// (it is first; it writes the new history) and of each pair where it is second (the same
// call with the same arguments, so the same bits as the owner's), summing in the reference
// order -- own pairs in slot order, then the others in slot order -- and then runs
// finish_particle.  No per-pair output buffer: each law is evaluated twice instead of its
// 80-byte result being written once and gathered twice.  Records and histories are read
// from rin / hin and written to rout / hout, so no lane sees a neighbour's updated state.
// Wall candidates of the last rebuild evaluate their wall contact here too (wall history
// per slot, in place), before the pair work, while few registers are live.
//
// Contact filter first: the partners' positions and radii come from the 16-byte FP32 copy
// cin (eight slots per cache line), several candidates at a time (independent loads in
// flight together), through maybe_touch, which is false only for spheres certainly apart;
// those pairs get the all-zero law result (history reset) at once.  The FP64 law then runs
// for every other pair, in order, reading the partner's record and the history: its own
// contact test decides, and a pair it finds out of contact gets the same all-zero result.  A
// pair out of contact contributes +0.0 (or -0.0) to every sum; the running sums start at +0.0
// and so are never -0.0, for which x + 0.0 == x, x + -0.0 == x and x - 0.0 == x bit for bit:
// skipping those terms or adding them keeps the reference order's result.
//
// A lane with at most kFast candidates on each side filters both sides in one pass (own and
// transposed loads in flight together) and its laws are balanced over the wave (below).
// Lanes with more candidates go through chunks of kChunk per side, own side first.
constexpr unsigned kChunk = 32;
constexpr unsigned kFast = 64;
constexpr unsigned kLook = 4;
// The block's own records are staged through LDS: read from rin and written to rout in
// 16-byte chunks by consecutive lanes (1 KB of contiguous memory per wave instruction instead
// of 64 scattered lines), each lane working on its row in between.  A row is the 128-byte
// record plus 16 bytes of padding, which spreads the lanes' 16-byte row accesses over all
// LDS banks.  Dynamic LDS: blockDim.x * kRow doubles, then blockDim.x words (k_step_lds).
constexpr unsigned kRow = 18;
// Dynamic LDS bytes of k_step for blocks of bd particles: the record rows and the zs words.
inline size_t k_step_lds(unsigned bd) { return (size_t)bd * (kRow * sizeof(double) + sizeof(unsigned)); }
// Own pairs of a block out of contact get their zero history in one cooperative pass over the
// block's contiguous own-pair range (consecutive lanes on consecutive pairs), steered by a
// contact bitmap in LDS; a block with more than kBits own pairs zeroes them lane by lane.
// zs[p] bit k records that history buffer k holds an exact zero for pair p.  zmode: bits 0-1
// the bit of the buffer written now (hout), bit 2 set when zs is valid -- then a pair out of
// contact whose bit is already set is not rewritten (most pairs of a loose bed stay out of
// contact for many steps); otherwise every pair out of contact is written and zs rebuilt.
// Both history buffers therefore always hold the exact history of every pair.
constexpr unsigned kBits = 8192;

__device__ __forceinline__ void zero_hist(Hist *h, unsigned p)
{
    double *HW = reinterpret_cast<double *>(h + p);
    st2(HW, 0.0, 0.0);
    st2(HW + 2, 0.0, 0.0);
}

// zs update of pair p for this step (con: in contact), given its current zs byte zc; writes
// the zero history if needed.
__device__ __forceinline__ void zs_pair_z(unsigned char *zs, Hist *hout, unsigned p, bool con,
                                          unsigned zmode, unsigned zc)
{
    const unsigned wbit = zmode & 3u;
    const bool lazy = (zmode & 4u) != 0u;
    const unsigned z = lazy ? zc : 0u;
    unsigned nz;
    if (con) {
        nz = z & ~wbit;
    } else {
        if (!(z & wbit)) zero_hist(hout, p);
        nz = z | wbit;
    }
    if (!lazy || nz != z) zs[p] = (unsigned char)nz;
}

__device__ __forceinline__ void zs_pair(unsigned char *zs, Hist *hout, unsigned p, bool con,
                                        unsigned zmode)
{
    zs_pair_z(zs, hout, p, con, zmode, (zmode & 4u) ? zs[p] : 0u);
}

// Contact-test outcome of own pair p (index k in the block's own-pair range): recorded in the
// block's bitmap (cooperative zeroing later) or applied to zs / the history at once.
__device__ __forceinline__ void own_outcome(bool coop, unsigned *cbits, unsigned k,
                                            unsigned char *zs, Hist *hout, unsigned p, bool con,
                                            unsigned zmode)
{
    if (coop) {
        if (con) atomicOr(&cbits[k >> 5], 1u << (k & 31u));
    } else {
        zs_pair(zs, hout, p, con, zmode);
    }
}

// Pops the next contact of a lane (own pairs first, in order, then transposed ones): its
// side, pair index and partner slot.
__device__ __forceinline__ void next_contact(unsigned long long &mo, unsigned long long &mt,
                                             unsigned p0, unsigned q0, const uint2 *ab,
                                             const unsigned *jt, const unsigned *jf, bool &own,
                                             unsigned &pp, unsigned &pb)
{
    own = mo != 0ull;
    if (own) {
        pp = p0 + (unsigned)__builtin_ctzll(mo);
        mo &= mo - 1ull;
        pb = ab[pp].y;
    } else {
        const unsigned k = q0 + (unsigned)__builtin_ctzll(mt);
        mt &= mt - 1ull;
        pp = jt[k];
        pb = jf[k];
    }
}

// Position of the k-th set bit (k counted from 0) of m, which has more than k set bits.
__device__ __forceinline__ unsigned nth_bit(unsigned long long m, unsigned k)
{
    unsigned pos = 0;
#pragma unroll
    for (unsigned w = 32; w; w >>= 1) {
        const unsigned cnt = (unsigned)__popcll((m >> pos) & ((1ull << w) - 1ull));
        if (k >= cnt) {
            k -= cnt;
            pos += w;
        }
    }
    return pos;
}

// Cross-lane reads of 64-bit values (two 32-bit lane permutes).
__device__ __forceinline__ unsigned long long shfl64(unsigned long long v, unsigned src)
{
    const unsigned lo = (unsigned)__shfl((int)(unsigned)v, (int)src, 64);
    const unsigned hi = (unsigned)__shfl((int)(unsigned)(v >> 32), (int)src, 64);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ double shfl_d(double v, int src)
{
    return __longlong_as_double((long long)shfl64((unsigned long long)__double_as_longlong(v),
                                                  (unsigned)src));
}

__device__ __forceinline__ d3 sel3(bool c, d3 u, d3 v)
{
    return d3{c ? u.x : v.x, c ? u.y : v.y, c ? u.z : v.z};
}

template <bool BATCHED, bool FLEN = false>
__global__ __launch_bounds__(kBlock, 4) void k_step(
    const PRec *rin, PRec *rout, const unsigned *bo, const unsigned *off, const unsigned *toff,
    const unsigned *jt, const unsigned *jf, const Hist *hin, Hist *hout, const unsigned *wflag,
    Hist *whist, const float4 *cin, float4 *cout, unsigned char *zs, unsigned zmode,
    unsigned n, const InstDev *inst, InstDev one, unsigned *contacti, double *fdbg, DevCtl *ctl,
    PRec *tape_rec, Hist *tape_hist, Hist *tape_whist, const unsigned *gate, unsigned fast_max,
    unsigned xchunk, unsigned *trig, unsigned *slot)
{
    extern __shared__ double lrec[];
    __shared__ unsigned cbits[kBits / 32];
    // one zs word per lane, after the record rows in the dynamic allocation (sized by the
    // block, so a 128-particle block takes 19,968 bytes and 8 blocks fit a CU's 160 KB)
    unsigned *zsl = reinterpret_cast<unsigned *>(lrec + blockDim.x * kRow);
    // Every load of the prologue is in bounds whether or not the step is gated off and
    // whatever the lane, so all of them are issued before any of their values is used: the
    // gate, the list offsets, the zs words and the block's records are in flight together.
    const unsigned gv = *gate;   // never null (a zero word for steps that are not queued)
    const unsigned np = ctl->npairs;
    const unsigned bd = blockDim.x, t = threadIdx.x;
    const unsigned s0 = xcd_chunk_block(blockIdx.x, gridDim.x, xchunk) * bd;
    const unsigned nb = s0 < n ? min(bd, n - s0) : 0u;
    const unsigned a = s0 + t;
    const bool live = t < nb;
    // the lane's list offsets and wall flag (off and toff hold n + 1 entries)
    const unsigned ai = live ? a : n;
    unsigned p0 = off[ai], p1 = off[ai + (live ? 1u : 0u)];
    unsigned q0 = toff[ai], q1 = toff[ai + (live ? 1u : 0u)];
    const unsigned wfa = wflag[live ? a : 0u];   // n >= 1 when the kernel runs
    const unsigned wf = live ? wfa : 0u;
    // the block's own pairs [pb0, pb1)
    const unsigned ob0 = off[min(s0, n)], ob1 = off[min(s0 + nb, n)];
    // the block's records: up to 8 16-byte chunks per lane, all loads in flight at once
    const double *G = reinterpret_cast<const double *>(rin + s0);
    double2 v[8];
#pragma unroll
    for (unsigned i = 0; i < 8; i++) {
        const unsigned k = t + i * bd;
        if (k < 8u * nb) v[i] = ld2(G + 2 * k);
    }
    const unsigned pb1 = nb ? min(ob1, np) : 0u;
    const unsigned pb0 = nb ? min(ob0, pb1) : 0u;
    const bool coop = pb1 - pb0 <= kBits;
    // zs bytes of the block's own pairs, one 4-byte word per lane from word pb0 / 4 (the
    // allocation is padded to whole words): the cooperative pass at the end reads them from
    // LDS; pairs past the words covered read zs directly
    const unsigned zw0 = pb0 >> 2;
    const bool zin = coop && zw0 + t < ((pb1 + 3u) >> 2);
    unsigned zword = 0;
    if (zin) zword = reinterpret_cast<const unsigned *>(zs)[zw0 + t];
    for (unsigned k = t; k < kBits / 32; k += bd) cbits[k] = 0u;
#pragma unroll
    for (unsigned i = 0; i < 8; i++) {
        const unsigned k = t + i * bd;
        if (k < 8u * nb) st2(lrec + (k >> 3) * kRow + 2 * (k & 7u), v[i].x, v[i].y);
    }
    if (zin) zsl[t] = zword;
    if (gv != 0u) {   // a gated-off step: uniform over the grid, no lane reaches the barrier
        if (blockIdx.x == 0 && t == 0) *trig = 1u;   // the gate carries on to the next step
        return;
    }
    __syncthreads();
    double *RA = lrec + t * kRow;
    unsigned flag = 0;
    // per-lane state of the pair work (zero on lanes that are not live)
    int hits = 0;
    d3 fw = d3{0.0, 0.0, 0.0}, tw = d3{0.0, 0.0, 0.0};
    double fx = 0.0, fy = 0.0, fz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
    unsigned c = 0;
    unsigned long long mo = 0ull, mt = 0ull;   // contacts found by the one-pass tests
    bool fast = false;
    if (live) {
        const unsigned inst_a = BATCHED ? reinterpret_cast<const uint2 *>(RA + 15)->x : 0u;
        const InstDev &IA = BATCHED ? inst[inst_a] : one;
        // wall contact first (its law needs many registers; few are live here)
        if (wf) hits = wall_slot<BATCHED, FLEN>(RA, IA, whist, a, tape_whist, fw, tw);
        if (p1 > np) p1 = np;
        if (q1 > np) q1 = np;
        const unsigned no = p1 > p0 ? p1 - p0 : 0u, nt = q1 > q0 ? q1 - q0 : 0u;
        if (tape_hist)   // the incoming histories of the own pairs, for the adjoint
            for (unsigned p = p0; p < p1; p++) {
                const double *H = reinterpret_cast<const double *>(hin + p);
                double *T = reinterpret_cast<double *>(tape_hist + p);
                double2 h0 = ld2(H), h1 = ld2(H + 2);
                st2(T, h0.x, h0.y);
                st2(T + 2, h1.x, h1.y);
            }
        fast = no <= fast_max && nt <= fast_max;
        if (fast) {
            // both sides' contact filters in one pass
            double2 a0 = ld2(RA), a1 = ld2(RA + 2);
            const float4 xa = pos32(a0.x, a0.y, a1.x, a1.y);
            const unsigned m = no > nt ? no : nt;
            // software-pipelined: the partner slots of group i + kLook are loaded while the
            // positions of group i are in flight, so a group waits for one load, not two
            // (the pair index jt is read by the law)
            unsigned b[kLook], f[kLook];
#pragma unroll
            for (unsigned k = 0; k < kLook; k++) {
                b[k] = k < no ? bo[p0 + k] : a;
                f[k] = k < nt ? jf[q0 + k] : a;
            }
            for (unsigned i = 0; i < m; i += kLook) {
                float4 cb[kLook], cf[kLook];
#pragma unroll
                for (unsigned k = 0; k < kLook; k++) {
                    cb[k] = cin[b[k]];
                    cf[k] = cin[f[k]];
                }
                const unsigned i2 = i + kLook;
#pragma unroll
                for (unsigned k = 0; k < kLook; k++) {
                    b[k] = i2 + k < no ? bo[p0 + i2 + k] : a;
                    f[k] = i2 + k < nt ? jf[q0 + i2 + k] : a;
                }
#pragma unroll
                for (unsigned k = 0; k < kLook; k++) {
                    if (i + k < no) {
                        const bool con = maybe_touch(xa, cb[k]);
                        if (con) mo |= 1ull << (i + k);
                        own_outcome(coop, cbits, p0 + i + k - pb0, zs, hout, p0 + i + k, con, zmode);
                    }
                    if (i + k < nt && maybe_touch(cf[k], xa)) mt |= 1ull << (i + k);
                }
            }
        }
    }
    // The laws of the one-pass lanes' contacts, balanced over the wave: contact g of the
    // wave (lanes in order, each lane's own pairs in order, then its transposed ones) is
    // evaluated in round g / 64 by lane g % 64, whatever lane owns it; each owner then adds
    // its results in contact order, so its sums are the reference order's bit for bit.  A
    // wave thus runs ceil(total / 64) laws instead of as many as its busiest lane has
    // contacts.  Every lane takes part (uniform control flow for the cross-lane reads).
    {
        const unsigned lane = __lane_id();
        const unsigned z = (unsigned)(__popcll(mo) + __popcll(mt));
        unsigned ex = z;   // inclusive scan over the wave, then exclusive
#pragma unroll
        for (unsigned d = 1; d < 64; d <<= 1) {
            const unsigned y = (unsigned)__shfl_up((int)ex, d, 64);
            if (lane >= d) ex += y;
        }
        const unsigned total = (unsigned)__builtin_amdgcn_readlane((int)ex, 63);
        ex -= z;
        const unsigned wbase = t - lane;   // the wave's first lane in the block
        for (unsigned r0 = 0; r0 < total; r0 += 64u) {
            const unsigned g = r0 + lane;
            // owner: the last lane whose first contact index is <= g
            unsigned L = 0;
#pragma unroll
            for (unsigned st = 32; st; st >>= 1) {
                const unsigned e = (unsigned)__shfl((int)ex, (int)(L + st), 64);
                if (e <= g) L += st;
            }
            const unsigned exL = (unsigned)__shfl((int)ex, (int)L, 64);
            const unsigned long long moL = shfl64(mo, L), mtL = shfl64(mt, L);
            const unsigned p0L = (unsigned)__shfl((int)p0, (int)L, 64);
            const unsigned q0L = (unsigned)__shfl((int)q0, (int)L, 64);
            const bool work = g < total;
            double rfx = 0.0, rfy = 0.0, rfz = 0.0, rtx = 0.0, rty = 0.0, rtz = 0.0;
            bool rc = false;
            if (work) {
                const unsigned j = g - exL;
                const unsigned co = (unsigned)__popcll(moL);
                const bool own = j < co;
                const unsigned idx = nth_bit(own ? moL : mtL, own ? j : j - co);
                unsigned pp, pb;
                if (own) {
                    pp = p0L + idx;
                    pb = bo[pp];
                } else {
                    pp = jt[q0L + idx];
                    pb = jf[q0L + idx];
                }
                // a transposed entry without a pair (jt past the list) takes no part
                const bool valid = pp < np;
                if (!valid) pp = 0u;
                const double *S = lrec + (wbase + L) * kRow;   // the owner's record
                const double *H = reinterpret_cast<const double *>(hin + pp);
                double2 h0 = ld2(H), h1 = ld2(H + 2);
                // the partner's record: from the block's LDS rows when the partner is one of
                // the block's particles (the same values as in rin: rows are updated only
                // after the barrier that ends the pair work), else from rin, through one
                // generic pointer so that both cases are the same loads with no branch.  The
                // batched form keeps the global read (its instantiation measured 25% slower
                // reading partners from LDS, while the plain one gains 4% on a poured bed and
                // nothing elsewhere).
                const double *B = (!BATCHED && pb - s0 < nb)
                                      ? static_cast<const double *>(lrec + (pb - s0) * kRow)
                                      : reinterpret_cast<const double *>(rin + pb);
                double2 b0 = ld2(B), b1 = ld2(B + 2), b2 = ld2(B + 4), b3 = ld2(B + 6),
                        b4 = ld2(B + 8), b5 = ld2(B + 10);
                double2 s0v = ld2(S), s1 = ld2(S + 2), s2 = ld2(S + 4), s3 = ld2(S + 6),
                        s4 = ld2(S + 8), s5 = ld2(S + 10);
                const d3 xs = d3{s0v.x, s0v.y, s1.x}, vs = d3{s2.x, s2.y, s3.x},
                         ws = d3{s4.x, s4.y, s5.x};
                const d3 xp = d3{b0.x, b0.y, b1.x}, vp = d3{b2.x, b2.y, b3.x},
                         wp = d3{b4.x, b4.y, b5.x};
                const double rs = s1.y, ms = s3.y, rp = b1.y, mp = b3.y;
                LawParams P = one.pp;
                double tdt = one.dt, rk = one.roll_k;
                if (BATCHED) {
                    const InstDev &IO = inst[reinterpret_cast<const uint2 *>(S + 15)->x];
                    P = IO.pp;
                    tdt = IO.dt;
                    rk = IO.roll_k;
                }
                d3 fa, ta, tb, ho;
                double fo;
                // first member = lower particle index: the owner on its own pairs
                rc = law_pair(sel3(own, xs, xp), sel3(own, xp, xs), sel3(own, vs, vp),
                              sel3(own, vp, vs), sel3(own, ws, wp), sel3(own, wp, ws),
                              own ? rs : rp, own ? rp : rs, own ? ms : mp, own ? mp : ms,
                              d3{h0.x, h0.y, h1.x}, tdt, P, fa, ta, tb, ho, fo,
                              roll_floor(rk, tdt, own ? s5.y : b5.y, own ? b5.y : s5.y),
                              FLEN);
                if (!valid) {
                    rc = false;
                    fa = ta = tb = d3{0.0, 0.0, 0.0};
                }
                if (own) {
                    double *HW = reinterpret_cast<double *>(hout + pp);
                    st2(HW, ho.x, ho.y);
                    st2(HW + 2, ho.z, fo);
                }
                // the owner's terms: +F and the first member's torque on its own pairs,
                // -F (x - y == x + (-y) bit for bit) and the second member's torque otherwise
                rfx = own ? fa.x : -fa.x;
                rfy = own ? fa.y : -fa.y;
                rfz = own ? fa.z : -fa.z;
                rtx = own ? ta.x : tb.x;
                rty = own ? ta.y : tb.y;
                rtz = own ? ta.z : tb.z;
            }
            // each lane collects its contacts of this round, in order
            const unsigned lo = ex > r0 ? ex : r0;
            const unsigned hi = min(ex + z, r0 + 64u);
            for (unsigned k = lo; __any(k < hi); k++) {
                const bool mine = k < hi;
                const int src = mine ? (int)(k - r0) : (int)lane;
                const double vfx = shfl_d(rfx, src), vfy = shfl_d(rfy, src), vfz = shfl_d(rfz, src);
                const double vtx = shfl_d(rtx, src), vty = shfl_d(rty, src), vtz = shfl_d(rtz, src);
                const int vc = __shfl((int)rc, src, 64);
                if (mine) {
                    fx += vfx; fy += vfy; fz += vfz;
                    tx += vtx; ty += vty; tz += vtz;
                    c += vc ? 1u : 0u;
                }
            }
        }
    }
    if (live && !fast) {
        const unsigned inst_a = BATCHED ? reinterpret_cast<const uint2 *>(RA + 15)->x : 0u;
        double2 a0 = ld2(RA), a1 = ld2(RA + 2);
        const float4 xa = pos32(a0.x, a0.y, a1.x, a1.y);
        for (unsigned base = p0; base < p1; base += kChunk) {
            const unsigned end = min(base + kChunk, p1);
            unsigned mask = 0;
            for (unsigned p = base; p < end; p += kLook) {
                unsigned b[kLook];
                float4 cb[kLook];
#pragma unroll
                for (unsigned i = 0; i < kLook; i++) b[i] = p + i < end ? bo[p + i] : a;
#pragma unroll
                for (unsigned i = 0; i < kLook; i++) cb[i] = cin[b[i]];
#pragma unroll
                for (unsigned i = 0; i < kLook; i++) {
                    if (p + i >= end) break;
                    const bool con = maybe_touch(xa, cb[i]);
                    if (con) mask |= 1u << (p + i - base);
                    own_outcome(coop, cbits, p + i - pb0, zs, hout, p + i, con, zmode);
                }
            }
            while (mask) {
                const unsigned p = base + (unsigned)__builtin_ctz(mask);
                mask &= mask - 1u;
                const double *B = reinterpret_cast<const double *>(rin + bo[p]);
                const double *H = reinterpret_cast<const double *>(hin + p);
                double2 h0 = ld2(H), h1 = ld2(H + 2);
                d3 fa, ta, tb, ho;
                double fo;
                bool con = pair_from_records<BATCHED, FLEN>(RA, B, d3{h0.x, h0.y, h1.x}, inst,
                                                      one.pp, one.dt, one.roll_k, inst_a, fa, ta, tb, ho,
                                                      fo);
                double *HW = reinterpret_cast<double *>(hout + p);
                st2(HW, ho.x, ho.y);
                st2(HW + 2, ho.z, fo);
                fx += fa.x; fy += fa.y; fz += fa.z;
                tx += ta.x; ty += ta.y; tz += ta.z;
                c += con ? 1u : 0u;
            }
        }
        for (unsigned base = q0; base < q1; base += kChunk) {
            const unsigned end = min(base + kChunk, q1);
            unsigned mask = 0;
            for (unsigned k = base; k < end; k += kLook) {
                unsigned f[kLook];
                float4 cf[kLook];
#pragma unroll
                for (unsigned i = 0; i < kLook; i++) f[i] = k + i < end ? jf[k + i] : a;
#pragma unroll
                for (unsigned i = 0; i < kLook; i++) cf[i] = cin[f[i]];
#pragma unroll
                for (unsigned i = 0; i < kLook; i++)
                    if (k + i < end && maybe_touch(cf[i], xa)) mask |= 1u << (k + i - base);
            }
            while (mask) {
                const unsigned k = base + (unsigned)__builtin_ctz(mask);
                mask &= mask - 1u;
                const unsigned q = jt[k];
                if (q >= np) continue;   // a transposed entry without a pair takes no part
                const double *F = reinterpret_cast<const double *>(rin + jf[k]);
                const double *H = reinterpret_cast<const double *>(hin + q);
                double2 h0 = ld2(H), h1 = ld2(H + 2);
                d3 fa, ta, tb, ho;
                double fo;
                bool con = pair_from_records<BATCHED, FLEN>(F, RA, d3{h0.x, h0.y, h1.x}, inst,
                                                      one.pp, one.dt, one.roll_k, inst_a, fa, ta, tb, ho,
                                                      fo);
                fx -= fa.x; fy -= fa.y; fz -= fa.z;
                tx += tb.x; ty += tb.y; tz += tb.z;
                c += con ? 1u : 0u;
            }
        }
    }
    // every wave has read its partners' LDS rows before any lane updates its own row
    if (!BATCHED) __syncthreads();
    if (live)
        flag = finish_particle<BATCHED, FLEN>(RA, RA, a, fx, fy, fz, tx, ty, tz, c, hits != 0, fw, tw,
                                        inst, one, contacti, fdbg, tape_rec, cout);
    __syncthreads();
    {
        double *GW = reinterpret_cast<double *>(rout + s0);
        double2 u[8];
#pragma unroll
        for (unsigned i = 0; i < 8; i++) {
            const unsigned k = t + i * bd;
            u[i] = k < 8u * nb ? ld2(lrec + (k >> 3) * kRow + 2 * (k & 7u)) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (unsigned i = 0; i < 8; i++) {
            const unsigned k = t + i * bd;
            if (k < 8u * nb) st2(GW + 2 * k, u[i].x, u[i].y);
        }
    }
    if (coop) {
        const unsigned zend = (zw0 + bd) << 2;   // first pair past the words in zsl
        for (unsigned k = t; k < pb1 - pb0; k += bd) {
            const unsigned p = pb0 + k;
            const unsigned z = p < zend ? (zsl[(p >> 2) - zw0] >> (8u * (p & 3u))) & 0xffu : zs[p];
            zs_pair_z(zs, hout, p, (cbits[k >> 5] >> (k & 31u)) & 1u, zmode, z);
        }
    }
    // rebuild trigger: one atomic per wave that saw a particle past the threshold (rare); a
    // queued step also reports it in its host-mapped outcome slot
    if (__any(flag) && __lane_id() == 0) {
        atomicOr(trig, 1u);
        if (slot) *reinterpret_cast<volatile unsigned *>(slot) = 1u;
    }
}

// A bounded busy period of `ticks` wall-clock ticks (one lane; the clock runs at a fixed
// rate, so the loop always ends).
__global__ void k_delay(unsigned long long ticks)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(2);
}

// Publish the control block to the host-mapped mirror (one lane; after the step kernels).
// With a queue slot, also reports the step's outcome there (0 ran, 1 ran and raised the
// trigger); a gated-off step leaves the slot as the host set it.
__global__ void k_publish(DevCtl *ctl, unsigned char *mirror, const unsigned *gate,
                          unsigned *slot)
{
    if (gated_off(gate)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->steps = ctl->steps + 1u;
        ctl->skip = ctl->need_rebuild;
        if (slot) *reinterpret_cast<volatile unsigned *>(slot) = ctl->need_rebuild;
        volatile unsigned *src = reinterpret_cast<volatile unsigned *>(ctl);
        unsigned *dst = reinterpret_cast<unsigned *>(mirror);
        for (unsigned k = 0; k < sizeof(DevCtl) / 4; k++) dst[k] = src[k];
    }
}

// Per-instance max(0, max z) and contact totals (the reference's host max over
// updatepositionD's block partials, particleSystem.cu:372-379, and outputporositycn).  Grid-
// stride over slots; each wave reduces its lanes of one instance and adds them with one
// atomic per quantity.  Max of non-negative doubles as ordered bit patterns and integer sums
// are exact, so the result does not depend on the order.  `acc` (u64[ninst] max-z bits, then
// u32[ninst] totals) must be zero on entry.
__global__ __launch_bounds__(kBlock) void k_inst_reduce(const PRec *rec, const unsigned *contacti,
                                                        unsigned n, unsigned long long *acc_z,
                                                        unsigned *acc_c, unsigned single)
{
    if (single) {
        // one instance: each lane folds its share of the grid-stride loop, then the block
        // reduces through cross-lane reads and LDS and adds one atomic pair (atomics on one
        // address serialize, so one pair per wave per iteration took ~180 us at 500K)
        __shared__ unsigned long long rz[kBlock / 64];
        __shared__ unsigned rc[kBlock / 64];
        unsigned long long zm = 0ull;
        unsigned cs = 0;
        for (unsigned s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
            const double z = rec[s].pz;
            const unsigned long long zb = (unsigned long long)__double_as_longlong(z > 0.0 ? z : 0.0);
            zm = zb > zm ? zb : zm;
            cs += contacti[s];
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const unsigned long long oz = __shfl_xor(zm, d, 64);
            zm = oz > zm ? oz : zm;
            cs += __shfl_xor(cs, d, 64);
        }
        if (__lane_id() == 0) {
            rz[threadIdx.x >> 6] = zm;
            rc[threadIdx.x >> 6] = cs;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (unsigned w = 1; w < kBlock / 64; w++) {
                zm = rz[w] > zm ? rz[w] : zm;
                cs += rc[w];
            }
            atomicMax(&acc_z[0], zm);
            atomicAdd(&acc_c[0], cs);
        }
        return;
    }
    for (unsigned base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        unsigned s = base + threadIdx.x;
        bool live = s < n;
        unsigned inst = live ? rec[s].inst : 0u;
        unsigned long long zb = 0ull;
        unsigned cc = 0;
        if (live) {
            double z = rec[s].pz;
            zb = (unsigned long long)__double_as_longlong(z > 0.0 ? z : 0.0);
            cc = contacti[s];
        }
        unsigned inst0 = __builtin_amdgcn_readfirstlane(inst);
        if (__all(!live || inst == inst0)) {
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                unsigned long long oz = __shfl_xor(zb, d, 64);
                zb = oz > zb ? oz : zb;
                cc += __shfl_xor(cc, d, 64);
            }
            if (__lane_id() == 0 && __any(live)) {
                atomicMax(&acc_z[inst0], zb);
                atomicAdd(&acc_c[inst0], cc);
            }
        } else if (live) {
            atomicMax(&acc_z[inst], zb);
            atomicAdd(&acc_c[inst], cc);
        }
    }
}

// ----------------------------------------------------------------------------------------
// Reference-layout kernels (drop-in entry points)
// ----------------------------------------------------------------------------------------

struct CompatGrid {
    unsigned g[3];
    double cell[3], origin[3];
};

__device__ __forceinline__ unsigned compat_hash(const double *p, const CompatGrid &G)
{
    unsigned x = cell_coord(p[0], G.origin[0], G.cell[0], G.g[0]);
    unsigned y = cell_coord(p[1], G.origin[1], G.cell[1], G.g[1]);
    unsigned z = cell_coord(p[2], G.origin[2], G.cell[2], G.g[2]);
    return z * G.g[1] * G.g[0] + y * G.g[0] + x;
}

__global__ __launch_bounds__(kBlock) void c_calc_hash(const double *pos, unsigned n, CompatGrid G,
                                                      uint2 *hash)
{
    unsigned i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    hash[i] = make_uint2(compat_hash(pos + 3 * (size_t)i, G), i);
}

// neighborarrayD (absent; call particleSystem.cu:85-96): candidates |xj - xi| < cub for
// unit-diameter spheres, in stencil order (rows z, then y ascending; in a row the cells x - 1,
// x, x + 1; in a cell its sorted entries), written into particle i's row of `width` slots;
// njgi / njli count the candidates with a larger / smaller index.  The search runs in sorted
// order: c_sorted_copy gathers the positions into sorted order once, c_cell_count and a scan
// give every cell's slot range, and c_neighbor_search stages each block's stencil rows in LDS
// (for_candidates_block) instead of gathering every candidate from the particle-index-ordered
// positions.

// row[from .. width) = v: 4-byte stores up to a 16-byte boundary, then 16-byte stores (a
// quarter of the store instructions for the tail of a mostly empty row), then the rest.
__device__ __forceinline__ void fill_row_tail(unsigned *row, unsigned from, unsigned width,
                                              unsigned v)
{
    unsigned m = from;
    while (m < width && ((reinterpret_cast<uintptr_t>(row + m) & 15u) != 0u)) row[m++] = v;
    for (; m + 4u <= width; m += 4u) *reinterpret_cast<uint4 *>(row + m) = make_uint4(v, v, v, v);
    for (; m < width; m++) row[m] = v;
}

// spos[s] = (position of the particle in sorted slot s, radius 0.5), sid[s] = its index.
__global__ __launch_bounds__(kBlock) void c_sorted_copy(const double *pos, const uint2 *sorted,
                                                        unsigned n, double4 *spos, unsigned *sid)
{
    const unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const unsigned i = sorted[s].y;
    const double *p = pos + 3 * (size_t)i;
    spos[s] = make_double4(p[0], p[1], p[2], 0.5);
    sid[s] = i;
}

// Particles per cell of the sorted hash (cnt zeroed beforehand); its exclusive scan gives
// every cell's first slot, cell c holding slots [pre[c], pre[c + 1]).
__global__ __launch_bounds__(kBlock) void c_cell_count(const uint2 *sorted, unsigned n,
                                                       unsigned ncells, unsigned *cnt)
{
    const unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    (void)run_count(cnt, min(sorted[s].x, ncells - 1u));
}

__global__ __launch_bounds__(kBlock) void c_neighbor_search(const double4 *spos,
                                                            const unsigned *sid,
                                                            const uint2 *sorted,
                                                            const unsigned *pre, unsigned n,
                                                            InstDev I, int width, unsigned *anei,
                                                            unsigned *njgi, unsigned *njli,
                                                            unsigned *overflow)
{
    __shared__ double4 tile[kTile + kWin];
    __shared__ unsigned tcid[kTile];
    __shared__ unsigned red[2 * (kBlock / 64)];
    const unsigned a = blockIdx.x * kBlock + threadIdx.x;   // sorted slot
    const bool live = a < n;
    const unsigned i = live ? sid[a] : 0u;
    int cnt = 0;
    unsigned ng = 0, nl = 0;
    for_candidates_block<false>(
        spos, sid, a, live, live ? sorted[a].x : 0u, I, pre, tile, tcid, red,
        [&](unsigned b, unsigned, unsigned, unsigned j) {
            if (b == a) return;   // the particle itself
            if (cnt < width) anei[(size_t)i * width + cnt] = j;
            cnt++;
            if (j > i) ng++;
            else nl++;
        },
        [&](int, unsigned, unsigned) {});
    if (!live) return;
    fill_row_tail(anei + (size_t)i * width, (unsigned)min(cnt, width), (unsigned)width, kEmpty);
    njgi[i] = ng;
    njli[i] = nl;
    if (cnt > width) atomicOr(overflow, 1u);
}

__global__ __launch_bounds__(kBlock) void c_add_arrays(unsigned *b, const unsigned *a, unsigned n)
{
    unsigned i = blockIdx.x * kBlock + threadIdx.x;   // addTwoArraysD, neiglist.cu:122-130
    if (i < n) b[i] += a[i];
}

// pairlistD (absent; call particleSystem.cu:172-190).  Aref row i holds i's own pair slots
// first (njgi entries, list order) and then the slots of pairs (j, i), j < i (njli entries,
// Anei order), so sumallforcei can tell the side of each entry from its position.
__global__ __launch_bounds__(kBlock) void c_pair_list(
    const unsigned *sjgi, const unsigned *oldsjgi, const unsigned *oldln, const unsigned *anei,
    const unsigned *njgi, const unsigned *njli, const double *olddispt, const double *oldfricp,
    int width, unsigned n, unsigned *aref, unsigned *lp, unsigned *ln, double *dispt,
    double *fricp)
{
    unsigned i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    unsigned cnt = njgi[i] + njli[i];
    unsigned start = sjgi[i] - njgi[i];
    unsigned k = 0, r = 0;
    unsigned os = i > 0 ? oldsjgi[i - 1] : 0u;
    unsigned oe = oldsjgi[i];
    // the old segment's partners, read once (kOldHeld of them held in registers; a longer
    // segment's rest is scanned in memory)
    constexpr unsigned kOldHeld = 16;
    unsigned ol[kOldHeld];
#pragma unroll
    for (unsigned t = 0; t < kOldHeld; t++) ol[t] = os + t < oe ? oldln[os + t] : kEmpty;
    // with cnt <= width every entry below cnt is written below (own slots, then transposed
    // marks); an overflowed row (neighborarray reported it) is cleared whole first
    fill_row_tail(aref + (size_t)i * width, cnt <= (unsigned)width ? cnt : 0u, (unsigned)width,
                  kEmpty);
    for (unsigned m = 0; m < cnt && m < (unsigned)width; m++) {
        unsigned j = anei[(size_t)i * width + m];
        if (j > i) {
            unsigned slot = start + k;
            lp[slot] = i;
            ln[slot] = j;
            aref[(size_t)i * width + k] = slot;
            k++;
            double h0 = 0.0, h1 = 0.0, h2 = 0.0, f = 0.0;
            // the first old entry of partner j, as a scan of the old segment finds it
            unsigned q = kEmpty;
#pragma unroll
            for (unsigned t = kOldHeld; t-- > 0;)
                if (ol[t] == j) q = os + t;
            if (q == kEmpty)
                for (unsigned t = os + kOldHeld; t < oe; t++)
                    if (oldln[t] == j) { q = t; break; }
            if (q != kEmpty) {
                h0 = olddispt[3 * (size_t)q];
                h1 = olddispt[3 * (size_t)q + 1];
                h2 = olddispt[3 * (size_t)q + 2];
                f = oldfricp[q];
            }
            dispt[3 * (size_t)slot] = h0;
            dispt[3 * (size_t)slot + 1] = h1;
            dispt[3 * (size_t)slot + 2] = h2;
            fricp[slot] = f;
        } else {
            // pair (j, i) sits in j's own segment of the new list, which j's lane fills with
            // its larger-index candidates in Anei order: its slot is the first in
            // [sjgi[j] - njgi[j], sjgi[j]) whose second member is i.  The segment is read
            // from the list the lanes write (Ln), so it is searched once the fill is done:
            // c_pair_list_partner does that in a second launch; here the entry is marked.
            unsigned pos = njgi[i] + r;
            if (pos < (unsigned)width) aref[(size_t)i * width + pos] = kEmpty - 1u - j;
            r++;
        }
    }
}

// Second pass of pairlistD: each entry (j, i), j < i, of Aref row i marked by c_pair_list
// (as kEmpty - 1 - j) gets the slot of pair (j, i) -- a search of j's own segment of Ln, a few
// contiguous entries (its larger-index candidates), instead of j's whole neighbour row.
__global__ __launch_bounds__(kBlock) void c_pair_list_partner(const unsigned *sjgi,
                                                              const unsigned *njgi,
                                                              const unsigned *njli, int width,
                                                              unsigned n, const unsigned *ln,
                                                              unsigned *aref)
{
    unsigned i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const unsigned g = njgi[i];
    const unsigned e = min(g + njli[i], (unsigned)width);
    for (unsigned m = g; m < e; m++) {
        const unsigned v = aref[(size_t)i * width + m];
        if (v == kEmpty || v < n) continue;
        const unsigned j = kEmpty - 1u - v;
        unsigned found = kEmpty;
        if (j < n) {
            const unsigned lo = sjgi[j] - njgi[j], hi = sjgi[j];
            for (unsigned q = lo; q < hi; q++)
                if (ln[q] == i) { found = q; break; }
        }
        aref[(size_t)i * width + m] = found;
    }
}

__device__ __forceinline__ d3 ld3(const double *p, size_t i)
{
    return d3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
__device__ __forceinline__ void st3(double *p, size_t i, d3 v)
{
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// calculateforcepairD (neiglist.cu:156-233).
__global__ __launch_bounds__(kBlock) void c_force_pair(
    unsigned np, const double *pos, const double *pdis, const double *angv, const double *rad,
    const double *rmass, const unsigned *lp, const unsigned *ln, unsigned *oldln, double *dispt,
    double *fricp, double *olddispt, double *oldfricp, double dt, LawParams P, double *forcepair,
    double *torqpairi, double *torqpairj, unsigned *contactpair)
{
    unsigned p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    unsigned i = lp[p], j = ln[p];
    // the contact test first, on the positions and radii alone (law_touch: the law's own test,
    // the same decision bit for bit); only a pair in contact gathers the displacements,
    // spins, masses and its history, and the others take the law's all-zero result
    const d3 xi = ld3(pos, i), xj = ld3(pos, j);
    const double ri = rad[i], rj = rad[j];
    d3 fa = d3{0.0, 0.0, 0.0}, ta = fa, tb = fa, ho = fa;
    double fo = 0.0;
    bool c = false;
    if (law_touch(xi, xj, ri, rj))
        c = law_pair(xi, xj, ld3(pdis, i), ld3(pdis, j), ld3(angv, i), ld3(angv, j), ri, rj,
                     rmass[i], rmass[j], ld3(dispt, p), dt, P, fa, ta, tb, ho, fo);
    st3(forcepair, p, fa);
    st3(torqpairi, p, ta);
    st3(torqpairj, p, tb);
    st3(dispt, p, ho);
    fricp[p] = fo;
    contactpair[p] = c ? 1u : 0u;
    st3(olddispt, p, ho);
    oldfricp[p] = fo;
    oldln[p] = j;
}

// sumallforceiD (absent; call particleSystem.cu:262-276): Aref row order, own pairs
// (+forcepair, torqpairi) then pairs where the particle is second (-forcepair, torqpairj).
__global__ __launch_bounds__(kBlock) void c_sum_forces(int n, int width, const unsigned *njgi,
                                                       const unsigned *njli,
                                                       const double *forcepair,
                                                       const double *torqpairi,
                                                       const double *torqpairj,
                                                       const unsigned *contactpair,
                                                       const unsigned *aref, double *forcei,
                                                       double *torquei, unsigned *contacti)
{
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
    unsigned c = 0;
    const unsigned g = njgi[i];
    const unsigned cnt = min(g + njli[i], (unsigned)width);
    // kSumAhead entries at a time: every load of a group in flight before the group's sums,
    // which then run in row order (x - y == x + (-y) bit for bit).  An entry whose contact
    // flag is 0 holds calculateforcepair's out-of-contact result, an exact zero force and
    // torques, so its force and torques are not read: the running sums start at +0.0 and
    // are never -0.0, and adding or subtracting a zero leaves them unchanged bit for bit (an
    // empty entry adds nothing either)
    constexpr unsigned kSumAhead = 8;
    for (unsigned m0 = 0; m0 < cnt; m0 += kSumAhead) {
        unsigned p[kSumAhead];
        d3 fp[kSumAhead], tp[kSumAhead];
        unsigned cp[kSumAhead];
#pragma unroll
        for (unsigned k = 0; k < kSumAhead; k++)
            p[k] = m0 + k < cnt ? aref[(size_t)i * width + m0 + k] : kEmpty;
#pragma unroll
        for (unsigned k = 0; k < kSumAhead; k++) cp[k] = p[k] != kEmpty ? contactpair[p[k]] : 0u;
#pragma unroll
        for (unsigned k = 0; k < kSumAhead; k++) {
            fp[k] = tp[k] = d3{0.0, 0.0, 0.0};
            if (cp[k]) {
                fp[k] = ld3(forcepair, p[k]);
                tp[k] = ld3(m0 + k < g ? torqpairi : torqpairj, p[k]);
            }
        }
#pragma unroll
        for (unsigned k = 0; k < kSumAhead; k++) {
            if (!cp[k]) continue;
            const bool own = m0 + k < g;
            f0 += own ? fp[k].x : -fp[k].x;
            f1 += own ? fp[k].y : -fp[k].y;
            f2 += own ? fp[k].z : -fp[k].z;
            t0 += tp[k].x;
            t1 += tp[k].y;
            t2 += tp[k].z;
            c += cp[k];
        }
    }
    st3(forcei, i, d3{f0, f1, f2});
    st3(torquei, i, d3{t0, t1, t2});
    contacti[i] = c;
}

// calculateforceiwD (neiglist.cu:238-306).
__global__ __launch_bounds__(kBlock) void c_wall_gravity(int n, WallGeom G, const double *pos,
                                                         const double *pdis, const double *angv,
                                                         const double *rad, const double *rmass,
                                                         double dt, unsigned np, LawParams P,
                                                         double *disptw, double *fricpw,
                                                         double *forcei, double *torquei)
{
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    d3 f = d3{0.0, 0.0, 0.0}, t = d3{0.0, 0.0, 0.0};
    if (np != 0u) { f = ld3(forcei, i); t = ld3(torquei, i); }
    double m = rmass[i];
    f.x = f.x + 0.0;
    f.y = f.y + 0.0;
    f.z = f.z + (-m);
    d3 fw, tw, hw;
    double fo;
    int hit = law_walls(G, ld3(pos, i), ld3(pdis, i), ld3(angv, i), rad[i], m, ld3(disptw, i),
                        dt, P, fw, tw, hw, fo);
    if (hit) {
        f = d3{f.x + fw.x, f.y + fw.y, f.z + fw.z};
        t = d3{t.x + tw.x, t.y + tw.y, t.z + tw.z};
        st3(disptw, i, hw);
        fricpw[i] = fo;
    } else {
        st3(disptw, i, d3{0.0, 0.0, 0.0});
        fricpw[i] = 0.0;
    }
    st3(forcei, i, f);
    st3(torquei, i, t);
}

// updatepositionD (neiglist.cu:308-409); the block's maximum height (as the bits of a
// non-negative double, whose order is the values' order) and its rebuild flag go to maxz and
// iall by one atomic each.
__global__ __launch_bounds__(kBlock) void c_update(int n, double *pos, double *pdis, double *angv,
                                                   const double *rmass, const double *inert,
                                                   const double *forcei, const double *torquei,
                                                   double *qi, double dt, double ann2,
                                                   unsigned *iall, unsigned long long *maxz)
{
    int i = blockIdx.x * kBlock + threadIdx.x;
    unsigned flag = 0;
    unsigned long long zb = 0ull;
    if (i < n) {
        double m = rmass[i], I = inert[i];
        d3 F = ld3(forcei, i), T = ld3(torquei, i);
        d3 v = ld3(pdis, i), x = ld3(pos, i), q = ld3(qi, i), w = ld3(angv, i);
        v.x = v.x + F.x * dt * dt / m;
        v.y = v.y + F.y * dt * dt / m;
        v.z = v.z + F.z * dt * dt / m;
        x = d3{x.x + v.x, x.y + v.y, x.z + v.z};
        q = d3{q.x + v.x, q.y + v.y, q.z + v.z};
        w.x = w.x + T.x * dt * dt / I;
        w.y = w.y + T.y * dt * dt / I;
        w.z = w.z + T.z * dt * dt / I;
        st3(pos, i, x);
        st3(pdis, i, v);
        st3(qi, i, q);
        st3(angv, i, w);
        double q2 = q.x * q.x + q.y * q.y + q.z * q.z;
        flag = q2 > ann2 ? 1u : 0u;
        double zz = x.z > 0.0 ? x.z : 0.0;
        zb = (unsigned long long)__double_as_longlong(zz);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        unsigned long long oz = __shfl_xor(zb, d, 64);
        zb = oz > zb ? oz : zb;
        flag |= __shfl_xor(flag, d, 64);
    }
    // one atomic per block, and none when the block cannot raise the maximum (the maximum
    // only grows, so a stale read only lets an atomic through)
    __shared__ unsigned long long bz[kBlock / 64];
    __shared__ unsigned bf[kBlock / 64];
    const unsigned w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        bz[w] = zb;
        bf[w] = flag;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (unsigned k = 1; k < blockDim.x / 64u; k++) {
            zb = bz[k] > zb ? bz[k] : zb;
            flag |= bf[k];
        }
        if (zb && zb > *reinterpret_cast<volatile unsigned long long *>(maxz)) atomicMax(maxz, zb);
        if (flag) atomicOr(iall, 1u);
    }
}

// outputporositycnD (absent; call particleSystem.cu:405-410): total contact count.
__global__ __launch_bounds__(kBlock) void c_contact_total(unsigned n, const unsigned *contacti,
                                                          unsigned *total)
{
    unsigned i = blockIdx.x * kBlock + threadIdx.x;
    unsigned c = i < n ? contacti[i] : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (__lane_id() == 0 && c) atomicAdd(total, c);
}

}  // namespace dem
