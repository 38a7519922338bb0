// Device-side data of the engine: the per-step descriptor the sweep kernels read, the control block that carries
// the sampling chain's state from step to step on the device, and the per-launch epilogue arguments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gs {

constexpr int kThreads = 256;       // threads per sweep block (4 waves of 64)
constexpr int kMaxGridLog2 = 14;    // at most 16384 blocks per sweep; beyond that each block loops over tiles
constexpr int kDefaultGridLog2 = 9;     // measured best for the width-30 sweeps (a few waves per SIMD suffice)
constexpr int kDefaultGridLog2One = 12; // plain sweeps of one-slot steps: one load per thread and tile, so more waves
                                        // (width 34, 1-slot norm passes: 691 -> 343 ms per sample at 2^12 blocks)
constexpr int kBigStepLog2 = 30;        // one-slot steps of at least 2^30 cosets take more blocks still:
constexpr int kDefaultGridLog2OneBig = 14;  // width 34: 1.231 -> 1.196 s per sample; width 30 prefers 2^12
constexpr int kDefaultTTGridLog2 = 14;  // gather sweeps of the two-table engine (latency-bound gathers: the
                                         // width-31 IBM sample 0.276 s at 2^9 blocks, 0.121 s at 2^14)
constexpr int kFenceBlocks = 64;    // sweeps of at most this many blocks reduce in their last block
constexpr int kPartStride = 8;      // doubles per block partial (write norms + fused next-step norms)
constexpr size_t kPartDoubles = size_t(kPartStride) << kMaxGridLog2;
constexpr int kCounters = 1;

// Canonical routing of the 2 * ntq terms of a step onto the outputs tau (after the host relabels the offsets):
//   kRoute1     m = 1: every term feeds tau 0
//   kRouteSplit m = 2: x = 0 terms feed tau 0, x = 1 terms feed tau 1 (growth, and q = 1 steps with x_v unmixed)
//   kRouteMixA  m = 2: offset j feeds tau (j & 1) with x = 0 and tau 1 - (j & 1) with x = 1
//   kRouteMixB  m = 2: offset j feeds tau (j & 1) with both x
enum Route { kRoute1 = 0, kRouteSplit = 1, kRouteMixA = 2, kRouteMixB = 3 };

// Sweep kinds.  kWriteGram / kWriteDirect store step k and, from the values just stored, reduce the norms of step
// k+1 (whose two slots per coset are both produced by one thread: the thread's cosets span the pairing vector).
enum Sweep { kGram = 0, kDirect = 1, kWrite = 2, kWriteGram = 3, kWriteDirect = 4,
             kWriteGram1 = 5 };   // kWriteGram1: a plain write folding in the Gram of a one-slot next step;
                                  // the engine launches it as its kWriteGram (same partials and chain logic)

// One step as the sweep kernels see it (device memory, read with uniform loads).  Coset coordinates are in the
// kernel's order: bits 0..lr-1 are unit physical bits (a run of consecutive slots), bit lr is the pairing coordinate
// when vg == 2, then the rest.
struct KStep {
    uint64_t cols[40];     // physical column of each coset coordinate
    uint64_t off[4];       // coset offsets, relabelled so that the routing is canonical
    uint64_t g;            // second output slot offset (m = 2)
    uint64_t pmask[4];     // parity masks over the coordinates: [0] output sign (rhoc), [1] next step's slot index,
                           // [2] next step's sign (fused sweeps), [3] next step's second slot-index bit (fused
                           // sweeps of a step followed by a four-slot step)
    double mre[2][2];      // M[b][x]
    double mim[2][2];
    uint32_t parmask;      // bit j: parity(qv_j & rho) of offset j
    int nk;                // number of coset coordinates
    int lr;                // log2 run length (cosets per thread along unit bits)
    int vg;                // 1, or 2 when coordinate lr is a per-thread coordinate too: the pairing coordinate of the
                           // next step (fused sweeps), or a second run per thread (plain sweeps of 1- and 2-slot steps);
                           // 4 when coordinates lr, lr+1 both pair the slots of a four-slot next step (fused sweeps)
    int lg_grid;           // log2 blocks
    int lg_iter;           // log2 tiles per block
    int ntq, m, route;
    int growth;            // q = 0 and m = 2 with x -> tau: branch norms follow from the table norm
    int route2;            // fused: next step's route
    int par2;              // fused: next step's par of its second offset
    int eps;               // fused: the partner of output tau is output tau ^ eps of the paired coset
    int eps2;              // fused, four-slot next step: the same for the second pairing coordinate
    int parmask2;          // fused, four-slot next step: its par bits of offsets 0..3
    int phig, nug;         // fused: slot-index (two bits for a four-slot next step) and sign parity of g
    int rev;               // walk the tiles of each block in reverse Gray order
    int xcd;               // give each XCD a contiguous eighth of the block coordinates (grids of 8 blocks or more)
    int sgg, sgk;          // kWriteGram1: next step's sign parity of g, and its constant (pmask[2]: per coordinate)
};

// One step of the two-table (gather-form) engine as its kernels see it: output cp of the new table gathers the inputs
// u = pdep(cp, pivmask) ^ kspan[t] of the old one (c = u & lowmask, x = the bit r of u).  pdep is linear in cp:
// output coordinate j goes to the physical bit dep[j].
struct TStep {
    uint64_t dep[40];      // 1 << (position of the j-th set bit of pivmask)
    uint64_t kspan[4];     // the 2^nker combinations of the kernel vectors
    uint64_t lowmask;      // (1 << r) - 1
    uint64_t xbit;         // 1 << r
    uint64_t rho;          // sign mask over c
    int rn;                // log2 outputs
    int nks;               // 2^nker
    int lg_grid;           // log2 blocks
    int lg_iter;           // log2 tiles per block
    // Partner outputs: cp and cp ^ pdelta gather the same entries (input address u ^ pD = u ^ xbit ^ kspan[ptd],
    // equal to u ^ kspan[ptd] below the x bit), so one thread takes both and loads them once; ph = the top bit of
    // pdelta, a tile bit, or -1 (no pairing).
    uint64_t pdelta, pD;
    int ph, ptd;
};

// A run of consecutive diagonal steps: one-slot, one-output steps with slot offset 0 over the same live table.  Step i
// of the run multiplies entry a by w0_i + (-1)^{s_i(a)} w1_i, s_i an affine parity of the address, so the run depends
// on an entry only through its pattern sigma = (s_0 .. s_{j-1}): one pass sums |v|^2 per pattern, one block makes all
// j decisions from those sums, one pass multiplies every entry by its pattern's product factor.  Coordinates: 0 the
// unit physical bit (one aligned run of 2 entries per thread), 1..8 the unit bits 2..256 (the threads), then the
// block coordinates, then the tile coordinates, chosen so that no tile coordinate changes sigma.  Every column but the
// first has bit 0 clear.
constexpr int kMaxRun = 8;              // steps per run: 2^8 sign patterns
constexpr int kRunBins = 1 << kMaxRun;
constexpr int kRunChunks = 64;          // first-stage reduction of the per-block pattern sums
struct RStep {
    uint64_t cols[40];                  // physical column of each coordinate
    uint8_t sig[40];                    // change of sigma (bit i: step i) when the coordinate is set
    double mre[kMaxRun][2][2];          // M[b][x] of each step of the run
    double mim[kMaxRun][2][2];
    int nk, j, k0;                      // coordinates, steps, first step
    int kappa;                          // sigma of the coordinate origin (bit i: par of step i)
    int lg_grid, lg_iter, xcd;
    // pair = 1: the next step is a two-slot step reading this table in pairs (x, x ^ o).  Tile coordinate 0 then has
    // column o (o ^ 1 when o is odd: element e of a run pairs with element e ^ 1 of the paired run, peps = 1) and both
    // passes walk it as an explicit pair of tiles; the product pass can sum the next step's Gram matrix, whose slot
    // index and sign are parities of the address (phim, num: bit c for coordinate c).
    int pair, peps, route2, par2;
    uint64_t phim, num;
};

// Chain state carried on the device between steps (one per engine).
struct Ctl {
    double prev;           // norm of the stored table (chain1 / run_steps bookkeeping)
    double nstored;        // sum of |stored entry|^2 of the rounded stored table (growth-step norms)
    double logp;           // chain: sum log2 p_bv; chain1: log2 Pr at the end; run: sum log2 cond
    double logscale;       // run_steps: stored table = true table * 2^-logscale
    double sc;             // scale of the pending write
    double w[4];           // M[bv][0]*sc, M[bv][1]*sc as re, im pairs (read by the write sweep)
    int bv;                // decision of the pending write
    int dead;              // a zero-probability branch was hit: later kernels do nothing
    int qidx;              // index of the next pre-drawn uniform (run_steps mode)
    int pad;
};

// Modes of the chain (how norms become a branch and bookkeeping).
enum ChainMode { kChain = 0, kChain1 = 1, kRun = 2 };

// What the last block of a sweep does after the sweep (kRoleStored: record the stored table's norm only -- the
// product pass of a run of diagonal steps, whose decisions and bookkeeping were made before it).
// kRoleStoredGram: the same, then the next step's decision from the Gram matrix the product pass summed.
enum EpiRole { kRoleNorm = 0, kRoleWrite = 1, kRoleStored = 2, kRoleStoredGram = 3 };
// Write role: what follows step k.
enum Next { kNextPublish = -2, kNextDecide = -1, kNextFusedGram = kGram, kNextFusedDirect = kDirect, kNextSweep = 3 };

struct Epi {
    int k, k1, ns, mode;
    int role;                  // EpiRole
    int next;                  // Next (write role)
    const int8_t* forced;      // this string's forced bits (nullptr: every step sampled)
    const double* U;           // this string's pre-drawn uniforms
    int8_t* bits;              // this string's output bits
    double* cond;              // this string's conditionals
    double* out;               // this string's scalars: logp, logscale, prev, dead
    unsigned* counter;         // blocks finished (in-kernel reduction; the last block resets it)
    const KStep* plain;        // plain descriptors (M, growth flags) indexed by step
    // The block partials a previous sweep left unreduced (pend_nblk 0: none).  Every block of this sweep reduces them
    // in the same fixed order and runs their chain step (step pend_k, its role and next) on a private copy of the
    // chain state read from the kernel's ctl; block 0 stores the result at ctl_out, the state the stream uses next.
    const double* pend_part;
    Ctl* ctl_out;
    unsigned pend_nblk;
    int pend_nq, pend_k, pend_role, pend_next, pend_sweep;
    int quiet;                 // 1: leave the per-string outputs (bits, cond, out) to another block
};

}  // namespace gs
