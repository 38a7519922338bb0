// Device-side data of the engine: the per-step descriptor the sweep kernels read, the control block that carries
// the sampling chain's state from step to step on the device, and the kernel entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gs {

constexpr int kThreads = 256;       // threads per sweep block (4 waves of 64)
constexpr int kMaxGridLog2 = 11;    // at most 2048 blocks per sweep; each loops over tiles
constexpr int kPartStride = 4;      // doubles per block partial

// Canonical routing of the 2 * ntq terms of a step onto the outputs tau (after the host relabels the offsets):
//   kRoute1     m = 1: every term feeds tau 0
//   kRouteSplit m = 2: x = 0 terms feed tau 0, x = 1 terms feed tau 1 (growth, and q = 1 steps with x_v unmixed)
//   kRouteMixA  m = 2: offset j feeds tau (j & 1) with x = 0 and tau 1 - (j & 1) with x = 1
//   kRouteMixB  m = 2: offset j feeds tau (j & 1) with both x
enum Route { kRoute1 = 0, kRouteSplit = 1, kRouteMixA = 2, kRouteMixB = 3 };

// Sweep kinds.
enum Sweep { kGram = 0, kDirect = 1, kWrite = 2 };

// One step as the sweep kernels see it (device memory, read with uniform loads).
struct KStep {
    uint64_t cols[40];     // Tkeep (physical column of each kept coordinate)
    uint64_t off[4];       // coset offsets, relabelled so that the routing is canonical
    uint64_t g;            // second output slot offset (m = 2)
    uint64_t rhoc;         // sign mask over the kept coordinates
    double mre[2][2];      // M[b][x]
    double mim[2][2];
    uint32_t parmask;      // bit j: parity(qv_j & rho) of offset j
    int nk;                // log2 number of cosets
    int v;                 // log2 cosets per thread (the low v kept columns are the unit bits 0..v-1)
    int lg_grid;           // log2 blocks
    int lg_iter;           // log2 tiles per block
    int ntq, m, route;
    int growth;            // q = 0 and m = 2 with x -> tau: branch norms follow from the table norm
};

// Chain state carried on the device between steps (one per engine).
struct Ctl {
    double prev;           // norm of the stored table (chain1 / run_steps bookkeeping)
    double nstored;        // sum of |stored entry|^2 of the rounded stored table (growth-step norms)
    double logp;           // chain: sum log2 p_bv; chain1: log2 Pr at the end; run: sum log2 cond
    double logscale;       // run_steps: stored table = true table * 2^-logscale
    double sc;             // scale of the pending write
    double w[4];           // M[bv][0]*sc, M[bv][1]*sc as re, im pairs (read by the write sweep)
    double amp[2];         // chain1: final single live entry Z
    int bv;                // decision of the pending write
    int dead;              // a zero-probability branch was hit: later kernels do nothing
    int qidx;              // index of the next pre-drawn uniform (run_steps mode)
    int pad;
};

// Modes of the chain (how the decide kernel turns norms into a branch and bookkeeping).
enum ChainMode { kChain = 0, kChain1 = 1, kRun = 2 };

}  // namespace gs
