// Batched scoring and adjoint: many forced strings of many plans (instances), one launch per step.
//
// The single-string engine runs one chain per stream; scoring thousands of strings that way is launch-bound on small
// steps and repeats every step's launch per string.  Here the arena of an instance holds B strings side by side,
// slot-major and batch-minor (entry (slot, b) at slot * B + b), so every gather of a coset's slot reads B consecutive
// complex numbers: coalesced whatever the plan's layout, with the string index as the wavefront lane.  Semantics are
// those of chain mode (gsamp2.cpp:56-70): one sweep per step, the branch of each lane forced by its string, outputs
// stored unnormalised, per-lane conditionals nb / (2 prev) and log2 Pr = log2(prev) - ns; the last live entry is the
// amplitude Z.
//
// One launch per step index k covers every instance with k < ns_i: the host lists work items (instance, lane group of
// 64 strings, a Gray-ordered range of 2^lg cosets) and each block takes one.  Within a block the 4 waves split the
// coset range; lane b of a wave is string group*64 + b.  Per-lane FP64 sums are reduced over the waves in LDS, then
// over the items of the (instance, lane group) by the last item to finish (agent-scope counter), in a fixed order.
// The sequence init -> steps -> finish can be captured once in a hipGraph and replayed for every batch.
//
// Adjoint.  Z = 1^T L_{ns-1} ... L_0 e_0 with L_k = w_k(0) A_k0 + w_k(1) A_k1 (w_k(x) = M_k[b_k, x]).  The gradient
// G_k(x) = dZ/dw_k(x) = lambda_k^T A_kx P_{k-1}, where P_{k-1} is the table entering step k and
// lambda_{k-1}^T = lambda_k^T L_k.  Per coset, G_k(x) += sum_tau lambda_k[out tau] A[tau][x] (the same routed signed
// slot sums as the forward sweep), and the backward butterfly writes lambda_{k-1}[slot j] =
// sum_x w(x) sign_{j,x} lambda_k[tau(j, x)] into the coset's own input slots: in place in a second arena, race free
// by the same ownership as the forward sweep.  The forward tables are needed in reverse order but the arena is
// overwritten, so the forward pass keeps compact copies ([coset][slot][lane]) of the input tables of every
// checkpoint-segment start, and of every step of the segment being back-propagated (recomputed from its checkpoint).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsamp_amd.h"
#include "gs_plan.h"

namespace {

constexpr int kLanes = 64;           // strings per wave (one lane group)
constexpr int kWaves = 4;            // waves per block, each on its own part of the coset range
constexpr int kRChunks = 64;         // chunks of the first reduction stage per lane group
constexpr int kBThreads = kLanes * kWaves;
constexpr int64_t kMinItems = 1024;  // work items (blocks) a step launch should have before tiles get coarser

struct BStep {
    uint64_t cols[40];
    uint64_t off[4];
    uint64_t g, rhoc;
    double mre[2][2], mim[2][2];
    int nk, ntq, m;
    int tau0[4], tau1[4], par[4];
};

struct BItem {
    uint64_t c0;         // first coset of the range (aligned to 2^lg)
    int inst;            // instance
    int group;           // lane group (strings group*64 .. group*64+63)
    int lg;              // log2 cosets in the range
    int index;           // index of this item among the items of (inst, group)
    int count;           // number of items of (inst, group) in this step
    int pad;
};

struct BInst {           // device view of one instance
    void* mem;           // [2^R][B] forward arena
    void* lam;           // [2^R][B] adjoint arena (gradient mode)
    const BStep* steps;  // [ns]
    const int8_t* bitsT; // [ns][B] forced bits, step-major
    double* prev;        // [B]
    int* dead;           // [B]
    double* cond;        // [B][ns]
    double* logp;        // [B]
    double* amp;         // [B][2]
    double* grad;        // [B][ns][2][2]: G_k(x) re, im (gradient mode)
    double* part;        // [groups][max items per group][64][4]
    double* cpart;       // [groups][kRChunks][64][4]: chunk sums of the item partials
    int B, ns, pad0, pad1;
};

// Per launch: what each instance's step sweep does besides the forward butterfly.
struct BLaunch {
    int k;
    int bookkeeping;     // 1: chain-mode conditionals / dead flags (scoring), 0: recomputation
    void* const* copy;   // [ninst] destination of a compact copy of this step's input table, or nullptr entries
};

typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));

// Arena entries: raw (non-temporal) loads and stores of the storage type, widening to FP64 where used.  The arena
// pointers come from the instance table in memory, so the compiler cannot tell their address space; the accesses
// name the global address space explicitly (global instead of flat instructions).
#define GS_GLOBAL __attribute__((address_space(1)))
template <class S>
struct Sto;
template <>
struct Sto<float2> {
    static __device__ __forceinline__ float2 ld(const float2* p) {
        const f2v t = __builtin_nontemporal_load((const GS_GLOBAL f2v*)(p));
        return make_float2(t.x, t.y);
    }
    static __device__ __forceinline__ void st(float2* p, float2 v) {
        const f2v t = {v.x, v.y};
        __builtin_nontemporal_store(t, (GS_GLOBAL f2v*)(p));
    }
    static __device__ __forceinline__ double2 wide(float2 v) { return make_double2(v.x, v.y); }
    static __device__ __forceinline__ float2 narrow(double2 v) { return make_float2(float(v.x), float(v.y)); }
};
template <>
struct Sto<double2> {
    static __device__ __forceinline__ double2 ld(const double2* p) {
        const d2v t = __builtin_nontemporal_load((const GS_GLOBAL d2v*)(p));
        return make_double2(t.x, t.y);
    }
    static __device__ __forceinline__ void st(double2* p, double2 v) {
        const d2v t = {v.x, v.y};
        __builtin_nontemporal_store(t, (GS_GLOBAL d2v*)(p));
    }
    static __device__ __forceinline__ double2 wide(double2 v) { return v; }
    static __device__ __forceinline__ double2 narrow(double2 v) { return v; }
};
// Kernels are instantiated per storage S, NQ (the most slots per coset among the instances' step k: 1, 2 or 4) and
// LU (log2 cosets a lane gathers at once: 2^LU * NQ loads in flight per lane).

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double norm2(double2 a) { return fma(a.x, a.x, a.y * a.y); }

// The step's coordinate columns in LDS, for the whole block (the Gray walk reads one per coset group; read through
// the instance table in memory, each would be a dependent memory round trip).  Every thread of the block calls it.
__device__ __forceinline__ const uint64_t* lds_cols(const uint64_t* __restrict__ cols, int nk, uint64_t* sc) {
    for (int j = threadIdx.x; j < nk; j += blockDim.x) sc[j] = cols[j];
    __syncthreads();
    return sc;
}

// Coset range of this wave: 2^lsub cosets c = cstart | (gray(t) << lu) | u, t < 2^(lsub - lu), u < 2^lu with
// lu = min(LU, lsub): binary inside a group of 2^lu cosets, whose bases are base(t) ^ ubase[u] (all gathered at once),
// Gray-ordered across groups, so the next group's base is one XOR away.
template <int LU>
struct Walk {
    static constexpr int U = 1 << LU;
    uint64_t cstart, base;    // base of cstart | (gray(t) << lu)
    int pc, lsub, lu;
    bool on;
    uint64_t ubase[U];
    int upc[U];
    __device__ Walk(const BItem& it, const uint64_t* __restrict__ cols, uint64_t rhoc, int nk, int wave) {
        const int lw = it.lg >= 2 ? 2 : it.lg;
        on = wave < (1 << lw);
        lsub = it.lg - lw;
        cstart = it.c0 | (uint64_t(wave) << lsub);
        base = 0;
        pc = 0;
        for (int j = lsub; j < nk; j++)
            if ((cstart >> j) & 1) {
                base ^= cols[j];
                pc ^= int((rhoc >> j) & 1);
            }
        lu = lsub < LU ? lsub : LU;
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint64_t a = 0;
            int p = 0;
#pragma unroll
            for (int j = 0; j < LU; j++)
                if (j < lu && ((u >> j) & 1)) {
                    a ^= cols[j];
                    p ^= int((rhoc >> j) & 1);
                }
            ubase[u] = a;
            upc[u] = p;
        }
    }
    __device__ uint64_t groups() const { return on ? uint64_t(1) << (lsub - lu) : 0; }
    __device__ bool valid(int u) const { return u < (1 << lu); }
    __device__ int own(int u) const { return u & ((1 << lu) - 1); }   // the coset whose addresses ubase[u] repeats
    __device__ void next(const uint64_t* __restrict__ cols, uint64_t rhoc, uint64_t t) {   // to group t >= 1
        const int j = lu + __builtin_ctzll(t);
        base ^= cols[j];
        pc ^= int((rhoc >> j) & 1);
    }
    __device__ uint64_t coset(uint64_t t, int u) const { return cstart | ((t ^ (t >> 1)) << lu) | uint64_t(u); }
};

// Per-lane sum of NQ values over the block's waves (fixed order), stored as this item's partial (plain stores: the
// per-(instance, lane group) reduction runs in batch_reduce_kernel after the step, at a launch boundary).
template <int NQ>
__device__ void lane_store(const BInst& I, const BItem& it, int wave, int lane, const double (&v)[NQ]) {
    __shared__ double sh[kWaves][kLanes][NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) sh[wave][lane][q] = v[q];
    __syncthreads();
    if (wave != 0) return;
    double* part = I.part + ((size_t(it.group) * size_t(it.count) + size_t(it.index)) * kLanes + lane) * 4;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double t = sh[0][lane][q];
        for (int w = 1; w < kWaves; w++) t += sh[w][lane][q];
        part[q] = t;
    }
}

// The fields of a step that the inner loops use, copied to registers once per block (read through a reference to
// global memory, they would be reloaded after every arena store the compiler cannot prove disjoint from them).
struct SR {
    uint64_t off[4], g, rhoc;
    int ntq, m;
    int tau0[4], tau1[4], par[4];
    __device__ explicit SR(const BStep& s) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            off[q] = s.off[q];
            tau0[q] = s.tau0[q];
            tau1[q] = s.tau1[q];
            par[q] = s.par[q];
        }
        g = s.g;
        rhoc = s.rhoc;
        ntq = s.ntq;
        m = s.m;
    }
};

// Routed signed slot sums A[tau][x] of one coset from its ntq slot values (ntq = NQ at compile time in the EX
// variants, which have no branch on it).
template <int NQ>
__device__ __forceinline__ void route(const SR& s, const double2 (&v)[NQ], int pc, int ntq, double2 (&A)[2][2]) {
    A[0][0] = A[0][1] = A[1][0] = A[1][1] = make_double2(-0.0, -0.0);   // x + (-0.0) = x: the first add folds away
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        if (q >= ntq) break;
        const double2 w = v[q];
        const double2 ws = (pc ^ s.par[q]) ? cneg(w) : w;
        if (s.tau0[q]) A[1][0] = cadd(A[1][0], w); else A[0][0] = cadd(A[0][0], w);   // uniform branches
        if (s.tau1[q]) A[1][1] = cadd(A[1][1], ws); else A[0][1] = cadd(A[0][1], ws);
    }
}

template <class S, int NQ, int LU, bool EX>
__global__ __launch_bounds__(kBThreads) void batch_step_kernel(const BInst* __restrict__ insts,
                                                               const BItem* __restrict__ items, BLaunch L) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    const int k = L.k;
    const BStep& s = I.steps[k];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int B = I.B;
    const int b = it.group * kLanes + lane;
    const int bv = I.bitsT[size_t(k) * B + b];
    const double2 w0 = make_double2(s.mre[bv][0], s.mim[bv][0]);
    const double2 w1 = make_double2(s.mre[bv][1], s.mim[bv][1]);
    S* __restrict__ mem = static_cast<S*>(I.mem);
    S* __restrict__ cp = L.copy ? static_cast<S*>(L.copy[it.inst]) : nullptr;
    double nb = 0;
    constexpr int U = 1 << LU;
    const SR r(s);
    __shared__ uint64_t scols[40];
    const uint64_t* cols = lds_cols(s.cols, s.nk, scols);
    Walk<LU> wk(it, cols, r.rhoc, s.nk, wave);
    const uint64_t ng = wk.groups();
    const int ntq = EX ? NQ : r.ntq;                      // EX: every instance has NQ slots at this step
    for (uint64_t t = 0; t < ng; t++) {
        if (t) wk.next(cols, r.rhoc, t);
        S v[U][NQ];
        // Every load of the group first, and unconditionally: slots beyond this instance's ntq re-read slot 0, and the
        // cosets u >= 2^lu of a short range re-read coset u mod 2^lu (ubase keeps only the low lu bits of u); both are
        // discarded.  Loads (or their only uses) under branches leave a load possibly in flight at the loop latch,
        // and the next group's load into the same registers then waits for all but the newest memory operation.
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < NQ; q++)
                v[u][q] = Sto<S>::ld(mem + (wk.base ^ wk.ubase[u] ^ (q < ntq ? r.off[q] : 0)) * B + b);
        if (cp)
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    if (q < ntq && wk.valid(u)) Sto<S>::st(cp + (wk.coset(t, u) * ntq + q) * B + b, v[u][q]);
#pragma unroll
        for (int u = 0; u < U; u++) {
            double2 vw[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) vw[q] = q < ntq ? Sto<S>::wide(v[u][q]) : make_double2(0, 0);
            double2 A[2][2];
            route(r, vw, wk.pc ^ wk.upc[u], ntq, A);
            const uint64_t a0 = wk.base ^ wk.ubase[u];
            const double2 o0 = cfma(w0, A[0][0], cmul(w1, A[0][1]));
            const double2 o1 = cfma(w0, A[1][0], cmul(w1, A[1][1]));
            if (!wk.valid(u)) continue;                     // stores and norms of the lane's own cosets only
            Sto<S>::st(mem + a0 * B + b, Sto<S>::narrow(o0));
            nb += norm2(o0);
            if (r.m == 2) {
                Sto<S>::st(mem + (a0 ^ r.g) * B + b, Sto<S>::narrow(o1));
                nb += norm2(o1);
            }
        }
    }
    if (!L.bookkeeping) return;
    const double val[1] = {nb};
    lane_store<1>(I, it, wave, lane, val);
}

// Put a compact copy of step k's input table back into the forward arena (restart of a checkpoint segment).
template <class S, int NQ, int LU, bool EX>
__global__ __launch_bounds__(kBThreads) void batch_restore_kernel(const BInst* __restrict__ insts,
                                                                  const BItem* __restrict__ items, BLaunch L) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    const BStep& s = I.steps[L.k];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int B = I.B;
    const int b = it.group * kLanes + lane;
    S* __restrict__ mem = static_cast<S*>(I.mem);
    const S* __restrict__ cp = static_cast<const S*>(L.copy[it.inst]);
    if (!cp) return;
    constexpr int U = 1 << LU;
    const SR r(s);
    __shared__ uint64_t scols[40];
    const uint64_t* cols = lds_cols(s.cols, s.nk, scols);
    Walk<LU> wk(it, cols, r.rhoc, s.nk, wave);
    const uint64_t ng = wk.groups();
    const int ntq = EX ? NQ : r.ntq;                      // EX: every instance has NQ slots at this step
    for (uint64_t t = 0; t < ng; t++) {
        if (t) wk.next(cols, r.rhoc, t);
        S v[U][NQ];
        // No branches around the loads or their uses (see batch_step_kernel): a coset u >= 2^lu of a short range is
        // coset u mod 2^lu, whose value it copies to the same address again.
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (q < ntq) v[u][q] = Sto<S>::ld(cp + (wk.coset(t, wk.own(u)) * ntq + q) * B + b);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (q < ntq) Sto<S>::st(mem + (wk.base ^ wk.ubase[u] ^ r.off[q]) * B + b, v[u][q]);
    }
}

// Backward step k: lambda_k (adjoint arena, at the coset's outputs) and the compact copy of P_{k-1} give
// G_k(x) += sum_tau lambda[tau] A[tau][x]; lambda_{k-1} goes to the coset's input slots.
template <class S, int NQ, int LU, bool EX>
__global__ __launch_bounds__(kBThreads) void batch_back_kernel(const BInst* __restrict__ insts,
                                                               const BItem* __restrict__ items, BLaunch L) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    const int k = L.k;
    const BStep& s = I.steps[k];
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const int B = I.B;
    const int b = it.group * kLanes + lane;
    const int bv = I.bitsT[size_t(k) * B + b];
    const double2 w0 = make_double2(s.mre[bv][0], s.mim[bv][0]);
    const double2 w1 = make_double2(s.mre[bv][1], s.mim[bv][1]);
    S* __restrict__ lam = static_cast<S*>(I.lam);
    const S* __restrict__ cp = static_cast<const S*>(L.copy[it.inst]);
    double2 g0 = make_double2(0, 0), g1 = g0;
    constexpr int U = 1 << LU;
    const SR r(s);
    __shared__ uint64_t scols[40];
    const uint64_t* cols = lds_cols(s.cols, s.nk, scols);
    Walk<LU> wk(it, cols, r.rhoc, s.nk, wave);
    const uint64_t ng = wk.groups();
    const int ntq = EX ? NQ : r.ntq;                      // EX: every instance has NQ slots at this step
    for (uint64_t t = 0; t < ng; t++) {
        if (t) wk.next(cols, r.rhoc, t);
        S v[U][NQ], l0[U], l1[U];
        // Every load of the group first, with no branches around the loads or their uses (see batch_step_kernel): a
        // coset u >= 2^lu of a short range reads coset u mod 2^lu and only its stores and gradient terms are dropped;
        // a one-output step (g = 0) reads its output slot twice.
        const uint64_t gq = r.m == 2 ? r.g : 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t a0 = wk.base ^ wk.ubase[u];
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (q < ntq) v[u][q] = Sto<S>::ld(cp + (wk.coset(t, wk.own(u)) * ntq + q) * B + b);
            l0[u] = Sto<S>::ld(lam + a0 * B + b);
            l1[u] = Sto<S>::ld(lam + (a0 ^ gq) * B + b);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t a0 = wk.base ^ wk.ubase[u];
            const int pc = wk.pc ^ wk.upc[u];
            double2 vw[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) vw[q] = q < ntq ? Sto<S>::wide(v[u][q]) : make_double2(0, 0);
            double2 A[2][2];
            route(r, vw, pc, ntq, A);
            const double2 m0 = Sto<S>::wide(l0[u]);
            const double2 m1 = r.m == 2 ? Sto<S>::wide(l1[u]) : make_double2(0, 0);
            double2 h0 = cfma(m0, A[0][0], g0);
            double2 h1 = cfma(m0, A[0][1], g1);
            if (r.m == 2) {
                h0 = cfma(m1, A[1][0], h0);
                h1 = cfma(m1, A[1][1], h1);
            }
            if (!wk.valid(u)) continue;                     // gradient terms and stores of the lane's own cosets only
            g0 = h0;
            g1 = h1;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                if (q >= ntq) break;
                const double2 lt0 = r.tau0[q] ? m1 : m0;
                const double2 lt1 = r.tau1[q] ? m1 : m0;
                const double2 x1 = cmul(w1, lt1);
                const double2 out = cfma(w0, lt0, (pc ^ r.par[q]) ? cneg(x1) : x1);
                Sto<S>::st(lam + (a0 ^ r.off[q]) * B + b, Sto<S>::narrow(out));
            }
        }
    }
    const double val[4] = {g0.x, g0.y, g1.x, g1.y};
    lane_store<4>(I, it, wave, lane, val);
}

// After step k, in two fixed-order stages: batch_chunk_kernel sums the items' partials (cnt[inst] items per lane
// group) in kRChunks chunks of consecutive items, one block of kWaves waves per (lane group, chunk) -- wave w takes
// items w, w + kWaves, ... of the chunk -- and batch_reduce_kernel sums the chunks per (instance, lane), then runs
// chain1 bookkeeping (mode 0: cond, prev, dead; gsamp2.cpp:67) or stores the gradient entry G_k (mode 1).
__host__ __device__ inline int chunk_items(int count) { return (count + kRChunks - 1) / kRChunks; }

__global__ __launch_bounds__(kBThreads) void batch_chunk_kernel(const BInst* __restrict__ insts,
                                                               const int* __restrict__ cnt, int mode) {
    const int inst = blockIdx.y, group = blockIdx.x / kRChunks, chunk = blockIdx.x % kRChunks;
    const int lane = threadIdx.x & (kLanes - 1), wave = threadIdx.x / kLanes;
    const BInst& I = insts[inst];
    const int count = cnt[inst];
    if (count == 0 || group >= I.B / kLanes) return;
    const int per = chunk_items(count);
    const int i0 = chunk * per, i1 = min(count, i0 + per);
    if (i0 >= i1) return;
    const int nq = mode == 0 ? 1 : 4;
    const double* gp = I.part + size_t(group) * size_t(count) * kLanes * 4;
    double tot[4] = {0, 0, 0, 0};
    for (int i = i0 + wave; i < i1; i += kWaves)
        for (int q = 0; q < nq; q++) tot[q] += gp[(size_t(i) * kLanes + lane) * 4 + q];
    __shared__ double sh[kWaves][kLanes][4];
    for (int q = 0; q < 4; q++) sh[wave][lane][q] = tot[q];
    __syncthreads();
    if (wave != 0) return;
    double* out = I.cpart + ((size_t(group) * kRChunks + chunk) * kLanes + lane) * 4;
    for (int q = 0; q < nq; q++) {
        double t = sh[0][lane][q];
        for (int w = 1; w < kWaves; w++) t += sh[w][lane][q];
        out[q] = t;
    }
}

__global__ __launch_bounds__(kLanes) void batch_reduce_kernel(const BInst* __restrict__ insts,
                                                              const int* __restrict__ cnt, int k, int mode) {
    const int inst = blockIdx.y, group = blockIdx.x, lane = threadIdx.x;
    const BInst& I = insts[inst];
    const int count = cnt[inst];
    if (count == 0 || group >= I.B / kLanes) return;
    const int b = group * kLanes + lane;
    const int nq = mode == 0 ? 1 : 4;
    double tot[4] = {0, 0, 0, 0};
    const int nchunks = (count + chunk_items(count) - 1) / chunk_items(count);
    const double* gp = I.cpart + size_t(group) * kRChunks * kLanes * 4;
    for (int c = 0; c < nchunks; c++)
        for (int q = 0; q < nq; q++) tot[q] += gp[(size_t(c) * kLanes + lane) * 4 + q];
    const int ns = I.ns;
    if (mode == 1) {
        double* G = I.grad + (size_t(b) * ns + k) * 4;
        for (int q = 0; q < 4; q++) G[q] = tot[q];
        return;
    }
    if (!I.dead[b]) {
        I.cond[size_t(b) * ns + k] = tot[0] / (2 * I.prev[b]);
        I.prev[b] = tot[0];
        if (tot[0] <= 0) {
            for (int q = k; q < ns; q++) I.cond[size_t(b) * ns + q] = 0;
            I.dead[b] = 1;
        }
    }
}

template <class S>
__global__ void batch_init_kernel(const BInst* __restrict__ insts, int ninst) {
    const int i = blockIdx.x;
    if (i >= ninst) return;
    const BInst& I = insts[i];
    S* mem = static_cast<S*>(I.mem);
    for (int b = threadIdx.x; b < I.B; b += blockDim.x) {
        Sto<S>::st(mem + b, Sto<S>::narrow(make_double2(1.0, 0.0)));
        I.prev[b] = 1.0;
        I.dead[b] = 0;
    }
}

template <class S>
__global__ void batch_finish_kernel(const BInst* __restrict__ insts, int ninst) {
    const int i = blockIdx.x;
    if (i >= ninst) return;
    const BInst& I = insts[i];
    const S* mem = static_cast<const S*>(I.mem);
    for (int b = threadIdx.x; b < I.B; b += blockDim.x) {
        I.logp[b] = I.dead[b] ? -INFINITY : log2(I.prev[b]) - I.ns;
        const double2 z = Sto<S>::wide(Sto<S>::ld(mem + b));
        I.amp[2 * b] = z.x;
        I.amp[2 * b + 1] = z.y;
    }
}

// lambda at the end of the chain: dZ/d(final single entry) = 1
template <class S>
__global__ void batch_lambda_init_kernel(const BInst* __restrict__ insts, int ninst) {
    const int i = blockIdx.x;
    if (i >= ninst) return;
    const BInst& I = insts[i];
    S* lam = static_cast<S*>(I.lam);
    for (int b = threadIdx.x; b < I.B; b += blockDim.x) Sto<S>::st(lam + b, Sto<S>::narrow(make_double2(1.0, 0.0)));
}

thread_local std::string g_berr;

int bfail(int code, const std::string& m) {
    g_berr = m;
    return code;
}

}  // namespace

struct gs_plan {
    gs::Plan p;
};

struct gs_batch {
    int device = 0;
    int storage = GS_C64;
    size_t esz = 8;
    hipStream_t stream = nullptr;
    std::vector<gs::Plan> plans;
    std::vector<int> B;
    std::vector<BInst> hinst;
    std::vector<std::vector<BStep>> hsteps;
    BInst* dinst = nullptr;
    std::vector<void*> allocs;
    int maxns = 0;
    // per step: work items (device) and their count; tuning: log2 cosets per item per step
    std::vector<BItem*> ditems;
    std::vector<int> nitems;
    std::vector<std::vector<int>> lg;     // [inst][step]
    int lg_default = 8;                   // smallest tile (sizes the partial buffers); steps start at 2^9
    std::vector<int> nq;                  // per step: most slots per coset among the instances (kernel variant)
    std::vector<int> nq_exact;            // per step: 1 if every instance at the step has nq slots per coset
    int* dcnt = nullptr;                  // device [maxns][ninst]: items per lane group of instance i at step k
    int maxgroups = 1;
    int lu = 2;                           // log2 cosets a lane gathers at once
    bool items_dirty = false;             // tiles changed since the work lists were built
    int xcd_items = 1;                    // work lists ordered so each XCD takes consecutive items (2: coset-major)
    double auto_budget = -1;              // mem_budget argument of the last automatic segment choice
    // graph replay
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> step_ev;      // event after each step (latency)
    double last_ms = 0;
    std::vector<double> step_ms;          // end of step k, ms after the start (last call)
    // gradient mode
    int seg = 0;                          // steps per checkpoint segment in use
    std::vector<void*> galloc;            // adjoint arena, copies
    std::vector<void**> dcopy;            // per step: device array [ninst] of copy destinations (forward, last segment)
    std::vector<std::vector<void*>> ckpt; // [segment][inst]
    std::vector<std::vector<void*>> stack;// [slot in segment][inst]
    void** dptrs = nullptr;               // device scratch for per-launch pointer arrays
    double grad_fwd_ms = 0, grad_ms = 0;
};

namespace {

template <class T>
T* dalloc(gs_batch* b, size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)) != hipSuccess) return nullptr;
    b->allocs.push_back(p);
    return static_cast<T*>(p);
}

int build_items(gs_batch* bt) {
    for (auto* p : bt->ditems) (void)hipFree(p);
    const size_t ni = bt->plans.size();
    std::vector<int> cnt(size_t(bt->maxns) * ni, 0);
    bt->maxgroups = 1;
    for (size_t i = 0; i < ni; i++) bt->maxgroups = std::max(bt->maxgroups, bt->B[i] / kLanes);
    bt->ditems.assign(size_t(bt->maxns), nullptr);
    bt->nitems.assign(size_t(bt->maxns), 0);
    bt->nq.assign(size_t(bt->maxns), 1);
    bt->nq_exact.assign(size_t(bt->maxns), 0);
    for (int k = 0; k < bt->maxns; k++) {
        // Tiles set per step index (by hand or by the autotuner, often on a larger batch) are made finer for this
        // step, down to 2^lg_default cosets, while the step has fewer than kMinItems items: one item per block, and
        // a launch of a few hundred blocks leaves most of the chip idle while each wave walks a long coset range.
        int shrink = 0;
        for (;; shrink++) {
            int64_t tot = 0;
            bool finer = false;
            for (size_t i = 0; i < bt->plans.size(); i++) {
                const gs::Plan& P = bt->plans[i];
                if (k >= P.ns) continue;
                const int nk = P.steps[size_t(k)].nk;
                const int lg = std::min(nk, std::max(bt->lg[i][size_t(k)] - shrink, std::min(nk, bt->lg_default)));
                tot += int64_t(bt->B[i] / kLanes) << (nk - lg);
                finer = finer || lg > std::min(nk, bt->lg_default);
            }
            if (tot >= kMinItems || !finer) break;
        }
        std::vector<BItem> items;
        for (size_t i = 0; i < bt->plans.size(); i++) {
            const gs::Plan& P = bt->plans[i];
            if (k >= P.ns) continue;
            const int nk = P.steps[size_t(k)].nk;
            const int lg = std::min(nk, std::max(bt->lg[i][size_t(k)] - shrink, std::min(nk, bt->lg_default)));
            const int count = 1 << (nk - lg);
            cnt[size_t(k) * ni + i] = count;
            const int ng = bt->B[i] / kLanes;
            for (int o = 0; o < ng * count; o++) {
                // group-major (every coset range of lane group 0, then group 1, ...) or, with xcd_items == 2,
                // coset-major (all lane groups of a coset range next to each other: whole slot rows)
                const int g = bt->xcd_items == 2 ? o % ng : o / count;
                const int q = bt->xcd_items == 2 ? o / ng : o % count;
                BItem it{};
                it.c0 = uint64_t(q) << lg;
                it.inst = int(i);
                it.group = g;
                it.lg = lg;
                it.index = q;
                it.count = count;
                items.push_back(it);
            }
        }
        if (bt->xcd_items && items.size() >= 8 && items.size() % 8 == 0) {
            // blocks are dealt round-robin to the 8 XCDs (block i on XCD i mod 8): list position b = (j mod n/8) * 8
            // + j / (n/8) gives XCD x the consecutive items x n/8 .. (x+1) n/8 - 1 (adjacent coset ranges)
            const size_t n = items.size(), chunk = n / 8;
            std::vector<BItem> order(n);
            for (size_t j = 0; j < n; j++) order[(j % chunk) * 8 + j / chunk] = items[j];
            items.swap(order);
        }
        int nqk = 1;
        for (size_t i = 0; i < bt->plans.size(); i++)
            if (k < bt->plans[i].ns) nqk = std::max(nqk, bt->plans[i].steps[size_t(k)].ntq);
        bt->nq[size_t(k)] = nqk;
        int exact = 1;
        for (size_t i = 0; i < bt->plans.size(); i++)
            if (k < bt->plans[i].ns && bt->plans[i].steps[size_t(k)].ntq != nqk) exact = 0;
        bt->nq_exact[size_t(k)] = exact;
        bt->nitems[size_t(k)] = int(items.size());
        void* d = nullptr;
        if (hipMalloc(&d, std::max<size_t>(items.size() * sizeof(BItem), 16)) != hipSuccess)
            return bfail(11, "work-item allocation failed");
        if (hipMemcpy(d, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice) != hipSuccess)
            return bfail(17, "work-item upload failed");
        bt->ditems[size_t(k)] = static_cast<BItem*>(d);
    }
    if (!bt->dcnt && hipMalloc(&bt->dcnt, std::max<size_t>(cnt.size() * sizeof(int), 16)) != hipSuccess) {
        bt->dcnt = nullptr;
        return bfail(11, "count table allocation failed");
    }
    if (hipMemcpy(bt->dcnt, cnt.data(), cnt.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(17, "count table upload failed");
    return 0;
}

// Largest number of items any (instance, group) may have in a step at the smallest tile gs_batch_set_tile accepts
// (2^lg_default cosets): sizes the partial buffers.
int max_items(const gs_batch* bt, size_t i) {
    int m = 1;
    for (int k = 0; k < bt->plans[i].ns; k++) {
        const int nk = bt->plans[i].steps[size_t(k)].nk;
        m = std::max(m, 1 << (nk - std::min(nk, bt->lg_default)));
    }
    return m;
}

// Instantiation of kernel K<S, NQ, LU, EX> for NQ in {1, 2, 4}, LU in 0..3, EX (every instance has NQ slots at the
// step) false or true.
#define GS_BV(K, S, nq, lu)                                                                                        \
    case (nq) * 8 + (lu) * 2: return reinterpret_cast<const void*>(&K<S, nq, lu, false>);                          \
    case (nq) * 8 + (lu) * 2 + 1: return reinterpret_cast<const void*>(&K<S, nq, lu, true>);
#define GS_BATCH_VARIANT(K, S, nq, lu, ex)                                                                        \
    [&]() -> const void* {                                                                                         \
        switch ((nq) * 8 + (lu) * 2 + ((ex) ? 1 : 0)) {                                                           \
            GS_BV(K, S, 1, 0) GS_BV(K, S, 1, 1) GS_BV(K, S, 1, 2) GS_BV(K, S, 1, 3)                                 \
            GS_BV(K, S, 2, 0) GS_BV(K, S, 2, 1) GS_BV(K, S, 2, 2) GS_BV(K, S, 2, 3)                                 \
            GS_BV(K, S, 4, 0) GS_BV(K, S, 4, 1) GS_BV(K, S, 4, 2) GS_BV(K, S, 4, 3)                                 \
        }                                                                                                          \
        return nullptr;                                                                                            \
    }()

enum BKernel { kBStep, kBRestore, kBBack };

template <class S>
int launch_batch(gs_batch* bt, BKernel kind, int k, BLaunch L) {
    const int n = bt->nitems[size_t(k)];
    if (n == 0) return 0;
    const int nq = bt->nq[size_t(k)], lu = bt->lu;
    const bool ex = bt->nq_exact[size_t(k)] != 0;
    const void* fn = kind == kBStep ? GS_BATCH_VARIANT(batch_step_kernel, S, nq, lu, ex)
                   : kind == kBRestore ? GS_BATCH_VARIANT(batch_restore_kernel, S, nq, lu, ex)
                                       : GS_BATCH_VARIANT(batch_back_kernel, S, nq, lu, ex);
    if (!fn) return bfail(16, "no kernel variant");
    const BInst* insts = bt->dinst;
    const BItem* items = bt->ditems[size_t(k)];
    void* args[] = {&insts, &items, &L};
    return hipLaunchKernel(fn, dim3(n), dim3(kBThreads), args, 0, bt->stream) == hipSuccess ? 0
                                                                                      : bfail(17, "step launch failed");
}

// Per-(instance, lane) reduction of step k's item partials: mode 0 chain bookkeeping, mode 1 gradient entry.
int launch_reduce(gs_batch* bt, int k, int mode) {
    if (bt->nitems[size_t(k)] == 0) return 0;
    hipLaunchKernelGGL(batch_chunk_kernel, dim3(bt->maxgroups * kRChunks, int(bt->plans.size())), dim3(kBThreads), 0,
                       bt->stream, bt->dinst, bt->dcnt + size_t(k) * bt->plans.size(), mode);
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(bt->maxgroups, int(bt->plans.size())), dim3(kLanes), 0, bt->stream,
                       bt->dinst, bt->dcnt + size_t(k) * bt->plans.size(), k, mode);
    return hipGetLastError() == hipSuccess ? 0 : bfail(17, "reduce launch failed");
}

template <class S>
int launch_step(gs_batch* bt, int k, int bookkeeping, void* const* copy) {
    if (int rc = launch_batch<S>(bt, kBStep, k, BLaunch{k, bookkeeping, copy})) return rc;
    return bookkeeping ? launch_reduce(bt, k, 0) : 0;
}

template <class S>
int enqueue_t(gs_batch* bt, bool record_steps) {
    const int ninst = int(bt->plans.size());
    hipLaunchKernelGGL(batch_init_kernel<S>, dim3(ninst), dim3(256), 0, bt->stream, bt->dinst, ninst);
    for (int k = 0; k < bt->maxns; k++) {
        if (int rc = launch_step<S>(bt, k, 1, nullptr)) return rc;
        if (record_steps && bt->nitems[size_t(k)] && hipEventRecord(bt->step_ev[size_t(k)], bt->stream) != hipSuccess)
            return bfail(17, "event record failed");
    }
    hipLaunchKernelGGL(batch_finish_kernel<S>, dim3(ninst), dim3(256), 0, bt->stream, bt->dinst, ninst);
    if (hipGetLastError() != hipSuccess) return bfail(17, "batch launch failed");
    return 0;
}

int enqueue(gs_batch* bt, bool record_steps) {
    return bt->storage == GS_C64 ? enqueue_t<float2>(bt, record_steps) : enqueue_t<double2>(bt, record_steps);
}

void drop_graph(gs_batch* bt) {
    if (bt->gexec) (void)hipGraphExecDestroy(bt->gexec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    bt->gexec = nullptr;
    bt->graph = nullptr;
}

int build_graph(gs_batch* bt) {
    drop_graph(bt);
    if (hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bfail(17, "stream capture failed");
    const int rc = enqueue(bt, false);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(bt->stream, &g) != hipSuccess || rc) return rc ? rc : bfail(17, "stream capture failed");
    bt->graph = g;
    if (hipGraphInstantiate(&bt->gexec, g, nullptr, nullptr, 0) != hipSuccess) return bfail(17, "graph instantiate failed");
    return 0;
}

int upload_bits(gs_batch* bt, const int8_t* const* bits) {
    for (size_t i = 0; i < bt->plans.size(); i++) {
        const int B = bt->B[i], ns = bt->plans[i].ns;
        std::vector<int8_t> t(size_t(B) * ns);
        for (int b = 0; b < B; b++)
            for (int k = 0; k < ns; k++) {
                const int8_t v = bits[i][size_t(b) * ns + k];
                if (v != 0 && v != 1) return bfail(16, "bits must be 0 or 1");
                t[size_t(k) * B + b] = v;
            }
        if (hipMemcpyAsync(const_cast<int8_t*>(bt->hinst[i].bitsT), t.data(), t.size(), hipMemcpyHostToDevice, bt->stream) != hipSuccess ||
            hipStreamSynchronize(bt->stream) != hipSuccess)
            return bfail(17, "bits upload failed");
    }
    return 0;
}

// Entries of the compact copy of step k's input table of instance i.
size_t copy_entries(const gs_batch* bt, size_t i, int k) {
    const gs::StepInfo& s = bt->plans[i].steps[size_t(k)];
    return (size_t(1) << s.nk) * size_t(s.ntq) * size_t(bt->B[i]);
}

void free_grad(gs_batch* bt) {
    for (void* p : bt->galloc) (void)hipFree(p);
    bt->galloc.clear();
    bt->ckpt.clear();
    bt->stack.clear();
    bt->seg = 0;
}

// Allocate the adjoint arenas and the copies for segments of `seg` steps.
int alloc_grad(gs_batch* bt, int seg) {
    free_grad(bt);
    const size_t ni = bt->plans.size();
    auto galloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
        bt->galloc.push_back(p);
        return p;
    };
    for (size_t i = 0; i < ni; i++) {
        BInst& I = bt->hinst[i];
        if (!I.lam) {
            I.lam = dalloc<char>(bt, (size_t(1) << bt->plans[i].R) * size_t(bt->B[i]) * bt->esz);
            I.grad = dalloc<double>(bt, size_t(bt->B[i]) * bt->plans[i].ns * 4);
            if (!I.lam || !I.grad) return bfail(11, "adjoint arena allocation failed");
        }
    }
    const int nseg = (bt->maxns + seg - 1) / seg;
    bt->ckpt.assign(size_t(nseg), std::vector<void*>(ni, nullptr));
    bt->stack.assign(size_t(seg), std::vector<void*>(ni, nullptr));
    for (int sgi = 0; sgi < nseg; sgi++)
        for (size_t i = 0; i < ni; i++) {
            const int k = sgi * seg;
            if (k >= bt->plans[i].ns) continue;
            if (!(bt->ckpt[size_t(sgi)][i] = galloc(copy_entries(bt, i, k) * bt->esz))) return bfail(11, "checkpoint allocation failed");
        }
    for (int j = 0; j < seg; j++)
        for (size_t i = 0; i < ni; i++) {
            size_t mx = 0;
            for (int sgi = 0; sgi < nseg; sgi++) {
                const int k = sgi * seg + j;
                if (k < bt->plans[i].ns) mx = std::max(mx, copy_entries(bt, i, k));
            }
            if (mx && !(bt->stack[size_t(j)][i] = galloc(mx * bt->esz))) return bfail(11, "segment stack allocation failed");
        }
    if (hipMemcpy(bt->dinst, bt->hinst.data(), bt->hinst.size() * sizeof(BInst), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(17, "instance upload failed");
    bt->seg = seg;
    return 0;
}

// Bytes of adjoint memory for segments of `seg` steps.
size_t grad_bytes(const gs_batch* bt, int seg) {
    size_t tot = 0;
    const int nseg = (bt->maxns + seg - 1) / seg;
    for (size_t i = 0; i < bt->plans.size(); i++) {
        for (int sgi = 0; sgi < nseg; sgi++)
            if (sgi * seg < bt->plans[i].ns) tot += copy_entries(bt, i, sgi * seg);
        for (int j = 0; j < seg; j++) {
            size_t mx = 0;
            for (int sgi = 0; sgi < nseg; sgi++)
                if (sgi * seg + j < bt->plans[i].ns) mx = std::max(mx, copy_entries(bt, i, sgi * seg + j));
            tot += mx;
        }
    }
    return tot * bt->esz;
}

// Steps per checkpoint segment: among the segment lengths whose copies fit the budget (mem_budget <= 0: 70 % of the
// device memory free now, plus what the current copies hold), the one that re-sweeps the fewest entries (every
// segment but the last is recomputed once); the smallest footprint when none fits.
int choose_segment(const gs_batch* bt, double mem_budget) {
    if (mem_budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
        mem_budget = 0.7 * double(fr) + (bt->seg ? double(grad_bytes(bt, bt->seg)) : 0.0);
    }
    int best = -1, smallest = 1;
    double best_cost = 0, small_bytes = -1;
    for (int sg = 1; sg <= bt->maxns; sg++) {
        const double bytes = double(grad_bytes(bt, sg));
        if (small_bytes < 0 || bytes < small_bytes) {
            small_bytes = bytes;
            smallest = sg;
        }
        if (bytes > mem_budget) continue;
        const int last_start = ((bt->maxns - 1) / sg) * sg;
        double cost = 0;
        for (size_t i = 0; i < bt->plans.size(); i++)
            for (int k = 0; k < std::min(bt->plans[i].ns, last_start); k++) cost += double(copy_entries(bt, i, k));
        if (best < 0 || cost <= best_cost) {
            best = sg;
            best_cost = cost;
        }
    }
    return best < 0 ? smallest : best;
}

template <class S>
int grad_t(gs_batch* bt) {
    const int ninst = int(bt->plans.size());
    const int seg = bt->seg, ns = bt->maxns;
    const int nseg = (ns + seg - 1) / seg;
    const int last = nseg - 1;
    // per-launch pointer arrays live in a device scratch of [maxns][ninst] (forward) and [seg][ninst] (recompute/back)
    const size_t rows = size_t(ns + seg + nseg);
    std::vector<void*> host(rows * size_t(ninst), nullptr);
    auto fwd_ptr = [&](int k) { return bt->dptrs + size_t(k) * ninst; };
    auto rec_ptr = [&](int j) { return bt->dptrs + size_t(ns + j) * ninst; };
    auto ck_ptr = [&](int s) { return bt->dptrs + size_t(ns + seg + s) * ninst; };
    // forward: checkpoint the input of every segment start; keep every input of the last segment
    for (int k = 0; k < ns; k++)
        for (int i = 0; i < ninst; i++)
            host[size_t(k) * ninst + i] = k / seg == last ? bt->stack[size_t(k - last * seg)][size_t(i)]
                                        : (k % seg == 0 ? bt->ckpt[size_t(k / seg)][size_t(i)] : nullptr);
    for (int j = 0; j < seg; j++)
        for (int i = 0; i < ninst; i++) host[size_t(ns + j) * ninst + i] = bt->stack[size_t(j)][size_t(i)];
    for (int s = 0; s < nseg; s++)
        for (int i = 0; i < ninst; i++) host[size_t(ns + seg + s) * ninst + i] = bt->ckpt[size_t(s)][size_t(i)];
    if (hipMemcpy(bt->dptrs, host.data(), sizeof(void*) * host.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(17, "pointer upload failed");
    hipEvent_t e0 = bt->ev0, e1 = bt->ev1, em;
    if (hipEventCreate(&em) != hipSuccess) return bfail(17, "event creation failed");
    (void)hipEventRecord(e0, bt->stream);
    hipLaunchKernelGGL(batch_init_kernel<S>, dim3(ninst), dim3(256), 0, bt->stream, bt->dinst, ninst);
    for (int k = 0; k < ns; k++)
        if (int rc = launch_step<S>(bt, k, 1, fwd_ptr(k))) return rc;
    hipLaunchKernelGGL(batch_finish_kernel<S>, dim3(ninst), dim3(256), 0, bt->stream, bt->dinst, ninst);
    (void)hipEventRecord(em, bt->stream);
    hipLaunchKernelGGL(batch_lambda_init_kernel<S>, dim3(ninst), dim3(256), 0, bt->stream, bt->dinst, ninst);
    for (int s = last; s >= 0; s--) {
        const int k0 = s * seg, k1 = std::min(ns, k0 + seg);
        if (s != last) {                                    // recompute the segment from its checkpoint
            if (int rc = launch_batch<S>(bt, kBRestore, k0, BLaunch{k0, 0, ck_ptr(s)})) return rc;
            for (int k = k0; k < k1; k++)
                if (int rc = launch_step<S>(bt, k, 0, rec_ptr(k - k0))) return rc;
        }
        for (int k = k1 - 1; k >= k0; k--) {
            if (int rc = launch_batch<S>(bt, kBBack, k, BLaunch{k, 0, rec_ptr(k - k0)})) return rc;
            if (int rc = launch_reduce(bt, k, 1)) return rc;
        }
    }
    (void)hipEventRecord(e1, bt->stream);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
        (void)hipEventDestroy(em);
        return bfail(17, "gradient run failed");
    }
    float a = 0, c = 0;
    (void)hipEventElapsedTime(&a, e0, em);
    (void)hipEventElapsedTime(&c, e0, e1);
    bt->grad_fwd_ms = a;
    bt->grad_ms = c;
    (void)hipEventDestroy(em);
    return 0;
}

}  // namespace

extern "C" {

const char* gs_batch_last_error(void) { return g_berr.c_str(); }

int gs_batch_create(const gs_plan* const* plans, const int* lanes, int nplans, int storage, int device, gs_batch** out) {
    if (!plans || !lanes || nplans <= 0 || !out) return bfail(16, "bad argument");
    if (storage != GS_C64 && storage != GS_C128) return bfail(16, "storage must be GS_C64 or GS_C128");
    auto* bt = new gs_batch;
    bt->device = device;
    bt->storage = storage;
    bt->esz = storage == GS_C64 ? 8 : 16;
    bt->lu = storage == GS_C64 ? 2 : 1;
    auto bail = [&](int rc) {
        for (void* p : bt->allocs) (void)hipFree(p);
        for (auto* p : bt->ditems) (void)hipFree(p);
        delete bt;
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(bfail(17, "hipSetDevice failed"));
    for (int i = 0; i < nplans; i++) {
        if (!plans[i] || lanes[i] <= 0 || lanes[i] % kLanes) return bail(bfail(16, "lanes must be a positive multiple of 64"));
        bt->plans.push_back(plans[i]->p);
        bt->B.push_back(lanes[i]);
        bt->maxns = std::max(bt->maxns, plans[i]->p.ns);
    }
    bt->lg.resize(size_t(nplans));
    for (int i = 0; i < nplans; i++) bt->lg[size_t(i)].assign(size_t(bt->plans[size_t(i)].ns), bt->lg_default + 1);
    if (hipStreamCreate(&bt->stream) != hipSuccess || hipEventCreate(&bt->ev0) != hipSuccess ||
        hipEventCreate(&bt->ev1) != hipSuccess)
        return bail(bfail(17, "stream creation failed"));
    bt->step_ev.resize(size_t(bt->maxns));
    for (auto& ev : bt->step_ev)
        if (hipEventCreate(&ev) != hipSuccess) return bail(bfail(17, "event creation failed"));
    for (int i = 0; i < nplans; i++) {
        const gs::Plan& P = bt->plans[size_t(i)];
        const int B = lanes[i];
        BInst I{};
        I.B = B;
        I.ns = P.ns;
        I.mem = dalloc<char>(bt, (size_t(1) << P.R) * size_t(B) * bt->esz);
        if (!I.mem) return bail(bfail(11, "arena allocation failed"));
        std::vector<BStep> hs(size_t(P.ns));
        for (int k = 0; k < P.ns; k++) {
            const gs::StepInfo& si = P.steps[size_t(k)];
            BStep& s = hs[size_t(k)];
            memset(&s, 0, sizeof s);
            for (int j = 0; j < si.nk; j++) s.cols[j] = si.cols[size_t(j)];
            for (int j = 0; j < si.ntq; j++) {
                s.off[j] = si.off[j];
                s.tau0[j] = si.tau0[j];
                s.tau1[j] = si.tau1[j];
                s.par[j] = si.par[j];
            }
            s.g = si.g;
            s.rhoc = si.rhoc;
            for (int bb = 0; bb < 2; bb++)
                for (int x = 0; x < 2; x++) {
                    s.mre[bb][x] = si.M[bb][x].real();
                    s.mim[bb][x] = si.M[bb][x].imag();
                }
            s.nk = si.nk;
            s.ntq = si.ntq;
            s.m = si.m;
        }
        BStep* ds = dalloc<BStep>(bt, hs.size());
        if (!ds || hipMemcpy(ds, hs.data(), hs.size() * sizeof(BStep), hipMemcpyHostToDevice) != hipSuccess)
            return bail(bfail(17, "plan upload failed"));
        bt->hsteps.push_back(hs);
        I.steps = ds;
        I.bitsT = dalloc<int8_t>(bt, size_t(P.ns) * B);
        I.prev = dalloc<double>(bt, size_t(B));
        I.dead = dalloc<int>(bt, size_t(B));
        I.cond = dalloc<double>(bt, size_t(B) * P.ns);
        I.logp = dalloc<double>(bt, size_t(B));
        I.amp = dalloc<double>(bt, size_t(B) * 2);
        I.part = dalloc<double>(bt, size_t(B / kLanes) * size_t(max_items(bt, size_t(i))) * kLanes * 4);
        I.cpart = dalloc<double>(bt, size_t(B / kLanes) * kRChunks * kLanes * 4);
        if (!I.bitsT || !I.prev || !I.dead || !I.cond || !I.logp || !I.amp || !I.part || !I.cpart)
            return bail(bfail(11, "batch buffer allocation failed"));
        bt->hinst.push_back(I);
    }
    bt->dinst = dalloc<BInst>(bt, bt->hinst.size());
    if (!bt->dinst || hipMemcpy(bt->dinst, bt->hinst.data(), bt->hinst.size() * sizeof(BInst), hipMemcpyHostToDevice) != hipSuccess)
        return bail(bfail(17, "instance upload failed"));
    bt->dptrs = dalloc<void*>(bt, size_t(4 * bt->maxns + 4) * size_t(nplans));
    if (!bt->dptrs) return bail(bfail(11, "allocation failed"));
    if (int rc = build_items(bt)) return bail(rc);
    if (hipDeviceSynchronize() != hipSuccess) return bail(bfail(17, "device synchronisation failed"));
    *out = bt;
    return 0;
}

void gs_batch_destroy(gs_batch* bt) {
    if (!bt) return;
    (void)hipSetDevice(bt->device);
    (void)hipStreamSynchronize(bt->stream);
    drop_graph(bt);
    free_grad(bt);
    for (void* p : bt->allocs) (void)hipFree(p);
    for (auto* p : bt->ditems) (void)hipFree(p);
    (void)hipFree(bt->dcnt);
    for (auto& ev : bt->step_ev) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(bt->ev0);
    (void)hipEventDestroy(bt->ev1);
    (void)hipStreamDestroy(bt->stream);
    delete bt;
}

// Set the log2 coset range per work item of instance i, step k (k < 0: all steps); rebuilds the work lists.
int gs_batch_set_tile(gs_batch* bt, int i, int k, int lg) {
    if (!bt || i < 0 || i >= int(bt->plans.size()) || lg < 0 || lg > 40) return bfail(16, "bad argument");
    const int ns = bt->plans[size_t(i)].ns;
    if (k >= ns) return bfail(16, "bad step");
    for (int kk = (k < 0 ? 0 : k); kk < (k < 0 ? ns : k + 1); kk++) {
        // the partial buffer was sized for the tiles chosen at creation; only coarser tiles are allowed later
        const int nk = bt->plans[size_t(i)].steps[size_t(kk)].nk;
        if (lg < std::min(nk, bt->lg_default)) return bfail(16, "tile below the default size");
        bt->lg[size_t(i)][size_t(kk)] = lg;
    }
    drop_graph(bt);
    bt->items_dirty = true;                               // rebuilt once, before the next run
    return 0;
}

// Work-list order: 0 round-robin over the XCDs (items group-major), 1 (default) each XCD takes consecutive
// group-major items, 2 each XCD takes consecutive coset-major items (all lane groups of a coset range together).
// Results do not depend on it.  Measured on one MI355X (uniform 2^10-coset tiles): n=60 x 1024 lanes forward
// 85.3 / 74.6 / 78.5 ms, gradient 218.5 / 204.2 / 211.8 ms; n=70 band 26 x 512 lanes 363.6 / 366.3 / 368.8 ms.
// The lists are rebuilt before the next run.
int gs_batch_set_xcd_items(gs_batch* bt, int on) {
    if (!bt) return bfail(16, "bad argument");
    if (on < 0 || on > 2) return bfail(16, "xcd item order must be 0, 1 or 2");
    bt->xcd_items = on;
    drop_graph(bt);
    bt->items_dirty = true;
    return 0;
}

// log2 cosets each lane gathers at once (0..3; default 2 for complex64, 1 for complex128).
int gs_batch_set_unroll(gs_batch* bt, int lu) {
    if (!bt || lu < 0 || lu > 3) return bfail(16, "unroll must be 0..3");
    bt->lu = lu;
    drop_graph(bt);
    return 0;
}

int gs_batch_use_graph(gs_batch* bt, int on) {
    if (!bt) return bfail(16, "null batch");
    bt->use_graph = on != 0;
    if (!bt->use_graph) drop_graph(bt);
    return 0;
}

// Replace instance i's measurement matrices: M [ns][2][2] complex as re, im (M[k][b][x]); plan structure unchanged.
int gs_batch_set_matrices(gs_batch* bt, int i, const double* M) {
    if (!bt || !M || i < 0 || i >= int(bt->plans.size())) return bfail(16, "bad argument");
    auto& hs = bt->hsteps[size_t(i)];
    for (size_t k = 0; k < hs.size(); k++)
        for (int b = 0; b < 2; b++)
            for (int x = 0; x < 2; x++) {
                hs[k].mre[b][x] = M[((k * 2 + size_t(b)) * 2 + size_t(x)) * 2];
                hs[k].mim[b][x] = M[((k * 2 + size_t(b)) * 2 + size_t(x)) * 2 + 1];
            }
    (void)hipSetDevice(bt->device);
    if (hipMemcpy(const_cast<BStep*>(bt->hinst[size_t(i)].steps), hs.data(), hs.size() * sizeof(BStep),
                  hipMemcpyHostToDevice) != hipSuccess)
        return bfail(17, "matrix upload failed");
    return 0;
}

// Score one batch: bits[i] is instance i's [B_i][ns_i] forced strings (step order, 0/1); log2p[i] [B_i],
// cond[i] [B_i][ns_i] and amp[i] [B_i][2] (may be NULL) receive chain-mode outputs.  Returns after the device is done.
int gs_batch_score(gs_batch* bt, const int8_t* const* bits, double* const* log2p, double* const* cond,
                   double* const* amp) {
    if (!bt || !bits || !log2p) return bfail(16, "bad argument");
    if (hipSetDevice(bt->device) != hipSuccess) return bfail(17, "hipSetDevice failed");
    if (bt->items_dirty) {
        if (int rc = build_items(bt)) return rc;
        bt->items_dirty = false;
    }
    if (int rc = upload_bits(bt, bits)) return rc;
    if (hipEventRecord(bt->ev0, bt->stream) != hipSuccess) return bfail(17, "event record failed");
    if (bt->use_graph) {
        if (!bt->gexec)
            if (int rc = build_graph(bt)) return rc;
        if (hipGraphLaunch(bt->gexec, bt->stream) != hipSuccess) return bfail(17, "graph launch failed");
    } else {
        if (int rc = enqueue(bt, true)) return rc;
    }
    if (hipEventRecord(bt->ev1, bt->stream) != hipSuccess || hipEventSynchronize(bt->ev1) != hipSuccess)
        return bfail(17, "batch failed");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, bt->ev0, bt->ev1);
    bt->last_ms = ms;
    bt->step_ms.assign(size_t(bt->maxns), -1.0);
    if (!bt->use_graph)
        for (int k = 0; k < bt->maxns; k++) {
            float t = 0;
            if (bt->nitems[size_t(k)] && hipEventElapsedTime(&t, bt->ev0, bt->step_ev[size_t(k)]) == hipSuccess)
                bt->step_ms[size_t(k)] = t;
        }
    for (size_t i = 0; i < bt->plans.size(); i++) {
        const BInst& I = bt->hinst[i];
        const int B = bt->B[i], ns = bt->plans[i].ns;
        if (hipMemcpy(log2p[i], I.logp, sizeof(double) * size_t(B), hipMemcpyDeviceToHost) != hipSuccess ||
            (cond && cond[i] && hipMemcpy(cond[i], I.cond, sizeof(double) * size_t(B) * ns, hipMemcpyDeviceToHost) != hipSuccess) ||
            (amp && amp[i] && hipMemcpy(amp[i], I.amp, sizeof(double) * size_t(B) * 2, hipMemcpyDeviceToHost) != hipSuccess))
            return bfail(17, "readback failed");
    }
    return 0;
}

// Gradient of every string's amplitude with respect to its per-step weights: G[i] [B_i][ns_i][2] complex (re, im)
// with G_k(x) = dZ / dM_k[b_k, x]; also log2p[i] and amp[i] (Z).  seg = steps per checkpoint segment (0: the largest
// whose copies fit in mem_budget bytes, all steps when they fit).  Returns the segment length used in *seg_used.
int gs_batch_grad(gs_batch* bt, const int8_t* const* bits, double* const* log2p, double* const* amp, double* const* G,
                  int seg, double mem_budget, int* seg_used) {
    if (!bt || !bits || !log2p || !G) return bfail(16, "bad argument");
    if (hipSetDevice(bt->device) != hipSuccess) return bfail(17, "hipSetDevice failed");
    if (bt->items_dirty) {
        if (int rc = build_items(bt)) return rc;
        bt->items_dirty = false;
    }
    if (seg <= 0) {                                       // automatic: chosen once per budget argument
        if (bt->seg && bt->auto_budget == mem_budget) seg = bt->seg;
        else seg = choose_segment(bt, mem_budget);
        bt->auto_budget = mem_budget;
    } else {
        bt->auto_budget = -1;
    }
    seg = std::min(std::max(seg, 1), bt->maxns);
    if (seg != bt->seg)
        if (int rc = alloc_grad(bt, seg)) return rc;
    if (seg_used) *seg_used = seg;
    if (int rc = upload_bits(bt, bits)) return rc;
    if (int rc = bt->storage == GS_C64 ? grad_t<float2>(bt) : grad_t<double2>(bt)) return rc;
    for (size_t i = 0; i < bt->plans.size(); i++) {
        const BInst& I = bt->hinst[i];
        const int B = bt->B[i], ns = bt->plans[i].ns;
        if (hipMemcpy(log2p[i], I.logp, sizeof(double) * size_t(B), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(G[i], I.grad, sizeof(double) * size_t(B) * ns * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            (amp && amp[i] && hipMemcpy(amp[i], I.amp, sizeof(double) * size_t(B) * 2, hipMemcpyDeviceToHost) != hipSuccess))
            return bfail(17, "readback failed");
    }
    return 0;
}

// Device time of the last gs_batch_score (ms); step_end_ms[k] (may be NULL) the time from its start to the end of
// step k (without a graph; -1 where no instance has step k or when the batch ran as a graph).
double gs_batch_last_ms(const gs_batch* bt, double* step_end_ms) {
    if (!bt) return 0;
    if (step_end_ms)
        for (int k = 0; k < bt->maxns; k++) step_end_ms[k] = bt->step_ms.empty() ? -1 : bt->step_ms[size_t(k)];
    return bt->last_ms;
}

// Forward part and total device time (ms) of the last gs_batch_grad.
double gs_batch_grad_ms(const gs_batch* bt, double* forward_ms) {
    if (!bt) return 0;
    if (forward_ms) *forward_ms = bt->grad_fwd_ms;
    return bt->grad_ms;
}

}  // extern "C"
