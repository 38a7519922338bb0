// Engine: device arena, per-step launch plan, the on-device sampling chain, and the C ABI (include/gsamp_amd.h).
//
// Per step k of a string the host enqueues, on the engine's stream and with no host synchronisation:
//   [norm sweep of step k]  ->  decide kernel  ->  write sweep of step k
// The decide kernel (one block) reduces the sweeps' per-block partials in a fixed order (deterministic), finishes the
// bookkeeping of step k-1's write (chain1 / run_steps conditionals, the stored table's norm) and makes step k's
// decision on the device: the branch (forced, or U_k < p1 with U_k drawn on the host beforehand -- the reference
// draws one uniform per sampled step from a stream that does not depend on p1, gsamp2.cpp:47, :86, :108), log2 p, and
// the scale of the write.  Growth steps (fresh slot, x -> tau) need no norm sweep: both branch norms are
// (|M_b0|^2 + |M_b1|^2) times the norm of the stored table, which the previous write sweep returns.
//
// Modes (gsamp2.cpp): chain (sample, sample2, prob) :39-55; chain1 (chain) :56-70; run_steps (hybrid, prefix,
// resume) :77-120; dumps :121-125, :162-208.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/gsamp_amd.h"
#include "gs_device.h"
#include "gs_plan.h"

namespace gs {
void* sweep_kernel_for(int storage, int lr, int vg, int ntq, int route, int sweep, int route2);
int plain_lr(int storage);
int fused_lr(int storage);
void* decide_kernel_ptr();
void* publish_kernel_ptr();
void* reduce_kernel_ptr(int sweep);
void* small_run_kernel_ptr(int storage);
int small_run_threads();
void* reduce1_kernel_ptr();
void* run_hist_kernel_ptr(int storage);
void* run_apply_kernel_ptr(int storage);
void* run_apply_gram_kernel_ptr(int storage, int route2);
void* reduce_stored_gram_kernel_ptr();
void* run_chunk_kernel_ptr();
void* run_decide_kernel_ptr();
}  // namespace gs

using namespace gs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(17, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

// ------------------------------------------------------------------------------------------------------------
// small kernels

template <class S>
__device__ S to_storage(double2 v);
template <>
__device__ float2 to_storage<float2>(double2 v) { return make_float2(float(v.x), float(v.y)); }
template <>
__device__ double2 to_storage<double2>(double2 v) { return v; }

template <class S>
__global__ void init_kernel(S* mem, Ctl* ctl) {
    mem[0] = to_storage<S>(make_double2(1.0, 0.0));
    Ctl c{};
    c.prev = 1;
    c.nstored = 1;
    *ctl = c;
}

__global__ void init_ctl_kernel(Ctl* ctl, double prev, double logscale) {
    Ctl c{};
    c.prev = prev;
    c.nstored = prev;
    c.logscale = logscale;
    *ctl = c;
}

__device__ uint64_t base_of(const KStep& s, uint64_t c) {
    uint64_t a = 0;
    for (int j = 0; c; j++, c >>= 1)
        if (c & 1) a ^= s.cols[j];
    return a;
}

// Live table of step K in logical order: out[c * ntq + t] = mem[base(c) ^ offs[t]] (offsets sorted ascending).
template <class S>
__global__ void gather_kernel(const KStep* __restrict__ sk, const S* __restrict__ mem, const uint64_t* __restrict__ offs,
                              int ntq, int64_t n, double2* __restrict__ out) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const uint64_t c = uint64_t(i / ntq);
        const int t = int(i % ntq);
        const uint64_t a = (sk ? base_of(*sk, c) : 0) ^ offs[t];
        const S v = mem[a];
        out[i] = make_double2(double(v.x), double(v.y));
    }
}

// acc[c * ntq + t] += w * tab[(c ^ cshift) * ntq + perm[t]]
__global__ void slice_accumulate_kernel(double2* __restrict__ acc, const double2* __restrict__ tab, double w,
                                        uint64_t cshift, const int* __restrict__ perm, int ntq, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const uint64_t c = uint64_t(i / ntq);
        const int t = int(i % ntq);
        const double2 v = tab[int64_t(c ^ cshift) * ntq + perm[t]];
        acc[i] = make_double2(fma(w, v.x, acc[i].x), fma(w, v.y, acc[i].y));
    }
}

// mem[base(c) ^ offs[t]] = acc[c * ntq + t]; per-block partial of sum |acc|^2 (FP64, before rounding).
template <class S>
__global__ __launch_bounds__(kThreads) void scatter_kernel(const KStep* __restrict__ sk, S* __restrict__ mem,
                                                           const uint64_t* __restrict__ offs, int ntq, int64_t n,
                                                           const double2* __restrict__ acc, double* __restrict__ part) {
    double nrm = 0;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const uint64_t c = uint64_t(i / ntq);
        const int t = int(i % ntq);
        const uint64_t a = (sk ? base_of(*sk, c) : 0) ^ offs[t];
        const double2 v = acc[i];
        mem[a] = to_storage<S>(v);
        nrm = fma(v.x, v.x, fma(v.y, v.y, nrm));
    }
    for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o, 64);
    __shared__ double sh[kThreads / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < kThreads / 64; w++) t += sh[w];
        part[size_t(blockIdx.x) * kPartStride] = t;
    }
}

// Fixed-order sum of the first value of nparts partials (one thread; used once per resume).
__global__ void sum_parts_kernel(const double* __restrict__ part, int nparts, double* __restrict__ out) {
    double t = 0;
    for (int i = 0; i < nparts; i++) t += part[size_t(i) * kPartStride];
    out[0] = t;
}

// Plain streaming copy (calibration of hardware byte counters against a kernel of known traffic).
__global__ __launch_bounds__(kThreads) void copy_probe_kernel(const float4* __restrict__ a, float4* __restrict__ b,
                                                              int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        b[i] = a[i];
}

// Race check of one step's coset layout (the debug mode of gs_race_check), with the arena as an owner map of 64-bit
// words (word index = slot * stride).  Coset c' (c' < 2^nk, in the layout's own coordinates) has base address
// XOR_j c'_j cols[j]; it reads base ^ off[t] (t < ntq) and writes base, and base ^ g when m = 2.
//   mode 0: clear every slot the layout reads or writes;
//   mode 1: claim each read slot for its coset (owner c' + 1); a slot already claimed counts in cnt[0];
//   mode 2: mark each written slot (bit 63); a slot marked twice counts in cnt[1], a slot claimed by another coset's
//           read in cnt[2];
//   mode 3: a run of diagonal steps (rp): mark each slot its passes visit (bit 62); a slot visited twice counts in
//           cnt[1], a slot no coset of the step claimed (outside the step's table) in cnt[2].
// cnt[3] counts the slots checked.
constexpr uint64_t kOwnWrite = uint64_t(1) << 63, kOwnVisit = uint64_t(1) << 62;
__global__ void race_kernel(const KStep* __restrict__ sp, const RStep* __restrict__ rp, unsigned long long* own,
                            int stride, int mode, unsigned long long* cnt) {
    const int nk = mode == 3 ? rp->nk : sp->nk;
    const uint64_t* cols = mode == 3 ? rp->cols : sp->cols;
    unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; c < (uint64_t(1) << nk);
         c += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t base = 0;
        for (int j = 0; j < nk; j++)
            if ((c >> j) & 1) base ^= cols[j];
        if (mode == 3) {
            const unsigned long long old = atomicOr(own + base * stride, (unsigned long long)kOwnVisit);
            c1 += (old & kOwnVisit) != 0;
            c2 += (old & ~(kOwnVisit | kOwnWrite)) == 0;
            c3++;
            continue;
        }
        const KStep& s = *sp;
        if (mode == 0) {
            for (int t = 0; t < s.ntq; t++) own[(base ^ s.off[t]) * stride] = 0;
            own[base * stride] = 0;
            if (s.m == 2) own[(base ^ s.g) * stride] = 0;
        } else if (mode == 1) {
            for (int t = 0; t < s.ntq; t++) {
                const unsigned long long old = atomicCAS(own + (base ^ s.off[t]) * stride, 0ull, c + 1);
                c0 += old != 0;
                c3++;
            }
        } else {
            for (int t = 0; t < s.m; t++) {
                const uint64_t a = t ? base ^ s.g : base;
                const unsigned long long old = atomicOr(own + a * stride, (unsigned long long)kOwnWrite);
                c1 += (old & kOwnWrite) != 0;
                const unsigned long long who = old & ~(kOwnWrite | kOwnVisit);
                c2 += who != 0 && who != c + 1;
                c3++;
            }
        }
    }
    if (c0) atomicAdd(cnt + 0, c0);
    if (c1) atomicAdd(cnt + 1, c1);
    if (c2) atomicAdd(cnt + 2, c2);
    if (c3) atomicAdd(cnt + 3, c3);
}

int route_and_relabel(const StepInfo& si, KStep& ks, std::string& err) {
    // canonical routing (see Route): reorder the offsets so that the term routing is one of four fixed patterns
    const int n = si.ntq;
    int order[kMaxOffsets];
    for (int j = 0; j < n; j++) order[j] = j;
    int route = -1;
    if (si.m == 1) {
        route = kRoute1;
    } else {
        bool split = true;
        for (int j = 0; j < n; j++) split &= si.tau0[j] == 0 && si.tau1[j] == 1;
        if (split) {
            route = kRouteSplit;
        } else {
            std::vector<int> z, o;
            for (int j = 0; j < n; j++) (si.tau0[j] ? o : z).push_back(j);
            if (z.size() != o.size()) { err = "unsupported term routing"; return 15; }
            for (size_t i = 0; i < z.size(); i++) {
                order[2 * i] = z[i];
                order[2 * i + 1] = o[i];
            }
            bool mixa = true, mixb = true;
            for (int jj = 0; jj < n; jj++) {
                const int j = order[jj];
                mixa &= si.tau1[j] == 1 - (jj & 1);
                mixb &= si.tau1[j] == (jj & 1);
            }
            if (mixa) route = kRouteMixA;
            else if (mixb) route = kRouteMixB;
            else { err = "unsupported term routing"; return 15; }
        }
    }
    ks.route = route;
    ks.parmask = 0;
    for (int jj = 0; jj < n; jj++) {
        ks.off[jj] = si.off[order[jj]];
        if (si.par[order[jj]]) ks.parmask |= 1u << jj;
    }
    return 0;
}

}  // namespace

// ------------------------------------------------------------------------------------------------------------
// engine

struct gs_plan {
    Plan p;
};

struct gs_engine {
    Plan plan;
    int storage = GS_C64;
    int device = 0;
    size_t esz = 8;                 // bytes per arena entry
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    // per step: the plain descriptor (norm and write sweeps) and, where the next step's norms can be folded into
    // this step's write, the fused descriptor (coordinates relabelled so that each thread owns the pairs)
    std::vector<KStep> hplain, hfused;
    std::vector<KStep> hlogical;    // plain steps in the plan's own coset coordinates (prefix/resume dump order)
    std::vector<char> fusable;
    KStep* dplain = nullptr;
    KStep* dfused = nullptr;
    KStep* dlogical = nullptr;
    struct Launch {
        void* gram = nullptr;
        void* direct = nullptr;
        void* write = nullptr;
        void* wgram = nullptr;      // fused write + next step's Gram
        void* wdirect = nullptr;    // fused write + next step's direct norms
        int grid = 1;
        int fgrid = 1;
    };
    std::vector<Launch> launch;
    double* part = nullptr;         // per-block partials of the running sweep (part_buf[0])
    double* part_buf[2] = {nullptr, nullptr};   // alternate between sweeps: one may hold partials still unreduced
    int cur_part = 0;               // the buffer the next sweep writes
    unsigned* counter = nullptr;    // blocks finished in the running sweep
    Ctl* ctl = nullptr;             // the current chain state (one of ctl_base[0..1])
    Ctl* ctl_base = nullptr;        // two chain states: a sweep that reduces pending partials writes the other one
    int cur_ctl = 0;
    // block partials the last sweep left unreduced (pend_nblk 0: none): the next sweep reduces them itself, or a
    // one-block kernel does before anything else runs
    int fold_r = -1;                // sweeps of steps of at most 2^fold_r live slots reduce the partials the previous
                                    // sweep left (-1: none; every sweep measured slower on the width-26 plan: 4.560
                                    // against 4.484 ms per sample)
    const double* pend_part = nullptr;
    unsigned pend_nblk = 0;
    int pend_sweep = 0;
    Epi pend_epi{};
    // per-string buffers, kChunk strings
    static constexpr int kChunk = 64;
    int8_t* d_forced = nullptr;
    double* d_U = nullptr;
    int8_t* d_bits = nullptr;
    double* d_cond = nullptr;
    double* d_out = nullptr;        // [kChunk][4]
    unsigned char* d_amp = nullptr; // [kChunk][16]
    int8_t* h_forced = nullptr;     // pinned mirrors
    double* h_U = nullptr;
    int8_t* h_bits = nullptr;
    double* h_cond = nullptr;
    double* h_out = nullptr;
    unsigned char* h_amp = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    bool allow_fusion = true;
    bool allow_fused4 = true;       // fold a four-slot next step's norms into the write (two pairing coordinates)
    int grid_cap = kDefaultGridLog2;  // log2 of the most blocks a sweep launches
    int grid_cap1 = kDefaultGridLog2One;  // ... a plain sweep of a one-slot step
    int grid_cap1_big = kDefaultGridLog2OneBig;  // ... of a one-slot step with at least 2^kBigStepLog2 cosets
    int small_nk = 10;              // runs of steps with nk <= small_nk go to one single-block launch (-1: never)
    int walk = 0;                   // tile order: 0 forward, 1 reverse on odd steps, 2 reverse on every step
    int fused_run = -1;             // log2 run of the fused layout (-1: the storage's default)
    int xcd_blocks = 1;             // XCD-contiguous block coordinates in every sweep
    int pair_ntq = 0;               // plain sweeps of steps with at most this many slots per coset take two runs per thread
    int pair_far = 0;               // ... the second one at the top coordinate (1) or the next unit bit (0)
    // runs of diagonal steps (RStep): descriptors, the run starting at each step (-1: none), the pattern sums
    std::vector<RStep> hrun;
    std::vector<int> run_at;
    RStep* drun = nullptr;
    size_t drun_cap = 0;            // descriptors the device array holds
    double* hpart = nullptr;        // [blocks][kRunBins] pattern sums of pass 1
    double* hchunk = nullptr;       // [kRunChunks][kRunBins]
    double2* rfac = nullptr;        // [kRunBins] product factors for pass 2
    int run_max = kMaxRun;          // steps per run (< 2: runs off)
    int run_min_nk = 20;            // runs only of steps with at least 2^run_min_nk cosets
    // optional per-sweep timing (bench / profiling)
    bool profile = false;
    std::vector<hipEvent_t> pev;
    std::vector<int> pev_tag;       // step * 8 + sweep
    std::vector<double> prof_ms;    // [ns][kNumSweeps]
    std::vector<long> prof_n;       // [ns][kNumSweeps]
};

namespace {

constexpr int kNumSweeps = 7;        // gram, direct, write, write+gram, write+direct, run pattern sums, run product

int ns_of(const gs_engine* e) { return e->plan.ns; }

// Partial sums per block of each sweep kind (Gram 4, direct 2, write 2, write + Gram 6, write + direct 4).
int partials_of(int sweep) { return sweep == kGram ? 4 : sweep == kWriteGram ? 6 : sweep == kWriteDirect ? 4 : 2; }

// Reduce the partials the last sweep left (one-block kernel) before a launch that does not reduce them itself.
int flush_pending(gs_engine* e) {
    if (!e->pend_nblk) return 0;
    const double* cpart = e->pend_part;
    unsigned nblk = e->pend_nblk;
    Ctl* ctl = e->ctl;
    Epi ep = e->pend_epi;
    int sw = e->pend_sweep;
    e->pend_nblk = 0;
    void* rargs[] = {&cpart, &nblk, &ctl, &ep, &sw};
    HIP_TRY(hipLaunchKernel(reduce_kernel_ptr(sw), dim3(1), dim3(kThreads), rargs, 0, e->stream));
    return 0;
}

int launch_sweep(gs_engine* e, int k, int sweep, const Epi& epi) {
    const auto& L = e->launch[size_t(k)];
    void* fn = sweep == kGram ? L.gram : sweep == kDirect ? L.direct : sweep == kWrite ? L.write
             : sweep == kWriteGram ? L.wgram : L.wdirect;
    if (!fn) return fail(15, "no kernel for step " + std::to_string(k));
    const bool fused = sweep == kWriteGram || sweep == kWriteDirect;
    const KStep* sp = (fused ? e->dfused : e->dplain) + k;
    const int grid = fused ? L.fgrid : L.grid;
    void* mem = e->arena;
    Epi ep = epi;
    const KStep& hs = (fused ? e->hfused : e->hplain)[size_t(k)];
    const int live_r = hs.nk + (hs.ntq == 4 ? 2 : hs.ntq == 2 ? 1 : 0);
    if (e->pend_nblk && (e->fold_r < 0 || live_r > e->fold_r))
        if (int rc = flush_pending(e)) return rc;
    Ctl* ctl = e->ctl;                              // the state this launch starts from
    if (e->pend_nblk) {                             // this sweep reduces the previous one's partials itself
        ep.pend_part = e->pend_part;
        ep.pend_nblk = e->pend_nblk;
        ep.pend_nq = partials_of(e->pend_sweep);
        ep.pend_k = e->pend_epi.k;
        ep.pend_role = e->pend_epi.role;
        ep.pend_next = e->pend_epi.next;
        ep.pend_sweep = e->pend_sweep;
        e->cur_ctl ^= 1;
        e->ctl = e->ctl_base + e->cur_ctl;
        ep.ctl_out = e->ctl;
        e->pend_nblk = 0;
    }
    double* part = e->part_buf[e->cur_part];
    void* args[] = {&sp, &mem, &part, &ctl, &ep};
    hipEvent_t a = nullptr, b = nullptr;
    if (e->profile) {
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, e->stream));
    }
    HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(kThreads), args, 0, e->stream));
    if (e->profile) {                   // the sweep kernel alone (the reduce kernel, if any, is not timed here)
        HIP_TRY(hipEventRecord(b, e->stream));
        e->pev.push_back(a);
        e->pev.push_back(b);
        e->pev_tag.push_back(k * 8 + sweep);
    }
    if (grid > kFenceBlocks) {          // the blocks only stored their partials: the next sweep reduces them
        e->pend_part = part;
        e->pend_nblk = unsigned(grid);
        e->pend_sweep = sweep;
        e->pend_epi = ep;
        e->pend_epi.pend_nblk = 0;
        e->cur_part ^= 1;
        if (e->fold_r < 0) return flush_pending(e);
    }
    return 0;
}

int collect_profile(gs_engine* e) {
    for (size_t i = 0; i < e->pev_tag.size(); i++) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e->pev[2 * i], e->pev[2 * i + 1]));
        const int k = e->pev_tag[i] / 8, sw = e->pev_tag[i] % 8;
        e->prof_ms[size_t(k) * kNumSweeps + sw] += ms;
        e->prof_n[size_t(k) * kNumSweeps + sw] += 1;
        (void)hipEventDestroy(e->pev[2 * i]);
        (void)hipEventDestroy(e->pev[2 * i + 1]);
    }
    e->pev.clear();
    e->pev_tag.clear();
    return 0;
}

int launch_small(gs_engine* e, void* fn, Epi epi, int k) {
    if (int rc = flush_pending(e)) return rc;
    Ctl* ctl = e->ctl;
    if (k >= 0) {
        void* args[] = {&ctl, &epi, &k};
        HIP_TRY(hipLaunchKernel(fn, dim3(1), dim3(1), args, 0, e->stream));
    } else {
        void* args[] = {&ctl, &epi};
        HIP_TRY(hipLaunchKernel(fn, dim3(1), dim3(1), args, 0, e->stream));
    }
    return 0;
}

int launch_init(gs_engine* e) {
    Ctl* ctl = e->ctl;
    if (e->storage == GS_C64) {
        float2* mem = static_cast<float2*>(e->arena);
        hipLaunchKernelGGL(init_kernel<float2>, dim3(1), dim3(1), 0, e->stream, mem, ctl);
    } else {
        double2* mem = static_cast<double2*>(e->arena);
        hipLaunchKernelGGL(init_kernel<double2>, dim3(1), dim3(1), 0, e->stream, mem, ctl);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// The run of diagonal steps that starts at step k and ends by k1, or -1.
int run_from(const gs_engine* e, int k, int k1) {
    if (k >= int(e->run_at.size())) return -1;
    const int ri = e->run_at[size_t(k)];
    return ri >= 0 && k + e->hrun[size_t(ri)].j <= k1 ? ri : -1;
}

// A run of diagonal steps: pattern sums (one pass), their reduction, the run's decisions and bookkeeping (one block),
// the product pass; its stored norm goes to the chain state.
int launch_run(gs_engine* e, int ri, const Epi& base, bool gram) {
    if (int rc = flush_pending(e)) return rc;
    const RStep& rs = e->hrun[size_t(ri)];
    const RStep* rp = e->drun + ri;
    void* mem = e->arena;
    Ctl* ctl = e->ctl;
    double* hp = e->hpart;
    const double* hpc = e->hpart;
    double* hc = e->hchunk;
    const double* hcc = e->hchunk;
    double2* fac = e->rfac;
    const double2* facc = e->rfac;
    double* part = e->part;
    unsigned nblk = 1u << rs.lg_grid;
    unsigned G = std::min<unsigned>(nblk, kRunChunks);
    auto timed = [&](int tag, auto&& launch) -> int {
        hipEvent_t a = nullptr, b = nullptr;
        if (e->profile) {
            HIP_TRY(hipEventCreate(&a));
            HIP_TRY(hipEventCreate(&b));
            HIP_TRY(hipEventRecord(a, e->stream));
        }
        if (int rc = launch()) return rc;
        if (e->profile) {
            HIP_TRY(hipEventRecord(b, e->stream));
            e->pev.push_back(a);
            e->pev.push_back(b);
            e->pev_tag.push_back(rs.k0 * 8 + tag);
        }
        return 0;
    };
    if (int rc = timed(5, [&]() -> int {
            void* args[] = {&rp, &mem, &hp, &ctl};
            HIP_TRY(hipLaunchKernel(run_hist_kernel_ptr(e->storage), dim3(nblk), dim3(kThreads), args, 0, e->stream));
            return 0;
        }))
        return rc;
    {
        void* args[] = {&hpc, &nblk, &hc, &ctl};
        HIP_TRY(hipLaunchKernel(run_chunk_kernel_ptr(), dim3(G), dim3(kRunBins), args, 0, e->stream));
    }
    Epi ed = base;
    ed.k = rs.k0;
    ed.role = kRoleWrite;
    {
        void* args[] = {&rp, &hcc, &G, &ctl, &ed, &fac};
        HIP_TRY(hipLaunchKernel(run_decide_kernel_ptr(), dim3(1), dim3(kRunBins), args, 0, e->stream));
    }
    Epi ea = base;
    ea.k = rs.k0 + rs.j - 1;
    ea.role = gram ? kRoleStoredGram : kRoleStored;
    void* apply = gram ? run_apply_gram_kernel_ptr(e->storage, rs.route2) : run_apply_kernel_ptr(e->storage);
    if (!apply) return fail(15, "no product-pass kernel for the run at step " + std::to_string(rs.k0));
    if (int rc = timed(6, [&]() -> int {
            void* args[] = {&rp, &mem, &facc, &part, &ctl, &ea};
            HIP_TRY(hipLaunchKernel(apply, dim3(nblk), dim3(kThreads), args, 0, e->stream));
            return 0;
        }))
        return rc;
    if (nblk > unsigned(kFenceBlocks)) {
        const double* cpart = part;
        int sw = kWrite;
        void* rargs[] = {&cpart, &nblk, &ctl, &ea, &sw};
        HIP_TRY(hipLaunchKernel(gram ? reduce_stored_gram_kernel_ptr() : reduce1_kernel_ptr(), dim3(1), dim3(kThreads),
                                rargs, 0, e->stream));
    }
    return 0;
}

// Which norms step k needs in a mode (-1: none -- forced, chain1, or a growth step).
int norm_for(const gs_engine* e, int mode, int k, int forced_bit, bool forced_string) {
    const KStep& s = e->hplain[size_t(k)];
    if (mode == kChain1) return -1;
    if (mode == kRun && forced_bit >= 0) return -1;
    if (s.growth) return -1;
    if (mode == kChain && forced_string) return kDirect;
    return kGram;
}

// Enqueue steps k0..k1-1 of one string.  forced_h: host copy of the string's forced bits (nullptr when every step is
// sampled); device pointers address this string's slot.  Per step: [norm sweep, when the previous write could not
// fold it in] -> write sweep, whose last block does the bookkeeping and decides the next step.
int enqueue_run(gs_engine* e, int mode, int k0, int k1, const int8_t* forced_h, const int8_t* forced_d, const double* U_d,
                int8_t* bits_d, double* cond_d, double* out_d) {
    Epi base{};
    base.k1 = k1;
    base.ns = e->plan.ns;
    base.mode = mode;
    base.forced = forced_d;
    base.U = U_d;
    base.bits = bits_d;
    base.cond = cond_d;
    base.out = out_d;
    base.counter = e->counter;
    base.plain = e->dplain;
    if (k0 == k1) return launch_small(e, publish_kernel_ptr(), base, -1);
    const bool fstr = forced_h != nullptr && mode == kChain;
    auto fb = [&](int k) { return forced_h ? int(forced_h[k]) : -1; };
    bool decided = false;
    for (int k = k0; k < k1; k++) {
        int ke = k;                                       // a run of two or more small steps: one launch
        while (e->small_nk >= 0 && ke < k1 && e->hplain[size_t(ke)].nk <= e->small_nk) ke++;
        if (const int ri = run_from(e, k, k1); ri >= 0) {   // a run of diagonal steps decides its own steps
            const RStep& rs = e->hrun[size_t(ri)];
            const int kn = rs.k0 + rs.j;            // a paired run's product pass can also decide the next step
            const bool gram = rs.pair && e->allow_fusion && kn < k1 && norm_for(e, mode, kn, fb(kn), fstr) == kGram;
            if (int rc = launch_run(e, ri, base, gram)) return rc;
            decided = gram;
            k += rs.j - 1;
            continue;
        }
        if (ke - k >= 2) {
            if (int rc = flush_pending(e)) return rc;
            const KStep* sp = e->dplain;
            void* mem = e->arena;
            Ctl* ctl = e->ctl;
            Epi ep = base;
            int a = k, b = ke, d = decided ? 1 : 0;
            void* args[] = {&sp, &mem, &ctl, &ep, &a, &b, &d};
            HIP_TRY(hipLaunchKernel(small_run_kernel_ptr(e->storage), dim3(1), dim3(small_run_threads()), args, 0,
                                    e->stream));
            decided = false;
            k = ke - 1;
            continue;
        }
        if (!decided) {
            const int nm = norm_for(e, mode, k, fb(k), fstr);
            if (nm >= 0) {
                Epi ep = base;
                ep.k = k;
                ep.role = kRoleNorm;
                if (int rc = launch_sweep(e, k, nm, ep)) return rc;
            } else {
                if (int rc = launch_small(e, decide_kernel_ptr(), base, k)) return rc;
            }
        }
        Epi ep = base;
        ep.k = k;
        ep.role = kRoleWrite;
        int sweep = kWrite;
        if (k + 1 == k1) {
            ep.next = kNextPublish;
        } else {
            const int nn = norm_for(e, mode, k + 1, fb(k + 1), fstr);
            const auto& L = e->launch[size_t(k)];
            if (run_from(e, k + 1, k1) >= 0) {
                ep.next = kNextSweep;
            } else if (nn < 0) {
                ep.next = kNextDecide;
            } else if (e->allow_fusion && e->fusable[size_t(k)] && (nn == kGram ? L.wgram : L.wdirect)) {
                ep.next = nn;
                sweep = nn == kGram ? kWriteGram : kWriteDirect;
            } else {
                ep.next = kNextSweep;
            }
        }
        if (int rc = launch_sweep(e, k, sweep, ep)) return rc;
        decided = ep.next != kNextSweep;
    }
    return flush_pending(e);
}

// ---- GF(2) helpers over u64 vectors (host)

// coef such that XOR of basis[i] over the bits i of coef == x; false if x is not in the span.
bool gf2_solve(const std::vector<uint64_t>& basis, uint64_t x, uint64_t& coef) {
    std::vector<std::pair<uint64_t, uint64_t>> rows;          // (vector, combination), distinct top bits
    for (size_t i = 0; i < basis.size(); i++) {
        uint64_t v = basis[i], c = uint64_t(1) << i;
        for (const auto& r : rows)
            if ((v >> (63 - __builtin_clzll(r.first))) & 1) {
                v ^= r.first;
                c ^= r.second;
            }
        if (!v) return false;                                     // basis not independent
        rows.emplace_back(v, c);
        std::sort(rows.begin(), rows.end(), [](const auto& p, const auto& q) { return p.first > q.first; });
    }
    uint64_t c = 0;
    for (const auto& r : rows)
        if ((x >> (63 - __builtin_clzll(r.first))) & 1) {
            x ^= r.first;
            c ^= r.second;
        }
    coef = c;
    return x == 0;
}

int parity64(uint64_t x) { return __builtin_parityll(x); }

// GF(2) echelon set of coordinate-space vectors (nk <= 40 bits): add() keeps a vector when it raises the rank.
struct Echelon {
    std::vector<uint64_t> rows;                     // reduced, distinct leading bits
    bool add(uint64_t v) {
        for (uint64_t r : rows)
            if ((v >> (63 - __builtin_clzll(r))) & 1) v ^= r;
        if (!v) return false;
        rows.push_back(v);
        std::sort(rows.begin(), rows.end(), [](uint64_t a, uint64_t b) { return a > b; });
        return true;
    }
};

// A new basis of a step's coset coordinates (vectors in the plan's coordinate space, bit j = coordinate j): `lead`
// first, then the coordinates whose physical columns are the unit bits 2^from, 2^(from+1), ... for as long as each is
// in the live span and independent, then the plan's own coordinates where they raise the rank.  The sweep is
// invariant under such a relabelling c = P c' (columns T P, sign mask P^T rhoc): it only reorders the cosets, and
// putting the low physical bits first makes each thread's cosets one aligned run.  Returns false when `lead` is
// dependent.
bool coord_basis(const std::vector<uint64_t>& cols, const std::vector<uint64_t>& lead, int from,
                 std::vector<uint64_t>& basis) {
    const int nk = int(cols.size());
    Echelon ech;
    basis.clear();
    for (uint64_t v : lead) {
        if (!ech.add(v)) return false;
        basis.push_back(v);
    }
    for (int i = from; i < 64 && int(basis.size()) < nk; i++) {
        uint64_t u = 0;
        if (!gf2_solve(cols, uint64_t(1) << i, u) || !ech.add(u)) break;
        basis.push_back(u);
    }
    for (int j = 0; j < nk && int(basis.size()) < nk; j++)
        if (ech.add(uint64_t(1) << j)) basis.push_back(uint64_t(1) << j);
    return int(basis.size()) == nk;
}

// Columns and sign mask of a step in the coordinates `basis`.
void apply_basis(const std::vector<uint64_t>& cols, uint64_t rhoc, const std::vector<uint64_t>& basis, KStep& ks) {
    memset(ks.cols, 0, sizeof ks.cols);
    ks.pmask[0] = 0;
    for (size_t i = 0; i < basis.size(); i++) {
        uint64_t col = 0;
        for (size_t j = 0; j < cols.size(); j++)
            if ((basis[i] >> j) & 1) col ^= cols[j];
        ks.cols[i] = col;
        ks.pmask[0] |= uint64_t(parity64(basis[i] & rhoc)) << i;
    }
}

// Number of leading columns equal to 1, 2, 4, ...
int unit_lead(const KStep& ks) {
    int u = 0;
    while (u < ks.nk && ks.cols[u] == (uint64_t(1) << u)) u++;
    return u;
}

void set_geometry(KStep& ks, int lv, int cap) {
    const int lt = lv + 8;
    if (ks.nk <= lt) {
        ks.lg_grid = 0;
        ks.lg_iter = 0;
    } else {
        const int tiles = ks.nk - lt;
        ks.lg_grid = std::min(tiles, cap);
        ks.lg_iter = tiles - ks.lg_grid;
    }
}

// Fused descriptor of step k (its write also reduces step k+1's norms), or false when step k+1 is not a
// two-slot step or the pairing vector cannot become a thread coordinate.
bool build_fused(const gs_engine* e, int k, KStep& f) {
    const Plan& P = e->plan;
    if (k + 1 >= P.ns) return false;
    const StepInfo& s0 = P.steps[size_t(k)];
    const KStep& p0 = e->hplain[size_t(k)];
    const KStep& p1 = e->hplain[size_t(k + 1)];
    if (p1.ntq != 2 || p1.growth) return false;
    const int nk = s0.nk, lrf = e->fused_run >= 0 ? e->fused_run : fused_lr(e->storage);
    // coordinates of the first lrf unit physical bits (one aligned run per thread)
    std::vector<uint64_t> lead;
    for (int i = 0; i < lrf; i++) {
        uint64_t u = 0;
        if (i >= nk || !gf2_solve(s0.cols, uint64_t(1) << i, u)) return false;
        lead.push_back(u);
    }
    // live space after step k: kept columns, then g
    std::vector<uint64_t> live(s0.cols.begin(), s0.cols.end());
    if (s0.m == 2) live.push_back(s0.g);
    const uint64_t tprime = p1.off[1];
    uint64_t coef = 0;
    if (!gf2_solve(live, tprime, coef)) return false;
    const uint64_t cmask = nk >= 64 ? ~uint64_t(0) : (uint64_t(1) << nk) - 1;
    const uint64_t gamma = coef & cmask;
    const int eps = s0.m == 2 ? int((coef >> nk) & 1) : 0;
    // coordinates: the unit-bit run, gamma (the pairing coordinate), then further unit bits and the rest
    lead.push_back(gamma);
    std::vector<uint64_t> basis;
    if (!coord_basis(s0.cols, lead, lrf, basis)) return false;
    // slot index (phi) and sign (nu) of the next step as functionals of a physical address in the live space
    const StepInfo& s1 = P.steps[size_t(k + 1)];
    std::vector<uint64_t> nb(s1.cols.begin(), s1.cols.end());
    nb.push_back(tprime);
    auto phinu = [&](uint64_t x, int& phi, int& nu) {
        uint64_t c2 = 0;
        if (!gf2_solve(nb, x, c2)) return false;
        phi = int((c2 >> s1.nk) & 1);
        const uint64_t c1 = s1.nk >= 64 ? c2 : c2 & ((uint64_t(1) << s1.nk) - 1);
        nu = parity64(c1 & s1.rhoc);
        return true;
    };
    std::vector<int> phic(static_cast<size_t>(nk)), nuc(static_cast<size_t>(nk));
    for (int j = 0; j < nk; j++)
        if (!phinu(s0.cols[size_t(j)], phic[size_t(j)], nuc[size_t(j)])) return false;
    int phig = 0, nug = 0;
    if (s0.m == 2 && !phinu(s0.g, phig, nug)) return false;
    f = p0;
    memset(f.cols, 0, sizeof f.cols);
    f.pmask[0] = f.pmask[1] = f.pmask[2] = f.pmask[3] = 0;
    for (int i = 0; i < nk; i++) {
        const uint64_t b = basis[size_t(i)];
        uint64_t col = 0;
        int pr = 0, pp = 0, pn = 0;
        for (int j = 0; j < nk; j++)
            if ((b >> j) & 1) {
                col ^= s0.cols[size_t(j)];
                pp ^= phic[size_t(j)];
                pn ^= nuc[size_t(j)];
            }
        pr = parity64(b & s0.rhoc);
        f.cols[i] = col;
        f.pmask[0] |= uint64_t(pr) << i;
        f.pmask[1] |= uint64_t(pp) << i;
        f.pmask[2] |= uint64_t(pn) << i;
    }
    f.lr = lrf;
    f.vg = 2;
    f.route2 = p1.route;
    f.par2 = int((p1.parmask >> 1) & 1);
    f.eps = eps;
    f.phig = phig;
    f.nug = nug;
    set_geometry(f, lrf + 1, e->grid_cap);
    return true;
}

// Fused descriptor of step k when step k+1 is a four-slot step (offsets 0, o1, o2, o1 ^ o2): coordinates lr and lr+1
// become gamma1 and gamma2, the coset parts of o1 and o2 in the live space after step k (o_i = cols . gamma_i ^ eps_i
// g), so a thread's cosets c, c ^ gamma1, c ^ gamma2, c ^ gamma1 ^ gamma2 hold all four slots of the next-step cosets
// their outputs belong to: output tau of coset c ^ gamma(m) is slot j0 ^ m of the next coset holding output tau ^
// eps(m) of c.  Slot-index bits and sign of the next step are parities of the stored address, as in build_fused.
bool build_fused4(const gs_engine* e, int k, KStep& f) {
    const Plan& P = e->plan;
    if (k + 1 >= P.ns || !e->allow_fused4) return false;
    const StepInfo& s0 = P.steps[size_t(k)];
    const KStep& p0 = e->hplain[size_t(k)];
    const KStep& p1 = e->hplain[size_t(k + 1)];
    if (p1.ntq != 4 || p1.growth || p1.off[0] != 0 || p1.off[3] != (p1.off[1] ^ p1.off[2])) return false;
    const int nk = s0.nk, lrf = e->fused_run >= 0 ? e->fused_run : fused_lr(e->storage);
    std::vector<uint64_t> lead;
    for (int i = 0; i < lrf; i++) {
        uint64_t u = 0;
        if (i >= nk || !gf2_solve(s0.cols, uint64_t(1) << i, u)) return false;
        lead.push_back(u);
    }
    std::vector<uint64_t> live(s0.cols.begin(), s0.cols.end());
    if (s0.m == 2) live.push_back(s0.g);
    const uint64_t cmask = nk >= 64 ? ~uint64_t(0) : (uint64_t(1) << nk) - 1;
    int epsv[2] = {0, 0};
    for (int i = 0; i < 2; i++) {
        uint64_t coef = 0;
        if (!gf2_solve(live, p1.off[1 + i], coef)) return false;
        lead.push_back(coef & cmask);
        epsv[i] = s0.m == 2 ? int((coef >> nk) & 1) : 0;
    }
    std::vector<uint64_t> basis;
    if (!coord_basis(s0.cols, lead, lrf, basis)) return false;
    const StepInfo& s1 = P.steps[size_t(k + 1)];
    std::vector<uint64_t> nb(s1.cols.begin(), s1.cols.end());
    nb.push_back(p1.off[1]);
    nb.push_back(p1.off[2]);
    auto phinu = [&](uint64_t x, int& phi, int& nu) {      // phi: the two slot-index bits
        uint64_t c2 = 0;
        if (!gf2_solve(nb, x, c2)) return false;
        phi = int((c2 >> s1.nk) & 3);
        const uint64_t c1 = s1.nk >= 64 ? c2 : c2 & ((uint64_t(1) << s1.nk) - 1);
        nu = parity64(c1 & s1.rhoc);
        return true;
    };
    std::vector<int> phic(static_cast<size_t>(nk)), nuc(static_cast<size_t>(nk));
    for (int j = 0; j < nk; j++)
        if (!phinu(s0.cols[size_t(j)], phic[size_t(j)], nuc[size_t(j)])) return false;
    int phig = 0, nug = 0;
    if (s0.m == 2 && !phinu(s0.g, phig, nug)) return false;
    f = p0;
    memset(f.cols, 0, sizeof f.cols);
    f.pmask[0] = f.pmask[1] = f.pmask[2] = f.pmask[3] = 0;
    for (int i = 0; i < nk; i++) {
        const uint64_t b = basis[size_t(i)];
        uint64_t col = 0;
        int pp = 0, pn = 0;
        for (int j = 0; j < nk; j++)
            if ((b >> j) & 1) {
                col ^= s0.cols[size_t(j)];
                pp ^= phic[size_t(j)];
                pn ^= nuc[size_t(j)];
            }
        f.cols[i] = col;
        f.pmask[0] |= uint64_t(parity64(b & s0.rhoc)) << i;
        f.pmask[1] |= uint64_t(pp & 1) << i;
        f.pmask[2] |= uint64_t(pn) << i;
        f.pmask[3] |= uint64_t((pp >> 1) & 1) << i;
    }
    f.lr = lrf;
    f.vg = 4;
    f.route2 = p1.route;
    f.parmask2 = int(p1.parmask);
    f.eps = epsv[0];
    f.eps2 = epsv[1];
    f.phig = phig;
    f.nug = nug;
    set_geometry(f, lrf + 2, e->grid_cap);
    return true;
}

// Fused descriptor of step k when step k+1 is a one-slot, one-output step (kWriteGram1): every value step k's plain
// write stores is then the one slot of one coset of step k+1, whose x = 1 sign (-1)^(pc ^ par) is an affine function
// of the stored address, so the write also sums step k+1's Gram matrix (G00 = G11 = sum |v|^2, Re G01 = signed sum).
// Requires the live table after step k to be exactly step k+1's slot space.
bool build_fused1(const gs_engine* e, int k, KStep& f) {
    const Plan& P = e->plan;
    if (k + 1 >= P.ns) return false;
    const StepInfo& s0 = P.steps[size_t(k)];
    const StepInfo& s1 = P.steps[size_t(k + 1)];
    const KStep& p0 = e->hplain[size_t(k)];
    if (s1.ntq != 1 || s1.m != 1 || s1.growth || s1.tau0[0] != 0 || s1.tau1[0] != 0) return false;
    if (p0.vg != 1 || p0.lr != 1) return false;
    if (s0.nk + (s0.m == 2 ? 1 : 0) != s1.nk) return false;
    // parity(c1 & rhoc1) of the step-(k+1) coordinates c1 of an address in its slot space (linear in the address)
    auto sign_of = [&](uint64_t x, int& out) {
        uint64_t c = 0;
        if (!gf2_solve(s1.cols, x, c)) return false;
        out = parity64(c & s1.rhoc);
        return true;
    };
    int kap = 0;
    if (!sign_of(s1.off[0], kap)) return false;
    f = p0;
    f.pmask[1] = f.pmask[2] = 0;
    for (int j = 0; j < p0.nk; j++) {
        int b = 0;
        if (!sign_of(p0.cols[j], b)) return false;            // a live column outside the next step's slot space
        f.pmask[2] |= uint64_t(b) << j;
    }
    int sg = 0;
    if (s0.m == 2 && !sign_of(s0.g, sg)) return false;
    f.sgg = sg;
    f.sgk = kap ^ s1.par[0];                                  // slot a = base1(c1) ^ off: sign of c1 = sign(a) ^ sign(off)
    return true;
}

// Basis of the null space of the rows (GF(2) masks over nk coordinates).
std::vector<uint64_t> gf2_kernel(std::vector<uint64_t> rows, int nk) {
    std::vector<int> piv;
    size_t r = 0;
    for (int col = 0; col < nk && r < rows.size(); col++) {
        size_t p = r;
        while (p < rows.size() && !((rows[p] >> col) & 1)) p++;
        if (p == rows.size()) continue;
        std::swap(rows[r], rows[p]);
        for (size_t i = 0; i < rows.size(); i++)
            if (i != r && ((rows[i] >> col) & 1)) rows[i] ^= rows[r];
        piv.push_back(col);
        r++;
    }
    std::vector<uint64_t> ker;
    for (int f = 0; f < nk; f++) {
        if (std::find(piv.begin(), piv.end(), f) != piv.end()) continue;
        uint64_t v = uint64_t(1) << f;
        for (size_t i = 0; i < r; i++)
            if ((rows[i] >> f) & 1) v |= uint64_t(1) << piv[i];
        ker.push_back(v);
    }
    return ker;
}

// Descriptor of the run of j diagonal steps from step k0 (see RStep), in the coordinates of step k0's table; false
// when the unit bits 1 .. 256 are not all in the live table or too few sign-neutral tile coordinates exist.
bool make_run(const gs_engine* e, int k0, int j, RStep& rs, bool try_pair) {
    const Plan& P = e->plan;
    const StepInfo& s0 = P.steps[size_t(k0)];
    const std::vector<uint64_t>& C0 = s0.cols;
    const int nk = s0.nk;
    // sign of step i at the address C0 c: parity(D_i & c) ^ par_i
    std::vector<uint64_t> D(static_cast<size_t>(j));
    int kappa = 0;
    for (int i = 0; i < j; i++) {
        const StepInfo& si = P.steps[size_t(k0 + i)];
        uint64_t row = 0;
        for (int q = 0; q < nk; q++) {
            uint64_t cq = 0;
            if (!gf2_solve(si.cols, C0[size_t(q)], cq)) return false;
            row |= uint64_t(parity64(cq & si.rhoc)) << q;
        }
        D[size_t(i)] = row;
        kappa |= (si.par[0] & 1) << i;
    }
    auto col_of = [&](uint64_t u) {
        uint64_t a = 0;
        for (int q = 0; q < nk; q++)
            if ((u >> q) & 1) a ^= C0[size_t(q)];
        return a;
    };
    Echelon ech;
    std::vector<uint64_t> lowv;                       // coordinates 0..8: the unit physical bits 1 .. 256
    for (int b = 0; b < 9; b++) {
        uint64_t u = 0;
        if (!gf2_solve(C0, uint64_t(1) << b, u) || !ech.add(u)) return false;
        lowv.push_back(u);
    }
    const int tiles = nk - 9;
    const int cap = nk >= kBigStepLog2 ? e->grid_cap1_big : e->grid_cap1;
    const int lg_grid = std::min(tiles, cap), lg_iter = tiles - lg_grid;
    // pairing with the next step when it is a two-slot step reading this table in pairs (x, x ^ o): tile coordinate 0
    // gets column o (o ^ 1 when o is odd), and the next step's slot index and sign become parities per coordinate
    uint64_t pairv = 0;
    int peps = 0;
    bool pair = false;
    const StepInfo* s1 = k0 + j < P.ns ? &P.steps[size_t(k0 + j)] : nullptr;
    const KStep* p1 = k0 + j < P.ns ? &e->hplain[size_t(k0 + j)] : nullptr;
    if (try_pair && lg_iter >= 1 && p1 && p1->ntq == 2 && !p1->growth && p1->off[0] == 0) {
        uint64_t u = 0;
        if (gf2_solve(C0, p1->off[1], u)) {
            if (col_of(u) & 1) {
                u ^= lowv[0];
                peps = 1;
            }
            if (ech.add(u)) {
                pairv = u;
                pair = true;
            }
        }
    }
    // tile coordinates: in the null space of every D_i, with an even physical column
    std::vector<uint64_t> ker = gf2_kernel(D, nk), tilev;
    uint64_t odd = 0;
    for (uint64_t v : ker) {
        if (col_of(v) & 1) {
            if (!odd) {
                odd = v;
                continue;
            }
            v ^= odd;
        }
        if (int(tilev.size()) < lg_iter - (pair ? 1 : 0) && ech.add(v)) tilev.push_back(v);
    }
    if (int(tilev.size()) < lg_iter - (pair ? 1 : 0)) return false;
    std::vector<uint64_t> blockv;                     // block coordinates: the plan's where they raise the rank
    for (int q = 0; q < nk && int(blockv.size()) < lg_grid; q++) {
        uint64_t u = uint64_t(1) << q;
        if (col_of(u) & 1) u ^= lowv[0];
        if (ech.add(u)) blockv.push_back(u);
    }
    if (int(blockv.size()) != lg_grid) return false;
    memset(&rs, 0, sizeof rs);
    std::vector<uint64_t> order(lowv);
    order.insert(order.end(), blockv.begin(), blockv.end());
    if (pair) order.push_back(pairv);
    order.insert(order.end(), tilev.begin(), tilev.end());
    for (size_t c = 0; c < order.size(); c++) {
        rs.cols[c] = col_of(order[c]);
        int sg = 0;
        for (int i = 0; i < j; i++) sg |= parity64(D[size_t(i)] & order[c]) << i;
        rs.sig[c] = uint8_t(sg);
    }
    for (int i = 0; i < j; i++)
        for (int b = 0; b < 2; b++)
            for (int x = 0; x < 2; x++) {
                rs.mre[i][b][x] = P.steps[size_t(k0 + i)].M[b][x].real();
                rs.mim[i][b][x] = P.steps[size_t(k0 + i)].M[b][x].imag();
            }
    rs.nk = nk;
    rs.j = j;
    rs.k0 = k0;
    rs.kappa = kappa;
    rs.lg_grid = lg_grid;
    rs.lg_iter = lg_iter;
    rs.xcd = e->xcd_blocks;
    if (pair) {                                       // next step's slot index and sign of each coordinate's column
        std::vector<uint64_t> nb(s1->cols.begin(), s1->cols.end());
        nb.push_back(p1->off[1]);
        bool ok = true;
        for (size_t c = 0; c < order.size() && ok; c++) {
            uint64_t c2 = 0;
            if (!gf2_solve(nb, rs.cols[c], c2)) {
                ok = false;
                break;
            }
            const uint64_t c1 = s1->nk >= 64 ? c2 : c2 & ((uint64_t(1) << s1->nk) - 1);
            rs.phim |= ((c2 >> s1->nk) & 1) << c;
            rs.num |= uint64_t(parity64(c1 & s1->rhoc)) << c;
        }
        if (ok) {
            rs.pair = 1;
            rs.peps = peps;
            rs.route2 = p1->route;
            rs.par2 = int((p1->parmask >> 1) & 1);
        } else {
            return false;
        }
    }
    // check the pattern bookkeeping against each step's own sign on pseudo-random coordinate vectors
    uint64_t x = 0x9E3779B97F4A7C15ull ^ uint64_t(k0);
    for (int t = 0; t < 64; t++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        uint64_t a = 0;
        int sg = kappa;
        for (int c = 0; c < nk; c++)
            if ((x >> c) & 1) {
                a ^= rs.cols[c];
                sg ^= rs.sig[c];
            }
        for (int i = 0; i < j; i++) {
            const StepInfo& si = P.steps[size_t(k0 + i)];
            uint64_t ci = 0;
            if (!gf2_solve(si.cols, a, ci) || (parity64(ci & si.rhoc) ^ si.par[0] ^ (sg >> i)) & 1) return false;
        }
    }
    return true;
}

// Runs of at most run_max consecutive diagonal steps (one slot, one output, slot offset 0, the same live table) of
// at least 2^run_min_nk cosets.
void build_runs(gs_engine* e) {
    const Plan& P = e->plan;
    e->hrun.clear();
    e->run_at.assign(size_t(std::max(P.ns, 1)), -1);
    if (e->run_max < 2) return;
    auto diag = [&](int k) {
        const StepInfo& si = P.steps[size_t(k)];
        return si.ntq == 1 && si.m == 1 && !si.growth && si.off[0] == 0 && si.tau0[0] == 0 && si.tau1[0] == 0 &&
               si.nk >= std::max(10, e->run_min_nk);
    };
    auto same_table = [&](int a, int b) {
        if (P.steps[size_t(a)].nk != P.steps[size_t(b)].nk) return false;
        Echelon ech;
        for (uint64_t c : P.steps[size_t(a)].cols) ech.add(c);
        for (uint64_t c : P.steps[size_t(b)].cols)
            if (ech.add(c)) return false;
        return true;
    };
    for (int k = 0; k < P.ns;) {
        if (!diag(k)) {
            k++;
            continue;
        }
        int end = k + 1;
        while (end < P.ns && end - k < e->run_max && diag(end) && same_table(k, end)) end++;
        RStep rs;
        if (end - k >= 2 && (make_run(e, k, end - k, rs, true) || make_run(e, k, end - k, rs, false))) {
            e->run_at[size_t(k)] = int(e->hrun.size());
            e->hrun.push_back(rs);
            k = end;
        } else {
            k++;
        }
    }
}

int build_steps(gs_engine* e) {
    const Plan& P = e->plan;
    e->hplain.assign(size_t(P.ns), KStep{});
    e->hlogical.assign(size_t(P.ns), KStep{});
    e->hfused.assign(size_t(P.ns), KStep{});
    e->fusable.assign(size_t(P.ns), 0);
    e->launch.assign(size_t(P.ns), gs_engine::Launch{});
    const int lrmax = plain_lr(e->storage);
    for (int k = 0; k < P.ns; k++) {
        const StepInfo& si = P.steps[size_t(k)];
        KStep& ks = e->hplain[size_t(k)];
        memset(&ks, 0, sizeof ks);
        for (int j = 0; j < si.nk; j++) ks.cols[j] = si.cols[size_t(j)];
        ks.g = si.g;
        ks.pmask[0] = si.rhoc;
        for (int b = 0; b < 2; b++)
            for (int x = 0; x < 2; x++) {
                ks.mre[b][x] = si.M[b][x].real();
                ks.mim[b][x] = si.M[b][x].imag();
            }
        ks.nk = si.nk;
        ks.ntq = si.ntq;
        ks.m = si.m;
        ks.growth = si.growth ? 1 : 0;
        ks.vg = 1;
        std::string err;
        if (int rc = route_and_relabel(si, ks, err)) return fail(rc, err + " at step " + std::to_string(k));
        e->hlogical[size_t(k)] = ks;                     // plan coordinates (dump order)
        // coordinates with the unit physical bits first: a run of 2^lrmax cosets per thread when 1 .. 2^(lrmax-1)
        // are in the live span
        std::vector<uint64_t> basis;
        coord_basis(si.cols, {}, 0, basis);
        apply_basis(si.cols, si.rhoc, basis, ks);
        ks.lr = unit_lead(ks) >= lrmax ? lrmax : 0;
        // Optionally (off by default) steps with few slots per coset give each thread a second run (coordinate lr
        // becomes a thread coordinate), the top coordinate (pair_far: a distant region of the table) or the next unit
        // bit.  Measured on one MI355X: adjacent pairs make the width-30 samples slower (4.47 -> 4.57 ms) and the
        // width-34 ones faster at 2^9 blocks (2.08 -> 1.95 s) but slower at 2^12 (1.66 -> 1.90 s), distant pairs
        // slower everywhere (width 34: 3.6 s); a one-slot step gets its loads in flight from more blocks instead.
        if (ks.lr == lrmax && ks.lr > 0 && ks.ntq <= e->pair_ntq && ks.nk >= ks.lr + 13) {
            ks.vg = 2;
            if (e->pair_far) {
                const int a = ks.lr, b = ks.nk - 1;
                std::swap(ks.cols[a], ks.cols[b]);
                if (((ks.pmask[0] >> a) ^ (ks.pmask[0] >> b)) & 1) ks.pmask[0] ^= (uint64_t(1) << a) | (uint64_t(1) << b);
            }
        }
        set_geometry(ks, ks.lr + (ks.vg == 2 ? 1 : 0),
                     ks.ntq != 1 ? e->grid_cap : ks.nk >= kBigStepLog2 ? e->grid_cap1_big : e->grid_cap1);
        auto& L = e->launch[size_t(k)];
        L.grid = 1 << ks.lg_grid;
        L.gram = sweep_kernel_for(e->storage, ks.lr, ks.vg, ks.ntq, ks.route, kGram, 0);
        L.direct = sweep_kernel_for(e->storage, ks.lr, ks.vg, ks.ntq, ks.route, kDirect, 0);
        L.write = sweep_kernel_for(e->storage, ks.lr, ks.vg, ks.ntq, ks.route, kWrite, 0);
        if (!L.gram || !L.direct || !L.write) return fail(15, "no kernel for step " + std::to_string(k));
    }
    for (int k = 0; k < P.ns; k++) {
        e->hplain[size_t(k)].rev = e->walk == 2 || (e->walk == 1 && (k & 1));
        e->hplain[size_t(k)].xcd = e->xcd_blocks;
    }
    for (int k = 0; k + 1 < P.ns; k++) {
        KStep f;
        auto& L = e->launch[size_t(k)];
        if (build_fused(e, k, f)) {                       // next step has two slots: pairing layout
            L.wgram = sweep_kernel_for(e->storage, f.lr, 2, f.ntq, f.route, kWriteGram, f.route2);
            L.wdirect = sweep_kernel_for(e->storage, f.lr, 2, f.ntq, f.route, kWriteDirect, f.route2);
            if (!L.wgram || !L.wdirect) {
                L.wgram = L.wdirect = nullptr;
                continue;
            }
        } else if (build_fused4(e, k, f)) {               // next step has four slots: two pairing coordinates
            L.wgram = sweep_kernel_for(e->storage, f.lr, 4, f.ntq, f.route, kWriteGram, f.route2);
            L.wdirect = sweep_kernel_for(e->storage, f.lr, 4, f.ntq, f.route, kWriteDirect, f.route2);
            if (!L.wgram || !L.wdirect) {
                L.wgram = L.wdirect = nullptr;
                continue;
            }
        } else if (build_fused1(e, k, f)) {               // next step has one slot: the plain layout (Gram form only)
            L.wgram = sweep_kernel_for(e->storage, f.lr, 1, f.ntq, f.route, kWriteGram1, 0);
            L.wdirect = nullptr;
            if (!L.wgram) continue;
        } else {
            continue;
        }
        L.fgrid = 1 << f.lg_grid;
        e->hfused[size_t(k)] = f;
        e->fusable[size_t(k)] = 1;
    }
    build_runs(e);
    return 0;
}

// Copy the step and run descriptors to the device, growing the run arrays (and allocating the pattern-sum buffers the
// first time the plan has a run).
int upload_descriptors(gs_engine* e, bool logical) {
    const size_t ns = size_t(e->plan.ns);
    if (ns) {
        HIP_TRY(hipMemcpy(e->dplain, e->hplain.data(), sizeof(KStep) * ns, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->dfused, e->hfused.data(), sizeof(KStep) * ns, hipMemcpyHostToDevice));
        if (logical) HIP_TRY(hipMemcpy(e->dlogical, e->hlogical.data(), sizeof(KStep) * ns, hipMemcpyHostToDevice));
    }
    if (e->hrun.empty()) return 0;
    if (e->hrun.size() > e->drun_cap) {
        (void)hipFree(e->drun);
        e->drun = nullptr;
        e->drun_cap = 0;
        if (hipMalloc(&e->drun, sizeof(RStep) * e->hrun.size()) != hipSuccess)
            return fail(11, "allocation of the run descriptors failed");
        e->drun_cap = e->hrun.size();
    }
    if (!e->hpart) {
        if (hipMalloc(&e->hpart, sizeof(double) * (size_t(kRunBins) << kMaxGridLog2)) != hipSuccess ||
            hipMalloc(&e->hchunk, sizeof(double) * kRunChunks * kRunBins) != hipSuccess ||
            hipMalloc(&e->rfac, sizeof(double2) * kRunBins) != hipSuccess)
            return fail(11, "allocation of the run buffers failed");
    }
    HIP_TRY(hipMemcpy(e->drun, e->hrun.data(), sizeof(RStep) * e->hrun.size(), hipMemcpyHostToDevice));
    return 0;
}

void free_engine(gs_engine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    (void)hipFree(e->arena);
    (void)hipFree(e->dplain);
    (void)hipFree(e->dfused);
    (void)hipFree(e->dlogical);
    (void)hipFree(e->part_buf[0]);
    (void)hipFree(e->counter);
    (void)hipFree(e->ctl_base);
    (void)hipFree(e->d_forced);
    (void)hipFree(e->d_U);
    (void)hipFree(e->d_bits);
    (void)hipFree(e->d_cond);
    (void)hipFree(e->d_out);
    (void)hipFree(e->d_amp);
    (void)hipFree(e->drun);
    (void)hipFree(e->hpart);
    (void)hipFree(e->hchunk);
    (void)hipFree(e->rfac);
    (void)hipHostFree(e->h_forced);
    (void)hipHostFree(e->h_U);
    (void)hipHostFree(e->h_bits);
    (void)hipHostFree(e->h_cond);
    (void)hipHostFree(e->h_out);
    (void)hipHostFree(e->h_amp);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// Device time of the enqueued work between begin and end (engines without events -- the scratch-backed prob entry
// point -- only drain the stream).
int begin_timing(gs_engine* e) {
    if (!e->ev0) return 0;
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    return 0;
}

int end_timing(gs_engine* e) {
    if (!e->ev1) {
        HIP_TRY(hipStreamSynchronize(e->stream));
        return 0;
    }
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    HIP_TRY(hipEventSynchronize(e->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_ms += ms;
    if (e->profile) return collect_profile(e);
    return 0;
}

std::complex<double> read_amp(const gs_engine* e, int slot) {
    const unsigned char* p = e->h_amp + size_t(slot) * 16;
    if (e->storage == GS_C64) {
        float v[2];
        memcpy(v, p, 8);
        return {double(v[0]), double(v[1])};
    }
    double v[2];
    memcpy(v, p, 16);
    return {v[0], v[1]};
}

// Copy arena entry 0 into the string's amplitude slot (ordered on the stream).
int enqueue_amp(gs_engine* e, int slot) {
    HIP_TRY(hipMemcpyAsync(e->d_amp + size_t(slot) * 16, e->arena, e->esz, hipMemcpyDeviceToDevice, e->stream));
    return 0;
}

int check_bits(const int8_t* b, int64_t n, bool allow_q) {
    for (int64_t i = 0; i < n; i++)
        if (!(b[i] == 0 || b[i] == 1 || (allow_q && b[i] == -1)))
            return fail(16, "bit values must be 0 or 1" + std::string(allow_q ? " or -1" : ""));
    return 0;
}

// Forced-string modes (prob, chain): strings are processed in chunks of kChunk.
int run_forced(gs_engine* e, int mode, const int8_t* bits, int64_t nstr, double* log2p, double* cond, double* amp) {
    const int ns = ns_of(e);
    if (int rc = check_bits(bits, nstr * ns, false)) return rc;
    HIP_TRY(hipSetDevice(e->device));
    e->last_ms = 0;
    for (int64_t s0 = 0; s0 < nstr; s0 += gs_engine::kChunk) {
        const int cnt = int(std::min<int64_t>(gs_engine::kChunk, nstr - s0));
        memcpy(e->h_forced, bits + s0 * ns, size_t(cnt) * ns);
        HIP_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        if (int rc = begin_timing(e)) return rc;
        for (int i = 0; i < cnt; i++) {
            if (int rc = launch_init(e)) return rc;
            if (int rc = enqueue_run(e, mode, 0, ns, e->h_forced + size_t(i) * ns, e->d_forced + size_t(i) * ns, e->d_U,
                                     e->d_bits + size_t(i) * ns, e->d_cond + size_t(i) * ns, e->d_out + size_t(i) * 4))
                return rc;
            if (amp)
                if (int rc = enqueue_amp(e, i)) return rc;
        }
        if (int rc = end_timing(e)) return rc;
        HIP_TRY(hipMemcpyAsync(e->h_out, e->d_out, size_t(cnt) * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        if (cond)
            HIP_TRY(hipMemcpyAsync(e->h_cond, e->d_cond, size_t(cnt) * ns * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        if (amp) HIP_TRY(hipMemcpyAsync(e->h_amp, e->d_amp, size_t(cnt) * 16, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (int i = 0; i < cnt; i++) {
            log2p[s0 + i] = e->h_out[i * 4];
            if (cond) memcpy(cond + (s0 + i) * ns, e->h_cond + size_t(i) * ns, size_t(ns) * sizeof(double));
            if (amp) {                  // a string of probability 0 stopped early: its amplitude is 0
                const std::complex<double> z = e->h_out[i * 4 + 3] != 0 ? std::complex<double>() : read_amp(e, i);
                amp[2 * (s0 + i)] = z.real();
                amp[2 * (s0 + i) + 1] = z.imag();
            }
        }
    }
    return 0;
}

}  // namespace

// ------------------------------------------------------------------------------------------------------------
// C ABI

extern "C" {

const char* gs_last_error(void) { return g_err.c_str(); }

int gs_plan_load(const char* path, gs_plan** out) {
    if (!path || !out) return fail(16, "null argument");
    auto* p = new gs_plan;
    std::string err;
    if (int rc = parse_plan_file(path, p->p, err)) {
        delete p;
        return fail(rc, err);
    }
    *out = p;
    return 0;
}

int gs_plan_parse(const char* text, gs_plan** out) {
    if (!text || !out) return fail(16, "null argument");
    auto* p = new gs_plan;
    std::string err;
    if (int rc = parse_plan_text(text, p->p, err)) {
        delete p;
        return fail(rc, err);
    }
    *out = p;
    return 0;
}

int gs_plan_info(const gs_plan* plan, int* n, int* ns, int* R) {
    if (!plan) return fail(16, "null plan");
    if (n) *n = plan->p.n;
    if (ns) *ns = plan->p.ns;
    if (R) *R = plan->p.R;
    return 0;
}

void gs_plan_destroy(gs_plan* plan) { delete plan; }

int gs_engine_create(const gs_plan* plan, int storage, int device, gs_engine** out) {
    if (!plan || !out) return fail(16, "null argument");
    if (storage != GS_C64 && storage != GS_C128) return fail(16, "storage must be GS_C64 or GS_C128");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(17, "no HIP device");
    if (device < 0 || device >= ndev) return fail(16, "device index out of range");
    auto* e = new gs_engine;
    e->plan = plan->p;
    e->storage = storage;
    e->device = device;
    e->esz = storage == GS_C64 ? 8 : 16;
    auto bail = [&](int rc) {
        free_engine(e);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(fail(17, "hipSetDevice failed"));
    if (int rc = build_steps(e)) return bail(rc);
    const int ns = e->plan.ns;
    const size_t bytes = (size_t(1) << e->plan.R) * e->esz;
    if (hipMalloc(&e->arena, bytes) != hipSuccess) {
        (void)hipGetLastError();
        e->arena = nullptr;
        return bail(fail(11, "allocation of " + std::to_string(bytes) + " bytes failed"));
    }
    auto dev_alloc = [&](void** p, size_t n) { return hipMalloc(p, std::max<size_t>(n, 16)) == hipSuccess; };
    auto host_alloc = [&](void** p, size_t n) { return hipHostMalloc(p, std::max<size_t>(n, 16), 0) == hipSuccess; };
    const size_t C = gs_engine::kChunk;
    const size_t nsz = size_t(std::max(ns, 1));
    bool ok = dev_alloc((void**)&e->dplain, sizeof(KStep) * nsz) && dev_alloc((void**)&e->dfused, sizeof(KStep) * nsz) &&
              dev_alloc((void**)&e->dlogical, sizeof(KStep) * nsz) &&
              dev_alloc((void**)&e->part_buf[0], 2 * sizeof(double) * kPartDoubles) &&
              dev_alloc((void**)&e->counter, sizeof(unsigned) * kCounters) &&
              dev_alloc((void**)&e->ctl_base, 2 * sizeof(Ctl)) && dev_alloc((void**)&e->d_forced, C * nsz) &&
              dev_alloc((void**)&e->d_U, sizeof(double) * C * nsz) && dev_alloc((void**)&e->d_bits, C * nsz) &&
              dev_alloc((void**)&e->d_cond, sizeof(double) * C * nsz) &&
              dev_alloc((void**)&e->d_out, sizeof(double) * C * 4) && dev_alloc((void**)&e->d_amp, 16 * C) &&
              host_alloc((void**)&e->h_forced, C * nsz) && host_alloc((void**)&e->h_U, sizeof(double) * C * nsz) &&
              host_alloc((void**)&e->h_bits, C * nsz) && host_alloc((void**)&e->h_cond, sizeof(double) * C * nsz) &&
              host_alloc((void**)&e->h_out, sizeof(double) * C * 4) && host_alloc((void**)&e->h_amp, 16 * C);
    if (!ok) return bail(fail(11, "allocation of engine buffers failed"));
    e->part_buf[1] = e->part_buf[0] + kPartDoubles;
    e->part = e->part_buf[0];
    e->ctl = e->ctl_base;
    if (hipStreamCreate(&e->stream) != hipSuccess || hipEventCreate(&e->ev0) != hipSuccess ||
        hipEventCreate(&e->ev1) != hipSuccess)
        return bail(fail(17, "stream/event creation failed"));
    if (int rc = upload_descriptors(e, true)) return bail(rc);
    if (hipMemset(e->counter, 0, sizeof(unsigned) * kCounters) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return bail(fail(17, "counter initialisation failed"));
    e->prof_ms.assign(nsz * kNumSweeps, 0.0);
    e->prof_n.assign(nsz * kNumSweeps, 0);
    *out = e;
    return 0;
}

void gs_engine_destroy(gs_engine* eng) { free_engine(eng); }

int gs_prob(gs_engine* eng, const int8_t* bits, int64_t nstr, double* log2p) {
    if (!eng || (nstr > 0 && (!bits || !log2p)) || nstr < 0) return fail(16, "bad argument");
    return run_forced(eng, kChain, bits, nstr, log2p, nullptr, nullptr);
}

int gs_chain(gs_engine* eng, const int8_t* bits, int64_t nstr, double* log2p, double* cond, double* amp_reim) {
    if (!eng || (nstr > 0 && (!bits || !log2p || !cond)) || nstr < 0) return fail(16, "bad argument");
    return run_forced(eng, kChain1, bits, nstr, log2p, cond, amp_reim);
}

int gs_sample(gs_engine* eng, uint64_t seed, int64_t first, int64_t count, int seeding, int8_t* bits, double* log2p) {
    if (!eng || count < 0 || first < 0 || (count > 0 && (!bits || !log2p))) return fail(16, "bad argument");
    if (seeding != GS_SEED_PLAIN && seeding != GS_SEED_SEQ) return fail(16, "bad seeding");
    gs_engine* e = eng;
    const int ns = ns_of(e);
    HIP_TRY(hipSetDevice(e->device));
    e->last_ms = 0;
    for (int64_t s0 = 0; s0 < count; s0 += gs_engine::kChunk) {
        const int cnt = int(std::min<int64_t>(gs_engine::kChunk, count - s0));
        for (int i = 0; i < cnt; i++) {
            const unsigned long long idx = (unsigned long long)(first + s0 + i);
            std::mt19937_64 rng;
            if (seeding == GS_SEED_PLAIN) {
                rng.seed(seed + idx);
            } else {
                std::seed_seq sq{unsigned(seed & 0xffffffffu), unsigned(seed >> 32), unsigned(idx & 0xffffffffu),
                                 unsigned(idx >> 32), 0x9E3779B9u};
                rng.seed(sq);
            }
            std::uniform_real_distribution<double> U(0, 1);
            for (int k = 0; k < ns; k++) e->h_U[size_t(i) * ns + k] = U(rng);
        }
        HIP_TRY(hipMemcpyAsync(e->d_U, e->h_U, sizeof(double) * size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        if (int rc = begin_timing(e)) return rc;
        for (int i = 0; i < cnt; i++) {
            if (int rc = launch_init(e)) return rc;
            if (int rc = enqueue_run(e, kChain, 0, ns, nullptr, nullptr, e->d_U + size_t(i) * ns,
                                     e->d_bits + size_t(i) * ns, e->d_cond + size_t(i) * ns, e->d_out + size_t(i) * 4))
                return rc;
        }
        if (int rc = end_timing(e)) return rc;
        HIP_TRY(hipMemcpyAsync(e->h_bits, e->d_bits, size_t(cnt) * ns, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(e->h_out, e->d_out, size_t(cnt) * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        memcpy(bits + s0 * ns, e->h_bits, size_t(cnt) * ns);
        for (int i = 0; i < cnt; i++) log2p[s0 + i] = e->h_out[i * 4];
    }
    return 0;
}

int gs_hybrid(gs_engine* eng, const int8_t* bits, int64_t nstr, uint64_t seed, int64_t first_line, int8_t* out_bits,
              double* log2p, double* cond) {
    if (!eng || nstr < 0 || (nstr > 0 && (!bits || !out_bits || !log2p || !cond))) return fail(16, "bad argument");
    gs_engine* e = eng;
    const int ns = ns_of(e);
    if (int rc = check_bits(bits, nstr * ns, true)) return rc;
    HIP_TRY(hipSetDevice(e->device));
    e->last_ms = 0;
    for (int64_t s0 = 0; s0 < nstr; s0 += gs_engine::kChunk) {
        const int cnt = int(std::min<int64_t>(gs_engine::kChunk, nstr - s0));
        memcpy(e->h_forced, bits + s0 * ns, size_t(cnt) * ns);
        for (int i = 0; i < cnt; i++) {
            std::mt19937_64 rng(seed + (unsigned long long)(first_line + s0 + i));
            std::uniform_real_distribution<double> U(0, 1);
            int q = 0;
            for (int k = 0; k < ns; k++)
                if (e->h_forced[size_t(i) * ns + k] < 0) e->h_U[size_t(i) * ns + q++] = U(rng);
        }
        HIP_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->d_U, e->h_U, sizeof(double) * size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        if (int rc = begin_timing(e)) return rc;
        for (int i = 0; i < cnt; i++) {
            if (int rc = launch_init(e)) return rc;
            if (int rc = enqueue_run(e, kRun, 0, ns, e->h_forced + size_t(i) * ns, e->d_forced + size_t(i) * ns,
                                     e->d_U + size_t(i) * ns, e->d_bits + size_t(i) * ns, e->d_cond + size_t(i) * ns,
                                     e->d_out + size_t(i) * 4))
                return rc;
        }
        if (int rc = end_timing(e)) return rc;
        HIP_TRY(hipMemcpyAsync(e->h_bits, e->d_bits, size_t(cnt) * ns, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(e->h_cond, e->d_cond, size_t(cnt) * ns * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(e->h_out, e->d_out, size_t(cnt) * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        memcpy(out_bits + s0 * ns, e->h_bits, size_t(cnt) * ns);
        memcpy(cond + s0 * ns, e->h_cond, size_t(cnt) * ns * sizeof(double));
        for (int i = 0; i < cnt; i++) log2p[s0 + i] = e->h_out[i * 4];
    }
    return 0;
}

int gs_prefix(gs_engine* eng, const int8_t* bits, int K, double* table_reim, int64_t table_cap, int* nk, int* ntq,
              double* logscale, double* log2p_prefix) {
    if (!eng || !bits || !nk || !ntq) return fail(16, "bad argument");
    gs_engine* e = eng;
    const int ns = ns_of(e);
    if (K < 0 || K > ns) return fail(16, "K out of range");
    for (int k = 0; k < K; k++)
        if (bits[k] == -1) return fail(8, "'?' before K in prefix mode");
    if (int rc = check_bits(bits, K, false)) return rc;
    std::vector<uint64_t> offs = K < ns ? e->plan.steps[size_t(K)].sorted_offsets() : std::vector<uint64_t>{0};
    *nk = K < ns ? e->plan.steps[size_t(K)].nk : 0;
    *ntq = int(offs.size());
    if (!table_reim) return 0;
    const int64_t n = (int64_t(1) << *nk) * *ntq;
    if (table_cap < n) return fail(16, "table capacity too small");
    HIP_TRY(hipSetDevice(e->device));
    e->last_ms = 0;
    memcpy(e->h_forced, bits, size_t(ns));
    for (int k = K; k < ns; k++) e->h_forced[k] = 0;
    HIP_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(ns), hipMemcpyHostToDevice, e->stream));
    if (int rc = begin_timing(e)) return rc;
    if (int rc = launch_init(e)) return rc;
    if (int rc = enqueue_run(e, kRun, 0, K, e->h_forced, e->d_forced, e->d_U, e->d_bits, e->d_cond, e->d_out)) return rc;
    if (int rc = end_timing(e)) return rc;
    void* doffs = nullptr;
    void* dtab = nullptr;
    HIP_TRY(hipMalloc(&doffs, sizeof(uint64_t) * offs.size()));
    if (hipMalloc(&dtab, sizeof(double2) * size_t(n)) != hipSuccess) {
        (void)hipFree(doffs);
        return fail(11, "allocation of the dump buffer failed");
    }
    HIP_TRY(hipMemcpy(doffs, offs.data(), sizeof(uint64_t) * offs.size(), hipMemcpyHostToDevice));
    const KStep* sk = K < ns ? e->dlogical + K : nullptr;
    const int grid = int(std::min<int64_t>((n + kThreads - 1) / kThreads, 65536));
    if (e->storage == GS_C64)
        hipLaunchKernelGGL(gather_kernel<float2>, dim3(grid), dim3(kThreads), 0, e->stream, sk,
                           static_cast<const float2*>(e->arena), static_cast<const uint64_t*>(doffs), *ntq, n,
                           static_cast<double2*>(dtab));
    else
        hipLaunchKernelGGL(gather_kernel<double2>, dim3(grid), dim3(kThreads), 0, e->stream, sk,
                           static_cast<const double2*>(e->arena), static_cast<const uint64_t*>(doffs), *ntq, n,
                           static_cast<double2*>(dtab));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(e->h_out, e->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(table_reim, dtab, sizeof(double2) * size_t(n), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    (void)hipFree(doffs);
    (void)hipFree(dtab);
    if (logscale) *logscale = e->h_out[1];
    if (log2p_prefix) *log2p_prefix = e->h_out[0];
    return 0;
}

int gs_resume(gs_engine* eng, const int8_t* bits, int K, int nslices, const double* const* tables,
              const double* logscales, const uint64_t* cshift, const uint64_t* tqshift, uint64_t seed, int64_t line,
              int8_t* out_bits, double* log2p_full, double* log2p_prefix, double* cond) {
    if (!eng || !bits || nslices < 1 || !tables || !logscales || !out_bits || !log2p_full || !log2p_prefix || !cond)
        return fail(16, "bad argument");
    gs_engine* e = eng;
    const int ns = ns_of(e);
    if (K < 0 || K > ns) return fail(16, "K out of range");
    if (int rc = check_bits(bits, ns, true)) return rc;
    const int nk = K < ns ? e->plan.steps[size_t(K)].nk : 0;
    std::vector<uint64_t> offs = K < ns ? e->plan.steps[size_t(K)].sorted_offsets() : std::vector<uint64_t>{0};
    const int ntq = int(offs.size());
    const int64_t n = (int64_t(1) << nk) * ntq;
    // twist: per slice, a shift of the kept-coordinate index and a permutation of the sorted offsets
    std::vector<int> perms(size_t(nslices) * ntq);
    for (int s = 0; s < nslices; s++) {
        if (cshift && (nk < 64 ? (cshift[s] >> nk) != 0 : false))
            return fail(14, "twist cshift exceeds the kept coordinates");
        for (int t = 0; t < ntq; t++) {
            const uint64_t want = offs[size_t(t)] ^ (tqshift ? tqshift[s] : 0);
            auto it = std::find(offs.begin(), offs.end(), want);
            if (it == offs.end()) return fail(14, "twist tqshift does not map the coset offsets onto themselves");
            perms[size_t(s) * ntq + t] = int(it - offs.begin());
        }
    }
    double L = -1e300;
    for (int s = 0; s < nslices; s++) L = std::max(L, logscales[s]);
    HIP_TRY(hipSetDevice(e->device));
    e->last_ms = 0;
    void *dacc = nullptr, *dtab = nullptr, *doffs = nullptr, *dperm = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dacc);
        (void)hipFree(dtab);
        (void)hipFree(doffs);
        (void)hipFree(dperm);
    };
    if (hipMalloc(&dacc, sizeof(double2) * size_t(n)) != hipSuccess ||
        hipMalloc(&dtab, sizeof(double2) * size_t(n)) != hipSuccess ||
        hipMalloc(&doffs, sizeof(uint64_t) * offs.size()) != hipSuccess ||
        hipMalloc(&dperm, sizeof(int) * perms.size()) != hipSuccess) {
        cleanup();
        return fail(11, "allocation of the resume buffers failed");
    }
    int rc = 0;
    const int grid = int(std::min<int64_t>((n + kThreads - 1) / kThreads, 2048));
    do {
        if (hipMemsetAsync(dacc, 0, sizeof(double2) * size_t(n), e->stream) != hipSuccess ||
            hipMemcpy(doffs, offs.data(), sizeof(uint64_t) * offs.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dperm, perms.data(), sizeof(int) * perms.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(17, "resume upload failed");
            break;
        }
        for (int s = 0; s < nslices && !rc; s++) {
            if (hipMemcpyAsync(dtab, tables[s], sizeof(double2) * size_t(n), hipMemcpyHostToDevice, e->stream) != hipSuccess) {
                rc = fail(17, "resume upload failed");
                break;
            }
            const double w = std::exp2(logscales[s] - L);
            hipLaunchKernelGGL(slice_accumulate_kernel, dim3(grid), dim3(kThreads), 0, e->stream,
                               static_cast<double2*>(dacc), static_cast<const double2*>(dtab), w,
                               cshift ? cshift[s] : 0ull, static_cast<const int*>(dperm) + size_t(s) * ntq, ntq, n);
        }
        if (rc) break;
        const KStep* sk = K < ns ? e->dlogical + K : nullptr;
        if (e->storage == GS_C64)
            hipLaunchKernelGGL(scatter_kernel<float2>, dim3(grid), dim3(kThreads), 0, e->stream, sk,
                               static_cast<float2*>(e->arena), static_cast<const uint64_t*>(doffs), ntq, n,
                               static_cast<const double2*>(dacc), e->part);
        else
            hipLaunchKernelGGL(scatter_kernel<double2>, dim3(grid), dim3(kThreads), 0, e->stream, sk,
                               static_cast<double2*>(e->arena), static_cast<const uint64_t*>(doffs), ntq, n,
                               static_cast<const double2*>(dacc), e->part);
        hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(1), 0, e->stream, e->part, grid, e->d_out + 4);
        double norm = 0;
        if (hipMemcpyAsync(&e->h_out[4], e->d_out + 4, sizeof(double), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess) {
            rc = fail(17, "resume reduction failed");
            break;
        }
        norm = e->h_out[4];
        const double lp_prefix = -K + std::log2(norm) + 2 * L;
        // forced bits and uniforms for steps K..ns-1
        memcpy(e->h_forced, bits, size_t(ns));
        std::mt19937_64 rng(seed + (unsigned long long)line);
        std::uniform_real_distribution<double> U(0, 1);
        int q = 0;
        for (int k = K; k < ns; k++)
            if (e->h_forced[k] < 0) e->h_U[q++] = U(rng);
        if (hipMemcpyAsync(e->d_forced, e->h_forced, size_t(ns), hipMemcpyHostToDevice, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->d_U, e->h_U, sizeof(double) * size_t(std::max(q, 1)), hipMemcpyHostToDevice, e->stream) != hipSuccess) {
            rc = fail(17, "resume upload failed");
            break;
        }
        hipLaunchKernelGGL(init_ctl_kernel, dim3(1), dim3(1), 0, e->stream, e->ctl, norm, L);
        if ((rc = begin_timing(e))) break;
        if ((rc = enqueue_run(e, kRun, K, ns, e->h_forced, e->d_forced, e->d_U, e->d_bits, e->d_cond, e->d_out))) break;
        if ((rc = end_timing(e))) break;
        if (hipMemcpyAsync(e->h_bits, e->d_bits, size_t(ns), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->h_cond, e->d_cond, size_t(ns) * sizeof(double), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->h_out, e->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess) {
            rc = fail(17, "resume readback failed");
            break;
        }
        for (int k = 0; k < ns; k++) out_bits[k] = k < K ? bits[k] : e->h_bits[k];
        for (int k = K; k < ns; k++) cond[k] = e->h_cond[k];
        *log2p_prefix = lp_prefix;
        *log2p_full = lp_prefix + e->h_out[0];
    } while (false);
    cleanup();
    return rc;
}

double gs_last_device_ms(const gs_engine* eng) { return eng ? eng->last_ms : 0.0; }

// Copy `bytes` (a multiple of 16) between two fresh device buffers with one launch of a plain streaming kernel; *ms
// receives its duration.  Not a reference mode: it calibrates byte counters (read `bytes`, write `bytes`).
int gs_copy_probe(int device, int64_t bytes, double* ms) {
    if (bytes <= 0 || bytes % 16) return fail(16, "bytes must be a positive multiple of 16");
    HIP_TRY(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, size_t(bytes)) != hipSuccess || hipMalloc(&b, size_t(bytes)) != hipSuccess) {
        (void)hipFree(a);
        return fail(11, "copy probe allocation failed");
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipMemset(a, 1, size_t(bytes)));
    HIP_TRY(hipDeviceSynchronize());
    const int64_t n = bytes / 16;
    HIP_TRY(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(copy_probe_kernel, dim3(2048), dim3(kThreads), 0, nullptr, static_cast<const float4*>(a),
                       static_cast<float4*>(b), n);
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

// Profiling hooks (not part of the reference's interface): per-step, per-sweep kernel time from HIP events recorded
// around every sweep launch on the engine's stream.
int gs_profile_enable(gs_engine* eng, int on) {
    if (!eng) return fail(16, "null engine");
    eng->profile = on != 0;
    std::fill(eng->prof_ms.begin(), eng->prof_ms.end(), 0.0);
    std::fill(eng->prof_n.begin(), eng->prof_n.end(), 0);
    return 0;
}

int gs_profile_read(const gs_engine* eng, double* ms, int64_t* launches) {
    if (!eng) return fail(16, "null engine");
    for (size_t i = 0; i < eng->prof_ms.size(); i++) {
        if (ms) ms[i] = eng->prof_ms[i];
        if (launches) launches[i] = eng->prof_n[i];
    }
    return 0;
}

// Launch geometry of step k, out[10]: plain log2 run, plain blocks, ntq, route, fusable, fused blocks, next route,
// plain log2 tiles per block, fused log2 run and fused log2 tiles per block (-1 when the step does not fuse).
int gs_step_geometry(const gs_engine* eng, int k, int* out) {
    if (!eng || k < 0 || k >= eng->plan.ns || !out) return fail(16, "bad step");
    const KStep& s = eng->hplain[size_t(k)];
    const KStep& f = eng->hfused[size_t(k)];
    out[0] = s.lr;
    out[1] = 1 << s.lg_grid;
    out[2] = s.ntq;
    out[3] = s.route;
    out[4] = eng->fusable[size_t(k)];
    out[5] = eng->fusable[size_t(k)] ? 1 << f.lg_grid : 0;
    out[6] = eng->fusable[size_t(k)] ? f.route2 : 0;
    out[7] = s.lg_iter;
    out[8] = eng->fusable[size_t(k)] ? f.lr : -1;
    out[9] = eng->fusable[size_t(k)] ? f.lg_iter : -1;
    out[10] = s.vg;
    out[11] = eng->fusable[size_t(k)] ? f.vg : -1;
    return 0;
}

// Cap the blocks of one sweep at 2^lg (1..14; default 9, and for the plain sweeps of one-slot steps 12, or 14 from
// 2^30 cosets on); tiles beyond the cap loop inside each block.  Sets every cap to lg (one_slot_only: the one-slot
// caps alone); lg = 0 restores the defaults of the caps it sets.  Rebuilds and uploads the step descriptors
// (synchronous).
static int set_grid_caps(gs_engine* eng, int lg, bool one_slot_only) {
    if (!eng) return fail(16, "null engine");
    if (lg < 0 || lg > kMaxGridLog2) return fail(16, "grid log2 must be in 0.." + std::to_string(kMaxGridLog2));
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    if (!one_slot_only) eng->grid_cap = lg ? lg : kDefaultGridLog2;
    eng->grid_cap1 = lg ? lg : kDefaultGridLog2One;
    eng->grid_cap1_big = lg ? lg : kDefaultGridLog2OneBig;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, true);
}

int gs_set_grid_log2(gs_engine* eng, int lg) { return set_grid_caps(eng, lg, false); }
int gs_set_grid_log2_one_slot(gs_engine* eng, int lg) { return set_grid_caps(eng, lg, true); }

// Coset-coordinate columns of step k as the sweeps use them (fused: the fused layout), into cols[0 .. cap-1];
// returns nk, or -1 (no such step / layout).
int gs_step_columns(const gs_engine* eng, int k, int fused, uint64_t* cols, int cap) {
    if (!eng || k < 0 || k >= eng->plan.ns || !cols) return -1;
    if (fused && !eng->fusable[size_t(k)]) return -1;
    const KStep& s = fused ? eng->hfused[size_t(k)] : eng->hplain[size_t(k)];
    for (int i = 0; i < s.nk && i < cap; i++) cols[i] = s.cols[i];
    return s.nk;
}

// log2 run length of the fused layout (complex64: 1 or 2; complex128: 0; -1 restores the default).  Rebuilds the
// step descriptors.
int gs_set_fused_run(gs_engine* eng, int lr) {
    if (!eng) return fail(16, "null engine");
    if (lr < -1 || lr > (eng->storage == GS_C64 ? 2 : 0)) return fail(16, "run length not available for this storage");
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    eng->fused_run = lr;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, false);
}

// Plain sweeps of steps with at most ntq_max slots per coset (0: none, at most 2) take two runs per thread, the second
// at the top coordinate (far = 1) or at the next unit bit (far = 0); results do not depend on it.  Rebuilds the step
// descriptors.
int gs_set_plain_pairs(gs_engine* eng, int ntq_max, int far) {
    if (!eng) return fail(16, "null engine");
    if (ntq_max < 0 || ntq_max > 2 || far < 0 || far > 1) return fail(16, "ntq_max must be 0..2 and far 0 or 1");
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    eng->pair_ntq = ntq_max;
    eng->pair_far = far;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, false);
}

// Give each XCD a contiguous eighth of every sweep's block coordinates (on = 1) or deal them round-robin (0); results
// do not depend on it beyond the order of the FP64 block sums.  Rebuilds the step descriptors.
int gs_set_xcd_blocks(gs_engine* eng, int on) {
    if (!eng) return fail(16, "null engine");
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    eng->xcd_blocks = on ? 1 : 0;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, false);
}

// Tile order of the sweeps: 0 forward, 1 reverse on odd steps, 2 reverse on every step (results do not depend on it).
int gs_set_walk(gs_engine* eng, int mode) {
    if (!eng) return fail(16, "null engine");
    if (mode < 0 || mode > 2) return fail(16, "walk must be 0, 1 or 2");
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    eng->walk = mode;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, false);
}

// Steps with at most 2^nk cosets run, in runs of two or more, in one single-block launch (nk -1: never; default 10).
int gs_set_small_steps(gs_engine* eng, int nk) {
    if (!eng) return fail(16, "null engine");
    if (nk < -1 || nk > 20) return fail(16, "small-step limit must be -1..20");
    eng->small_nk = nk;
    return 0;
}

// Runs of at most max_steps consecutive diagonal steps (2..8; 0 or 1: off; default 8) of at least 2^min_nk cosets
// (10..40; default 20) take one pattern-sum pass and one product pass instead of one pass per step.  Rebuilds the
// descriptors (synchronous).
int gs_set_diag_runs(gs_engine* eng, int max_steps, int min_nk) {
    if (!eng) return fail(16, "null engine");
    if (max_steps < 0 || max_steps > kMaxRun || min_nk < 10 || min_nk > kMaxKeep)
        return fail(16, "max_steps must be 0.." + std::to_string(kMaxRun) + " and min_nk 10..40");
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipStreamSynchronize(eng->stream));
    eng->run_max = max_steps;
    eng->run_min_nk = min_nk;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, false);
}

// The run of diagonal steps that step k belongs to: out[0] its first step, out[1] its length, out[2] log2 blocks of
// its passes, out[3] log2 tiles per block, out[4] 1 when its product pass can sum the next step's Gram matrix (all -1
// when step k is in no run).
int gs_step_run(const gs_engine* eng, int k, int* out) {
    if (!eng || k < 0 || k >= eng->plan.ns || !out) return fail(16, "bad step");
    out[0] = out[1] = out[2] = out[3] = out[4] = -1;
    for (const RStep& r : eng->hrun)
        if (k >= r.k0 && k < r.k0 + r.j) {
            out[0] = r.k0;
            out[1] = r.j;
            out[2] = r.lg_grid;
            out[3] = r.lg_iter;
            out[4] = r.pair;
        }
    return 0;
}

// Each sweep reduces the block partials the previous sweep left (on = 1), only the sweeps of steps with at most 2^on
// live slots do (on >= 2), or a one-block kernel reduces them between the two launches (0, default: every block
// repeating the reduction and chain step costs more than the launch it saves); results are bitwise the same.
int gs_set_fold_reduce(gs_engine* eng, int on) {
    if (!eng) return fail(16, "null engine");
    if (on < 0 || on > 64) return fail(16, "fold setting must be 0..64");
    eng->fold_r = on == 0 ? -1 : on == 1 ? 64 : on;
    return 0;
}

// Race check of step k's layout on the device (debug mode; overwrites the arena, which serves as the owner map):
// layout 0 the plain sweeps' coordinates, 1 the fused sweep's (the step must fold the next step's norms), 2 the plan's
// own coordinates, 3 the run of diagonal steps starting at k (its passes' slots against the step's own table).
// counts[4]: slots read by two cosets, slots written twice (visited twice by a run's passes), writes onto another
// coset's input (run slots outside the step's table), slots checked.
int gs_race_check(gs_engine* eng, int k, int layout, int64_t* counts) {
    if (!eng || k < 0 || k >= eng->plan.ns || layout < 0 || layout > 3 || !counts) return fail(16, "bad argument");
    gs_engine* e = eng;
    if (layout == 1 && !e->fusable[size_t(k)]) return fail(16, "step " + std::to_string(k) + " has no fused layout");
    const int ri = layout == 3 ? e->run_at[size_t(k)] : -1;
    if (layout == 3 && ri < 0) return fail(16, "no run of diagonal steps starts at step " + std::to_string(k));
    HIP_TRY(hipSetDevice(e->device));
    const KStep* sp = (layout == 1 ? e->dfused : layout == 2 ? e->dlogical : e->dplain) + k;
    const RStep* rp = ri >= 0 ? e->drun + ri : nullptr;
    unsigned long long* own = static_cast<unsigned long long*>(e->arena);
    const int stride = e->storage == GS_C64 ? 1 : 2;
    void* dcnt = nullptr;
    HIP_TRY(hipMalloc(&dcnt, 4 * sizeof(unsigned long long)));
    unsigned long long h[4] = {0, 0, 0, 0};
    int rc = 0;
    do {
        if (hipMemcpy(dcnt, h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) { rc = fail(17, "race check upload"); break; }
        unsigned long long* cnt = static_cast<unsigned long long*>(dcnt);
        const int modes[3] = {0, 1, layout == 3 ? 3 : 2};
        for (int m : modes) {
            const KStep* msp = sp;
            if (m < 3 && layout == 3) msp = e->dplain + k;   // a run is checked against the step's plain table
            hipLaunchKernelGGL(race_kernel, dim3(4096), dim3(256), 0, e->stream, msp, rp, own, stride, m, cnt);
        }
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(e->stream) != hipSuccess ||
            hipMemcpy(h, dcnt, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = fail(17, "race check failed");
            break;
        }
    } while (false);
    (void)hipFree(dcnt);
    for (int i = 0; i < 4; i++) counts[i] = int64_t(h[i]);
    return rc;
}

// Turn the folding of the next step's norms into the previous write off or on (for measurements; default on).
int gs_set_fusion(gs_engine* eng, int on) {
    if (!eng) return fail(16, "null engine");
    eng->allow_fusion = on != 0;
    return 0;
}

// Turn the folding of a four-slot next step's norms into the previous write off or on (default on).  Rebuilds the
// descriptors.
int gs_set_fused4(gs_engine* eng, int on) {
    if (!eng) return fail(16, "null engine");
    eng->allow_fused4 = on != 0;
    if (int rc = build_steps(eng)) return rc;
    return upload_descriptors(eng, false);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------------------
// prob with no device allocation (gs_prob_noalloc).  Every device buffer gs_prob's engine allocates -- the arena,
// the step descriptors, the chain state, the sweeps' block partials and counter, the per-string staging -- is
// carved out of `scratch`, one file-scope device array per device that lives until the process ends.  Host code
// never uses the array's name as an address: the device address comes from hipGetSymbolAddress, looked up once per
// device with that device current.  Calls that overlap on one device share the array and race.

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;     // 64 MiB per device
constexpr int kMaxScratchDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxScratchDevices];              // device address of the array on each device (once found)

// Device address of the scratch array on `device` (which must be current).
int scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !p) return fail(17, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        g_scratch_addr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_addr[device]);
    return 0;
}

__global__ void clear_words_kernel(unsigned* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

extern "C" int gs_prob_noalloc(const gs_plan* plan, int storage, int device, const int8_t* bits, int64_t nstr,
                               double* log2p) {
    if (!plan || nstr < 0 || (nstr > 0 && (!bits || !log2p))) return fail(16, "bad argument");
    if (storage != GS_C64 && storage != GS_C128) return fail(16, "storage must be GS_C64 or GS_C128");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(17, "no HIP device");
    if (device < 0 || device >= ndev || device >= kMaxScratchDevices) return fail(16, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    gs_engine e;                                       // host-side plan of the launches; no device memory of its own
    e.run_max = 0;                                     // no runs of diagonal steps (their pattern sums need 32 MB)
    e.plan = plan->p;
    e.storage = storage;
    e.device = device;
    e.esz = storage == GS_C64 ? 8 : 16;
    const int ns = e.plan.ns;
    if (e.plan.R > 40) return fail(16, "arena of 2^" + std::to_string(e.plan.R) + " entries exceeds the scratch buffer");
    // carve the buffer (256-byte aligned pieces)
    const size_t C = gs_engine::kChunk, nsz = size_t(std::max(ns, 1));
    size_t used = 0;
    auto take = [&](size_t n) {
        const size_t at = used;
        used += (std::max<size_t>(n, 16) + 255) & ~size_t(255);
        return at;
    };
    const size_t o_ctl = take(2 * sizeof(Ctl)), o_counter = take(sizeof(unsigned) * kCounters),
                 o_part = take(2 * sizeof(double) * kPartDoubles), o_out = take(sizeof(double) * C * 4),
                 o_amp = take(16 * C), o_forced = take(C * nsz), o_U = take(sizeof(double) * C * nsz),
                 o_bits = take(C * nsz), o_cond = take(sizeof(double) * C * nsz), o_plain = take(sizeof(KStep) * nsz),
                 o_fused = take(sizeof(KStep) * nsz), o_arena = take((size_t(1) << e.plan.R) * e.esz);
    if (used > kScratchBytes)
        return fail(16, "the plan needs " + std::to_string(used) + " bytes of scratch; gs_prob_noalloc holds " +
                            std::to_string(kScratchBytes) + " (arena 2^R entries: R <= 22 in complex64, 21 in complex128)");
    unsigned char* base = nullptr;
    if (int rc = scratch_address(device, &base)) return rc;
    if (int rc = build_steps(&e)) return rc;
    e.ctl_base = e.ctl = reinterpret_cast<Ctl*>(base + o_ctl);
    e.counter = reinterpret_cast<unsigned*>(base + o_counter);
    e.part = e.part_buf[0] = reinterpret_cast<double*>(base + o_part);
    e.part_buf[1] = e.part_buf[0] + kPartDoubles;
    e.d_out = reinterpret_cast<double*>(base + o_out);
    e.d_amp = base + o_amp;
    e.d_forced = reinterpret_cast<int8_t*>(base + o_forced);
    e.d_U = reinterpret_cast<double*>(base + o_U);
    e.d_bits = reinterpret_cast<int8_t*>(base + o_bits);
    e.d_cond = reinterpret_cast<double*>(base + o_cond);
    e.dplain = reinterpret_cast<KStep*>(base + o_plain);
    e.dfused = reinterpret_cast<KStep*>(base + o_fused);
    e.arena = base + o_arena;
    // host staging in ordinary (pageable) memory; the null stream; no timing events
    std::vector<int8_t> h_forced(C * nsz);
    std::vector<double> h_out(C * 4), h_cond(C * nsz);
    std::vector<unsigned char> h_amp(16 * C);
    e.h_forced = h_forced.data();
    e.h_out = h_out.data();
    e.h_cond = h_cond.data();
    e.h_amp = h_amp.data();
    if (ns > 0) {
        HIP_TRY(hipMemcpy(e.dplain, e.hplain.data(), sizeof(KStep) * size_t(ns), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e.dfused, e.hfused.data(), sizeof(KStep) * size_t(ns), hipMemcpyHostToDevice));
    }
    // the sweeps' block counter must start at 0 (a kernel fill, not a memset node: callers may capture this)
    hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, e.stream, e.counter, kCounters);
    HIP_TRY(hipGetLastError());
    return run_forced(&e, kChain, bits, nstr, log2p, nullptr, nullptr);
}
