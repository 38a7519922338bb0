// Two-table (gather-form) engine: the reference's gsamp.cpp sampler and dp.cpp amplitude engine on the device.
//
// Each step reads the table of step k (2^r entries) and writes the next one (2^rn entries) into a second buffer: output
// cp gathers the 2^nker inputs u = pdep(cp, pivmask) ^ kspan[t] (gsamp.cpp:21-29).  Writes are contiguous, reads are
// gathered, and the two tables cost 2 x 2^maxr entries.  The chain logic -- branch norms summed directly, the decision
// U_k < p1 from pre-drawn uniforms, chain1 conditionals nb / (2 prev), run_steps rescaling -- is the in-place engine's
// (the same device functions), so a sample is one stream of kernels with no host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/gsamp_amd.h"
#include "gs_device.h"

using namespace gs;

namespace gs {
void* gather_kernel_ptr(int storage, int sweep, int nks);
void* table_one_kernel_ptr(int storage);
void* decide_kernel_ptr();
void* publish_kernel_ptr();
void* reduce_kernel_ptr(int sweep);
}  // namespace gs

namespace {

thread_local std::string g_tterr;

int ttfail(int code, const std::string& m) {
    g_tterr = m;
    return code;
}

#define TT_TRY(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return ttfail(17, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

struct TTStep {
    int v, r, rn;
    uint64_t rho, pivmask;
    std::vector<uint64_t> ker;
    std::complex<double> M[2][2];
};

__global__ void tt_init_ctl_kernel(Ctl* ctl, double prev, double logscale) {
    Ctl c{};
    c.prev = prev;
    c.nstored = prev;
    c.logscale = logscale;
    *ctl = c;
}

}  // namespace

struct gs_tt_plan {
    int n = 0, ns = 0, maxr = 0;
    bool with_m = false;
    std::vector<TTStep> steps;
};

struct gs_tt_engine {
    gs_tt_plan plan;
    int storage = GS_C64, device = 0;
    size_t esz = 8;
    hipStream_t stream = nullptr;
    void* tab[2] = {nullptr, nullptr};   // the two tables, 2^maxr entries each
    TStep* dsteps = nullptr;
    KStep* dplain = nullptr;             // M of each step (read by the decision)
    std::vector<TStep> hsteps;
    double* part = nullptr;
    unsigned* counter = nullptr;
    Ctl* ctl = nullptr;
    static constexpr int kChunk = 64;
    int8_t* d_forced = nullptr;
    double* d_U = nullptr;
    int8_t* d_bits = nullptr;
    double* d_cond = nullptr;
    double* d_out = nullptr;
    double* d_w = nullptr;               // amplitude engine: per-step weights of one slice [ns][4]
    unsigned char* d_amp = nullptr;      // [kChunk][16]
    int8_t* h_forced = nullptr;
    double* h_U = nullptr;
    int8_t* h_bits = nullptr;
    double* h_cond = nullptr;
    double* h_out = nullptr;
    double* h_w = nullptr;
    unsigned char* h_amp = nullptr;
    int cur = 0;                         // table holding the current step's input after the last enqueue
    int grid_cap = kDefaultTTGridLog2;   // log2 of the most blocks a gather sweep launches
};

namespace {

// Parse the gsamp plan (with_m: `n ns maxr`, per step `v r rn rho pivmask nker ker... M[8 doubles]`,
// gsamp.cpp:96-107) or the dp plan (no M, dp.cpp:16-25).  Exit codes: 2 header, 3 step head, 4 kernel, 5 M; 15 for a
// step the engine cannot run (widths above 40, pivmask weight not rn, indices outside the tables).
int parse_tt(FILE* f, bool with_m, gs_tt_plan& P) {
    if (fscanf(f, "%d %d %d", &P.n, &P.ns, &P.maxr) != 3) return ttfail(2, "bad plan header");
    if (P.ns < 0 || P.maxr < 0 || P.maxr > 40) return ttfail(15, "plan widths must be at most 40");
    P.with_m = with_m;
    P.steps.resize(size_t(P.ns));
    for (int k = 0; k < P.ns; k++) {
        TTStep& s = P.steps[size_t(k)];
        unsigned long long rho, pm;
        int nk;
        if (fscanf(f, "%d %d %d %llu %llu %d", &s.v, &s.r, &s.rn, &rho, &pm, &nk) != 6)
            return ttfail(3, "bad step " + std::to_string(k));
        s.rho = rho;
        s.pivmask = pm;
        for (int i = 0; i < nk; i++) {
            unsigned long long x;
            if (fscanf(f, "%llu", &x) != 1) return ttfail(4, "bad kernel vector at step " + std::to_string(k));
            s.ker.push_back(x);
        }
        if (with_m)
            for (int b = 0; b < 2; b++)
                for (int x = 0; x < 2; x++) {
                    double re, im;
                    if (fscanf(f, "%lf %lf", &re, &im) != 2) return ttfail(5, "bad M at step " + std::to_string(k));
                    s.M[b][x] = {re, im};
                }
        if (s.r < 0 || s.rn < 0 || s.r > P.maxr || s.rn > P.maxr || nk < 0 || nk > 2 ||
            __builtin_popcountll(s.pivmask) != s.rn || (s.pivmask >> (s.r + 1)) != 0 || (s.rho >> s.r) != 0)
            return ttfail(15, "step " + std::to_string(k) + " is outside what the engine runs");
        for (uint64_t x : s.ker)
            if (x >> (s.r + 1)) return ttfail(15, "kernel vector outside the table at step " + std::to_string(k));
    }
    return 0;
}

void free_tt(gs_tt_engine* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (void* p : {e->tab[0], e->tab[1], (void*)e->dsteps, (void*)e->dplain, (void*)e->part, (void*)e->counter,
                    (void*)e->ctl, (void*)e->d_forced, (void*)e->d_U, (void*)e->d_bits, (void*)e->d_cond,
                    (void*)e->d_out, (void*)e->d_w, (void*)e->d_amp})
        (void)hipFree(p);
    for (void* p : {(void*)e->h_forced, (void*)e->h_U, (void*)e->h_bits, (void*)e->h_cond, (void*)e->h_out,
                    (void*)e->h_w, (void*)e->h_amp})
        (void)hipHostFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int launch_one(gs_tt_engine* e, void* fn, Epi epi, int k) {
    Ctl* ctl = e->ctl;
    if (k >= 0) {
        void* args[] = {&ctl, &epi, &k};
        TT_TRY(hipLaunchKernel(fn, dim3(1), dim3(1), args, 0, e->stream));
    } else {
        void* args[] = {&ctl, &epi};
        TT_TRY(hipLaunchKernel(fn, dim3(1), dim3(1), args, 0, e->stream));
    }
    return 0;
}

// One gather sweep of step k reading table `in` (norm pass) or reading `in` and writing `1 - in` (write pass).
int launch_gather(gs_tt_engine* e, int k, int sweep, int in, const Epi& epi, const double* wext) {
    const TStep* sp = e->dsteps + k;
    const void* old = e->tab[in];
    void* nw = e->tab[1 - in];
    double* part = e->part;
    Ctl* ctl = e->ctl;
    Epi ep = epi;
    const double* w = wext;
    void* args[] = {&sp, &old, &nw, &part, &ctl, &ep, &w};
    const int grid = 1 << e->hsteps[size_t(k)].lg_grid;
    void* fn = gather_kernel_ptr(e->storage, sweep, e->hsteps[size_t(k)].nks);
    if (!fn) return ttfail(15, "no gather kernel for step " + std::to_string(k));
    TT_TRY(hipLaunchKernel(fn, dim3(grid), dim3(kThreads), args, 0, e->stream));
    if (grid > kFenceBlocks) {
        const double* cpart = part;
        unsigned nblk = unsigned(grid);
        int sw = sweep;
        void* rargs[] = {&cpart, &nblk, &ctl, &ep, &sw};
        TT_TRY(hipLaunchKernel(reduce_kernel_ptr(sweep), dim3(1), dim3(kThreads), rargs, 0, e->stream));
    }
    return 0;
}

// Norms a step needs (gsamp.cpp chain: both branch norms directly; chain1 and forced run_steps steps: none).
int tt_norm(int mode, int forced_bit) {
    if (mode == kChain1) return -1;
    if (mode == kRun && forced_bit >= 0) return -1;
    return kDirect;
}

// Steps k0..k1-1 of one string from table e->cur; leaves the final table in e->cur.
int tt_enqueue(gs_tt_engine* e, int mode, int k0, int k1, const int8_t* forced_h, const int8_t* forced_d,
               const double* U_d, int8_t* bits_d, double* cond_d, double* out_d, const double* wext) {
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
    if (k0 == k1) return launch_one(e, publish_kernel_ptr(), base, -1);
    auto fb = [&](int k) { return forced_h ? int(forced_h[k]) : -1; };
    bool decided = false;
    for (int k = k0; k < k1; k++) {
        if (!decided && !wext) {
            const int nm = tt_norm(mode, fb(k));
            Epi ep = base;
            ep.k = k;
            ep.role = kRoleNorm;
            if (nm >= 0) {
                if (int rc = launch_gather(e, k, kDirect, e->cur, ep, nullptr)) return rc;
            } else {
                if (int rc = launch_one(e, decide_kernel_ptr(), base, k)) return rc;
            }
        }
        Epi ep = base;
        ep.k = k;
        ep.role = kRoleWrite;
        if (k + 1 == k1) ep.next = kNextPublish;
        else if (wext) ep.next = kNextSweep;              // amplitude engine: weights given, no decision
        else if (tt_norm(mode, fb(k + 1)) < 0) ep.next = kNextDecide;
        else ep.next = kNextSweep;
        if (int rc = launch_gather(e, k, kWrite, e->cur, ep, wext)) return rc;
        e->cur = 1 - e->cur;
        decided = ep.next == kNextDecide;
    }
    return 0;
}

int tt_init(gs_tt_engine* e, double prev, double logscale) {
    e->cur = 0;
    void* t0 = e->tab[0];
    void* args[] = {&t0};
    TT_TRY(hipLaunchKernel(table_one_kernel_ptr(e->storage), dim3(1), dim3(1), args, 0, e->stream));
    hipLaunchKernelGGL(tt_init_ctl_kernel, dim3(1), dim3(1), 0, e->stream, e->ctl, prev, logscale);
    TT_TRY(hipGetLastError());
    return 0;
}

int check_tt_bits(const int8_t* b, int64_t n, bool allow_q) {
    for (int64_t i = 0; i < n; i++)
        if (!(b[i] == 0 || b[i] == 1 || (allow_q && b[i] == -1))) return ttfail(16, "bits must be 0, 1" + std::string(allow_q ? " or -1" : ""));
    return 0;
}

// Final entry of the current table of string slot i (the amplitude after the last step).
int enqueue_amp(gs_tt_engine* e, int slot) {
    TT_TRY(hipMemcpyAsync(e->d_amp + size_t(slot) * 16, e->tab[e->cur], e->esz, hipMemcpyDeviceToDevice, e->stream));
    return 0;
}

std::complex<double> read_amp(const gs_tt_engine* e, int slot) {
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

int run_forced(gs_tt_engine* e, int mode, const int8_t* bits, int64_t nstr, double* log2p, double* cond, double* amp) {
    const int ns = e->plan.ns;
    if (int rc = check_tt_bits(bits, nstr * ns, false)) return rc;
    TT_TRY(hipSetDevice(e->device));
    for (int64_t s0 = 0; s0 < nstr; s0 += gs_tt_engine::kChunk) {
        const int cnt = int(std::min<int64_t>(gs_tt_engine::kChunk, nstr - s0));
        memcpy(e->h_forced, bits + s0 * ns, size_t(cnt) * ns);
        TT_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        for (int i = 0; i < cnt; i++) {
            if (int rc = tt_init(e, 1.0, 0.0)) return rc;
            if (int rc = tt_enqueue(e, mode, 0, ns, e->h_forced + size_t(i) * ns, e->d_forced + size_t(i) * ns,
                                    e->d_U, e->d_bits + size_t(i) * ns, e->d_cond + size_t(i) * ns,
                                    e->d_out + size_t(i) * 4, nullptr))
                return rc;
            if (amp)
                if (int rc = enqueue_amp(e, i)) return rc;
        }
        TT_TRY(hipMemcpyAsync(e->h_out, e->d_out, size_t(cnt) * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        if (cond)
            TT_TRY(hipMemcpyAsync(e->h_cond, e->d_cond, size_t(cnt) * ns * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        if (amp) TT_TRY(hipMemcpyAsync(e->h_amp, e->d_amp, size_t(cnt) * 16, hipMemcpyDeviceToHost, e->stream));
        TT_TRY(hipStreamSynchronize(e->stream));
        for (int i = 0; i < cnt; i++) {
            log2p[s0 + i] = e->h_out[i * 4];
            if (cond) memcpy(cond + (s0 + i) * ns, e->h_cond + size_t(i) * ns, size_t(ns) * sizeof(double));
            if (amp) {
                const std::complex<double> z = read_amp(e, i);
                amp[2 * (s0 + i)] = z.real();
                amp[2 * (s0 + i) + 1] = z.imag();
            }
        }
    }
    return 0;
}

// Blocks and tiles per block of every step's gather sweep (2^rn outputs, one per thread and tile), and its partner
// direction: an output-space vector pdelta whose image pD = pdep(pdelta) is xbit ^ kspan[q] for some q.  Output
// u0 ^ pD then gathers the entries (u0 ^ kspan[t ^ q]) & lowmask, the same as output u0: the x bit is not an address
// bit of the old table.  Used when the top bit of pdelta is a tile bit above the first (the kernel's groups of 4
// tiles pair j with j + 2 along pdelta).
void tt_geometry(gs_tt_engine* e) {
    for (auto& t : e->hsteps) {
        const int tiles = std::max(0, t.rn - 8);
        t.lg_grid = std::min(tiles, e->grid_cap);
        t.lg_iter = tiles - t.lg_grid;
        t.ph = -1;
        t.pdelta = t.pD = 0;
        t.ptd = 0;
        if (t.lg_iter < 2) continue;
        const int jit = 8 + t.lg_grid;
        for (int q = 0; q < t.nks; q++) {
            const uint64_t D = t.xbit ^ t.kspan[q];
            uint64_t delta = 0, rest = D;
            for (int j = 0; j < t.rn; j++)
                if (D & t.dep[j]) {
                    delta |= uint64_t(1) << j;
                    rest &= ~t.dep[j];
                }
            if (rest || !delta) continue;             // D outside the outputs' image
            const int h = 63 - __builtin_clzll(delta);
            if (h >= jit + 1) {
                t.ph = h;
                t.pdelta = delta;
                t.pD = D;
                t.ptd = q;
            }
            break;
        }
    }
}

}  // namespace

extern "C" {

int gs_tt_plan_load(const char* path, int with_m, gs_tt_plan** out) {
    if (!path || !out) return ttfail(16, "null argument");
    FILE* f = fopen(path, "r");
    if (!f) return ttfail(2, std::string("cannot open ") + path);
    auto* P = new gs_tt_plan;
    const int rc = parse_tt(f, with_m != 0, *P);
    fclose(f);
    if (rc) {
        delete P;
        return rc;
    }
    *out = P;
    return 0;
}

int gs_tt_plan_info(const gs_tt_plan* plan, int* n, int* ns, int* maxr) {
    if (!plan) return ttfail(16, "null plan");
    if (n) *n = plan->n;
    if (ns) *ns = plan->ns;
    if (maxr) *maxr = plan->maxr;
    return 0;
}

void gs_tt_plan_destroy(gs_tt_plan* plan) { delete plan; }

int gs_tt_engine_create(const gs_tt_plan* plan, int storage, int device, gs_tt_engine** out) {
    if (!plan || !out) return ttfail(16, "null argument");
    if (storage != GS_C64 && storage != GS_C128) return ttfail(16, "storage must be GS_C64 or GS_C128");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ttfail(17, "no HIP device");
    if (device < 0 || device >= ndev) return ttfail(16, "device index out of range");
    auto* e = new gs_tt_engine;
    e->plan = *plan;
    e->storage = storage;
    e->device = device;
    e->esz = storage == GS_C64 ? 8 : 16;
    auto bail = [&](int rc) {
        free_tt(e);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(ttfail(17, "hipSetDevice failed"));
    const int ns = e->plan.ns;
    const size_t nsz = size_t(std::max(ns, 1));
    e->hsteps.assign(nsz, TStep{});
    std::vector<KStep> hplain(nsz, KStep{});
    for (int k = 0; k < ns; k++) {
        const TTStep& s = e->plan.steps[size_t(k)];
        TStep& t = e->hsteps[size_t(k)];
        int j = 0;
        for (int b = 0; b < 64; b++)
            if ((s.pivmask >> b) & 1) t.dep[j++] = uint64_t(1) << b;
        t.kspan[0] = 0;
        t.nks = 1;
        for (uint64_t x : s.ker) {
            for (int i = 0; i < t.nks; i++) t.kspan[t.nks + i] = t.kspan[i] ^ x;
            t.nks *= 2;
        }
        t.lowmask = (uint64_t(1) << s.r) - 1;
        t.xbit = uint64_t(1) << s.r;
        t.rho = s.rho;
        t.rn = s.rn;
        for (int b = 0; b < 2; b++)
            for (int x = 0; x < 2; x++) {
                hplain[size_t(k)].mre[b][x] = s.M[b][x].real();
                hplain[size_t(k)].mim[b][x] = s.M[b][x].imag();
            }
    }
    tt_geometry(e);
    const size_t tbytes = (size_t(1) << e->plan.maxr) * e->esz;
    for (int i = 0; i < 2; i++)
        if (hipMalloc(&e->tab[i], tbytes) != hipSuccess) {
            (void)hipGetLastError();
            e->tab[i] = nullptr;
            return bail(ttfail(11, "allocation of 2 x " + std::to_string(tbytes) + " bytes failed"));
        }
    const size_t C = gs_tt_engine::kChunk;
    auto dev_alloc = [&](void** p, size_t n) { return hipMalloc(p, std::max<size_t>(n, 16)) == hipSuccess; };
    auto host_alloc = [&](void** p, size_t n) { return hipHostMalloc(p, std::max<size_t>(n, 16), 0) == hipSuccess; };
    bool ok = dev_alloc((void**)&e->dsteps, sizeof(TStep) * nsz) && dev_alloc((void**)&e->dplain, sizeof(KStep) * nsz) &&
              dev_alloc((void**)&e->part, sizeof(double) * kPartDoubles) &&
              dev_alloc((void**)&e->counter, sizeof(unsigned) * kCounters) && dev_alloc((void**)&e->ctl, sizeof(Ctl)) &&
              dev_alloc((void**)&e->d_forced, C * nsz) && dev_alloc((void**)&e->d_U, sizeof(double) * C * nsz) &&
              dev_alloc((void**)&e->d_bits, C * nsz) && dev_alloc((void**)&e->d_cond, sizeof(double) * C * nsz) &&
              dev_alloc((void**)&e->d_out, sizeof(double) * C * 4) && dev_alloc((void**)&e->d_w, sizeof(double) * 4 * nsz) &&
              dev_alloc((void**)&e->d_amp, 16 * C) && host_alloc((void**)&e->h_forced, C * nsz) &&
              host_alloc((void**)&e->h_U, sizeof(double) * C * nsz) && host_alloc((void**)&e->h_bits, C * nsz) &&
              host_alloc((void**)&e->h_cond, sizeof(double) * C * nsz) &&
              host_alloc((void**)&e->h_out, sizeof(double) * C * 4) && host_alloc((void**)&e->h_w, sizeof(double) * 4 * nsz) &&
              host_alloc((void**)&e->h_amp, 16 * C);
    if (!ok) return bail(ttfail(11, "allocation of engine buffers failed"));
    if (hipStreamCreate(&e->stream) != hipSuccess) return bail(ttfail(17, "stream creation failed"));
    if (hipMemcpy(e->dsteps, e->hsteps.data(), sizeof(TStep) * nsz, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->dplain, hplain.data(), sizeof(KStep) * nsz, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(e->counter, 0, sizeof(unsigned) * kCounters) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return bail(ttfail(17, "plan upload failed"));
    *out = e;
    return 0;
}

void gs_tt_engine_destroy(gs_tt_engine* eng) { free_tt(eng); }

int gs_tt_set_grid_log2(gs_tt_engine* eng, int lg) {
    if (!eng) return ttfail(16, "null engine");
    if (lg < 0 || lg > kMaxGridLog2) return ttfail(16, "grid log2 must be in 0.." + std::to_string(kMaxGridLog2));
    if (hipSetDevice(eng->device) != hipSuccess || hipStreamSynchronize(eng->stream) != hipSuccess)
        return ttfail(17, "device synchronisation failed");
    eng->grid_cap = lg ? lg : kDefaultTTGridLog2;
    tt_geometry(eng);
    if (hipMemcpy(eng->dsteps, eng->hsteps.data(), sizeof(TStep) * eng->hsteps.size(), hipMemcpyHostToDevice) !=
        hipSuccess)
        return ttfail(17, "step upload failed");
    return 0;
}

int gs_tt_prob(gs_tt_engine* eng, const int8_t* bits, int64_t nstr, double* log2p) {
    if (!eng || !eng->plan.with_m || nstr < 0 || (nstr > 0 && (!bits || !log2p))) return ttfail(16, "bad argument");
    return run_forced(eng, kChain, bits, nstr, log2p, nullptr, nullptr);
}

int gs_tt_chain(gs_tt_engine* eng, const int8_t* bits, int64_t nstr, double* log2p, double* cond, double* amp_reim) {
    if (!eng || !eng->plan.with_m || nstr < 0 || (nstr > 0 && (!bits || !log2p))) return ttfail(16, "bad argument");
    return run_forced(eng, kChain1, bits, nstr, log2p, cond, amp_reim);
}

int gs_tt_sample(gs_tt_engine* e, uint64_t seed, int64_t first, int64_t count, int8_t* bits, double* log2p) {
    if (!e || !e->plan.with_m || count < 0 || first < 0 || (count > 0 && (!bits || !log2p))) return ttfail(16, "bad argument");
    const int ns = e->plan.ns;
    TT_TRY(hipSetDevice(e->device));
    for (int64_t s0 = 0; s0 < count; s0 += gs_tt_engine::kChunk) {
        const int cnt = int(std::min<int64_t>(gs_tt_engine::kChunk, count - s0));
        for (int i = 0; i < cnt; i++) {
            std::mt19937_64 rng(seed + (unsigned long long)(first + s0 + i));          // gsamp.cpp:115
            std::uniform_real_distribution<double> U(0, 1);
            for (int k = 0; k < ns; k++) e->h_U[size_t(i) * ns + k] = U(rng);
        }
        TT_TRY(hipMemcpyAsync(e->d_U, e->h_U, sizeof(double) * size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        for (int i = 0; i < cnt; i++) {
            if (int rc = tt_init(e, 1.0, 0.0)) return rc;
            if (int rc = tt_enqueue(e, kChain, 0, ns, nullptr, nullptr, e->d_U + size_t(i) * ns,
                                    e->d_bits + size_t(i) * ns, e->d_cond + size_t(i) * ns, e->d_out + size_t(i) * 4,
                                    nullptr))
                return rc;
        }
        TT_TRY(hipMemcpyAsync(e->h_bits, e->d_bits, size_t(cnt) * ns, hipMemcpyDeviceToHost, e->stream));
        TT_TRY(hipMemcpyAsync(e->h_out, e->d_out, size_t(cnt) * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        TT_TRY(hipStreamSynchronize(e->stream));
        memcpy(bits + s0 * ns, e->h_bits, size_t(cnt) * ns);
        for (int i = 0; i < cnt; i++) log2p[s0 + i] = e->h_out[i * 4];
    }
    return 0;
}

int gs_tt_hybrid(gs_tt_engine* e, const int8_t* bits, int64_t nstr, uint64_t seed, int64_t first_line,
                 int8_t* out_bits, double* log2p, double* cond) {
    if (!e || !e->plan.with_m || nstr < 0 || (nstr > 0 && (!bits || !out_bits || !log2p || !cond)))
        return ttfail(16, "bad argument");
    const int ns = e->plan.ns;
    if (int rc = check_tt_bits(bits, nstr * ns, true)) return rc;
    TT_TRY(hipSetDevice(e->device));
    for (int64_t s0 = 0; s0 < nstr; s0 += gs_tt_engine::kChunk) {
        const int cnt = int(std::min<int64_t>(gs_tt_engine::kChunk, nstr - s0));
        memcpy(e->h_forced, bits + s0 * ns, size_t(cnt) * ns);
        for (int i = 0; i < cnt; i++) {                     // gsamp.cpp:121-122: one generator per line
            std::mt19937_64 rng(seed + (unsigned long long)(first_line + s0 + i));
            std::uniform_real_distribution<double> U(0, 1);
            int q = 0;
            for (int k = 0; k < ns; k++)
                if (e->h_forced[size_t(i) * ns + k] < 0) e->h_U[size_t(i) * ns + q++] = U(rng);
        }
        TT_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        TT_TRY(hipMemcpyAsync(e->d_U, e->h_U, sizeof(double) * size_t(cnt) * ns, hipMemcpyHostToDevice, e->stream));
        for (int i = 0; i < cnt; i++) {
            if (int rc = tt_init(e, 1.0, 0.0)) return rc;
            if (int rc = tt_enqueue(e, kRun, 0, ns, e->h_forced + size_t(i) * ns, e->d_forced + size_t(i) * ns,
                                    e->d_U + size_t(i) * ns, e->d_bits + size_t(i) * ns, e->d_cond + size_t(i) * ns,
                                    e->d_out + size_t(i) * 4, nullptr))
                return rc;
        }
        TT_TRY(hipMemcpyAsync(e->h_bits, e->d_bits, size_t(cnt) * ns, hipMemcpyDeviceToHost, e->stream));
        TT_TRY(hipMemcpyAsync(e->h_cond, e->d_cond, size_t(cnt) * ns * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        TT_TRY(hipMemcpyAsync(e->h_out, e->d_out, size_t(cnt) * 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        TT_TRY(hipStreamSynchronize(e->stream));
        for (int i = 0; i < cnt; i++) {
            for (int k = 0; k < ns; k++) {
                const int8_t f = e->h_forced[size_t(i) * ns + k];
                out_bits[(s0 + i) * ns + k] = f >= 0 ? f : e->h_bits[size_t(i) * ns + k];
            }
            log2p[s0 + i] = e->h_out[i * 4];
            memcpy(cond + (s0 + i) * ns, e->h_cond + size_t(i) * ns, size_t(ns) * sizeof(double));
        }
    }
    return 0;
}

// prefix (gsamp.cpp:124-135): forced steps 0..K-1 of one string (run_steps semantics), then the contiguous table read
// by step K (2^r entries, r = st[K].r; K == ns: the amplitude) as complex doubles, its logscale and log2 Pr(prefix).
// table_reim = NULL: only *r_out is set (no steps run).
int gs_tt_prefix(gs_tt_engine* e, const int8_t* bits, int K, double* table_reim, int64_t table_cap, int* r_out,
                 double* logscale, double* log2p_prefix) {
    if (!e || !e->plan.with_m || !bits || !r_out || K < 0 || K > e->plan.ns) return ttfail(16, "bad argument");
    const int ns = e->plan.ns;
    for (int k = 0; k < K; k++)
        if (bits[k] != 0 && bits[k] != 1) return ttfail(8, "'?' before K");
    const int r = K < ns ? e->plan.steps[size_t(K)].r : 0;
    *r_out = r;
    if (!table_reim) return 0;
    const int64_t n = int64_t(1) << r;
    if (table_cap < n || !logscale || !log2p_prefix) return ttfail(16, "table capacity below 2^r");
    TT_TRY(hipSetDevice(e->device));
    memcpy(e->h_forced, bits, size_t(ns));
    TT_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(ns), hipMemcpyHostToDevice, e->stream));
    if (int rc = tt_init(e, 1.0, 0.0)) return rc;
    if (int rc = tt_enqueue(e, kRun, 0, K, e->h_forced, e->d_forced, e->d_U, e->d_bits, e->d_cond, e->d_out, nullptr))
        return rc;
    std::vector<unsigned char> raw(size_t(n) * e->esz);
    TT_TRY(hipMemcpyAsync(raw.data(), e->tab[e->cur], raw.size(), hipMemcpyDeviceToHost, e->stream));
    TT_TRY(hipMemcpyAsync(e->h_out, e->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    TT_TRY(hipStreamSynchronize(e->stream));
    for (int64_t i = 0; i < n; i++) {
        if (e->storage == GS_C64) {
            float v[2];
            memcpy(v, raw.data() + size_t(i) * 8, 8);
            table_reim[2 * i] = v[0];
            table_reim[2 * i + 1] = v[1];
        } else {
            memcpy(table_reim + 2 * i, raw.data() + size_t(i) * 16, 16);
        }
    }
    *log2p_prefix = e->h_out[0];
    *logscale = e->h_out[1];
    return 0;
}

// resume (gsamp.cpp:136-148) for one line: the slices' step-K tables summed with weights 2^(ls_s - max ls) into the
// first table (stored in the table type), then steps K..ns-1 with '?' (-1) sampled from mt19937_64(seed + line).
int gs_tt_resume(gs_tt_engine* e, const int8_t* bits, int K, int nslices, const double* const* tables,
                 const double* logscales, uint64_t seed, int64_t line, int8_t* out_bits, double* log2p_full,
                 double* log2p_prefix, double* cond) {
    if (!e || !e->plan.with_m || !bits || K < 0 || K > e->plan.ns || nslices <= 0 || !tables || !logscales ||
        !out_bits || !log2p_full || !log2p_prefix || !cond)
        return ttfail(16, "bad argument");
    const int ns = e->plan.ns;
    if (int rc = check_tt_bits(bits, ns, true)) return rc;
    const int r = K < ns ? e->plan.steps[size_t(K)].r : 0;
    const int64_t n = int64_t(1) << r;
    double L = -1e300;
    for (int s = 0; s < nslices; s++) L = std::max(L, logscales[s]);
    std::vector<unsigned char> raw(size_t(n) * e->esz);
    double norm = 0;
    for (int64_t i = 0; i < n; i++) {
        std::complex<double> acc = 0;
        for (int s = 0; s < nslices; s++)
            acc += std::complex<double>(tables[s][2 * i], tables[s][2 * i + 1]) * std::exp2(logscales[s] - L);
        if (e->storage == GS_C64) {
            const float v[2] = {float(acc.real()), float(acc.imag())};
            memcpy(raw.data() + size_t(i) * 8, v, 8);
        } else {
            const double v[2] = {acc.real(), acc.imag()};
            memcpy(raw.data() + size_t(i) * 16, v, 16);
        }
        norm += std::norm(acc);
    }
    const double lp_prefix = -K + std::log2(norm) + 2 * L;
    TT_TRY(hipSetDevice(e->device));
    memcpy(e->h_forced, bits, size_t(ns));
    std::mt19937_64 rng(seed + (unsigned long long)line);
    std::uniform_real_distribution<double> U(0, 1);
    int q = 0;
    for (int k = K; k < ns; k++)
        if (bits[k] < 0) e->h_U[q++] = U(rng);
    e->cur = 0;
    TT_TRY(hipMemcpyAsync(e->tab[0], raw.data(), raw.size(), hipMemcpyHostToDevice, e->stream));
    TT_TRY(hipMemcpyAsync(e->d_forced, e->h_forced, size_t(ns), hipMemcpyHostToDevice, e->stream));
    TT_TRY(hipMemcpyAsync(e->d_U, e->h_U, sizeof(double) * size_t(std::max(q, 1)), hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(tt_init_ctl_kernel, dim3(1), dim3(1), 0, e->stream, e->ctl, norm, L);
    TT_TRY(hipGetLastError());
    if (int rc = tt_enqueue(e, kRun, K, ns, e->h_forced, e->d_forced, e->d_U, e->d_bits, e->d_cond, e->d_out, nullptr))
        return rc;
    TT_TRY(hipMemcpyAsync(e->h_bits, e->d_bits, size_t(ns), hipMemcpyDeviceToHost, e->stream));
    TT_TRY(hipMemcpyAsync(e->h_cond, e->d_cond, size_t(ns) * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    TT_TRY(hipMemcpyAsync(e->h_out, e->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    TT_TRY(hipStreamSynchronize(e->stream));
    for (int k = 0; k < ns; k++) out_bits[k] = k < K ? bits[k] : e->h_bits[k];
    for (int k = K; k < ns; k++) cond[k] = e->h_cond[k];
    *log2p_prefix = lp_prefix;
    *log2p_full = lp_prefix + e->h_out[0];
    return 0;
}

// dp.cpp:33-62: for each job, sum over its slices of pre * Z(w_slice), Z(w) = the final entry after every step with
// the per-vertex weights w_v(0), w_v(1) (slice weights: [m][4] as w0re w0im w1re w1im, vertex-indexed).
int gs_tt_dp(gs_tt_engine* e, int64_t njobs, const int* nslices, const double* pre, const double* w, double* out) {
    if (!e || njobs < 0 || (njobs > 0 && (!nslices || !pre || !w || !out))) return ttfail(16, "bad argument");
    const int ns = e->plan.ns, m = e->plan.n;
    for (const TTStep& s : e->plan.steps)
        if (s.v < 0 || s.v >= m) return ttfail(16, "step vertex outside the job's weights");
    TT_TRY(hipSetDevice(e->device));
    size_t sl = 0;
    for (int64_t j = 0; j < njobs; j++) {
        std::complex<double> tot = 0;
        for (int q = 0; q < nslices[j]; q++, sl++) {
            const double* ws = w + sl * size_t(m) * 4;
            for (int k = 0; k < ns; k++) {
                const int v = e->plan.steps[size_t(k)].v;
                for (int c = 0; c < 4; c++) e->h_w[size_t(k) * 4 + c] = ws[size_t(v) * 4 + c];
            }
            TT_TRY(hipMemcpyAsync(e->d_w, e->h_w, sizeof(double) * 4 * size_t(std::max(ns, 1)), hipMemcpyHostToDevice,
                                  e->stream));
            if (int rc = tt_init(e, 1.0, 0.0)) return rc;
            if (int rc = tt_enqueue(e, kChain1, 0, ns, nullptr, nullptr, e->d_U, e->d_bits, e->d_cond, e->d_out, e->d_w))
                return rc;
            if (int rc = enqueue_amp(e, 0)) return rc;
            TT_TRY(hipMemcpyAsync(e->h_out, e->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
            TT_TRY(hipMemcpyAsync(e->h_amp, e->d_amp, 16, hipMemcpyDeviceToHost, e->stream));
            TT_TRY(hipStreamSynchronize(e->stream));
            const std::complex<double> z = e->h_out[3] != 0 ? std::complex<double>(0) : read_amp(e, 0);
            tot += std::complex<double>(pre[2 * sl], pre[2 * sl + 1]) * z;
        }
        out[2 * j] = tot.real();
        out[2 * j + 1] = tot.imag();
    }
    return 0;
}

const char* gs_tt_last_error(void) { return g_tterr.c_str(); }

}  // extern "C"
