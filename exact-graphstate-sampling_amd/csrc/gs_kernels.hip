// Sweep kernels of the in-place cut-rank dynamic program on CDNA4 (gfx950).
//
// One step reads, for every coset representative c < 2^nk, the ntq live slots base(c) ^ off_j (base is GF(2)-linear
// in c: XOR of the kept columns selected by the bits of c) and produces for each measured bit b and output tau
//   out[b][tau] = sum_{terms} M[b][x] (-1)^{x (parity(c & rhoc) ^ par_j)} mem[base(c) ^ off_j]
// (gsamp2.cpp:32-38).  The sums are regrouped as out[b][tau] = M[b][0] A[tau][0] + M[b][1] A[tau][1] with
// A[tau][x] the signed slot sums routed to (tau, x); all arithmetic is FP64, the arena is complex<float> (c64) or
// complex<double> (c128), and a store rounds once (gsamp2.cpp:35, :52).
//
// Three sweeps share the addressing:
//   kGram   -> per block, the 2x2 Gram matrix G[x][x'] = sum_c sum_tau A[tau][x] conj(A[tau][x']); the branch norms
//              n_b = sum_{x,x'} M[b][x] conj(M[b][x']) G[x][x'] are formed by the decide kernel (sampling steps);
//   kDirect -> per block, n_0 and n_1 summed directly as |out[b][tau]|^2 (the reference's norm pass, used where a
//              branch with a tiny probability must keep full relative precision: forced strings);
//   kWrite  -> stores out[bv][tau] * sc in place (tau 0 at base, tau 1 at base ^ g) and returns, per block, the FP64
//              norm of the written values and the norm of the rounded stored values.
//
// Layout and addressing.  A block of 256 threads handles a tile of 256 * V consecutive cosets per iteration; thread t
// owns cosets tile*256V + t*V + i, i < V.  The low v = log2 V kept columns are the unit bits 0..v-1 (the host only
// picks V > 1 when they are), so a thread's V cosets have bases ag ^ i for one ag, and the slot group
// {ag ^ off_j ^ i : i < V} is one aligned run of V entries: one 16-byte (or two) vector load per slot, coalesced
// across the wave whenever consecutive threads differ only in low physical bits.  XOR with the low bits of ag ^ off_j
// permutes the run; the permutation is undone in registers.  Each thread writes exactly the slots it read, so the
// sweep is race free in place (the planner's invariant: a coset's outputs land in its own slots, or in fresh slots
// for growth steps).
//
// Tiles are enumerated in Gray-code order inside a block, so moving to the next tile XORs one kept column into the
// base and flips at most one sign bit: address generation costs O(1) scalar work per tile.
#include "gs_device.h"

namespace gs {

// ---------------------------------------------------------------------------------------------------------------
// storage traits

template <class S>
struct Cx;

template <>
struct Cx<float2> {
    static __device__ __forceinline__ double2 widen(float2 a) { return make_double2(double(a.x), double(a.y)); }
    static __device__ __forceinline__ float2 narrow(double2 a) { return make_float2(float(a.x), float(a.y)); }
};

template <>
struct Cx<double2> {
    static __device__ __forceinline__ double2 widen(double2 a) { return a; }
    static __device__ __forceinline__ double2 narrow(double2 a) { return a; }
};

// V consecutive arena entries at an aligned index, as 16-byte vector accesses.
template <class S, int V>
struct Run;

template <>
struct Run<float2, 1> {
    static __device__ __forceinline__ void load(const float2* m, uint64_t a, float2 (&e)[1]) { e[0] = m[a]; }
    static __device__ __forceinline__ void store(float2* m, uint64_t a, const float2 (&e)[1]) { m[a] = e[0]; }
};
template <>
struct Run<float2, 2> {
    static __device__ __forceinline__ void load(const float2* m, uint64_t a, float2 (&e)[2]) {
        const float4 x = *reinterpret_cast<const float4*>(m + a);
        e[0] = make_float2(x.x, x.y);
        e[1] = make_float2(x.z, x.w);
    }
    static __device__ __forceinline__ void store(float2* m, uint64_t a, const float2 (&e)[2]) {
        *reinterpret_cast<float4*>(m + a) = make_float4(e[0].x, e[0].y, e[1].x, e[1].y);
    }
};
template <>
struct Run<float2, 4> {
    static __device__ __forceinline__ void load(const float2* m, uint64_t a, float2 (&e)[4]) {
        const float4 x = *reinterpret_cast<const float4*>(m + a);
        const float4 y = *reinterpret_cast<const float4*>(m + a + 2);
        e[0] = make_float2(x.x, x.y);
        e[1] = make_float2(x.z, x.w);
        e[2] = make_float2(y.x, y.y);
        e[3] = make_float2(y.z, y.w);
    }
    static __device__ __forceinline__ void store(float2* m, uint64_t a, const float2 (&e)[4]) {
        *reinterpret_cast<float4*>(m + a) = make_float4(e[0].x, e[0].y, e[1].x, e[1].y);
        *reinterpret_cast<float4*>(m + a + 2) = make_float4(e[2].x, e[2].y, e[3].x, e[3].y);
    }
};
template <>
struct Run<double2, 1> {
    static __device__ __forceinline__ void load(const double2* m, uint64_t a, double2 (&e)[1]) { e[0] = m[a]; }
    static __device__ __forceinline__ void store(double2* m, uint64_t a, const double2 (&e)[1]) { m[a] = e[0]; }
};
template <>
struct Run<double2, 2> {
    static __device__ __forceinline__ void load(const double2* m, uint64_t a, double2 (&e)[2]) {
        e[0] = m[a];
        e[1] = m[a + 1];
    }
    static __device__ __forceinline__ void store(double2* m, uint64_t a, const double2 (&e)[2]) {
        m[a] = e[0];
        m[a + 1] = e[1];
    }
};

// e[i] <- e[i ^ p] for p < V (per-lane conditional swaps; an involution, used for loads and stores alike).
template <class S, int V>
__device__ __forceinline__ void permute(S (&e)[V], unsigned p) {
    if constexpr (V >= 2) {
        const bool s1 = p & 1;
#pragma unroll
        for (int i = 0; i < V; i += 2) {
            const S a = e[i], b = e[i + 1];
            e[i] = s1 ? b : a;
            e[i + 1] = s1 ? a : b;
        }
    }
    if constexpr (V >= 4) {
        const bool s2 = p & 2;
#pragma unroll
        for (int i = 0; i < V; i++) {
            if (i & 2) continue;
            const S a = e[i], b = e[i + 2];
            e[i] = s2 ? b : a;
            e[i + 2] = s2 ? a : b;
        }
    }
}

__device__ __forceinline__ double2 negate_if(double2 a, int s) {
    const unsigned long long m = (unsigned long long)(s & 1) << 63;
    return make_double2(__longlong_as_double(__double_as_longlong(a.x) ^ m),
                        __longlong_as_double(__double_as_longlong(a.y) ^ m));
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// a * b + c (complex, FP64)
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double norm2(double2 a) { return fma(a.x, a.x, a.y * a.y); }

// output index of offset j's x = 0 / x = 1 term under each canonical route
template <int ROUTE>
__device__ __host__ constexpr int tau_x0(int j) {
    return ROUTE == kRouteMixA || ROUTE == kRouteMixB ? (j & 1) : 0;
}
template <int ROUTE>
__device__ __host__ constexpr int tau_x1(int j) {
    return ROUTE == kRoute1 ? 0 : ROUTE == kRouteSplit ? 1 : ROUTE == kRouteMixA ? 1 - (j & 1) : (j & 1);
}

// Sum over the wave, then over the block's 4 waves in a fixed order; thread 0 writes NA doubles.
template <int NA>
__device__ __forceinline__ void block_sum_store(double (&acc)[NA], double* out) {
#pragma unroll
    for (int q = 0; q < NA; q++)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o, 64);
    __shared__ double sh[kThreads / 64][NA];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NA; q++) sh[wave][q] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NA; q++) {
            double t = sh[0][q];
#pragma unroll
            for (int w = 1; w < kThreads / 64; w++) t += sh[w][q];
            out[q] = t;
        }
}

template <class S, int V, int NTQ, int ROUTE, int SWEEP>
__global__ __launch_bounds__(kThreads) void sweep_kernel(const KStep* __restrict__ sp, S* __restrict__ mem,
                                                         double* __restrict__ part, const Ctl* __restrict__ ctl) {
    constexpr int MO = ROUTE == kRoute1 ? 1 : 2;
    constexpr int LV = V == 1 ? 0 : V == 2 ? 1 : 2;
    constexpr int NA = SWEEP == kGram ? 4 : 2;
    if (ctl->dead) return;
    const KStep& s = *sp;
    const int nk = s.nk;
    const int tid = threadIdx.x;
    const int lt = LV + 8;                         // log2 cosets per tile

    // thread part of the base address and of the sign parity (bits LV .. LV+7 of c)
    uint64_t a_thr = 0;
    int p_thr = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int j = LV + b;
        if (j < nk && ((tid >> b) & 1)) {
            a_thr ^= s.cols[j];
            p_thr ^= int((s.rhoc >> j) & 1);
        }
    }
    const bool active = nk >= lt || (uint64_t(tid) << LV) < (uint64_t(1) << nk);
    // block part (bits lt .. lt+lg_grid-1)
    uint64_t a_blk = 0;
    int p_blk = 0;
    for (int b = 0; b < s.lg_grid; b++)
        if ((blockIdx.x >> b) & 1) {
            a_blk ^= s.cols[lt + b];
            p_blk ^= int((s.rhoc >> (lt + b)) & 1);
        }
    const int jit = lt + s.lg_grid;                // first column of the iteration index
    const int niter = 1 << s.lg_iter;

    double2 w0 = make_double2(0, 0), w1 = make_double2(0, 0);
    double2 m00 = w0, m01 = w0, m10 = w0, m11 = w0;
    if constexpr (SWEEP == kWrite) {
        w0 = make_double2(ctl->w[0], ctl->w[1]);
        w1 = make_double2(ctl->w[2], ctl->w[3]);
    }
    if constexpr (SWEEP == kDirect) {
        m00 = make_double2(s.mre[0][0], s.mim[0][0]);
        m01 = make_double2(s.mre[0][1], s.mim[0][1]);
        m10 = make_double2(s.mre[1][0], s.mim[1][0]);
        m11 = make_double2(s.mre[1][1], s.mim[1][1]);
    }
    uint64_t off[NTQ];
#pragma unroll
    for (int j = 0; j < NTQ; j++) off[j] = s.off[j];
    const uint32_t parmask = s.parmask;
    const uint64_t g = s.g;

    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; q++) acc[q] = 0;

    uint64_t a_it = 0;
    int p_it = 0;
    for (int it = 0; it < niter; it++) {
        if (it) {                                   // Gray code: one column changes
            const int b = __builtin_ctz(unsigned(it));
            a_it ^= s.cols[jit + b];
            p_it ^= int((s.rhoc >> (jit + b)) & 1);
        }
        if (!active) continue;
        const uint64_t ag = a_thr ^ a_blk ^ a_it;   // coset i of this thread has base ag ^ i
        const int pbase = p_thr ^ p_blk ^ p_it;
        S e[NTQ][V];
#pragma unroll
        for (int j = 0; j < NTQ; j++) {
            const uint64_t a = ag ^ off[j];
            Run<S, V>::load(mem, a & ~uint64_t(V - 1), e[j]);
            permute<S, V>(e[j], unsigned(a) & (V - 1));
        }
        S o0[V], o1[V];
#pragma unroll
        for (int i = 0; i < V; i++) {
            const int pc = pbase ^ __builtin_parity(unsigned(i) & unsigned(s.rhoc));
            double2 A[MO][2];
#pragma unroll
            for (int t = 0; t < MO; t++) A[t][0] = A[t][1] = make_double2(0, 0);
#pragma unroll
            for (int j = 0; j < NTQ; j++) {
                const double2 wv = Cx<S>::widen(e[j][i]);
                A[tau_x0<ROUTE>(j)][0] = cadd(A[tau_x0<ROUTE>(j)][0], wv);
                A[tau_x1<ROUTE>(j)][1] = cadd(A[tau_x1<ROUTE>(j)][1], negate_if(wv, pc ^ int(parmask >> j)));
            }
            if constexpr (SWEEP == kGram) {
#pragma unroll
                for (int t = 0; t < MO; t++) {
                    const double2 a = A[t][0], b = A[t][1];
                    acc[0] += norm2(a);
                    acc[1] += norm2(b);
                    acc[2] = fma(a.x, b.x, fma(a.y, b.y, acc[2]));      // Re a conj(b)
                    acc[3] = fma(a.y, b.x, fma(-a.x, b.y, acc[3]));     // Im a conj(b)
                }
            } else if constexpr (SWEEP == kDirect) {
#pragma unroll
                for (int t = 0; t < MO; t++) {
                    acc[0] += norm2(cfma(m00, A[t][0], cmul(m01, A[t][1])));
                    acc[1] += norm2(cfma(m10, A[t][0], cmul(m11, A[t][1])));
                }
            } else {
                const double2 v0 = cfma(w0, A[0][0], cmul(w1, A[0][1]));
                const S s0 = Cx<S>::narrow(v0);
                o0[i] = s0;
                acc[0] += norm2(v0);
                acc[1] += norm2(Cx<S>::widen(s0));
                if constexpr (MO == 2) {
                    const double2 v1 = cfma(w0, A[1][0], cmul(w1, A[1][1]));
                    const S s1 = Cx<S>::narrow(v1);
                    o1[i] = s1;
                    acc[0] += norm2(v1);
                    acc[1] += norm2(Cx<S>::widen(s1));
                }
            }
        }
        if constexpr (SWEEP == kWrite) {
            permute<S, V>(o0, unsigned(ag) & (V - 1));
            Run<S, V>::store(mem, ag & ~uint64_t(V - 1), o0);
            if constexpr (MO == 2) {
                const uint64_t a1 = ag ^ g;
                permute<S, V>(o1, unsigned(a1) & (V - 1));
                Run<S, V>::store(mem, a1 & ~uint64_t(V - 1), o1);
            }
        }
    }
    block_sum_store<NA>(acc, part + size_t(blockIdx.x) * kPartStride);
}

// ---------------------------------------------------------------------------------------------------------------
// dispatch table

using SweepFn = void (*)(const KStep*, void*, double*, const Ctl*);

template <class S, int V, int NTQ, int ROUTE, int SWEEP>
void* sweep_ptr() {
    return reinterpret_cast<void*>(&sweep_kernel<S, V, NTQ, ROUTE, SWEEP>);
}

template <class S, int V, int NTQ, int SWEEP>
void* by_route(int route) {
    switch (route) {
        case kRoute1: return sweep_ptr<S, V, NTQ, kRoute1, SWEEP>();
        case kRouteSplit: return sweep_ptr<S, V, NTQ, kRouteSplit, SWEEP>();
        case kRouteMixA: if constexpr (NTQ >= 2) return sweep_ptr<S, V, NTQ, kRouteMixA, SWEEP>(); else return nullptr;
        case kRouteMixB: if constexpr (NTQ >= 2) return sweep_ptr<S, V, NTQ, kRouteMixB, SWEEP>(); else return nullptr;
    }
    return nullptr;
}

template <class S, int V, int SWEEP>
void* by_ntq(int ntq, int route) {
    switch (ntq) {
        case 1: return by_route<S, V, 1, SWEEP>(route);
        case 2: return by_route<S, V, 2, SWEEP>(route);
        case 4: return by_route<S, V, 4, SWEEP>(route);
    }
    return nullptr;
}

template <class S, int SWEEP>
void* by_v(int v, int ntq, int route) {
    if constexpr (sizeof(S) == 8) {
        switch (v) {
            case 0: return by_ntq<S, 1, SWEEP>(ntq, route);
            case 2: return by_ntq<S, 4, SWEEP>(ntq, route);
        }
    } else {
        switch (v) {
            case 0: return by_ntq<S, 1, SWEEP>(ntq, route);
            case 1: return by_ntq<S, 2, SWEEP>(ntq, route);
        }
    }
    return nullptr;
}

// Kernel for (storage: 0 = c64, 1 = c128, v, ntq, route, sweep); nullptr when that combination is not built.
void* sweep_kernel_for(int storage, int v, int ntq, int route, int sweep) {
    if (storage == 0) {
        switch (sweep) {
            case kGram: return by_v<float2, kGram>(v, ntq, route);
            case kDirect: return by_v<float2, kDirect>(v, ntq, route);
            case kWrite: return by_v<float2, kWrite>(v, ntq, route);
        }
    } else {
        switch (sweep) {
            case kGram: return by_v<double2, kGram>(v, ntq, route);
            case kDirect: return by_v<double2, kDirect>(v, ntq, route);
            case kWrite: return by_v<double2, kWrite>(v, ntq, route);
        }
    }
    return nullptr;
}

// v values built per storage (largest first): the host picks the largest the step's layout allows.
int preferred_v(int storage) { return storage == 0 ? 2 : 1; }

}  // namespace gs
