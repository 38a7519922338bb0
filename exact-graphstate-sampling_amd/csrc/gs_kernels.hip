// Sweep kernels and on-device chain logic of the in-place cut-rank dynamic program on CDNA4 (gfx950).
//
// One step reads, for every coset representative c < 2^nk, the ntq live slots base(c) ^ off_j (base is GF(2)-linear
// in c: XOR of the coordinate columns selected by the bits of c) and produces for each measured bit b and output tau
//   out[b][tau] = sum_{terms} M[b][x] (-1)^{x (parity(c & rhoc) ^ par_j)} mem[base(c) ^ off_j]
// (gsamp2.cpp:32-38).  The sums are regrouped as out[b][tau] = M[b][0] A[tau][0] + M[b][1] A[tau][1] with A[tau][x]
// the signed slot sums routed to (tau, x); all arithmetic is FP64, the arena is complex<float> (c64) or
// complex<double> (c128), and a store rounds once (gsamp2.cpp:35, :52).
//
// Sweeps (template SWEEP):
//   kGram        per block, the 2x2 Gram matrix G[x][x'] = sum_c sum_tau A[tau][x] conj(A[tau][x']); the branch norms
//                n_b = sum_{x,x'} M[b][x] conj(M[b][x']) G[x][x'] follow without recomputing both branches;
//   kDirect      per block, n_0 and n_1 summed directly as |out[b][tau]|^2 (full relative precision for forced
//                strings, the reference's norm pass);
//   kWrite       stores out[bv][tau] * sc in place (tau 0 at base, tau 1 at base ^ g) and returns the FP64 norm of the
//                written values and of the rounded stored values;
//   kWriteGram / kWriteDirect  kWrite, plus the Gram matrix / direct norms of the NEXT step computed from the values
//                just stored: the thread's cosets are chosen to span the next step's pairing vector, so both slots of
//                each next-step coset are in the thread's registers.  This removes the next step's norm pass (one
//                read of the table per sampled step).
//
// Layout.  A block of 256 threads handles a tile of 256 * V cosets per iteration; thread t owns the V = VR * VG
// cosets whose low coordinates run over 0..V-1.  Coordinates 0..log2(VR)-1 have unit physical columns, so the VR
// cosets' slots are one aligned run (a 16-byte vector access); with VG = 2, coordinate log2(VR) has an arbitrary
// column Gamma (a second run).  XOR with low offset bits permutes a run; the permutation is undone in registers.
// Each thread writes exactly the slots it read (or fresh slots of a growth step): race free in place.
//
// Tiles are enumerated in Gray-code order inside a block, so the next tile XORs one column into the base address and
// flips the parity bits by one mask: O(1) scalar work per tile.
//
// Chain logic.  The block partial sums of a sweep are reduced in a fixed order and the chain runs on the device: the
// bookkeeping of a write (chain1 / run_steps conditionals, the stored table's norm) and the decision of the next step
// (forced, or U_k < p1 with U_k drawn on the host: gsamp2.cpp:47, :86, :108 draw one uniform per sampled step from a
// stream that does not depend on p1).  Sweeps of a few blocks reduce in their last block (agent-scope counter);
// larger ones are followed by a one-block reduction kernel.  No host round trip.
//
// Arithmetic per coset.  The sign of an x = 1 term is (-1)^{pc} sigma_j with pc = parity(c & rhoc) (per coset) and
// sigma_j = (-1)^{par_j} (per step), so A[tau][1] = (-1)^{pc} sigma_{j0} B_tau with B_tau = sum_j sigma_{j0} sigma_j
// slot_j: sigma_{j0} goes into the step's coefficient and (-1)^{pc} flips the fresh product M_b1 B_tau (one XOR per
// component).  The fused sweeps reduce the next step's Gram matrix from per-route invariants of each next-step coset
// (a, b): |a|^2 + |b|^2, Re/Im a conj(b), |a|^2 - |b|^2, with the per-coset signs applied to scalars.
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

// Arena accesses carry the non-temporal hint: every sweep streams the table once, so its lines should not displace
// anything in L2, and written lines leave few dirty lines behind.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 ld_nt(const float2* p) {
    const f2v t = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
    return make_float2(t.x, t.y);
}
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
    const f2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<f2v*>(p));
}
__device__ __forceinline__ float4 ld_nt4(const float2* p) {
    const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ void st_nt4(float2* p, float2 a, float2 b) {
    const f4v t = {a.x, a.y, b.x, b.y};
    __builtin_nontemporal_store(t, reinterpret_cast<f4v*>(p));
}
__device__ __forceinline__ double2 ld_nt(const double2* p) {
    const d2v t = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    return make_double2(t.x, t.y);
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
    const d2v t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<d2v*>(p));
}

// VR consecutive arena entries at an aligned index, as 16-byte vector accesses.
template <class S, int VR>
struct Run;

template <>
struct Run<float2, 1> {
    static __device__ __forceinline__ void load(const float2* m, uint64_t a, float2 (&e)[1]) { e[0] = ld_nt(m + a); }
    static __device__ __forceinline__ void store(float2* m, uint64_t a, const float2 (&e)[1]) { st_nt(m + a, e[0]); }
};
template <>
struct Run<float2, 2> {
    static __device__ __forceinline__ void load(const float2* m, uint64_t a, float2 (&e)[2]) {
        const float4 x = ld_nt4(m + a);
        e[0] = make_float2(x.x, x.y);
        e[1] = make_float2(x.z, x.w);
    }
    static __device__ __forceinline__ void store(float2* m, uint64_t a, const float2 (&e)[2]) {
        st_nt4(m + a, e[0], e[1]);
    }
};
template <>
struct Run<float2, 4> {
    static __device__ __forceinline__ void load(const float2* m, uint64_t a, float2 (&e)[4]) {
        const float4 x = ld_nt4(m + a);
        const float4 y = ld_nt4(m + a + 2);
        e[0] = make_float2(x.x, x.y);
        e[1] = make_float2(x.z, x.w);
        e[2] = make_float2(y.x, y.y);
        e[3] = make_float2(y.z, y.w);
    }
    static __device__ __forceinline__ void store(float2* m, uint64_t a, const float2 (&e)[4]) {
        st_nt4(m + a, e[0], e[1]);
        st_nt4(m + a + 2, e[2], e[3]);
    }
};
template <>
struct Run<double2, 1> {
    static __device__ __forceinline__ void load(const double2* m, uint64_t a, double2 (&e)[1]) { e[0] = ld_nt(m + a); }
    static __device__ __forceinline__ void store(double2* m, uint64_t a, const double2 (&e)[1]) { st_nt(m + a, e[0]); }
};
template <>
struct Run<double2, 2> {
    static __device__ __forceinline__ void load(const double2* m, uint64_t a, double2 (&e)[2]) {
        e[0] = ld_nt(m + a);
        e[1] = ld_nt(m + a + 1);
    }
    static __device__ __forceinline__ void store(double2* m, uint64_t a, const double2 (&e)[2]) {
        st_nt(m + a, e[0]);
        st_nt(m + a + 1, e[1]);
    }
};

// e[i] <- e[i ^ p] for p < VR (per-lane conditional swaps; an involution, used for loads and stores alike).
template <class S, int VR>
__device__ __forceinline__ void permute(S (&e)[VR], unsigned p) {
    if constexpr (VR >= 2) {
        const bool s1 = p & 1;
#pragma unroll
        for (int i = 0; i < VR; i += 2) {
            const S a = e[i], b = e[i + 1];
            e[i] = s1 ? b : a;
            e[i + 1] = s1 ? a : b;
        }
    }
    if constexpr (VR >= 4) {
        const bool s2 = p & 2;
#pragma unroll
        for (int i = 0; i < VR; i++) {
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
// acc += |a|^2 as two fused multiply-adds
__device__ __forceinline__ void add_norm(double& acc, double2 a) { acc = fma(a.x, a.x, fma(a.y, a.y, acc)); }

// output index of offset j's x = 0 / x = 1 term under each canonical route
template <int ROUTE>
__device__ __host__ constexpr int tau_x0(int j) {
    return ROUTE == kRouteMixA || ROUTE == kRouteMixB ? (j & 1) : 0;
}
template <int ROUTE>
__device__ __host__ constexpr int tau_x1(int j) {
    return ROUTE == kRoute1 ? 0 : ROUTE == kRouteSplit ? 1 : ROUTE == kRouteMixA ? 1 - (j & 1) : (j & 1);
}

// Accumulate the routed signed slot sums of one coset: A[tau][x] += slot values (sign (-1)^{pc ^ par_j} on x = 1).
// The sums start from -0.0, the additive identity (x + -0.0 == x for every x), so a sum of one term compiles to no
// addition at all.
template <int ROUTE, int NTQ, int MO>
__device__ __forceinline__ void route_terms(double2 (&A)[MO][2], const double2 (&w)[NTQ], int pc, uint32_t parmask) {
#pragma unroll
    for (int t = 0; t < MO; t++) A[t][0] = A[t][1] = make_double2(-0.0, -0.0);
#pragma unroll
    for (int j = 0; j < NTQ; j++) {
        A[tau_x0<ROUTE>(j)][0] = cadd(A[tau_x0<ROUTE>(j)][0], w[j]);
        A[tau_x1<ROUTE>(j)][1] = cadd(A[tau_x1<ROUTE>(j)][1], negate_if(w[j], pc ^ int(parmask >> j)));
    }
}

template <int MO>
__device__ __forceinline__ void gram_acc(const double2 (&A)[MO][2], double* acc) {
#pragma unroll
    for (int t = 0; t < MO; t++) {
        const double2 a = A[t][0], b = A[t][1];
        add_norm(acc[0], a);
        add_norm(acc[1], b);
        acc[2] = fma(a.x, b.x, fma(a.y, b.y, acc[2]));      // Re a conj(b)
        acc[3] = fma(a.y, b.x, fma(-a.x, b.y, acc[3]));     // Im a conj(b)
    }
}

template <int MO>
__device__ __forceinline__ void direct_acc(const double2 (&A)[MO][2], const double2 (&M)[2][2], double* acc) {
#pragma unroll
    for (int t = 0; t < MO; t++) {
        add_norm(acc[0], cfma(M[0][0], A[t][0], cmul(M[0][1], A[t][1])));
        add_norm(acc[1], cfma(M[1][0], A[t][0], cmul(M[1][1], A[t][1])));
    }
}

// (-1)^s v for s in {0, 1} (sign bit flip)
__device__ __forceinline__ double flip(double v, uint32_t s) {
    return __hiloint2double(__double2hiint(v) ^ int(s << 31), __double2loint(v));
}
__device__ __forceinline__ double2 flip2(double2 v, uint32_t s) { return make_double2(flip(v.x, s), flip(v.y, s)); }

// First offset routed to (tau, x = 1), or -1.
template <int ROUTE, int NTQ>
__device__ __host__ constexpr int first_x1(int tau) {
    for (int j = 0; j < NTQ; j++)
        if (tau_x1<ROUTE>(j) == tau) return j;
    return -1;
}

// Per-launch write coefficients: w1t[tau] = sigma_{j0(tau)} w1 and rho[j] = sigma_{j0} sigma_j (j0 = first_x1 of
// offset j's x = 1 output).
template <int ROUTE, int NTQ, int MO>
__device__ __forceinline__ void write_coefs(double2 w1, uint32_t parmask, double2 (&w1t)[MO], double (&rho)[NTQ]) {
#pragma unroll
    for (int t = 0; t < MO; t++) {
        const int j0 = first_x1<ROUTE, NTQ>(t);
        w1t[t] = j0 >= 0 ? flip2(w1, (parmask >> j0) & 1) : w1;
    }
#pragma unroll
    for (int j = 0; j < NTQ; j++) {
        const int j0 = first_x1<ROUTE, NTQ>(tau_x1<ROUTE>(j));
        rho[j] = flip(1.0, ((parmask >> j) ^ (parmask >> j0)) & 1);
    }
}

// Output tau of one coset: w0 A[tau][0] + (-1)^pc w1t[tau] B_tau (the slot sums in offset order, as route_terms).
template <int ROUTE, int NTQ, int MO>
__device__ __forceinline__ double2 write_out(int t, const double2 (&x)[NTQ], uint32_t pc, double2 w0,
                                             const double2 (&w1t)[MO], const double (&rho)[NTQ]) {
    double2 a = make_double2(0, 0), b = a;
    bool h0 = false, h1 = false;
#pragma unroll
    for (int j = 0; j < NTQ; j++) {
        if (tau_x0<ROUTE>(j) == t) {
            a = h0 ? cadd(a, x[j]) : x[j];
            h0 = true;
        }
        if (tau_x1<ROUTE>(j) == t) {
            b = h1 ? make_double2(fma(rho[j], x[j].x, b.x), fma(rho[j], x[j].y, b.y)) : x[j];
            h1 = true;
        }
    }
    if (!h1) return cmul(w0, a);
    const double2 p = flip2(cmul(w1t[t], b), pc);
    return h0 ? cfma(w0, a, p) : p;
}

// Next-step Gram invariants of one coset (va, vb) in coset order; J swaps the roles of the two slots, sg is the
// coset's sign bit.  Accumulators: N = sum |va|^2 + |vb|^2, R = sum Re va conj(vb), SR = sum (-1)^sg Re, SI = sum
// (-1)^{sg ^ J} Im, SN = sum (-1)^sg N, SD = sum (-1)^{sg ^ J} (|va|^2 - |vb|^2).
enum { kgN, kgR, kgSR, kgSI, kgSN, kgSD, kgCount };
template <int ROUTE2>
__device__ __forceinline__ void gram_invariants(double2 va, double2 vb, uint32_t J, uint32_t sg, double (&g)[kgCount]) {
    if constexpr (ROUTE2 == kRouteSplit || ROUTE2 == kRouteMixA) {
        add_norm(g[kgN], va);
        add_norm(g[kgN], vb);
    } else {
        const double na = norm2(va), nb = norm2(vb);
        g[kgN] += na + nb;
        g[kgSN] += flip(na + nb, sg);
        g[kgSD] += flip(na - nb, sg ^ J);
    }
    if constexpr (ROUTE2 == kRouteSplit) {
        g[kgR] = fma(va.x, vb.x, fma(va.y, vb.y, g[kgR]));
    } else if constexpr (ROUTE2 == kRouteMixA || ROUTE2 == kRoute1) {
        const double re = fma(va.x, vb.x, va.y * vb.y), im = fma(va.y, vb.x, -va.x * vb.y);
        if constexpr (ROUTE2 == kRoute1) g[kgR] += re;
        g[kgSR] += flip(re, sg);
        g[kgSI] += flip(im, sg ^ J);
    }
}

// G00, G11, Re G01, Im G01 of the next step from the invariants (sigma1 = (-1)^{par} of its second slot).
template <int ROUTE2>
__device__ __forceinline__ void gram_assemble(const double (&g)[kgCount], int par, double* G) {
    const double s1 = par ? -1.0 : 1.0;
    if constexpr (ROUTE2 == kRouteSplit) {
        G[0] = g[kgN] + 2 * g[kgR];
        G[1] = g[kgN] + 2 * s1 * g[kgR];
        G[2] = 0;
        G[3] = 0;
    } else if constexpr (ROUTE2 == kRouteMixA) {
        G[0] = G[1] = g[kgN];
        G[2] = (1 + s1) * g[kgSR];
        G[3] = (s1 - 1) * g[kgSI];
    } else if constexpr (ROUTE2 == kRouteMixB) {
        G[0] = G[1] = g[kgN];
        G[2] = par ? g[kgSD] : g[kgSN];
        G[3] = 0;
    } else {
        G[0] = g[kgN] + 2 * g[kgR];
        G[1] = g[kgN] + 2 * s1 * g[kgR];
        G[2] = (par ? g[kgSD] : g[kgSN]) + (1 + s1) * g[kgSR];
        G[3] = (s1 - 1) * g[kgSI];
    }
}

// ---------------------------------------------------------------------------------------------------------------
// chain logic (one thread)

__device__ __forceinline__ double2 step_m(const KStep& s, int b, int x) {
    return make_double2(s.mre[b][x], s.mim[b][x]);
}

__device__ void publish(Ctl& c, const Epi& e) {
    if (e.mode == kChain1 && !c.dead) c.logp = log2(c.prev) - e.ns;
    if (e.quiet) return;
    e.out[0] = c.logp;
    e.out[1] = c.logscale;
    e.out[2] = c.prev;
    e.out[3] = c.dead;
}

// Bookkeeping of the write of step k (norm nb of the written FP64 values, nst of the rounded stored values).
__device__ void finish_write(Ctl& c, const Epi& e, int k, double nb, double nst) {
    c.nstored = nst;
    if (e.mode == kChain1) {                                   // gsamp2.cpp:67
        if (!e.quiet) e.cond[k] = nb / (2 * c.prev);
        c.prev = nb;
        if (nb <= 0) {
            if (!e.quiet)
                for (int q = k; q < e.ns; q++) e.cond[q] = 0;
            c.logp = -INFINITY;
            c.dead = 1;
        }
    } else if (e.mode == kRun) {                               // gsamp2.cpp:93-95
        c.logscale += 0.5 * log2(c.prev);
        const double cd = nb / 2;
        if (!e.quiet) e.cond[k] = cd;
        c.prev = nb;
        if (nb <= 0) {
            if (!e.quiet)
                for (int q = k; q < e.k1; q++) {
                    e.cond[q] = 0;
                    e.bits[q] = e.forced[q] < 0 ? 0 : e.forced[q];
                }
            c.logp = -INFINITY;
            c.dead = 1;
        } else {
            c.logp += log2(cd);
        }
    }
    if (c.dead) publish(c, e);
}

// Decision of step k from its norms: norm == kGram (red = G00, G11, Re G01, Im G01), kDirect (red = n0, n1), or -1
// (no norms: forced step, chain1, or a growth step whose norms follow from the stored table's norm).
__device__ void decide(Ctl& c, const Epi& e, int k, int norm, const double* red) {
    const KStep& s = e.plain[k];
    int bv = e.forced ? e.forced[k] : -1;
    double sc = 1.0;
    const bool need_p = (e.mode == kChain) || (e.mode == kRun && bv < 0);
    double n0 = 0, n1 = 0;
    if (need_p) {
        for (int b = 0; b < 2; b++) {
            const double2 m0 = step_m(s, b, 0), m1 = step_m(s, b, 1);
            double nbv;
            if (norm == kGram) {
                const double pr = m0.x * m1.x + m0.y * m1.y, pi = m0.y * m1.x - m0.x * m1.y;   // m0 conj(m1)
                nbv = norm2(m0) * red[0] + norm2(m1) * red[1] + 2 * (pr * red[2] - pi * red[3]);
            } else if (norm == kDirect) {
                nbv = red[b];
            } else {
                nbv = (norm2(m0) + norm2(m1)) * c.nstored;
            }
            (b ? n1 : n0) = nbv;
        }
    }
    if (e.mode == kChain) {                                    // gsamp2.cpp:46-49
        const double p1 = n1 / (n0 + n1);
        if (bv < 0) bv = e.U[k] < p1;
        const double pb = bv ? p1 : 1 - p1;
        if (pb <= 0) {
            c.logp = -INFINITY;
            c.dead = 1;
            publish(c, e);
            return;
        }
        c.logp += log2(pb);
        sc = 1.0 / sqrt(bv ? n1 : n0);
    } else if (e.mode == kChain1) {
        sc = 1.0;
    } else {                                                   // gsamp2.cpp:80-87
        sc = 1.0 / sqrt(c.prev);
        if (bv < 0) bv = e.U[c.qidx++] < n1 / (n0 + n1);
    }
    if (!e.quiet) e.bits[k] = int8_t(bv);
    c.bv = bv;
    c.sc = sc;
    const double2 m0 = step_m(s, bv, 0), m1 = step_m(s, bv, 1);
    c.w[0] = m0.x * sc;
    c.w[1] = m0.y * sc;
    c.w[2] = m1.x * sc;
    c.w[3] = m1.y * sc;
}

// Block-wide sum of v[NQ] over the NT threads in a fixed order (a butterfly inside each 64-lane wave, then the wave
// sums in wave order by one thread: two barriers instead of a log2(NT)-level tree); the result is in sh[0][*].
template <int NQ, int NT = kThreads>
__device__ __forceinline__ void block_tree_sum(double (&sh)[NT / 64][NQ], const double (&v)[NQ], int t) {
    double r[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        r[q] = v[q];
        for (int o = 32; o > 0; o >>= 1) r[q] += __shfl_xor(r[q], o, 64);
    }
    if ((t & 63) == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) sh[t >> 6][q] = r[q];
    __syncthreads();
    if (t == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double x = sh[0][q];
            for (int w = 1; w < NT / 64; w++) x += sh[w][q];
            sh[0][q] = x;
        }
    __syncthreads();
}

// Chain logic after a sweep, from its reduced sums (one thread).
template <int NQ>
__device__ void chain_step(const double (&red)[NQ], Ctl* ctl, const Epi& e, int sweep) {
    Ctl& c = *ctl;
    if constexpr (NQ == 1) {            // one sum: the product pass of a run (role kRoleStored), the stored norm
        c.nstored = red[0];
        return;
    } else if constexpr (NQ == 5) {     // a product pass that also summed the next step's Gram (kRoleStoredGram)
        c.nstored = red[0];
        if (!c.dead) decide(c, e, e.k + 1, kGram, red + 1);
        return;
    } else if (e.role == kRoleNorm) {
        decide(c, e, e.k, sweep, red);
    } else {
        finish_write(c, e, e.k, red[0], red[1]);
        if (!c.dead) {
            if (e.k + 1 == e.k1) publish(c, e);
            else if (e.next == kNextDecide) decide(c, e, e.k + 1, -1, nullptr);
            else if (e.next == kNextFusedGram) decide(c, e, e.k + 1, kGram, red + 2);
            else if (e.next == kNextFusedDirect) decide(c, e, e.k + 1, kDirect, red + 2);
        }
    }
}

// Fixed-order sum of the first NQ values of nblk block partials by one block; the result is in sh[0][*].
template <int NQ>
__device__ __forceinline__ void reduce_partials(double (&sh)[kThreads / 64][NQ], const double* part, unsigned nblk,
                                                int t, bool atomic_loads) {
    double r[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) r[q] = 0;
    for (unsigned i = t; i < nblk; i += kThreads)
#pragma unroll
        for (int q = 0; q < NQ; q++)
            r[q] += atomic_loads ? __hip_atomic_load(part + size_t(i) * kPartStride + q, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                                 : part[size_t(i) * kPartStride + q];
    block_tree_sum<NQ>(sh, r, t);
}

// End of a sweep block: the block's sums go to its partial slot.  Grids of at most kFenceBlocks blocks finish in
// the kernel: the last block to arrive (agent-scope counter) reduces and runs the chain.  Larger grids end here and
// sweep_reduce_kernel follows on the stream: an agent-scope release per block writes back the XCD's L2 each time
// (tens of ns per block), which a launch boundary does once.
template <int NQ>
__device__ void sweep_epilogue(double (&acc)[NQ], double* part, Ctl* ctl, const Epi& e, int sweep) {
    // block sum (wave butterfly, then the 4 waves in order)
#pragma unroll
    for (int q = 0; q < NQ; q++)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o, 64);
    __shared__ double sh[kThreads / 64][NQ];
    __shared__ int s_last;
    const int t = threadIdx.x;
    const unsigned nblk = gridDim.x;
    const bool inkernel = nblk <= unsigned(kFenceBlocks);
    if ((t & 63) == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) sh[t >> 6][q] = acc[q];
    __syncthreads();
    if (t == 0) {
        double* mine = part + size_t(blockIdx.x) * kPartStride;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double v = sh[0][q];
#pragma unroll
            for (int w = 1; w < kThreads / 64; w++) v += sh[w][q];
            if (inkernel) __hip_atomic_store(mine + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else mine[q] = v;
        }
        if (inkernel) {
            const unsigned done = __hip_atomic_fetch_add(e.counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            s_last = done == nblk - 1;
        }
    }
    if (!inkernel) return;
    __syncthreads();
    if (!s_last) return;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    reduce_partials<NQ>(sh, part, nblk, t, true);
    if (t != 0) return;
    double red[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) red[q] = sh[0][q];
    chain_step<NQ>(red, ctl, e, sweep);
    __hip_atomic_store(e.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduction and chain logic of a sweep of nblk > kFenceBlocks blocks (one block, launched after the sweep).
template <int NQ>
__global__ __launch_bounds__(kThreads) void sweep_reduce_kernel(const double* __restrict__ part, unsigned nblk,
                                                                Ctl* __restrict__ ctl, Epi e, int sweep) {
    if (ctl->dead) return;
    __shared__ double sh[kThreads / 64][NQ];
    const int t = threadIdx.x;
    reduce_partials<NQ>(sh, part, nblk, t, false);
    if (t != 0) return;
    double red[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) red[q] = sh[0][q];
    chain_step<NQ>(red, ctl, e, sweep);
}

// The chain step of block partials a previous sweep left unreduced, run at the start of the next sweep by every block
// (fixed-order sums: every block computes the same state); thread 0 of block 0 alone writes the string's outputs and
// stores the state at epi.ctl_out.  cl receives the state in every thread.
template <int NQ>
__device__ void pending_reduce_step(Ctl& cl, const Ctl* ctl, const Epi& epi, Ctl* shc) {
    __shared__ double sh[kThreads / 64][NQ];
    const int t = threadIdx.x;
    reduce_partials<NQ>(sh, epi.pend_part, epi.pend_nblk, t, false);
    if (t == 0) {
        Ctl c = *ctl;
        if (!c.dead) {
            double red[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) red[q] = sh[0][q];
            Epi ep = epi;
            ep.k = epi.pend_k;
            ep.role = epi.pend_role;
            ep.next = epi.pend_next;
            ep.quiet = blockIdx.x != 0;
            chain_step<NQ>(red, &c, ep, epi.pend_sweep);
        }
        *shc = c;
        if (blockIdx.x == 0) *epi.ctl_out = c;
    }
    __syncthreads();
    cl = *shc;
}

__device__ void pending_step(Ctl& cl, const Ctl* ctl, const Epi& epi) {
    __shared__ Ctl shc;
    if (epi.pend_nq == 2) pending_reduce_step<2>(cl, ctl, epi, &shc);
    else if (epi.pend_nq == 4) pending_reduce_step<4>(cl, ctl, epi, &shc);
    else pending_reduce_step<6>(cl, ctl, epi, &shc);
}

// ---------------------------------------------------------------------------------------------------------------
// the sweep

template <class S, int VR, int VG, int NTQ, int ROUTE, int SWEEP, int ROUTE2>
__global__ __launch_bounds__(kThreads) void sweep_kernel(const KStep* __restrict__ sp, S* __restrict__ mem,
                                                         double* __restrict__ part, Ctl* __restrict__ ctl, Epi epi) {
    constexpr int MO = ROUTE == kRoute1 ? 1 : 2;
    constexpr int MO2 = ROUTE2 == kRoute1 ? 1 : 2;
    constexpr int LR = VR == 1 ? 0 : VR == 2 ? 1 : 2;
    constexpr int LV = LR + (VG == 4 ? 2 : VG == 2 ? 1 : 0);
    constexpr int V = VR * VG;
    constexpr bool WRITES = SWEEP >= kWrite;
    constexpr bool FUSED = SWEEP == kWriteGram || SWEEP == kWriteDirect;
    constexpr bool GRAM1 = SWEEP == kWriteGram1;
    constexpr int NA = SWEEP == kGram ? 4 : SWEEP == kDirect ? 2 : SWEEP == kWrite ? 2
                     : SWEEP == kWriteGram || SWEEP == kWriteGram1 ? 6 : 4;
    const KStep& s = *sp;
    const int nk = s.nk;
    const int tid = threadIdx.x;
    constexpr int lt = LV + 8;                     // log2 cosets per tile

    // parity bits of a coordinate: bit 0 output sign, bit 1 next-step slot index, bit 2 next-step sign, bit 3 the
    // next step's second slot-index bit (four-slot next step)
    auto pb = [&](int j) -> int {
        int p = int((s.pmask[0] >> j) & 1) | (int((s.pmask[1] >> j) & 1) << 1) | (int((s.pmask[2] >> j) & 1) << 2);
        if constexpr (VG == 4) p |= int((s.pmask[3] >> j) & 1) << 3;
        return p;
    };
    uint64_t a_thr = 0;
    int p_thr = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int j = LV + b;
        if (j < nk && ((tid >> b) & 1)) {
            a_thr ^= s.cols[j];
            p_thr ^= pb(j);
        }
    }
    const bool active = nk >= lt || (uint64_t(tid) << LV) < (uint64_t(1) << nk);
    uint64_t a_blk = 0;
    int p_blk = 0;
    // s.xcd: blocks are dealt round-robin to the 8 XCDs, so block i's XCD is i mod 8; rotating the 3 low bits of the
    // block index to the top gives each XCD one contiguous eighth of the block coordinates (of each tile window)
    const unsigned bid = s.xcd && s.lg_grid >= 3
                             ? (blockIdx.x >> 3) | ((blockIdx.x & 7u) << (s.lg_grid - 3))
                             : blockIdx.x;
    for (int b = 0; b < s.lg_grid; b++)
        if ((bid >> b) & 1) {
            a_blk ^= s.cols[lt + b];
            p_blk ^= pb(lt + b);
        }
    const int jit = lt + s.lg_grid;
    const int niter = 1 << s.lg_iter;
    int plow[V];                                   // parity bits of the thread's low coordinates
#pragma unroll
    for (int i = 0; i < V; i++) {
        int p = 0;
#pragma unroll
        for (int b = 0; b < LV; b++)
            if ((i >> b) & 1) p ^= pb(b);
        plow[i] = p;
    }
    const uint64_t gam = VG >= 2 ? s.cols[LR] : 0, gam2 = VG == 4 ? s.cols[LR + 1] : 0;
    auto goff = [&](int ig) -> uint64_t { return ((ig & 1) ? gam : 0) ^ ((ig & 2) ? gam2 : 0); };

    double2 w0 = make_double2(0, 0), w1 = w0;
    double2 M[2][2] = {{w0, w0}, {w0, w0}};
    double2 w1t[MO];
    double rho[NTQ];
    if constexpr (SWEEP == kDirect || SWEEP == kWriteDirect) {
        const KStep& ms = SWEEP == kDirect ? epi.plain[epi.k] : epi.plain[epi.k + 1];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int x = 0; x < 2; x++) M[b][x] = make_double2(ms.mre[b][x], ms.mim[b][x]);
    }
    uint64_t off[NTQ];
#pragma unroll
    for (int j = 0; j < NTQ; j++) off[j] = s.off[j];
    const uint32_t parmask = s.parmask;
    const uint64_t g = s.g;
    const int eps = s.eps, phig = s.phig, nug = s.nug;
    const int sgg = s.sgg, sgk = s.sgk;
    const uint32_t par2mask = VG == 4 ? uint32_t(s.parmask2) : uint32_t(s.par2) << 1;
    const int eps2 = s.eps2;

    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; q++) acc[q] = 0;
    double gi[kgCount];                            // next-step Gram invariants (kWriteGram)
#pragma unroll
    for (int q = 0; q < kgCount; q++) gi[q] = 0;

    // Each tile is loaded at the top of its iteration, with no software pipeline: on this access pattern (four
    // regions per wave, each read and rewritten in place) issuing the next tile's loads early measured slower
    // (192.6 against 182.6 us per 2^25-coset fused sweep), as did a two-tile-deep pipeline and larger grids -- more
    // requests in flight cost more in the memory system than they hide.
    const uint64_t a_tb = a_thr ^ a_blk;
    S e[NTQ][VG][VR];
    auto load_tile = [&](uint64_t ag, S (&dst)[NTQ][VG][VR]) {
#pragma unroll
        for (int j = 0; j < NTQ; j++)
#pragma unroll
            for (int ig = 0; ig < VG; ig++)
                Run<S, VR>::load(mem, (ag ^ off[j] ^ goff(ig)) & ~uint64_t(VR - 1), dst[j][ig]);
    };
    // Tile index i runs 0 .. niter-1, or niter-1 .. 0 when s.rev (gray(i) ^ gray(i-1) = 1 << ctz(i)): the start of a
    // sweep then meets the end of the previous one, whose last lines may still be in the memory-side cache.
    const bool rev = s.rev != 0;
    uint64_t a_it = 0;
    int p_it = 0;
    if (rev && s.lg_iter > 0) {                     // gray(niter - 1) is the top bit alone
        a_it = s.cols[jit + s.lg_iter - 1];
        p_it = pb(jit + s.lg_iter - 1);
    }
    // Chain state of this launch: the kernel's ctl, or -- when the previous sweep left its block partials unreduced --
    // the chain step of those partials, run by every block on its own copy (the same fixed-order sums, so the same
    // result everywhere) while the first tile's loads are in flight; block 0 stores it at ctl_out.
    bool first_loaded = false;
    Ctl* cur = ctl;
    Ctl cl;
    if (epi.pend_nblk) {
        cur = epi.ctl_out;
        if (active) {
            load_tile(a_tb ^ a_it, e);
            first_loaded = true;
        }
        pending_step(cl, ctl, epi);
    } else {
        cl = *ctl;
    }
    if (cl.dead) return;
    if constexpr (WRITES) {
        w0 = make_double2(cl.w[0], cl.w[1]);
        w1 = make_double2(cl.w[2], cl.w[3]);
        write_coefs<ROUTE, NTQ, MO>(w1, s.parmask, w1t, rho);
    }
    for (int it = 0; it < niter; it++) {
        if (it > 0) {                               // Gray code: one coordinate changes
            const int b = __builtin_ctz(unsigned(rev ? niter - it : it));
            a_it ^= s.cols[jit + b];
            p_it ^= pb(jit + b);
        }
        const uint64_t ag = a_tb ^ a_it;
        const int pbase = p_thr ^ p_blk ^ p_it;
        if (!active) continue;
        if (it > 0 || !first_loaded) load_tile(ag, e);
#pragma unroll
        for (int j = 0; j < NTQ; j++)
#pragma unroll
            for (int ig = 0; ig < VG; ig++)
                permute<S, VR>(e[j][ig], unsigned(ag ^ off[j] ^ goff(ig)) & (VR - 1));
        S o[2][VG][VR];
#pragma unroll
        for (int ig = 0; ig < VG; ig++)
#pragma unroll
            for (int ir = 0; ir < VR; ir++) {
                const int i = ir + VR * ig;
                const int pc = (pbase ^ plow[i]) & 1;
                double2 wv[NTQ];
#pragma unroll
                for (int j = 0; j < NTQ; j++) wv[j] = Cx<S>::widen(e[j][ig][ir]);
                if constexpr (SWEEP == kGram || SWEEP == kDirect) {
                    double2 A[MO][2];
                    route_terms<ROUTE, NTQ, MO>(A, wv, pc, parmask);
                    if constexpr (SWEEP == kGram) gram_acc<MO>(A, acc);
                    else direct_acc<MO>(A, M, acc);
                } else {
#pragma unroll
                    for (int t = 0; t < MO; t++) {
                        const double2 v = write_out<ROUTE, NTQ, MO>(t, wv, uint32_t(pc), w0, w1t, rho);
                        const S sv = Cx<S>::narrow(v);
                        o[t][ig][ir] = sv;
                        // the written norm feeds chain1 / run_steps bookkeeping only (sampling renormalises by the
                        // decided branch's norm; accumulating it always is cheaper than selecting); the stored
                        // table's norm is read only by a growth step's analytic norms, and a fused write is never
                        // followed by a growth step
                        add_norm(acc[0], v);
                        if constexpr (GRAM1) {
                            // the next step reads this value as its coset's one slot: |v|^2 into N = G00 = G11 and,
                            // signed by that coset's (-1)^(pc ^ par), into Re G01 (Im G01 = 0)
                            const double nv = norm2(Cx<S>::widen(sv));
                            const uint32_t sg = uint32_t(((pbase ^ plow[i]) >> 2) ^ (t ? sgg : 0) ^ sgk) & 1;
                            acc[2] += nv;
                            acc[4] += flip(nv, sg);
                        } else if constexpr (!FUSED) {
                            add_norm(acc[1], Cx<S>::widen(sv));
                        }
                    }
                }
            }
        if constexpr (WRITES) {
            // stores go through a permuted copy; o stays in coset order for the fused norms
#pragma unroll
            for (int ig = 0; ig < VG; ig++)
#pragma unroll
                for (int t = 0; t < MO; t++) {
                    const uint64_t a = ag ^ goff(ig) ^ (t ? g : 0);
                    S so[VR];
#pragma unroll
                    for (int ir = 0; ir < VR; ir++) so[ir] = o[t][ig][ir];
                    permute<S, VR>(so, unsigned(a) & (VR - 1));
                    Run<S, VR>::store(mem, a & ~uint64_t(VR - 1), so);
                }
        }
        if constexpr (FUSED && VG == 4) {
            // four-slot next step: the next coset holding output (ir, ig=0, tau d) holds output (ir, ig=m, tau d ^
            // eps(m)) of the thread's coset m as its slot j0 ^ m, j0 the slot index of the first (both index bits and
            // the coset's sign are parities of the address)
#pragma unroll
            for (int ir = 0; ir < VR; ir++)
#pragma unroll
                for (int d = 0; d < MO; d++) {
                    const int pl = pbase ^ plow[ir];
                    const int j0 = (((pl >> 1) & 1) | (((pl >> 3) & 1) << 1)) ^ (d ? phig : 0);
                    const int sg = ((pl >> 2) ^ (d & nug)) & 1;
                    S a[4];
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const int tm = d ^ ((m & 1) ? eps : 0) ^ ((m & 2) ? eps2 : 0);
                        if constexpr (MO == 2) a[m] = tm ? o[1][m][ir] : o[0][m][ir];
                        else a[m] = o[0][m][ir];
                    }
                    // slot q = member q ^ j0, by selects (j0 is per lane)
                    const S b0 = (j0 & 1) ? a[1] : a[0], b1 = (j0 & 1) ? a[0] : a[1];
                    const S b2 = (j0 & 1) ? a[3] : a[2], b3 = (j0 & 1) ? a[2] : a[3];
                    double2 w2[4];
                    w2[0] = Cx<S>::widen((j0 & 2) ? b2 : b0);
                    w2[1] = Cx<S>::widen((j0 & 2) ? b3 : b1);
                    w2[2] = Cx<S>::widen((j0 & 2) ? b0 : b2);
                    w2[3] = Cx<S>::widen((j0 & 2) ? b1 : b3);
                    double2 A2[MO2][2];
                    route_terms<ROUTE2, 4, MO2>(A2, w2, sg, par2mask);
                    if constexpr (SWEEP == kWriteGram) gram_acc<MO2>(A2, acc + 2);
                    else direct_acc<MO2>(A2, M, acc + 2);
                }
        } else if constexpr (FUSED) {
            // next step's cosets: output (ir, ig=0, tau d) pairs with output (ir, ig=1, tau d ^ eps); the element with
            // slot index 0 is the next coset's base slot, the other its second slot.
#pragma unroll
            for (int ir = 0; ir < VR; ir++)
#pragma unroll
                for (int d = 0; d < MO; d++) {
                    const int pl = pbase ^ plow[ir];
                    const int jidx = ((pl >> 1) ^ (d & phig)) & 1;
                    const int sg = ((pl >> 2) ^ (d & nug)) & 1;
                    const S va = o[d][0][ir];
                    S vb;
                    if constexpr (MO == 2) vb = (d ^ eps) ? o[1][1][ir] : o[0][1][ir];
                    else vb = o[0][1][ir];
                    if constexpr (SWEEP == kWriteGram) {
                        gram_invariants<ROUTE2>(Cx<S>::widen(va), Cx<S>::widen(vb), uint32_t(jidx), uint32_t(sg), gi);
                    } else {
                        double2 w2[2];              // select in the stored type, then widen
                        w2[0] = Cx<S>::widen(jidx ? vb : va);
                        w2[1] = Cx<S>::widen(jidx ? va : vb);
                        double2 A2[MO2][2];
                        route_terms<ROUTE2, 2, MO2>(A2, w2, sg, par2mask);
                        direct_acc<MO2>(A2, M, acc + 2);
                    }
                }
        }
    }
    if constexpr (SWEEP == kWriteGram && VG == 2) gram_assemble<ROUTE2>(gi, s.par2, acc + 2);
    if constexpr (GRAM1) acc[3] = acc[2];
    sweep_epilogue<NA>(acc, part, cur, epi, GRAM1 ? int(kWriteGram) : SWEEP);
}

// ---------------------------------------------------------------------------------------------------------------
// single-thread chain kernels

__global__ void decide_kernel(Ctl* ctl, Epi e, int k) {
    Ctl& c = *ctl;
    if (!c.dead) decide(c, e, k, -1, nullptr);
}

__global__ void publish_kernel(Ctl* ctl, Epi e) { publish(*ctl, e); }

// ---------------------------------------------------------------------------------------------------------------
// A run of small steps in one block.  The head and tail of a plan are steps of a few thousand cosets, where a sweep
// costs far less than a launch; one block runs such a run step after step -- norm pass, decision, write pass,
// bookkeeping -- with block barriers in place of launch boundaries.  The arithmetic per coset is the sweep kernels'
// (FP64 routed slot sums, one rounding per store) and the chain functions are the same.

constexpr int kSmallThreads = 512;        // one block of 8 waves: 256 registers per lane, no spills (16 waves: 4.48
                                          // against 4.41 ms per width-26 sample)

// One pass of a small step by the block's threads (cosets tid, tid + kSmallThreads, ...): sums into acc (Gram: 4,
// direct: 2, write: written norm and stored norm).  The low kSmallLog2 coordinates of a coset are the thread's, the
// rest the iteration's.
constexpr int kSmallLog2 = 9;
static_assert(kSmallThreads == 1 << kSmallLog2, "one coset per thread and iteration");
template <class S, int ROUTE, int NTQ>
__device__ __forceinline__ void small_pass(const KStep& s, S* mem, int sweep, const double2 (&M)[2][2], double2 w0,
                                           double2 w1, double (&acc)[4]) {
    constexpr int MO = ROUTE == kRoute1 ? 1 : 2;
    constexpr int kHi = 5;          // iteration coordinates whose columns are read up front (steps up to 2^14 cosets)
    // the descriptor fields the pass needs, read once with independent loads
    const int nk = s.nk;
    uint64_t lo[kSmallLog2], hi[kHi], off[NTQ];
#pragma unroll
    for (int j = 0; j < kSmallLog2; j++) lo[j] = s.cols[j];
#pragma unroll
    for (int b = 0; b < kHi; b++) hi[b] = s.cols[kSmallLog2 + b];
#pragma unroll
    for (int j = 0; j < NTQ; j++) off[j] = s.off[j];
    const uint64_t g = s.g, pm = s.pmask[0];
    const uint32_t parmask = s.parmask;
    const uint64_t n = uint64_t(1) << nk;
    const unsigned t = threadIdx.x;
    uint64_t bt = 0;
#pragma unroll
    for (int j = 0; j < kSmallLog2; j++)
        if (j < nk && ((t >> j) & 1)) bt ^= lo[j];
    for (uint64_t c = t; c < n; c += kSmallThreads) {
        uint64_t base = bt;
#pragma unroll
        for (int b = 0; b < kHi; b++)
            if (kSmallLog2 + b < nk && ((c >> (kSmallLog2 + b)) & 1)) base ^= hi[b];
        for (int j = kSmallLog2 + kHi; j < nk; j++)
            if ((c >> j) & 1) base ^= s.cols[j];
        const int pc = __popcll(c & pm) & 1;
        double2 wv[NTQ];
#pragma unroll
        for (int j = 0; j < NTQ; j++) wv[j] = Cx<S>::widen(mem[base ^ off[j]]);
        double2 A[MO][2];
        route_terms<ROUTE, NTQ, MO>(A, wv, pc, parmask);
        if (sweep == kGram) {
            gram_acc<MO>(A, acc);
        } else if (sweep == kDirect) {
            direct_acc<MO>(A, M, acc);
        } else {
#pragma unroll
            for (int t = 0; t < MO; t++) {
                const double2 v = cfma(w0, A[t][0], cmul(w1, A[t][1]));
                const S sv = Cx<S>::narrow(v);
                mem[t ? base ^ g : base] = sv;
                add_norm(acc[0], v);
                add_norm(acc[1], Cx<S>::widen(sv));
            }
        }
    }
}

// Block-wide sum of v[4] for thread 0 alone (one barrier): a butterfly inside each wave, then thread 0 adds the waves
// in order -- the order of block_tree_sum, without its second barrier.
__device__ __forceinline__ void small_block_sum(double (&sh)[kSmallThreads / 64][4], const double (&v)[4], int t,
                                                double (&out)[4]) {
    double r[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        r[q] = v[q];
        for (int o = 32; o > 0; o >>= 1) r[q] += __shfl_xor(r[q], o, 64);
    }
    if ((t & 63) == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) sh[t >> 6][q] = r[q];
    __syncthreads();
    if (t == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            double x = sh[0][q];
            for (int w = 1; w < kSmallThreads / 64; w++) x += sh[w][q];
            out[q] = x;
        }
}

template <class S>
__device__ __forceinline__ void small_pass_any(const KStep& s, S* mem, int sweep, const double2 (&M)[2][2], double2 w0, double2 w1,
                               double (&acc)[4]) {
    switch (s.ntq * 4 + s.route) {
        case 1 * 4 + kRoute1: small_pass<S, kRoute1, 1>(s, mem, sweep, M, w0, w1, acc); break;
        case 1 * 4 + kRouteSplit: small_pass<S, kRouteSplit, 1>(s, mem, sweep, M, w0, w1, acc); break;
        case 2 * 4 + kRoute1: small_pass<S, kRoute1, 2>(s, mem, sweep, M, w0, w1, acc); break;
        case 2 * 4 + kRouteSplit: small_pass<S, kRouteSplit, 2>(s, mem, sweep, M, w0, w1, acc); break;
        case 2 * 4 + kRouteMixA: small_pass<S, kRouteMixA, 2>(s, mem, sweep, M, w0, w1, acc); break;
        case 2 * 4 + kRouteMixB: small_pass<S, kRouteMixB, 2>(s, mem, sweep, M, w0, w1, acc); break;
        case 4 * 4 + kRoute1: small_pass<S, kRoute1, 4>(s, mem, sweep, M, w0, w1, acc); break;
        case 4 * 4 + kRouteSplit: small_pass<S, kRouteSplit, 4>(s, mem, sweep, M, w0, w1, acc); break;
        case 4 * 4 + kRouteMixA: small_pass<S, kRouteMixA, 4>(s, mem, sweep, M, w0, w1, acc); break;
        case 4 * 4 + kRouteMixB: small_pass<S, kRouteMixB, 4>(s, mem, sweep, M, w0, w1, acc); break;
    }
}

// Steps k0 .. k1-1 of one chain (one block of kSmallThreads).  decided0: step k0's decision was already made (by the
// previous write's fused norms).  The norm choice per step is the host's (norm_for in the engine): chain1 and forced
// run steps need none, growth steps use the stored norm, forced chain strings use direct norms, sampled steps Gram.
template <class S>
__global__ __launch_bounds__(kSmallThreads) void small_run_kernel(const KStep* __restrict__ steps, S* __restrict__ mem,
                                                                  Ctl* __restrict__ ctl, Epi e, int k0, int k1,
                                                                  int decided0) {
    __shared__ double sh[kSmallThreads / 64][4];
    __shared__ int s_dead;
    __shared__ Ctl c;               // the chain state, in LDS while the run lasts
    const int t = threadIdx.x;
    if (t == 0) {
        c = *ctl;
        s_dead = c.dead;
    }
    __syncthreads();
    const bool fstr = e.forced != nullptr && e.mode == kChain;
    for (int k = k0; k < k1 && !s_dead; k++) {
        const KStep& s = steps[k];
        double2 M[2][2];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int x = 0; x < 2; x++) M[b][x] = make_double2(s.mre[b][x], s.mim[b][x]);
        const double2 zero = make_double2(0, 0);
        if (k > k0 || !decided0) {
            const int fb = e.forced ? int(e.forced[k]) : -1;
            int nm = kGram;
            if (e.mode == kChain1 || (e.mode == kRun && fb >= 0) || s.growth) nm = -1;
            else if (fstr) nm = kDirect;
            double acc[4] = {0, 0, 0, 0}, red[4];
            if (nm >= 0) small_pass_any<S>(s, mem, nm, M, zero, zero, acc);
            small_block_sum(sh, acc, t, red);
            if (t == 0) {
                decide(c, e, k, nm, nm >= 0 ? red : nullptr);
                s_dead = c.dead;
            }
            __syncthreads();
            if (s_dead) break;
        }
        const double2 w0 = make_double2(c.w[0], c.w[1]), w1 = make_double2(c.w[2], c.w[3]);
        double acc[4] = {0, 0, 0, 0}, red[4];
        small_pass_any<S>(s, mem, kWrite, M, w0, w1, acc);
        small_block_sum(sh, acc, t, red);
        if (t == 0) {
            finish_write(c, e, k, red[0], red[1]);
            if (!c.dead && k + 1 == e.k1) publish(c, e);
            s_dead = c.dead;
        }
        __syncthreads();
    }
    if (t == 0) *ctl = c;
}

// ---------------------------------------------------------------------------------------------------------------
// Runs of diagonal steps (RStep).  Pass 1 sums |v|^2 of the run's table per sign pattern sigma, pass 2 multiplies every
// entry by the product factor of its pattern; between them one block reduces the sums and makes the run's j decisions
// with the chain functions above, from the exact norms of each step's two branches:
//   n_b(step i) = sum_sigma H[sigma] |F_{i-1}(sigma)|^2 |M_b0 + (-1)^{sigma_i} M_b1|^2,
//   F_i(sigma) = F_{i-1}(sigma) (w0_i + (-1)^{sigma_i} w1_i),  (w0, w1) = the decided row of M_i times the step's scale,
// which is what the per-step sweeps compute one step at a time (the values are rounded to the storage type once per
// run instead of once per step).  All sums are in a fixed order: results are bitwise reproducible.

// Thread and block parts of a run pass: base address and sigma of the thread's coordinate origin.
__device__ __forceinline__ void run_origin(const RStep& r, uint64_t& a, int& sg, bool& active) {
    const int tid = threadIdx.x;
    a = 0;
    sg = r.kappa;
#pragma unroll
    for (int b = 0; b < 8; b++)
        if ((tid >> b) & 1) {
            a ^= r.cols[1 + b];
            sg ^= r.sig[1 + b];
        }
    active = r.nk >= 9 || (uint64_t(tid) << 1) < (uint64_t(1) << r.nk);
    const unsigned bid = r.xcd && r.lg_grid >= 3 ? (blockIdx.x >> 3) | ((blockIdx.x & 7u) << (r.lg_grid - 3))
                                                 : blockIdx.x;
    for (int b = 0; b < r.lg_grid; b++)
        if ((bid >> b) & 1) {
            a ^= r.cols[9 + b];
            sg ^= r.sig[9 + b];
        }
}

// Pass 1: per block, the sum of |v|^2 per pattern (2^j bins; threads own bins and scan the block's values in a fixed
// order) into hpart[block][kRunBins].
template <class S>
__global__ __launch_bounds__(kThreads) void run_hist_kernel(const RStep* __restrict__ rp, const S* __restrict__ mem,
                                                           double* __restrict__ hpart, const Ctl* __restrict__ ctl) {
    if (ctl->dead) return;
    const RStep& r = *rp;
    uint64_t a_tb;
    int sg;
    bool active;
    run_origin(r, a_tb, sg, active);
    const int jit = 9 + r.lg_grid;
    const int niter = 1 << r.lg_iter;
    __shared__ double sw[4 * kThreads];
    __shared__ int ss[4 * kThreads];
    const int t = threadIdx.x;
    int nv = 2;                                     // values per thread
    if (r.pair) {
        // tile pairs (a, a ^ gc): the pairing coordinate may change sigma, so the second tile's values go to their
        // own two patterns
        const uint64_t gc = r.cols[jit];
        const int sgp = r.sig[jit];
        double acc[4] = {0, 0, 0, 0};
        uint64_t a_it = 0;
        for (int i = 0; i < (niter >> 1); i++) {
            if (i > 0) a_it ^= r.cols[jit + 1 + __builtin_ctz(unsigned(i))];
            if (!active) continue;
            S e[2], h[2];
            Run<S, 2>::load(mem, a_tb ^ a_it, e);
            Run<S, 2>::load(mem, a_tb ^ a_it ^ gc, h);
            add_norm(acc[0], Cx<S>::widen(e[0]));
            add_norm(acc[1], Cx<S>::widen(e[1]));
            add_norm(acc[2], Cx<S>::widen(h[0]));
            add_norm(acc[3], Cx<S>::widen(h[1]));
        }
        for (int q = 0; q < 4; q++) {
            sw[4 * t + q] = acc[q];
            ss[4 * t + q] = sg ^ ((q & 1) ? r.sig[0] : 0) ^ ((q & 2) ? sgp : 0);
        }
        nv = 4;
    } else {
        double acc0 = 0, acc1 = 0;
        uint64_t a_it = 0;
        for (int it = 0; it < niter; it++) {
            if (it > 0) a_it ^= r.cols[jit + __builtin_ctz(unsigned(it))];
            if (!active) continue;
            S e[2];
            Run<S, 2>::load(mem, a_tb ^ a_it, e);
            add_norm(acc0, Cx<S>::widen(e[0]));
            add_norm(acc1, Cx<S>::widen(e[1]));
        }
        sw[2 * t] = acc0;
        sw[2 * t + 1] = acc1;
        ss[2 * t] = sg;
        ss[2 * t + 1] = sg ^ r.sig[0];
    }
    __syncthreads();
    double h = 0;
    for (int q = 0; q < nv * kThreads; q++)
        if (ss[q] == t) h += sw[q];
    hpart[size_t(blockIdx.x) * kRunBins + t] = h;
}

// First stage of the reduction: chunk g sums the pattern sums of blocks [g nblk / G, (g + 1) nblk / G) in order.
__global__ __launch_bounds__(kRunBins) void run_chunk_kernel(const double* __restrict__ hpart, unsigned nblk,
                                                            double* __restrict__ chunk, const Ctl* __restrict__ ctl) {
    if (ctl->dead) return;
    const unsigned G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
    const unsigned b0 = unsigned((uint64_t(g) * nblk) / G), b1 = unsigned((uint64_t(g + 1) * nblk) / G);
    double h = 0;
    for (unsigned b = b0; b < b1; b++) h += hpart[size_t(b) * kRunBins + t];
    chunk[size_t(g) * kRunBins + t] = h;
}

// The run's decisions and bookkeeping (one block of kRunBins threads, thread t = pattern t), then the product factors
// F_j(sigma) for pass 2.  Each step: the branch norms by a fixed-order block sum, decide() (forced bit, or U < p1),
// the written norm nb, finish_write(); the stored norm of pass 2 replaces nb as the growth-step norm afterwards.
__global__ __launch_bounds__(kRunBins) void run_decide_kernel(const RStep* __restrict__ rp,
                                                             const double* __restrict__ chunk, unsigned G,
                                                             Ctl* __restrict__ ctl, Epi e, double2* __restrict__ fac) {
    Ctl& c = *ctl;
    if (c.dead) return;
    const RStep& r = *rp;
    const int t = threadIdx.x;
    __shared__ double sh[kRunBins / 64][2];
    __shared__ int s_dead;
    double H = 0;
    for (unsigned g = 0; g < G; g++) H += chunk[size_t(g) * kRunBins + t];
    double2 F = make_double2(1.0, 0.0);
    double W = H;                                              // H |F|^2
    if (t == 0) s_dead = 0;
    for (int i = 0; i < r.j; i++) {
        const int k = r.k0 + i;
        const int s = (t >> i) & 1;
        double nb2[2];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const double2 m0 = make_double2(r.mre[i][b][0], r.mim[i][b][0]);
            const double2 m1 = make_double2(r.mre[i][b][1], r.mim[i][b][1]);
            nb2[b] = W * norm2(cadd(m0, negate_if(m1, s)));
        }
        block_tree_sum<2, kRunBins>(sh, nb2, t);
        if (t == 0) {
            const double red[2] = {sh[0][0], sh[0][1]};
            decide(c, e, k, kDirect, red);
            s_dead = c.dead;
        }
        __syncthreads();
        if (s_dead) return;
        const double2 w0 = make_double2(c.w[0], c.w[1]), w1 = make_double2(c.w[2], c.w[3]);
        F = cmul(F, cadd(w0, negate_if(w1, s)));
        W = H * norm2(F);
        const double one[2] = {W, 0.0};
        block_tree_sum<2, kRunBins>(sh, one, t);
        if (t == 0) {
            finish_write(c, e, k, sh[0][0], sh[0][0]);
            if (!c.dead && k + 1 == e.k1) publish(c, e);
            s_dead = c.dead;
        }
        __syncthreads();
        if (s_dead) return;
    }
    fac[t] = F;
}

// Next-step slot index (bit 0) and sign (bit 1) parities of coordinate c of a paired run.
__device__ __forceinline__ int run_pn(const RStep& r, int c) {
    return int((r.phim >> c) & 1) | (int((r.num >> c) & 1) << 1);
}

// Pass 2: every entry times its pattern's factor (FP64, one rounding), in place; the stored table's norm goes to the
// chain state (role kRoleStored).  GRAM (paired runs): the pass also sums the next step's Gram matrix from the values
// it stores -- run element e at address a pairs with element e ^ peps at a ^ gc, the two slots of one next-step coset
// -- and the chain decides that step (role kRoleStoredGram).
template <class S, bool GRAM, int ROUTE2>
__global__ __launch_bounds__(kThreads) void run_apply_kernel(const RStep* __restrict__ rp, S* __restrict__ mem,
                                                            const double2* __restrict__ fac, double* __restrict__ part,
                                                            Ctl* __restrict__ ctl, Epi epi) {
    constexpr int NA = GRAM ? 5 : 1;
    if (ctl->dead) return;
    const RStep& r = *rp;
    uint64_t a_tb;
    int sg;
    bool active;
    run_origin(r, a_tb, sg, active);
    const double2 f0 = fac[sg], f1 = fac[sg ^ r.sig[0]];
    const int jit = 9 + r.lg_grid;
    const int niter = 1 << r.lg_iter;
    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; q++) acc[q] = 0;
    if (GRAM || r.pair) {
        const uint64_t gc = r.cols[jit];
        const int sgp = r.sig[jit];
        const double2 g0 = fac[sg ^ sgp], g1 = fac[sg ^ sgp ^ r.sig[0]];
        int pn_tb = 0;                              // next-step parities of the thread's and block's coordinates
        if constexpr (GRAM) {
            const int tid = threadIdx.x;
#pragma unroll
            for (int b = 0; b < 8; b++)
                if ((tid >> b) & 1) pn_tb ^= run_pn(r, 1 + b);
            const unsigned bid = r.xcd && r.lg_grid >= 3 ? (blockIdx.x >> 3) | ((blockIdx.x & 7u) << (r.lg_grid - 3))
                                                         : blockIdx.x;
            for (int b = 0; b < r.lg_grid; b++)
                if ((bid >> b) & 1) pn_tb ^= run_pn(r, 9 + b);
        }
        const int pn0 = run_pn(r, 0), peps = r.peps;
        double gi[kgCount];
#pragma unroll
        for (int q = 0; q < kgCount; q++) gi[q] = 0;
        uint64_t a_it = 0;
        int pn_it = 0;
        for (int i = 0; i < (niter >> 1); i++) {
            if (i > 0) {
                const int c = jit + 1 + __builtin_ctz(unsigned(i));
                a_it ^= r.cols[c];
                if constexpr (GRAM) pn_it ^= run_pn(r, c);
            }
            if (!active) continue;
            const uint64_t a = a_tb ^ a_it, b = a ^ gc;
            S e[2], h[2];
            Run<S, 2>::load(mem, a, e);
            Run<S, 2>::load(mem, b, h);
            e[0] = Cx<S>::narrow(cmul(Cx<S>::widen(e[0]), f0));
            e[1] = Cx<S>::narrow(cmul(Cx<S>::widen(e[1]), f1));
            h[0] = Cx<S>::narrow(cmul(Cx<S>::widen(h[0]), g0));
            h[1] = Cx<S>::narrow(cmul(Cx<S>::widen(h[1]), g1));
            Run<S, 2>::store(mem, a, e);
            Run<S, 2>::store(mem, b, h);
            add_norm(acc[0], Cx<S>::widen(e[0]));
            add_norm(acc[0], Cx<S>::widen(e[1]));
            add_norm(acc[0], Cx<S>::widen(h[0]));
            add_norm(acc[0], Cx<S>::widen(h[1]));
            if constexpr (GRAM) {
#pragma unroll
                for (int el = 0; el < 2; el++) {
                    const int pe = pn_tb ^ pn_it ^ (el ? pn0 : 0);      // parities of element el's address
                    const S vb = (el ^ peps) ? h[1] : h[0];
                    gram_invariants<ROUTE2>(Cx<S>::widen(e[el]), Cx<S>::widen(vb), uint32_t(pe & 1),
                                            uint32_t((pe >> 1) & 1), gi);
                }
            }
        }
        if constexpr (GRAM) gram_assemble<ROUTE2>(gi, r.par2, acc + 1);
        sweep_epilogue<NA>(acc, part, ctl, epi, kWrite);
        return;
    }
    uint64_t a_it = 0;
    for (int it = 0; it < niter; it++) {
        if (it > 0) a_it ^= r.cols[jit + __builtin_ctz(unsigned(it))];
        if (!active) continue;
        const uint64_t a = a_tb ^ a_it;
        S e[2];
        Run<S, 2>::load(mem, a, e);
        e[0] = Cx<S>::narrow(cmul(Cx<S>::widen(e[0]), f0));
        e[1] = Cx<S>::narrow(cmul(Cx<S>::widen(e[1]), f1));
        Run<S, 2>::store(mem, a, e);
        add_norm(acc[0], Cx<S>::widen(e[0]));
        add_norm(acc[0], Cx<S>::widen(e[1]));
    }
    sweep_epilogue<NA>(acc, part, ctl, epi, kWrite);
}

// ---------------------------------------------------------------------------------------------------------------
// Two-table (gather-form) steps (gsamp.cpp:21-29, :39-55; dp.cpp:44-53).  Output cp gathers its 2^nker inputs; S0 sums
// the x = 0 terms, S1 the x = 1 terms with sign (-1)^{parity(c & rho)}; the outputs are M[b][0] S0 + M[b][1] S1.
// Norm pass (SWEEP = kDirect): both branch norms |M[b][0] S0 + M[b][1] S1|^2 summed directly (the reference's chain).
// Write pass (kWrite): w0 S0 + w1 S1 stored at nw[cp] (contiguous), with w from the chain's decision or, for the
// amplitude engine, from wext[4k..4k+3] (the job's per-vertex weights).  NKS = 2^nker inputs per output.  Output
// bits 0..7 are the thread index, then the block index, then the tile index (groups of 4 tiles per thread, see
// below; one tile at a time when a block has fewer than 4).
template <class S, int SWEEP, int NKS>
__global__ __launch_bounds__(kThreads) void gather_kernel(const TStep* __restrict__ tp, const S* __restrict__ old,
                                                          S* __restrict__ nw, double* __restrict__ part,
                                                          Ctl* __restrict__ ctl, Epi epi,
                                                          const double* __restrict__ wext) {
    if (ctl->dead) return;
    const TStep& s = *tp;
    const int tid = threadIdx.x;
    uint64_t a_thr = 0;
#pragma unroll
    for (int b = 0; b < 8; b++)
        if (b < s.rn && ((tid >> b) & 1)) a_thr ^= s.dep[b];
    const bool active = s.rn >= 8 || tid < (1 << s.rn);
    uint64_t a_blk = 0;
    for (int b = 0; b < s.lg_grid; b++)
        if ((blockIdx.x >> b) & 1) a_blk ^= s.dep[8 + b];
    const int jit = 8 + s.lg_grid;
    const int niter = 1 << s.lg_iter;
    double2 M[2][2], w0 = make_double2(0, 0), w1 = w0;
    if constexpr (SWEEP == kDirect) {
        const KStep& ms = epi.plain[epi.k];
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int x = 0; x < 2; x++) M[b][x] = make_double2(ms.mre[b][x], ms.mim[b][x]);
    } else {
        const double* w = wext ? wext + 4 * epi.k : ctl->w;
        w0 = make_double2(w[0], w[1]);
        w1 = make_double2(w[2], w[3]);
    }
    uint64_t ks[NKS];
#pragma unroll
    for (int t = 0; t < NKS; t++) ks[t] = s.kspan[t];
    const uint64_t lowmask = s.lowmask, xbit = s.xbit, rho = s.rho;
    double acc[2] = {0, 0};
    // One output from its NKS gathered inputs (u0: its address before the kernel span): S0, S1, then the branch norms
    // (norm pass) or the write at cp.
    auto output = [&](const S (&v)[NKS], uint64_t u0, uint64_t cp) {
        double2 S0 = make_double2(0, 0), S1 = S0;
#pragma unroll
        for (int t = 0; t < NKS; t++) {
            const uint64_t u = u0 ^ ks[t];
            const double2 x = Cx<S>::widen(v[t]);
            if (u & xbit) S1 = cadd(S1, negate_if(x, __popcll(u & lowmask & rho) & 1));
            else S0 = cadd(S0, x);
        }
        if constexpr (SWEEP == kDirect) {
            add_norm(acc[0], cfma(M[0][0], S0, cmul(M[0][1], S1)));
            add_norm(acc[1], cfma(M[1][0], S0, cmul(M[1][1], S1)));
        } else {
            const double2 o = cfma(w0, S0, cmul(w1, S1));
            const S so = Cx<S>::narrow(o);
            Run<S, 1>::store(nw, cp, {so});
            add_norm(acc[0], o);
            add_norm(acc[1], Cx<S>::widen(so));
        }
    };
    const uint64_t ab = a_thr ^ a_blk;
    const uint64_t cb = (uint64_t(blockIdx.x) << 8) | uint64_t(tid);
    constexpr int U = 4;        // tiles per iteration (every gather of the U tiles in flight at once)
    if (niter >= U) {
        // Tiles of a group: j & 1 selects output bit jit, j >> 1 the second direction -- the partner direction
        // pdelta when the step has one (tiles j = 2, 3 then reuse the loads of tiles 0, 1, permuted by ptd: each
        // table entry is read once instead of twice), else output bit jit + 1.  The other tile bits (jit + 1 ..
        // rn - 1 without h, the second direction's top bit) take a Gray walk over the groups.
        const bool share = s.ph >= 0;
        const int h = share ? s.ph : jit + 1;
        const uint64_t d0 = s.dep[jit];
        const uint64_t D = share ? s.pD : s.dep[jit + 1];
        const uint64_t dc = share ? s.pdelta : uint64_t(1) << (jit + 1);
        const int td = s.ptd;
        uint64_t a_g = 0, c_g = 0;
        for (int g = 0; g < niter / U; g++) {
            if (g) {
                int b = jit + 1 + __builtin_ctz(unsigned(g));
                if (b >= h) b++;
                a_g ^= s.dep[b];
                c_g ^= uint64_t(1) << b;
            }
            if (!active) continue;
            S v[U][NKS];
            if (share) {
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const uint64_t u0 = ab ^ a_g ^ (j ? d0 : 0);
#pragma unroll
                    for (int t = 0; t < NKS; t++) v[j][t] = old[(u0 ^ ks[t]) & lowmask];
#pragma unroll
                    for (int t = 0; t < NKS; t++) {           // v[j][t ^ td] by selects (no indexed registers)
                        S x = v[j][0];
#pragma unroll
                        for (int q = 1; q < NKS; q++)
                            if ((t ^ td) == q) x = v[j][q];
                        v[j + 2][t] = x;
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < U; j++) {
                    const uint64_t u0 = ab ^ a_g ^ ((j & 1) ? d0 : 0) ^ ((j & 2) ? D : 0);
#pragma unroll
                    for (int t = 0; t < NKS; t++) v[j][t] = old[(u0 ^ ks[t]) & lowmask];
                }
            }
#pragma unroll
            for (int j = 0; j < U; j++) {
                const uint64_t u0 = ab ^ a_g ^ ((j & 1) ? d0 : 0) ^ ((j & 2) ? D : 0);
                output(v[j], u0, (c_g | (uint64_t(j & 1) << jit) | cb) ^ ((j & 2) ? dc : 0));
            }
        }
    } else {
        uint64_t a_it = 0, c_it = 0;
        for (int it = 0; it < niter; it++) {
            if (it) {
                const int b = __builtin_ctz(unsigned(it));
                a_it ^= s.dep[jit + b];
                c_it ^= uint64_t(1) << (jit + b);
            }
            if (!active) continue;
            const uint64_t u0 = ab ^ a_it;
            S v[NKS];
#pragma unroll
            for (int t = 0; t < NKS; t++) v[t] = old[(u0 ^ ks[t]) & lowmask];
            output(v, u0, c_it | cb);
        }
    }
    sweep_epilogue<2>(acc, part, ctl, epi, SWEEP);
}

// Entry 0 of a table set to 1 (start of a chain).
template <class S>
__global__ void table_one_kernel(S* t) { t[0] = Cx<S>::narrow(make_double2(1.0, 0.0)); }

// ---------------------------------------------------------------------------------------------------------------
// dispatch table

template <class S, int VR, int VG, int NTQ, int ROUTE, int SWEEP, int ROUTE2>
void* kptr() {
    return reinterpret_cast<void*>(&sweep_kernel<S, VR, VG, NTQ, ROUTE, SWEEP, ROUTE2>);
}

template <class S, int VR, int VG, int NTQ, int ROUTE, int SWEEP>
void* by_route2(int route2) {
    if constexpr (SWEEP == kWriteGram || SWEEP == kWriteDirect) {
        switch (route2) {
            case kRoute1: return kptr<S, VR, VG, NTQ, ROUTE, SWEEP, kRoute1>();
            case kRouteSplit: return kptr<S, VR, VG, NTQ, ROUTE, SWEEP, kRouteSplit>();
            case kRouteMixA: return kptr<S, VR, VG, NTQ, ROUTE, SWEEP, kRouteMixA>();
            case kRouteMixB: return kptr<S, VR, VG, NTQ, ROUTE, SWEEP, kRouteMixB>();
        }
        return nullptr;
    } else {
        return kptr<S, VR, VG, NTQ, ROUTE, SWEEP, kRoute1>();
    }
}

template <class S, int VR, int VG, int NTQ, int SWEEP>
void* by_route(int route, int route2) {
    switch (route) {
        case kRoute1: return by_route2<S, VR, VG, NTQ, kRoute1, SWEEP>(route2);
        case kRouteSplit: return by_route2<S, VR, VG, NTQ, kRouteSplit, SWEEP>(route2);
        case kRouteMixA: if constexpr (NTQ >= 2) return by_route2<S, VR, VG, NTQ, kRouteMixA, SWEEP>(route2); else return nullptr;
        case kRouteMixB: if constexpr (NTQ >= 2) return by_route2<S, VR, VG, NTQ, kRouteMixB, SWEEP>(route2); else return nullptr;
    }
    return nullptr;
}

template <class S, int VR, int VG, int SWEEP>
void* by_ntq(int ntq, int route, int route2) {
    switch (ntq) {
        case 1: return by_route<S, VR, VG, 1, SWEEP>(route, route2);
        case 2: return by_route<S, VR, VG, 2, SWEEP>(route, route2);
        case 4: return by_route<S, VR, VG, 4, SWEEP>(route, route2);
    }
    return nullptr;
}

// Fused sweeps followed by a four-slot step: mixed routes on both sides (the only ones the planner produces there).
template <class S, int VR, int NTQ, int SWEEP>
void* fused4(int route, int route2) {
    if (route == kRouteMixA)
        return route2 == kRouteMixA ? kptr<S, VR, 4, NTQ, kRouteMixA, SWEEP, kRouteMixA>()
                                    : kptr<S, VR, 4, NTQ, kRouteMixA, SWEEP, kRouteMixB>();
    return route2 == kRouteMixA ? kptr<S, VR, 4, NTQ, kRouteMixB, SWEEP, kRouteMixA>()
                                : kptr<S, VR, 4, NTQ, kRouteMixB, SWEEP, kRouteMixB>();
}

template <class S, int SWEEP>
void* by_layout(int lr, int vg, int ntq, int route, int route2) {
    constexpr bool c64 = sizeof(S) == 8;
    if constexpr (SWEEP == kWriteGram1) {                 // plain layout with runs of 2 only
        return vg == 1 && lr == 1 ? by_ntq<S, 2, 1, SWEEP>(ntq, route, route2) : nullptr;
    } else
    if constexpr (SWEEP == kWriteGram || SWEEP == kWriteDirect) {
        if (vg == 4) {                           // four-slot next step: two- and four-slot steps, mixed routes only
            constexpr int VR4 = c64 ? 2 : 1;
            if (lr != (c64 ? 1 : 0) || (route != kRouteMixA && route != kRouteMixB) ||
                (route2 != kRouteMixA && route2 != kRouteMixB))
                return nullptr;
            if (ntq == 2) return fused4<S, VR4, 2, SWEEP>(route, route2);
            if (ntq == 4) return fused4<S, VR4, 4, SWEEP>(route, route2);
            return nullptr;
        }
        if (vg != 2) return nullptr;
        if constexpr (c64) {
            if (lr == 1) return by_ntq<S, 2, 2, SWEEP>(ntq, route, route2);
            if (lr == 2) return by_ntq<S, 4, 2, SWEEP>(ntq, route, route2);
            return nullptr;
        }
        else return lr == 0 ? by_ntq<S, 1, 2, SWEEP>(ntq, route, route2) : nullptr;
    } else {
        if (vg == 2) {                           // two runs per thread: plain sweeps of one- and two-slot steps
            if (lr != 1) return nullptr;
            if (ntq == 1) return by_route<S, 2, 2, 1, SWEEP>(route, route2);
            if (ntq == 2) return by_route<S, 2, 2, 2, SWEEP>(route, route2);
            return nullptr;
        }
        if (vg != 1) return nullptr;
        if (lr == 0) return by_ntq<S, 1, 1, SWEEP>(ntq, route, route2);
        return lr == 1 ? by_ntq<S, 2, 1, SWEEP>(ntq, route, route2) : nullptr;
    }
}

// Kernel for (storage 0 = c64 / 1 = c128, log2 run, vg, ntq, route, sweep, next route); nullptr when not built.
void* sweep_kernel_for(int storage, int lr, int vg, int ntq, int route, int sweep, int route2) {
    if (storage == 0) {
        switch (sweep) {
            case kGram: return by_layout<float2, kGram>(lr, vg, ntq, route, route2);
            case kDirect: return by_layout<float2, kDirect>(lr, vg, ntq, route, route2);
            case kWrite: return by_layout<float2, kWrite>(lr, vg, ntq, route, route2);
            case kWriteGram: return by_layout<float2, kWriteGram>(lr, vg, ntq, route, route2);
            case kWriteDirect: return by_layout<float2, kWriteDirect>(lr, vg, ntq, route, route2);
            case kWriteGram1: return by_layout<float2, kWriteGram1>(lr, vg, ntq, route, route2);
        }
    } else {
        switch (sweep) {
            case kGram: return by_layout<double2, kGram>(lr, vg, ntq, route, route2);
            case kDirect: return by_layout<double2, kDirect>(lr, vg, ntq, route, route2);
            case kWrite: return by_layout<double2, kWrite>(lr, vg, ntq, route, route2);
            case kWriteGram: return by_layout<double2, kWriteGram>(lr, vg, ntq, route, route2);
            case kWriteDirect: return by_layout<double2, kWriteDirect>(lr, vg, ntq, route, route2);
            case kWriteGram1: return by_layout<double2, kWriteGram1>(lr, vg, ntq, route, route2);
        }
    }
    return nullptr;
}

// log2 run length of the plain layout per storage, and of the fused layout (which adds the pairing coordinate).
// Runs of 2 for both storages: complex64 plain sweeps measured faster with runs of 2 than of 4 (c2: 4.58 -> 4.47 ms
// per sample; the 2^25-coset plain write 222 -> 200 us) and slower with runs of 1.
int plain_lr(int) { return 1; }
int fused_lr(int storage) { return storage == 0 ? 1 : 0; }

void* decide_kernel_ptr() { return reinterpret_cast<void*>(&decide_kernel); }
// Reduce kernel of a sweep kind (number of partial sums: Gram 4, direct 2, write 2, write + Gram 6, write + direct 4).
void* reduce_kernel_ptr(int sweep) {
    switch (sweep) {
        case kGram: return reinterpret_cast<void*>(&sweep_reduce_kernel<4>);
        case kDirect: return reinterpret_cast<void*>(&sweep_reduce_kernel<2>);
        case kWrite: return reinterpret_cast<void*>(&sweep_reduce_kernel<2>);
        case kWriteGram: return reinterpret_cast<void*>(&sweep_reduce_kernel<6>);
        case kWriteDirect: return reinterpret_cast<void*>(&sweep_reduce_kernel<4>);
    }
    return nullptr;
}
void* publish_kernel_ptr() { return reinterpret_cast<void*>(&publish_kernel); }
void* reduce1_kernel_ptr() { return reinterpret_cast<void*>(&sweep_reduce_kernel<1>); }
void* run_hist_kernel_ptr(int storage) {
    return storage == 0 ? reinterpret_cast<void*>(&run_hist_kernel<float2>)
                        : reinterpret_cast<void*>(&run_hist_kernel<double2>);
}
void* run_apply_kernel_ptr(int storage) {
    return storage == 0 ? reinterpret_cast<void*>(&run_apply_kernel<float2, false, kRoute1>)
                        : reinterpret_cast<void*>(&run_apply_kernel<double2, false, kRoute1>);
}
// The product pass of a paired run that also sums the next step's Gram matrix (next route route2).
template <class S>
void* run_apply_gram_for(int route2) {
    switch (route2) {
        case kRoute1: return reinterpret_cast<void*>(&run_apply_kernel<S, true, kRoute1>);
        case kRouteSplit: return reinterpret_cast<void*>(&run_apply_kernel<S, true, kRouteSplit>);
        case kRouteMixA: return reinterpret_cast<void*>(&run_apply_kernel<S, true, kRouteMixA>);
        case kRouteMixB: return reinterpret_cast<void*>(&run_apply_kernel<S, true, kRouteMixB>);
    }
    return nullptr;
}
void* run_apply_gram_kernel_ptr(int storage, int route2) {
    return storage == 0 ? run_apply_gram_for<float2>(route2) : run_apply_gram_for<double2>(route2);
}
void* reduce_stored_gram_kernel_ptr() { return reinterpret_cast<void*>(&sweep_reduce_kernel<5>); }
void* run_chunk_kernel_ptr() { return reinterpret_cast<void*>(&run_chunk_kernel); }
void* run_decide_kernel_ptr() { return reinterpret_cast<void*>(&run_decide_kernel); }
int small_run_threads() { return kSmallThreads; }
// Two-table kernels: sweep kDirect (norm pass) or kWrite; storage 0 = complex64, 1 = complex128.
template <class S, int NKS>
void* gather_for(int sweep) {
    return sweep == kDirect ? reinterpret_cast<void*>(&gather_kernel<S, kDirect, NKS>)
                            : reinterpret_cast<void*>(&gather_kernel<S, kWrite, NKS>);
}
template <class S>
void* gather_by_nks(int sweep, int nks) {
    switch (nks) {
        case 1: return gather_for<S, 1>(sweep);
        case 2: return gather_for<S, 2>(sweep);
        case 4: return gather_for<S, 4>(sweep);
    }
    return nullptr;
}
void* gather_kernel_ptr(int storage, int sweep, int nks) {
    return storage == 0 ? gather_by_nks<float2>(sweep, nks) : gather_by_nks<double2>(sweep, nks);
}
void* table_one_kernel_ptr(int storage) {
    return storage == 0 ? reinterpret_cast<void*>(&table_one_kernel<float2>)
                        : reinterpret_cast<void*>(&table_one_kernel<double2>);
}
void* small_run_kernel_ptr(int storage) {
    return storage == 0 ? reinterpret_cast<void*>(&small_run_kernel<float2>)
                        : reinterpret_cast<void*>(&small_run_kernel<double2>);
}

}  // namespace gs
