// gs_oracle -- CPU restatement of the reference's in-place executor (gsamp2.cpp), written for this project.
//
// TEST INFRASTRUCTURE ONLY.  It is the checker of the HIP engine (tests/, __graft_entry__.smoke()) and the
// CPU baseline leg of bench.py.  Nothing in the product links, loads or calls it.
//
// Same command line, plan format, exit codes and output lines as the reference binary:
//   gs_oracle [--c128] PLAN sample N SEED | sample2 N SEED | prob STRINGS | chain STRINGS | hybrid STRINGS SEED
//             | prefix STRINGS K OUTPREFIX | resume STRINGS K SEED DUMP1 [DUMP2 ...] [--twist FILE]
// --c128 stores the arena as complex<double> (the reference's -DGS_DOUBLE build); default complex<float>.
// With GS_ORACLE_TIMING=1 in the environment, each output line's compute time (excluding plan parsing and the
// arena allocation) is printed to stderr as "time <seconds>".
//
// What follows which reference lines:
//   plan parsing and its exit codes 2..7 ............ gsamp2.cpp:126-139
//   coset base address (GF(2)-linear in c) .......... gsamp2.cpp:29-31  (here: four 10-bit lookup tables)
//   per-coset outputs out[b][tau] in FP64 .......... gsamp2.cpp:32-38
//   chain (sample / sample2 / prob) ................. gsamp2.cpp:39-55, modes :149-161, :223
//   chain1 (chain mode, one pass per step) .......... gsamp2.cpp:56-70, mode :218-222
//   run_steps / hybrid .............................. gsamp2.cpp:71-120, mode :209-216
//   coset offsets, prefix dump, resume + twist ...... gsamp2.cpp:121-125, :162-208
//   arena allocation (2 MB aligned, huge pages) ..... gsamp2.cpp:140-147
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>
#include <sys/mman.h>

namespace {

using cd = std::complex<double>;

struct Term {
    uint64_t off;   // physical offset tq of the slot read, relative to the coset base
    int x, tau, par;
};

struct Step {
    int v = 0, nk = 0, m = 0;
    uint64_t g = 0, rhoc = 0;
    std::vector<uint64_t> cols;       // Tkeep
    std::vector<Term> terms;
    cd M[2][2];
    std::vector<uint64_t> lut;        // 4 x 1024: XOR of the columns selected by each 10-bit chunk of c
    uint64_t base(uint64_t c) const {
        const uint64_t* t = lut.data();
        return t[c & 1023] ^ t[1024 + ((c >> 10) & 1023)] ^ t[2048 + ((c >> 20) & 1023)] ^ t[3072 + ((c >> 30) & 1023)];
    }
};

struct Plan {
    int n = 0, ns = 0, R = 0;
    std::vector<Step> steps;
};

int load_plan(const char* path, Plan& p) {
    FILE* f = fopen(path, "r");
    if (!f) return 2;
    int rc = 0;
    if (fscanf(f, "%d %d %d", &p.n, &p.ns, &p.R) != 3) { fclose(f); return 2; }
    p.steps.resize(p.ns);
    for (Step& s : p.steps) {
        unsigned long long g, rhoc;
        int nterms;
        if (fscanf(f, "%d %d %d %llu %llu %d", &s.v, &s.nk, &s.m, &g, &rhoc, &nterms) != 6) { rc = 3; break; }
        s.g = g;
        s.rhoc = rhoc;
        s.cols.assign(std::max(s.nk, 0), 0);
        for (uint64_t& col : s.cols) {
            unsigned long long y;
            if (fscanf(f, "%llu", &y) != 1) { rc = 4; break; }
            col = y;
        }
        if (rc) break;
        s.terms.assign(std::max(nterms, 0), Term{});
        for (Term& t : s.terms) {
            unsigned long long y;
            if (fscanf(f, "%llu %d %d %d", &y, &t.x, &t.tau, &t.par) != 4) { rc = 5; break; }
            t.off = y;
        }
        if (rc) break;
        for (int b = 0; b < 2 && !rc; b++)
            for (int x = 0; x < 2; x++) {
                double re, im;
                if (fscanf(f, "%lf %lf", &re, &im) != 2) { rc = 6; break; }
                s.M[b][x] = cd(re, im);
            }
        if (rc) break;
        if (s.nk > 40) { rc = 7; break; }
        s.lut.assign(4 * 1024, 0);
        for (int chunk = 0; chunk < 4; chunk++)
            for (int v = 1; v < 1024; v++) {
                // lut[v] = lut[v without its lowest bit] ^ column of that bit
                int low = __builtin_ctz(v), j = 10 * chunk + low;
                uint64_t col = j < s.nk ? s.cols[j] : 0;
                s.lut[1024 * chunk + v] = s.lut[1024 * chunk + (v & (v - 1))] ^ col;
            }
    }
    fclose(f);
    return rc;
}

template <class S>
struct Executor {
    const Plan& P;
    S* mem;

    // out[b][tau] = sum over terms of M[b][x] * (-1)^{x (parity(c & rhoc) ^ par)} * mem[base ^ tq]
    inline void coset(const Step& s, uint64_t c, uint64_t base, cd out[2][2]) const {
        out[0][0] = out[0][1] = out[1][0] = out[1][1] = cd(0, 0);
        const int pc = __builtin_parityll(c & s.rhoc);
        for (const Term& t : s.terms) {
            cd a = cd(mem[base ^ t.off]);
            if (t.x && (pc ^ t.par)) a = -a;
            out[0][t.tau] += s.M[0][t.x] * a;
            out[1][t.tau] += s.M[1][t.x] * a;
        }
    }

    void norms(const Step& s, double& n0, double& n1) const {
        const int64_t N = int64_t(1) << s.nk;
        double a0 = 0, a1 = 0;
#pragma omp parallel for schedule(static) reduction(+ : a0, a1)
        for (int64_t c = 0; c < N; c++) {
            cd o[2][2];
            coset(s, uint64_t(c), s.base(uint64_t(c)), o);
            a0 += std::norm(o[0][0]) + std::norm(o[0][1]);
            a1 += std::norm(o[1][0]) + std::norm(o[1][1]);
        }
        n0 = a0;
        n1 = a1;
    }

    // Store branch bv scaled by sc; returns sum of |scaled value|^2 in FP64 (before the store rounds).
    double write(const Step& s, int bv, double sc) const {
        const int64_t N = int64_t(1) << s.nk;
        double nb = 0;
#pragma omp parallel for schedule(static) reduction(+ : nb)
        for (int64_t c = 0; c < N; c++) {
            cd o[2][2];
            const uint64_t base = s.base(uint64_t(c));
            coset(s, uint64_t(c), base, o);
            const cd w0 = o[bv][0] * sc, w1 = o[bv][1] * sc;
            mem[base] = S(w0);
            if (s.m == 2) mem[base ^ s.g] = S(w1);
            nb += std::norm(w0) + std::norm(w1);
        }
        return nb;
    }

    // gsamp2.cpp:39-55: both passes per step, table renormalised to 1; forced == nullptr draws every bit.
    double chain(const int* forced, int* bits, std::mt19937_64* rng) const {
        mem[0] = S(1);
        double logp = 0;
        for (int k = 0; k < P.ns; k++) {
            const Step& s = P.steps[k];
            double n0, n1;
            norms(s, n0, n1);
            const double p1 = n1 / (n0 + n1);
            int bv;
            if (forced) bv = forced[k];
            else {
                std::uniform_real_distribution<double> U(0, 1);
                bv = U(*rng) < p1;
            }
            const double pb = bv ? p1 : 1 - p1;
            if (pb <= 0) return -INFINITY;
            logp += std::log2(pb);
            bits[k] = bv;
            write(s, bv, 1.0 / std::sqrt(bv ? n1 : n0));
        }
        return logp;
    }

    // gsamp2.cpp:56-70: one pass per step; cond[k] = nb / (2 prev) since n0 + n1 = 2 prev.
    double chain1(const int* forced, double* cond) const {
        mem[0] = S(1);
        double prev = 1;
        for (int k = 0; k < P.ns; k++) {
            const double nb = write(P.steps[k], forced[k], 1.0);
            cond[k] = nb / (2 * prev);
            prev = nb;
            if (nb <= 0) {
                for (; k < P.ns; k++) cond[k] = 0;
                return -INFINITY;
            }
        }
        return std::log2(prev) - P.ns;
    }

    // gsamp2.cpp:77-98: steps k0..k1-1 on a stored table of norm *prev (true table = stored * 2^*logscale).
    double run_steps(int k0, int k1, const int* forced, int* bits, double* cond, std::mt19937_64* rng,
                     double* prev, double* logscale) const {
        double logp = 0;
        for (int k = k0; k < k1; k++) {
            const Step& s = P.steps[k];
            int bv = forced[k];
            const double sc = 1.0 / std::sqrt(*prev);
            if (bv < 0) {
                double n0, n1;
                norms(s, n0, n1);
                std::uniform_real_distribution<double> U(0, 1);
                bv = U(*rng) < n1 / (n0 + n1);
            }
            const double nb = write(s, bv, sc);
            *logscale += 0.5 * std::log2(*prev);
            cond[k] = nb / 2;
            bits[k] = bv;
            *prev = nb;
            if (nb <= 0) {
                for (; k < k1; k++) {
                    cond[k] = 0;
                    bits[k] = forced[k] < 0 ? 0 : forced[k];
                }
                return -INFINITY;
            }
            logp += std::log2(cond[k]);
        }
        return logp;
    }

    // gsamp2.cpp:99-120 (run_steps over the whole order from a fresh table, logscale unused).
    double hybrid(const int* forced, int* bits, double* cond, std::mt19937_64* rng) const {
        mem[0] = S(1);
        double prev = 1, logscale = 0;
        return run_steps(0, P.ns, forced, bits, cond, rng, &prev, &logscale);
    }

    // Sorted distinct term offsets of step K: the live table is {base_K(c) ^ tq}.  {0} past the end.
    std::vector<uint64_t> offsets(int K) const {
        std::vector<uint64_t> v;
        if (K >= P.ns) return {0};
        for (const Term& t : P.steps[K].terms) v.push_back(t.off);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        return v;
    }
};

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool timing() {
    const char* e = getenv("GS_ORACLE_TIMING");
    return e && e[0] == '1';
}

void print_bits(const std::vector<int>& b) {
    for (int x : b) putchar('0' + x);
}

template <class S>
int run(const Plan& P, int argc, char** argv) {
    const size_t bytes = (size_t(1) << P.R) * sizeof(S);
    void* raw = nullptr;
    if (posix_memalign(&raw, size_t(1) << 21, bytes) != 0 || !raw) {
        fprintf(stderr, "allocation of %zu bytes failed\n", bytes);
        return 11;
    }
#ifdef MADV_HUGEPAGE
    madvise(raw, bytes, MADV_HUGEPAGE);
#endif
    Executor<S> E{P, static_cast<S*>(raw)};
    const int ns = P.ns;
    std::vector<int> bits(ns);
    std::vector<double> cond(ns);
    const std::string mode = argv[2];
    const bool tm = timing();
    int rc = 0;
    if (mode == "sample" || mode == "sample2") {
        const long N = atol(argv[3]);
        const unsigned long long seed = strtoull(argv[4], nullptr, 10);
        for (long i = 0; i < N; i++) {
            std::mt19937_64 rng;
            if (mode == "sample") rng.seed(seed + (unsigned long long)i);
            else {
                std::seed_seq sq{unsigned(seed & 0xffffffffu), unsigned(seed >> 32), unsigned((unsigned long long)i & 0xffffffffu),
                                 unsigned((unsigned long long)i >> 32), 0x9E3779B9u};
                rng.seed(sq);
            }
            const double t0 = now();
            const double lp = E.chain(nullptr, bits.data(), &rng);
            if (tm) fprintf(stderr, "time %.6f\n", now() - t0);
            print_bits(bits);
            printf(" %.12f %llu\n", lp, seed + (unsigned long long)i);
            fflush(stdout);
        }
    } else if (mode == "prefix") {
        const int K = atoi(argv[4]);
        const char* outp = argv[5];
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); free(raw); return 1; }
        char line[4096];
        long li = 0;
        while (fgets(line, sizeof line, in)) {
            if ((int)strlen(line) < K) continue;
            std::vector<int> fb(ns, 0);
            for (int k = 0; k < K; k++) {
                if (line[k] == '?') { fprintf(stderr, "'?' before K in prefix mode\n"); fclose(in); free(raw); return 8; }
                fb[k] = line[k] - '0';
            }
            E.mem[0] = S(1);
            double prev = 1, logscale = 0;
            const double lp = E.run_steps(0, K, fb.data(), bits.data(), cond.data(), nullptr, &prev, &logscale);
            const int nk = K < ns ? P.steps[K].nk : 0;
            const std::vector<uint64_t> offs = E.offsets(K);
            const int ntq = (int)offs.size();
            char fn[1024];
            snprintf(fn, sizeof fn, "%s_%ld.bin", outp, li++);
            FILE* o = fopen(fn, "wb");
            fwrite(&nk, 4, 1, o);
            fwrite(&ntq, 4, 1, o);
            fwrite(&logscale, 8, 1, o);
            for (int64_t c = 0; c < (int64_t(1) << nk); c++)
                for (uint64_t tq : offs) {
                    S v = K < ns ? E.mem[P.steps[K].base(uint64_t(c)) ^ tq] : E.mem[0];
                    fwrite(&v, sizeof(S), 1, o);
                }
            fclose(o);
            printf("%ld %.12f %.6f\n", li - 1, lp, logscale);
            fflush(stdout);
        }
        fclose(in);
    } else if (mode == "resume") {
        int nsl = argc - 6;
        const char* twistf = nullptr;
        if (argc >= 8 && !strcmp(argv[argc - 2], "--twist")) { twistf = argv[argc - 1]; nsl -= 2; }
        const int K = atoi(argv[4]);
        const unsigned long long seed = strtoull(argv[5], nullptr, 10);
        const int nk = K < ns ? P.steps[K].nk : 0;
        const std::vector<uint64_t> offs = E.offsets(K);
        const int ntq = (int)offs.size();
        const int64_t Nt = (int64_t(1) << nk) * ntq;
        std::vector<std::vector<S>> tabs(nsl, std::vector<S>(Nt));
        std::vector<double> ls(nsl);
        std::vector<uint64_t> cshift(nsl, 0);
        std::vector<std::vector<int>> perm(nsl, std::vector<int>(ntq));
        for (auto& pm : perm)
            for (int t = 0; t < ntq; t++) pm[t] = t;
        if (twistf) {
            FILE* tf = fopen(twistf, "r");
            if (!tf) { fprintf(stderr, "missing twist file\n"); free(raw); return 12; }
            for (int s = 0; s < nsl; s++) {
                unsigned long long cs, ts;
                if (fscanf(tf, "%llu %llu", &cs, &ts) != 2) { fprintf(stderr, "bad twist line %d\n", s); fclose(tf); free(raw); return 13; }
                if (cs >> nk) { fprintf(stderr, "twist cshift exceeds the kept coordinates\n"); fclose(tf); free(raw); return 14; }
                cshift[s] = cs;
                for (int t = 0; t < ntq; t++) {
                    auto it = std::find(offs.begin(), offs.end(), offs[t] ^ ts);
                    if (it == offs.end()) {
                        fprintf(stderr, "twist tqshift %llu does not map the coset offsets of step %d onto themselves\n", ts, K);
                        fclose(tf); free(raw); return 14;
                    }
                    perm[s][t] = int(it - offs.begin());
                }
            }
            fclose(tf);
        }
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); free(raw); return 1; }
        char line[4096];
        long li = 0;
        while (fgets(line, sizeof line, in)) {
            if ((int)strlen(line) < ns) continue;
            std::vector<int> fb(ns);
            for (int k = 0; k < ns; k++) fb[k] = line[k] == '?' ? -1 : line[k] - '0';
            double L = -1e300;
            for (int s = 0; s < nsl; s++) {
                char fn[1024];
                snprintf(fn, sizeof fn, "%s_%ld.bin", argv[6 + s], li);
                FILE* o = fopen(fn, "rb");
                if (!o) { fprintf(stderr, "missing %s\n", fn); fclose(in); free(raw); return 9; }
                int nk2 = -1, ntq2 = -1;
                const bool ok = fread(&nk2, 4, 1, o) == 1 && nk2 == nk && fread(&ntq2, 4, 1, o) == 1 && ntq2 == ntq &&
                                fread(&ls[s], 8, 1, o) == 1 && fread(tabs[s].data(), sizeof(S), Nt, o) == (size_t)Nt;
                fclose(o);
                if (!ok) { fprintf(stderr, "bad dump %s\n", fn); fclose(in); free(raw); return 10; }
                L = std::max(L, ls[s]);
            }
            double norm = 0;
            for (int64_t c = 0; c < (int64_t(1) << nk); c++)
                for (int ti = 0; ti < ntq; ti++) {
                    cd acc = 0;
                    for (int s = 0; s < nsl; s++)
                        acc += cd(tabs[s][int64_t(uint64_t(c) ^ cshift[s]) * ntq + perm[s][ti]]) * std::exp2(ls[s] - L);
                    if (K < ns) E.mem[P.steps[K].base(uint64_t(c)) ^ offs[ti]] = S(acc);
                    else E.mem[0] = S(acc);
                    norm += std::norm(acc);
                }
            const double lp_prefix = -K + std::log2(norm) + 2 * L;
            double prev = norm, logscale = L;
            std::mt19937_64 r2(seed + (unsigned long long)li);
            const double lps = E.run_steps(K, ns, fb.data(), bits.data(), cond.data(), &r2, &prev, &logscale);
            for (int k = 0; k < ns; k++) putchar('0' + (k < K ? fb[k] : bits[k]));
            printf(" %.12f %.12f", lp_prefix + lps, lp_prefix);
            for (int k = K; k < ns; k++) printf(" %.9f", cond[k]);
            putchar('\n');
            fflush(stdout);
            li++;
        }
        fclose(in);
    } else if (mode == "hybrid") {
        const unsigned long long seed = strtoull(argv[4], nullptr, 10);
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); free(raw); return 1; }
        char line[4096];
        long li = 0;
        while (fgets(line, sizeof line, in)) {
            if ((int)strlen(line) < ns) continue;
            std::vector<int> fb(ns);
            for (int k = 0; k < ns; k++) fb[k] = line[k] == '?' ? -1 : line[k] - '0';
            std::mt19937_64 r2(seed + (unsigned long long)li++);
            const double t0 = now();
            const double lp = E.hybrid(fb.data(), bits.data(), cond.data(), &r2);
            if (tm) fprintf(stderr, "time %.6f\n", now() - t0);
            print_bits(bits);
            printf(" %.12f", lp);
            for (int k = 0; k < ns; k++) printf(" %.9f", cond[k]);
            putchar('\n');
            fflush(stdout);
        }
        fclose(in);
    } else {
        const bool ch = mode == "chain";
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); free(raw); return 1; }
        char line[4096];
        while (fgets(line, sizeof line, in)) {
            if ((int)strlen(line) < ns) continue;
            std::vector<int> fb(ns);
            for (int k = 0; k < ns; k++) fb[k] = line[k] - '0';
            const double t0 = now();
            if (ch) {
                const double lp = E.chain1(fb.data(), cond.data());
                if (tm) fprintf(stderr, "time %.6f\n", now() - t0);
                printf("%.12f", lp);
                for (int k = 0; k < ns; k++) printf(" %.9f", cond[k]);
                putchar('\n');
            } else {
                const double lp = E.chain(fb.data(), bits.data(), nullptr);
                if (tm) fprintf(stderr, "time %.6f\n", now() - t0);
                printf("%.12f\n", lp);
            }
            fflush(stdout);
        }
        fclose(in);
    }
    free(raw);
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    bool c128 = false;
    if (argc > 1 && !strcmp(argv[1], "--c128")) {
        c128 = true;
        argv++;
        argc--;
    }
    if (argc < 4) return 1;
    Plan P;
    const int rc = load_plan(argv[1], P);
    if (rc) return rc;
    return c128 ? run<std::complex<double>>(P, argc, argv) : run<std::complex<float>>(P, argc, argv);
}
