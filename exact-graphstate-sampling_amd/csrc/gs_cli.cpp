// gsamp2_amd -- command-line front end over the C ABI with the reference engine's contract (gsamp2.cpp:1-6, :149-225):
//   gsamp2_amd [--c128] [--device N] PLAN sample N SEED | sample2 N SEED | prob STRINGS | chain STRINGS
//              | hybrid STRINGS SEED | prefix STRINGS K OUTPREFIX | resume STRINGS K SEED DUMP1 [DUMP2 ...] [--twist FILE]
// Output lines, dump files and exit codes are those of the reference, so drivers that take a gsamp2-compatible
// binary run unchanged.  --c128 selects the complex<double> arena (the reference's -DGS_DOUBLE build).
// One difference: a strings-file line whose first ns characters are not bits ('?' where sampled) is an error
// (exit 16) here; the reference reads such characters as garbage bit values.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsamp_amd.h"

namespace {

int die(int rc) {
    const char* m = gs_last_error();
    if (m && *m) fprintf(stderr, "gsamp2_amd: %s\n", m);
    return rc;
}

void put_bits(const int8_t* b, int ns) {
    for (int k = 0; k < ns; k++) putchar('0' + b[k]);
}

// Lines of at least minlen characters (newline included, as strlen(line) < ns skips at gsamp2.cpp:220).
bool read_line(FILE* f, std::string& line) {
    char buf[4096];
    if (!fgets(buf, sizeof buf, f)) return false;
    line = buf;
    return true;
}

int to_bits(const std::string& line, int n, bool allow_q, int8_t* out) {
    for (int k = 0; k < n; k++) {
        const char ch = line[size_t(k)];
        if (ch == '0' || ch == '1') out[k] = int8_t(ch - '0');
        else if (allow_q && ch == '?') out[k] = -1;
        else {
            fprintf(stderr, "gsamp2_amd: character '%c' at position %d is not a bit\n", ch, k);
            return 16;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    int storage = GS_C64, device = 0;
    while (argc > 1 && argv[1][0] == '-' && argv[1][1] == '-') {
        if (!strcmp(argv[1], "--c128")) {
            storage = GS_C128;
            argv++;
            argc--;
        } else if (!strcmp(argv[1], "--device") && argc > 2) {
            device = atoi(argv[2]);
            argv += 2;
            argc -= 2;
        } else {
            break;
        }
    }
    if (argc < 4) return 1;
    gs_plan* plan = nullptr;
    if (int rc = gs_plan_load(argv[1], &plan)) return rc;
    int n = 0, ns = 0, R = 0;
    gs_plan_info(plan, &n, &ns, &R);
    gs_engine* eng = nullptr;
    if (int rc = gs_engine_create(plan, storage, device, &eng)) {
        if (rc == 11) fprintf(stderr, "allocation of %zu bytes failed\n", ((size_t)1 << R) * (storage ? 16 : 8));
        else die(rc);
        gs_plan_destroy(plan);
        return rc;
    }
    const std::string mode = argv[2];
    int rc = 0;
    std::vector<int8_t> bits(size_t(std::max(ns, 1))), obits(size_t(std::max(ns, 1)));
    std::vector<double> cond(size_t(std::max(ns, 1)));
    if (mode == "sample" || mode == "sample2") {
        const long N = atol(argv[3]);
        const unsigned long long seed = strtoull(argv[4], nullptr, 10);
        const int seeding = mode == "sample" ? GS_SEED_PLAIN : GS_SEED_SEQ;
        const long chunk = 16;
        std::vector<int8_t> b(size_t(chunk) * size_t(std::max(ns, 1)));
        std::vector<double> lp(static_cast<size_t>(chunk));
        for (long i0 = 0; i0 < N && !rc; i0 += chunk) {
            const long cnt = std::min(chunk, N - i0);
            if ((rc = gs_sample(eng, seed, i0, cnt, seeding, b.data(), lp.data()))) { die(rc); break; }
            for (long i = 0; i < cnt; i++) {
                put_bits(b.data() + size_t(i) * ns, ns);
                printf(" %.12f %llu\n", lp[size_t(i)], seed + (unsigned long long)(i0 + i));
            }
            fflush(stdout);
        }
    } else if (mode == "prefix") {
        const int K = atoi(argv[4]);
        const char* outp = argv[5];
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); rc = 1; }
        std::string line;
        long li = 0;
        std::vector<double> tab;
        while (!rc && read_line(in, line)) {
            if ((int)line.size() < K) continue;
            std::fill(bits.begin(), bits.end(), 0);
            for (int k = 0; k < K; k++)
                if (line[size_t(k)] == '?') { fprintf(stderr, "'?' before K in prefix mode\n"); rc = 8; break; }
            if (rc) break;
            if ((rc = to_bits(line, K, false, bits.data()))) break;
            int nk = 0, ntq = 0;
            double ls = 0, lp = 0;
            if ((rc = gs_prefix(eng, bits.data(), K, nullptr, 0, &nk, &ntq, nullptr, nullptr))) { die(rc); break; }
            const size_t nt = (size_t(1) << nk) * size_t(ntq);
            tab.resize(2 * nt);
            if ((rc = gs_prefix(eng, bits.data(), K, tab.data(), int64_t(nt), &nk, &ntq, &ls, &lp))) { die(rc); break; }
            char fn[1024];
            snprintf(fn, sizeof fn, "%s_%ld.bin", outp, li++);
            FILE* o = fopen(fn, "wb");
            if (!o) { fprintf(stderr, "cannot write %s\n", fn); rc = 1; break; }
            fwrite(&nk, 4, 1, o);
            fwrite(&ntq, 4, 1, o);
            fwrite(&ls, 8, 1, o);
            if (storage == GS_C64) {
                std::vector<float> f(2 * nt);
                for (size_t i = 0; i < 2 * nt; i++) f[i] = float(tab[i]);
                fwrite(f.data(), sizeof(float), f.size(), o);
            } else {
                fwrite(tab.data(), sizeof(double), tab.size(), o);
            }
            fclose(o);
            printf("%ld %.12f %.6f\n", li - 1, lp, ls);
            fflush(stdout);
        }
        if (in) fclose(in);
    } else if (mode == "resume") {
        int nsl = argc - 6;
        const char* twistf = nullptr;
        if (argc >= 8 && !strcmp(argv[argc - 2], "--twist")) {
            twistf = argv[argc - 1];
            nsl -= 2;
        }
        const int K = atoi(argv[4]);
        const unsigned long long seed = strtoull(argv[5], nullptr, 10);
        int nk = 0, ntq = 0;
        std::fill(bits.begin(), bits.end(), 0);
        if ((rc = gs_prefix(eng, bits.data(), K, nullptr, 0, &nk, &ntq, nullptr, nullptr))) return die(rc);
        const size_t nt = (size_t(1) << nk) * size_t(ntq);
        std::vector<unsigned long long> cs(size_t(std::max(nsl, 1)), 0), ts(size_t(std::max(nsl, 1)), 0);
        if (twistf) {
            FILE* tf = fopen(twistf, "r");
            if (!tf) { fprintf(stderr, "missing twist file\n"); return 12; }
            for (int s = 0; s < nsl; s++)
                if (fscanf(tf, "%llu %llu", &cs[size_t(s)], &ts[size_t(s)]) != 2) {
                    fprintf(stderr, "bad twist line %d\n", s);
                    fclose(tf);
                    return 13;
                }
            fclose(tf);
        }
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); rc = 1; }
        std::string line;
        long li = 0;
        std::vector<std::vector<double>> tabs(size_t(std::max(nsl, 1)), std::vector<double>(2 * nt));
        std::vector<const double*> tp(size_t(std::max(nsl, 1)));
        std::vector<double> lss(size_t(std::max(nsl, 1)));
        while (!rc && read_line(in, line)) {
            if ((int)line.size() < ns) continue;
            if ((rc = to_bits(line, ns, true, bits.data()))) break;
            for (int s = 0; s < nsl && !rc; s++) {
                char fn[1024];
                snprintf(fn, sizeof fn, "%s_%ld.bin", argv[6 + s], li);
                FILE* o = fopen(fn, "rb");
                if (!o) { fprintf(stderr, "missing %s\n", fn); rc = 9; break; }
                int nk2 = -1, ntq2 = -1;
                bool ok = fread(&nk2, 4, 1, o) == 1 && nk2 == nk && fread(&ntq2, 4, 1, o) == 1 && ntq2 == ntq &&
                          fread(&lss[size_t(s)], 8, 1, o) == 1;
                auto& T = tabs[size_t(s)];
                if (ok) {
                    if (storage == GS_C64) {
                        std::vector<float> f(2 * nt);
                        ok = fread(f.data(), sizeof(float), f.size(), o) == f.size();
                        for (size_t i = 0; ok && i < f.size(); i++) T[i] = double(f[i]);
                    } else {
                        ok = fread(T.data(), sizeof(double), T.size(), o) == T.size();
                    }
                }
                fclose(o);
                if (!ok) { fprintf(stderr, "bad dump %s\n", fn); rc = 10; break; }
                tp[size_t(s)] = T.data();
            }
            if (rc) break;
            double lpf = 0, lpp = 0;
            if ((rc = gs_resume(eng, bits.data(), K, nsl, tp.data(), lss.data(), twistf ? (const uint64_t*)cs.data() : nullptr,
                                twistf ? (const uint64_t*)ts.data() : nullptr, seed, li, obits.data(), &lpf, &lpp,
                                cond.data()))) {
                die(rc);
                break;
            }
            put_bits(obits.data(), ns);
            printf(" %.12f %.12f", lpf, lpp);
            for (int k = K; k < ns; k++) printf(" %.9f", cond[size_t(k)]);
            putchar('\n');
            fflush(stdout);
            li++;
        }
        if (in) fclose(in);
    } else if (mode == "hybrid") {
        const unsigned long long seed = strtoull(argv[4], nullptr, 10);
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); rc = 1; }
        std::string line;
        long li = 0;
        while (!rc && read_line(in, line)) {
            if ((int)line.size() < ns) continue;
            if ((rc = to_bits(line, ns, true, bits.data()))) break;
            double lp = 0;
            if ((rc = gs_hybrid(eng, bits.data(), 1, seed, li++, obits.data(), &lp, cond.data()))) { die(rc); break; }
            put_bits(obits.data(), ns);
            printf(" %.12f", lp);
            for (int k = 0; k < ns; k++) printf(" %.9f", cond[size_t(k)]);
            putchar('\n');
            fflush(stdout);
        }
        if (in) fclose(in);
    } else {
        const bool ch = mode == "chain";
        FILE* in = fopen(argv[3], "r");
        if (!in) { fprintf(stderr, "missing %s\n", argv[3]); rc = 1; }
        std::string line;
        while (!rc && read_line(in, line)) {
            if ((int)line.size() < ns) continue;
            if ((rc = to_bits(line, ns, false, bits.data()))) break;
            double lp = 0;
            if (ch) {
                if ((rc = gs_chain(eng, bits.data(), 1, &lp, cond.data(), nullptr))) { die(rc); break; }
                printf("%.12f", lp);
                for (int k = 0; k < ns; k++) printf(" %.9f", cond[size_t(k)]);
                putchar('\n');
            } else {
                if ((rc = gs_prob(eng, bits.data(), 1, &lp))) { die(rc); break; }
                printf("%.12f\n", lp);
            }
            fflush(stdout);
        }
        if (in) fclose(in);
    }
    gs_engine_destroy(eng);
    gs_plan_destroy(plan);
    return rc;
}
