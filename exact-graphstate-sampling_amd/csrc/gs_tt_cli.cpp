// gsamp_amd / dp_amd -- command-line front ends of the two-table engine with the reference's contracts:
//   gsamp_amd [--c128] [--device N] PLAN sample N SEED | prob STRINGS | chain STRINGS | hybrid STRINGS SEED
//                                        | prefix STRINGS K OUTPREFIX | resume STRINGS K SEED DUMPPREFIX...
//                                                                        (gsamp.cpp:1-5, :108-161)
//   dp_amd    [--c64] [--device N] PLAN JOBS                             (dp.cpp:1-4, :26-62)
// The program's file name selects the contract (a name ending in "dp_amd" is the amplitude engine).  Output lines and
// exit codes are the reference's; gsamp's tables are complex64 by default (the reference's only storage), dp's
// complex128 (the reference's default; --c64 is its -DDP_FLOAT build).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsamp_amd.h"

namespace {

int die(int rc) {
    const char* m = gs_tt_last_error();
    if (m && *m) fprintf(stderr, "%s\n", m);
    return rc;
}

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

// Lines of a strings file with at least n characters (shorter lines are skipped, gsamp.cpp:150).
int read_strings(const char* path, int n, bool allow_q, std::vector<int8_t>& bits, int64_t& count) {
    FILE* g = fopen(path, "r");
    if (!g) {
        fprintf(stderr, "missing %s\n", path);
        return 1;
    }
    char line[4096];
    count = 0;
    while (fgets(line, sizeof line, g)) {
        if (int(strlen(line)) < n) continue;
        for (int k = 0; k < n; k++) {
            const char ch = line[k];
            if (ch == '0' || ch == '1') bits.push_back(int8_t(ch - '0'));
            else if (allow_q && ch == '?') bits.push_back(-1);
            else {
                fprintf(stderr, "character '%c' at position %d is not a bit\n", ch, k);
                fclose(g);
                return 16;
            }
        }
        count++;
    }
    fclose(g);
    return 0;
}

int run_dp(int storage, int device, const char* plan_path, const char* jobs_path) {
    gs_tt_plan* plan = nullptr;
    if (int rc = gs_tt_plan_load(plan_path, 0, &plan)) return die(rc);
    int m = 0, ns = 0, maxr = 0;
    gs_tt_plan_info(plan, &m, &ns, &maxr);
    FILE* g = fopen(jobs_path, "r");
    int nj;
    if (!g || fscanf(g, "%d", &nj) != 1) return 5;
    gs_tt_engine* eng = nullptr;
    if (int rc = gs_tt_engine_create(plan, storage, device, &eng)) return die(rc);
    for (int j = 0; j < nj; j++) {                         // one job at a time, as the reference prints
        int nsl;
        if (fscanf(g, "%d", &nsl) != 1) return 6;
        std::vector<double> pre(2 * size_t(nsl)), w(size_t(nsl) * size_t(m) * 4);
        for (int sl = 0; sl < nsl; sl++) {
            if (fscanf(g, "%lf %lf", &pre[2 * size_t(sl)], &pre[2 * size_t(sl) + 1]) != 2) return 7;
            for (int q = 0; q < m; q++)
                for (int c = 0; c < 4; c++)
                    if (fscanf(g, "%lf", &w[(size_t(sl) * m + q) * 4 + c]) != 1) return 8;
        }
        double out[2];
        if (int rc = gs_tt_dp(eng, 1, &nsl, pre.data(), w.data(), out)) return die(rc);
        printf("%.17g %.17g\n", out[0], out[1]);
        fflush(stdout);
    }
    fclose(g);
    gs_tt_engine_destroy(eng);
    gs_tt_plan_destroy(plan);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const bool dp = ends_with(argv[0], "dp_amd");
    int storage = dp ? GS_C128 : GS_C64, device = 0;
    while (argc > 1 && argv[1][0] == '-' && argv[1][1] == '-') {
        if (!strcmp(argv[1], "--c128")) storage = GS_C128;
        else if (!strcmp(argv[1], "--c64")) storage = GS_C64;
        else if (!strcmp(argv[1], "--device") && argc > 2) {
            device = atoi(argv[2]);
            argv++;
            argc--;
        } else break;
        argv++;
        argc--;
    }
    if (dp) {
        if (argc < 3) {
            fprintf(stderr, "usage: dp plan.txt jobs.txt\n");
            return 1;
        }
        return run_dp(storage, device, argv[1], argv[2]);
    }
    if (argc < 4) {
        fprintf(stderr, "usage\n");
        return 1;
    }
    gs_tt_plan* plan = nullptr;
    if (int rc = gs_tt_plan_load(argv[1], 1, &plan)) return die(rc);
    int n = 0, ns = 0, maxr = 0;
    gs_tt_plan_info(plan, &n, &ns, &maxr);
    gs_tt_engine* eng = nullptr;
    if (int rc = gs_tt_engine_create(plan, storage, device, &eng)) return die(rc);
    const std::string mode = argv[2];
    int rc = 0;
    if (mode == "sample") {
        if (argc < 5) return 1;
        const long N = atol(argv[3]);
        const unsigned long long seed = strtoull(argv[4], nullptr, 10);
        std::vector<int8_t> bits(static_cast<size_t>(ns));
        double lp;
        for (long i = 0; i < N && !rc; i++) {
            if ((rc = gs_tt_sample(eng, seed, i, 1, bits.data(), &lp))) break;
            for (int k = 0; k < ns; k++) putchar('0' + bits[size_t(k)]);
            printf(" %.12f %llu\n", lp, seed + (unsigned long long)i);
            fflush(stdout);
        }
    } else if (mode == "prob" || mode == "chain") {
        std::vector<int8_t> bits;
        int64_t cnt = 0;
        if ((rc = read_strings(argv[3], ns, false, bits, cnt))) return rc;
        std::vector<double> lp(static_cast<size_t>(cnt)), cond(static_cast<size_t>(cnt) * ns);
        rc = mode == "chain" ? gs_tt_chain(eng, bits.data(), cnt, lp.data(), cond.data(), nullptr)
                             : gs_tt_prob(eng, bits.data(), cnt, lp.data());
        if (!rc)
            for (int64_t i = 0; i < cnt; i++) {
                printf("%.12f", lp[size_t(i)]);
                if (mode == "chain")
                    for (int k = 0; k < ns; k++) printf(" %.9f", cond[size_t(i) * ns + k]);
                putchar('\n');
            }
    } else if (mode == "hybrid") {
        if (argc < 5) return 1;
        const unsigned long long seed = strtoull(argv[4], nullptr, 10);
        std::vector<int8_t> bits;
        int64_t cnt = 0;
        if ((rc = read_strings(argv[3], ns, true, bits, cnt))) return rc;
        std::vector<int8_t> ob(bits.size());
        std::vector<double> lp(static_cast<size_t>(cnt)), cond(static_cast<size_t>(cnt) * ns);
        rc = gs_tt_hybrid(eng, bits.data(), cnt, seed, 0, ob.data(), lp.data(), cond.data());
        if (!rc)
            for (int64_t i = 0; i < cnt; i++) {
                for (int k = 0; k < ns; k++) putchar('0' + ob[size_t(i) * ns + k]);
                printf(" %.12f", lp[size_t(i)]);
                for (int k = 0; k < ns; k++) printf(" %.9f", cond[size_t(i) * ns + k]);
                putchar('\n');
            }
    } else if (mode == "prefix") {
        // gsamp.cpp:124-135: lines of at least K characters; the table read by step K to OUTPREFIX_<line>.bin
        if (argc < 6) return 1;
        const int K = atoi(argv[4]);
        if (K < 0 || K > ns) return 16;
        std::vector<int8_t> pre;
        int64_t cnt = 0;
        if ((rc = read_strings(argv[3], K, false, pre, cnt))) return rc == 16 ? 8 : rc;
        int r = 0;
        std::vector<int8_t> bits(static_cast<size_t>(ns), 0);
        if ((rc = gs_tt_prefix(eng, bits.data(), K, nullptr, 0, &r, nullptr, nullptr))) return die(rc);
        const int64_t nt = int64_t(1) << r;
        std::vector<double> tab(2 * static_cast<size_t>(nt));
        std::vector<float> tf(2 * static_cast<size_t>(nt));
        for (int64_t i = 0; i < cnt && !rc; i++) {
            std::copy(pre.begin() + i * K, pre.begin() + (i + 1) * K, bits.begin());
            double ls = 0, lp = 0;
            if ((rc = gs_tt_prefix(eng, bits.data(), K, tab.data(), nt, &r, &ls, &lp))) break;
            const std::string fn = std::string(argv[5]) + "_" + std::to_string(i) + ".bin";
            FILE* o = fopen(fn.c_str(), "wb");
            if (!o) return 9;
            const int one = 1;
            for (size_t q = 0; q < tf.size(); q++) tf[q] = float(tab[q]);
            fwrite(&r, 4, 1, o);
            fwrite(&one, 4, 1, o);
            fwrite(&ls, 8, 1, o);
            fwrite(tf.data(), sizeof(float), tf.size(), o);
            fclose(o);
            printf("%lld %.12f %.6f\n", (long long)i, lp, ls);
            fflush(stdout);
        }
    } else if (mode == "resume") {
        // gsamp.cpp:136-148: per line, sum the dumps DUMPPREFIX_s_<line>.bin and continue steps K..ns-1
        if (argc < 7) return 1;
        const int K = atoi(argv[4]);
        const unsigned long long seed = strtoull(argv[5], nullptr, 10);
        const int nsl = argc - 6;
        if (K < 0 || K > ns) return 16;
        std::vector<int8_t> bits;
        int64_t cnt = 0;
        if ((rc = read_strings(argv[3], ns, true, bits, cnt))) return rc;
        int r = 0;                                         // st[K].r: every dump must match it
        const std::vector<int8_t> zeros(static_cast<size_t>(ns), 0);
        if ((rc = gs_tt_prefix(eng, zeros.data(), K, nullptr, 0, &r, nullptr, nullptr))) return die(rc);
        std::vector<int8_t> ob(static_cast<size_t>(ns));
        std::vector<double> cond(static_cast<size_t>(ns));
        for (int64_t i = 0; i < cnt && !rc; i++) {
            std::vector<std::vector<double>> tabs(static_cast<size_t>(nsl));
            std::vector<const double*> ptrs(static_cast<size_t>(nsl));
            std::vector<double> ls(static_cast<size_t>(nsl));
            for (int s = 0; s < nsl; s++) {
                const std::string fn = std::string(argv[6 + s]) + "_" + std::to_string(i) + ".bin";
                FILE* o = fopen(fn.c_str(), "rb");
                if (!o) {
                    fprintf(stderr, "missing %s\n", fn.c_str());
                    return 9;
                }
                int r2 = 0, one = 0;
                if (fread(&r2, 4, 1, o) != 1 || fread(&one, 4, 1, o) != 1 || fread(&ls[size_t(s)], 8, 1, o) != 1 ||
                    r2 != r) {
                    fprintf(stderr, "bad dump %s\n", fn.c_str());
                    fclose(o);
                    return 10;
                }
                const size_t nt = size_t(1) << r2;
                std::vector<float> tf(2 * nt);
                if (fread(tf.data(), sizeof(float), tf.size(), o) != tf.size()) {
                    fprintf(stderr, "bad dump %s\n", fn.c_str());
                    fclose(o);
                    return 10;
                }
                fclose(o);
                tabs[size_t(s)].assign(tf.begin(), tf.end());
                ptrs[size_t(s)] = tabs[size_t(s)].data();
            }
            double lpf = 0, lpp = 0;
            const int8_t* row = bits.data() + i * ns;
            if ((rc = gs_tt_resume(eng, row, K, nsl, ptrs.data(), ls.data(), seed, i, ob.data(), &lpf, &lpp,
                                   cond.data())))
                break;
            for (int k = 0; k < ns; k++) putchar('0' + ob[size_t(k)]);
            printf(" %.12f %.12f", lpf, lpp);
            for (int k = K; k < ns; k++) printf(" %.9f", cond[size_t(k)]);
            putchar('\n');
            fflush(stdout);
        }
    } else {
        fprintf(stderr, "mode must be sample, prob, chain, hybrid, prefix or resume\n");
        rc = 16;
    }
    gs_tt_engine_destroy(eng);
    gs_tt_plan_destroy(plan);
    return rc ? die(rc) : 0;
}
