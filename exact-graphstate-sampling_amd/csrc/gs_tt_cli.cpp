// gsamp_amd / dp_amd -- command-line front ends of the two-table engine with the reference's contracts:
//   gsamp_amd [--c128] [--device N] PLAN sample N SEED | prob STRINGS | chain STRINGS | hybrid STRINGS SEED
//                                                                        (gsamp.cpp:1-5, :108-161)
//   dp_amd    [--c64] [--device N] PLAN JOBS                             (dp.cpp:1-4, :26-62)
// The program's file name selects the contract (a name ending in "dp_amd" is the amplitude engine).  Output lines and
// exit codes are the reference's; gsamp's tables are complex64 by default (the reference's only storage), dp's
// complex128 (the reference's default; --c64 is its -DDP_FLOAT build).
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
    } else {
        fprintf(stderr, "mode must be sample, prob, chain or hybrid\n");
        rc = 16;
    }
    gs_tt_engine_destroy(eng);
    gs_tt_plan_destroy(plan);
    return rc ? die(rc) : 0;
}
