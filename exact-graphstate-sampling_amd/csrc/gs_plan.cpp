// Plan parser (follows the record layout read at gsamp2.cpp:126-139) and structural validation.
#include "gs_plan.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <sstream>

namespace gs {

namespace {

// Minimal whitespace tokenizer over an in-memory buffer (fscanf-compatible for the plan's number formats).
class Tokens {
  public:
    explicit Tokens(const char* text) : p_(text) {}
    bool next_int(long long& out) {
        skip();
        if (!*p_) return false;
        char* end = nullptr;
        out = strtoll(p_, &end, 10);
        if (end == p_) return false;
        p_ = end;
        return true;
    }
    bool next_u64(unsigned long long& out) {
        skip();
        if (!*p_ || *p_ == '-') return false;
        char* end = nullptr;
        out = strtoull(p_, &end, 10);
        if (end == p_) return false;
        p_ = end;
        return true;
    }
    bool next_double(double& out) {
        skip();
        if (!*p_) return false;
        char* end = nullptr;
        out = strtod(p_, &end);
        if (end == p_) return false;
        p_ = end;
        return true;
    }

  private:
    void skip() {
        while (*p_ == ' ' || *p_ == '\n' || *p_ == '\t' || *p_ == '\r') p_++;
    }
    const char* p_;
};

int structure(StepInfo& s, std::string& err) {
    // distinct offsets in first-appearance order
    s.ntq = 0;
    for (const Term& t : s.terms) {
        int j = 0;
        while (j < s.ntq && s.off[j] != t.tq) j++;
        if (j == s.ntq) {
            if (s.ntq == kMaxOffsets) { err = "more than 4 coset offsets in one step"; return 15; }
            s.off[s.ntq++] = t.tq;
        }
    }
    if (!(s.ntq == 1 || s.ntq == 2 || s.ntq == 4) || (int)s.terms.size() != 2 * s.ntq) {
        err = "step terms are not a (coset offset, x) product";
        return 15;
    }
    bool seen0[kMaxOffsets] = {}, seen1[kMaxOffsets] = {};
    int parv[kMaxOffsets];
    for (int j = 0; j < kMaxOffsets; j++) parv[j] = -1;
    for (const Term& t : s.terms) {
        int j = 0;
        while (s.off[j] != t.tq) j++;
        if (t.x < 0 || t.x > 1 || t.tau < 0 || t.tau >= s.m || t.par < 0 || t.par > 1) {
            err = "term field out of range";
            return 15;
        }
        bool& seen = t.x ? seen1[j] : seen0[j];
        if (seen) { err = "duplicate (offset, x) term"; return 15; }
        seen = true;
        (t.x ? s.tau1[j] : s.tau0[j]) = t.tau;
        // par only matters with x = 1; the planner writes the same value for both x
        if (t.x) s.par[j] = t.par;
        parv[j] = t.par;
    }
    for (int j = 0; j < s.ntq; j++)
        if (!seen0[j] || !seen1[j]) { err = "missing (offset, x) term"; return 15; }
    s.growth = s.ntq == 1 && s.m == 2 && s.tau0[0] == 0 && s.tau1[0] == 1;
    return 0;
}

int parse(Tokens& tk, Plan& p, std::string& err) {
    long long n, ns, R;
    if (!tk.next_int(n) || !tk.next_int(ns) || !tk.next_int(R)) { err = "bad plan header"; return 2; }
    if (ns < 0 || R < 0 || R > 62) { err = "plan header out of range"; return 2; }
    p.n = int(n);
    p.ns = int(ns);
    p.R = int(R);
    p.steps.assign(size_t(ns), StepInfo{});
    for (long long k = 0; k < ns; k++) {
        StepInfo& s = p.steps[size_t(k)];
        long long v, nk, m, nt;
        unsigned long long g, rhoc;
        if (!tk.next_int(v) || !tk.next_int(nk) || !tk.next_int(m) || !tk.next_u64(g) || !tk.next_u64(rhoc) || !tk.next_int(nt)) {
            err = "bad step header at step " + std::to_string(k);
            return 3;
        }
        if (nk < 0 || nt < 0 || nt > 64 || nk > 1000) { err = "bad step header at step " + std::to_string(k); return 3; }
        s.v = int(v);
        s.nk = int(nk);
        s.m = int(m);
        s.g = g;
        s.rhoc = rhoc;
        s.cols.resize(size_t(nk));
        for (auto& c : s.cols) {
            unsigned long long y;
            if (!tk.next_u64(y)) { err = "bad Tkeep at step " + std::to_string(k); return 4; }
            c = y;
        }
        s.terms.resize(size_t(nt));
        for (auto& t : s.terms) {
            unsigned long long y;
            long long x, tau, par;
            if (!tk.next_u64(y) || !tk.next_int(x) || !tk.next_int(tau) || !tk.next_int(par)) {
                err = "bad term at step " + std::to_string(k);
                return 5;
            }
            t.tq = y;
            t.x = int(x);
            t.tau = int(tau);
            t.par = int(par);
        }
        for (int b = 0; b < 2; b++)
            for (int x = 0; x < 2; x++) {
                double re, im;
                if (!tk.next_double(re) || !tk.next_double(im)) { err = "bad matrix at step " + std::to_string(k); return 6; }
                s.M[b][x] = std::complex<double>(re, im);
            }
        if (s.nk > kMaxKeep) { err = "nk > 40 at step " + std::to_string(k); return 7; }
    }
    // Checks the kernels rely on (the reference trusts its planner; a bad plan here must not address outside the
    // arena): addresses below 2^R, m in {1, 2}, rhoc within the kept coordinates.
    const uint64_t lim = uint64_t(1) << p.R;
    for (int k = 0; k < p.ns; k++) {
        StepInfo& s = p.steps[size_t(k)];
        if (s.m != 1 && s.m != 2) { err = "m must be 1 or 2 at step " + std::to_string(k); return 15; }
        if (s.m == 1 && s.g != 0) { err = "g must be 0 when m = 1 at step " + std::to_string(k); return 15; }
        if (s.nk < 64 && (s.rhoc >> s.nk)) { err = "rhoc outside the kept coordinates at step " + std::to_string(k); return 15; }
        uint64_t span = s.g;
        for (uint64_t c : s.cols) span |= c;
        for (const Term& t : s.terms) span |= t.tq;
        if (span >= lim) { err = "address beyond 2^R at step " + std::to_string(k); return 15; }
        if (int rc = structure(s, err)) {
            err += " at step " + std::to_string(k);
            return rc;
        }
    }
    return 0;
}

}  // namespace

std::vector<uint64_t> StepInfo::sorted_offsets() const {
    std::vector<uint64_t> v(off, off + ntq);
    std::sort(v.begin(), v.end());
    return v;
}

int parse_plan_text(const char* text, Plan& p, std::string& err) {
    Tokens tk(text);
    return parse(tk, p, err);
}

int parse_plan_file(const char* path, Plan& p, std::string& err) {
    FILE* f = fopen(path, "rb");
    if (!f) { err = std::string("cannot open plan ") + path; return 2; }
    std::string buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) buf.append(chunk, got);
    fclose(f);
    return parse_plan_text(buf.c_str(), p, err);
}

uint64_t coset_base(const StepInfo& s, uint64_t c) {
    uint64_t a = 0;
    for (int j = 0; c; j++, c >>= 1)
        if (c & 1) a ^= s.cols[size_t(j)];
    return a;
}

}  // namespace gs
