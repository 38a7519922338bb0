// Host-side in-place plan: parser of the plan file and the per-step structure the device kernels consume.
//
// Plan file (whitespace separated, as written by inplace.write_inplace_plan and read by gsamp2.cpp:126-139):
//   n ns R
//   per step: v nk m g rhoc nterms  Tkeep[0..nk)  nterms x (tq x tau par)  M00re M00im M01re M01im M10re ... (M[b][x])
// Error codes are the reference's exit codes: 2 header, 3 step head, 4 Tkeep, 5 term, 6 matrix, 7 nk > 40; this
// engine adds 15 for a step whose terms are not the planner's (coset offset, x) product (see StepInfo).
#pragma once
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace gs {

constexpr int kMaxKeep = 40;     // nk limit (gsamp2.cpp:135)
constexpr int kMaxOffsets = 4;   // at most 2^2 slots per coset in any planner output

struct Term {
    uint64_t tq;
    int x, tau, par;
};

// One step, as parsed, plus the structured view the kernels use.
struct StepInfo {
    int v = 0, nk = 0, m = 0;
    uint64_t g = 0, rhoc = 0;
    std::vector<uint64_t> cols;        // Tkeep: physical column of each kept coordinate
    std::vector<Term> terms;           // as in the file
    std::complex<double> M[2][2];      // M[b][x]
    // Structured view.  The planner enumerates Q x {0, 1}: for each distinct coset offset j (in file order) one
    // term with x = 0 and one with x = 1, both with the same par.
    int ntq = 0;                       // number of distinct offsets: 1, 2 or 4
    uint64_t off[kMaxOffsets] = {};    // offsets in first-appearance order
    int tau0[kMaxOffsets] = {}, tau1[kMaxOffsets] = {}, par[kMaxOffsets] = {};
    bool growth = false;               // q = 0, m = 2, x -> tau: both branch norms are |M_b0|^2+|M_b1|^2 times the table norm
    std::vector<uint64_t> sorted_offsets() const;
};

struct Plan {
    int n = 0, ns = 0, R = 0;
    std::vector<StepInfo> steps;
};

// Parse a plan file (path) or plan text.  Returns 0 or an error code; err receives a message.
int parse_plan_file(const char* path, Plan& p, std::string& err);
int parse_plan_text(const char* text, Plan& p, std::string& err);

// Physical address of coset c: XOR of cols[j] over the set bits j of c.
uint64_t coset_base(const StepInfo& s, uint64_t c);

}  // namespace gs
