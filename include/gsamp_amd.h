/* gsamp_amd -- C ABI of the MI355X in-place cut-rank engine (libgsamp_amd.so).
 *
 * The reference engine has no library interface: it is the process contract of gsamp2.cpp, i.e.
 *   gsamp2 <plan> sample|sample2 N SEED | prob STRINGS | chain STRINGS | hybrid STRINGS SEED
 *          | prefix STRINGS K OUT | resume STRINGS K SEED DUMP... [--twist FILE]        (gsamp2.cpp:1-6, :149-225)
 * Each entry point below replaces one of those modes; the comment of each names the reference lines whose
 * semantics it keeps.  The bundled CLI (bin/gsamp2_amd) is a byte-compatible front end over this ABI.
 *
 * Conventions
 *   - Bit strings are int8 arrays in STEP order (character k of a reference strings-file line), row-major
 *     [nstr][ns]; values 0, 1, and -1 for '?' where a mode samples.
 *   - All pointers are host pointers owned by the caller; every call is synchronous.
 *   - Return value: 0 on success, else the reference's exit code for the same failure (2..7 plan parse,
 *     8 '?' before K, 9/10 missing/bad dump, 11 allocation, 12..14 twist) or 15 (plan step structure not
 *     supported), 16 (bad argument), 17 (HIP runtime error).  gs_last_error() gives the message of the last
 *     failure on the calling thread.
 *   - A zero-probability branch gives log2 Pr = -inf and zero conditionals from that step on, as the reference.
 *   - One engine owns one device arena of 2^R entries and one HIP stream; engines are not re-entrant, separate
 *     host threads may drive separate engines.
 */
#ifndef GSAMP_AMD_H
#define GSAMP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gs_plan gs_plan;
typedef struct gs_engine gs_engine;

enum { GS_C64 = 0, GS_C128 = 1 };          /* arena storage: complex<float> / complex<double> (-DGS_DOUBLE) */
enum { GS_SEED_PLAIN = 0, GS_SEED_SEQ = 1 };/* sample: mt19937_64(seed+i) / sample2: seed_seq{seed,i,0x9E3779B9} */

/* Parse an in-place plan file (gsamp2.cpp:126-139; exit codes 2..7) or the same text from memory. */
int gs_plan_load(const char* path, gs_plan** out);
int gs_plan_parse(const char* text, gs_plan** out);
/* Plan header: qubits n, steps ns, arena bits R. */
int gs_plan_info(const gs_plan* plan, int* n, int* ns, int* R);
void gs_plan_destroy(gs_plan* plan);

/* Allocate the 2^R arena on a HIP device and upload the plan (gsamp2.cpp:140-147; 11 when it does not fit). */
int gs_engine_create(const gs_plan* plan, int storage, int device, gs_engine** out);
void gs_engine_destroy(gs_engine* eng);

/* prob mode (gsamp2.cpp:217-224 with chain() at :39-55): log2 Pr of each forced string, two passes per step,
 * table renormalised every step.  bits: [nstr][ns] in {0,1}. */
int gs_prob(gs_engine* eng, const int8_t* bits, int64_t nstr, double* log2p);

/* chain mode (gsamp2.cpp:56-70, :218-222): one pass per step; log2p[i], cond[i][ns] = Pr(b_k | b_<k).
 * amp_reim (may be NULL) receives [nstr][2]: Z(w_b) = sum_x prod_j M_j[b_j,x_j] (-1)^{x^T A x}, the last live
 * arena entry, so that log2 Pr = log2|Z|^2 - ns (up to the global phase the graph form drops); 0 for a string that
 * meets a zero-probability branch. */
int gs_chain(gs_engine* eng, const int8_t* bits, int64_t nstr, double* log2p, double* cond, double* amp_reim);

/* sample / sample2 modes (gsamp2.cpp:149-161): samples first..first+count-1, each from its own generator
 * (seeding GS_SEED_PLAIN: mt19937_64(seed + i); GS_SEED_SEQ: std::seed_seq as gsamp2.cpp:159).
 * bits: [count][ns] output; log2p: [count]. */
int gs_sample(gs_engine* eng, uint64_t seed, int64_t first, int64_t count, int seeding, int8_t* bits, double* log2p);

/* hybrid mode (gsamp2.cpp:99-120, :209-216): bits [nstr][ns] with -1 = '?' (sampled from the exact
 * conditional); line i uses mt19937_64(seed + first_line + i).  out_bits [nstr][ns], log2p [nstr],
 * cond [nstr][ns]. */
int gs_hybrid(gs_engine* eng, const int8_t* bits, int64_t nstr, uint64_t seed, int64_t first_line,
              int8_t* out_bits, double* log2p, double* cond);

/* prefix mode (gsamp2.cpp:162-175): run forced steps 0..K-1 of one string (bits [ns], entries < K used) and
 * return the live table of step K in logical order (c-major, sorted coset offsets of step K; {0} past the end)
 * as complex doubles table_reim [2 * 2^nk * ntq] (capacity table_cap complex entries; the reference's dump
 * stores the arena's type, the caller narrows), plus nk, ntq, logscale (true = stored * 2^logscale) and
 * log2 Pr(prefix).  Pass table_reim = NULL to query nk and ntq only (no steps are run). */
int gs_prefix(gs_engine* eng, const int8_t* bits, int K, double* table_reim, int64_t table_cap, int* nk, int* ntq,
              double* logscale, double* log2p_prefix);

/* resume mode (gsamp2.cpp:176-208) for one line: sum nslices dumped step-K tables (each [2^nk * ntq] complex
 * doubles in logical order, with logscales ls[s]) with weights 2^(ls_s - max ls), optionally translated by
 * the twist (cshift[s], tqshift[s]; NULL = none), scatter into the arena, then run steps K..ns-1 with '?'
 * sampled from mt19937_64(seed + line).  out_bits [ns] (prefix copied from bits), log2p_full, log2p_prefix,
 * cond [ns] (entries K..ns-1 written). */
int gs_resume(gs_engine* eng, const int8_t* bits, int K, int nslices, const double* const* tables,
              const double* logscales, const uint64_t* cshift, const uint64_t* tqshift, uint64_t seed, int64_t line,
              int8_t* out_bits, double* log2p_full, double* log2p_prefix, double* cond);

/* ---- batched scoring (no counterpart in the reference: chain mode, gsamp2.cpp:56-70, for many strings and many
 * plans at once).  Instance i holds lanes[i] strings (a multiple of 64) in an arena of 2^R_i x lanes[i] complex64
 * entries, slot-major and batch-minor; one launch per step index covers every instance. */
typedef struct gs_batch gs_batch;
int gs_batch_create(const gs_plan* const* plans, const int* lanes, int nplans, int storage, int device, gs_batch** out);
void gs_batch_destroy(gs_batch* batch);
/* bits[i]: instance i's [lanes_i][ns_i] forced strings; log2p[i] [lanes_i], cond[i] [lanes_i][ns_i] and
 * amp[i] [lanes_i][2] (cond, amp or their entries may be NULL) receive chain-mode outputs. */
int gs_batch_score(gs_batch* batch, const int8_t* const* bits, double* const* log2p, double* const* cond,
                   double* const* amp);
/* Replay the step sequence as one captured hipGraph (on) or launch it kernel by kernel (off, default). */
int gs_batch_use_graph(gs_batch* batch, int on);
/* log2 cosets per work item of instance i at step k (k < 0: every step); not below the creation default. */
int gs_batch_set_tile(gs_batch* batch, int i, int k, int lg);
/* Work-list order: 0 round-robin over the XCDs, 1 (default) each XCD takes consecutive items (lane group by lane
   group), 2 each XCD takes consecutive coset ranges with all their lane groups. */
int gs_batch_set_xcd_items(gs_batch* batch, int on);
/* log2 cosets each lane gathers at once (0..3; default 2 for complex64, 1 for complex128). */
int gs_batch_set_unroll(gs_batch* batch, int lu);
/* Device time of the last gs_batch_score (ms); step_end_ms (may be NULL) [max ns]: end of each step after the start. */
double gs_batch_last_ms(const gs_batch* batch, double* step_end_ms);
/* Replace instance i's measurement matrices M[ns][2][2] (complex as re, im pairs; M[k][b][x]). */
int gs_batch_set_matrices(gs_batch* batch, int i, const double* M);
/* Adjoint (no counterpart in the reference): G[i] [lanes_i][ns_i][2] complex, G_k(x) = dZ/dM_k[b_k, x] for every
 * string, with checkpoint-and-recompute of the in-place table every seg steps (seg 0: the segment length that
 * recomputes least while its copies fit in mem_budget bytes; mem_budget <= 0: 70 % of the free device memory); also
 * log2p[i] and amp[i] (may be NULL).  *seg_used receives the segment. */
int gs_batch_grad(gs_batch* batch, const int8_t* const* bits, double* const* log2p, double* const* amp,
                  double* const* G, int seg, double mem_budget, int* seg_used);
/* Device time (ms) of the last gs_batch_grad; *forward_ms (may be NULL) its first forward pass. */
double gs_batch_grad_ms(const gs_batch* batch, double* forward_ms);
const char* gs_batch_last_error(void);

/* ---- two-table (gather-form) engine: the reference's gsamp.cpp sampler and dp.cpp amplitude engine.  Two tables of
 * 2^maxr entries; step k's table is read (gathered) and the next one written (contiguous), gsamp.cpp:21-29. */
typedef struct gs_tt_plan gs_tt_plan;
typedef struct gs_tt_engine gs_tt_engine;
/* gsamp plan (with_m = 1: per step `v r rn rho pivmask nker ker... M`, gsamp.cpp:96-107) or dp plan (with_m = 0,
 * dp.cpp:16-25).  Exit codes: 2 header, 3 step head, 4 kernel vector, 5 M, 15 a step outside what the engine runs. */
int gs_tt_plan_load(const char* path, int with_m, gs_tt_plan** out);
int gs_tt_plan_info(const gs_tt_plan* plan, int* n, int* ns, int* maxr);
void gs_tt_plan_destroy(gs_tt_plan* plan);
int gs_tt_engine_create(const gs_tt_plan* plan, int storage, int device, gs_tt_engine** out);
void gs_tt_engine_destroy(gs_tt_engine* eng);
/* prob (gsamp.cpp:39-55 forced, both branch norms per step), chain (chain1, gsamp.cpp:59-70; amp_reim may be NULL),
 * sample (mt19937_64(seed + first + i), gsamp.cpp:111-117), hybrid ('?' = -1 sampled, gsamp.cpp:118-123). */
int gs_tt_prob(gs_tt_engine* eng, const int8_t* bits, int64_t nstr, double* log2p);
int gs_tt_chain(gs_tt_engine* eng, const int8_t* bits, int64_t nstr, double* log2p, double* cond, double* amp_reim);
int gs_tt_sample(gs_tt_engine* eng, uint64_t seed, int64_t first, int64_t count, int8_t* bits, double* log2p);
int gs_tt_hybrid(gs_tt_engine* eng, const int8_t* bits, int64_t nstr, uint64_t seed, int64_t first_line,
                 int8_t* out_bits, double* log2p, double* cond);
/* prefix (gsamp.cpp:124-135): forced steps 0..K-1, then the contiguous table of step K (2^r complex doubles in
 * table_reim, capacity table_cap entries; r = st[K].r, 0 when K = ns) with its logscale and log2 Pr(prefix);
 * table_reim = NULL only sets *r.  resume (gsamp.cpp:136-148): the nslices tables summed with weights
 * 2^(ls_s - max ls), then steps K..ns-1 with '?' sampled from mt19937_64(seed + line). */
int gs_tt_prefix(gs_tt_engine* eng, const int8_t* bits, int K, double* table_reim, int64_t table_cap, int* r,
                 double* logscale, double* log2p_prefix);
int gs_tt_resume(gs_tt_engine* eng, const int8_t* bits, int K, int nslices, const double* const* tables,
                 const double* logscales, uint64_t seed, int64_t line, int8_t* out_bits, double* log2p_full,
                 double* log2p_prefix, double* cond);
/* dp (dp.cpp:33-62): out[j] = sum over job j's nslices[j] slices of pre * Z(w_slice); pre [total slices][2],
 * w [total slices][m][4] (w0re w0im w1re w1im per vertex), out [njobs][2]. */
int gs_tt_dp(gs_tt_engine* eng, int64_t njobs, const int* nslices, const double* pre, const double* w, double* out);
const char* gs_tt_last_error(void);
/* Cap every gather sweep at 2^lg blocks (1..14; 0 restores the default 14); tiles beyond the cap loop inside
   each block.  No counterpart in the reference (a tuning knob of this engine). */
int gs_tt_set_grid_log2(gs_tt_engine* eng, int lg);

/* prob mode without any device allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same inputs, outputs, results and kernels as gs_prob on an engine of the same plan and storage, but the arena,
 * step descriptors, chain state and staging buffers live in one file-scope device array of 64 MiB on each device,
 * held until the process ends; plans that need more (an arena of 2^R entries with R > 22 in complex64, R > 21 in
 * complex128) are rejected with code 16.  Runs on the null stream of `device` and returns when it is done.  Every
 * call on a device shares that device's array: calls that overlap on one device race. */
int gs_prob_noalloc(const gs_plan* plan, int storage, int device, const int8_t* bits, int64_t nstr, double* log2p);

/* Timing of the device work of the last call, in milliseconds (HIP events around the step kernels). */
double gs_last_device_ms(const gs_engine* eng);

/* ---- measurement and tuning (no counterpart in the reference; results do not depend on them). */
/* Per-sweep timing: enable/disable, then read ms[ns][7] and launches[ns][7] accumulated since enabling
 * (sweeps: 0 Gram norms, 1 direct norms, 2 write, 3 write + next Gram, 4 write + next direct, 5 pattern sums and
 * 6 product pass of a run of diagonal steps, counted at the run's first step). */
int gs_profile_enable(gs_engine* eng, int on);
int gs_profile_read(const gs_engine* eng, double* ms, int64_t* launches);
/* Launch geometry of step k, out[12]: log2 cosets per run, blocks, slots per coset, route, fusable, blocks of the
 * fused sweep, next step's route, log2 tiles per block, log2 run of the fused layout, its log2 tiles per block
 * (those two -1 when the step does not fuse), runs per thread of the plain sweeps (1 or 2), runs per thread of the
 * fused sweep (2: next step has two slots; 1: the plain layout folding a one-slot next step's Gram; -1: no fusion). */
int gs_step_geometry(const gs_engine* eng, int k, int* out);
/* Coset-coordinate columns (physical address vectors) of step k as the plain (fused = 0) or fused sweep uses them,
 * into cols[0 .. cap-1]; returns nk, or -1 when the step or layout does not exist. */
int gs_step_columns(const gs_engine* eng, int k, int fused, uint64_t* cols, int cap);
/* Fold the next step's norm pass into the previous write (on, default) or run it as its own sweep (off). */
int gs_set_fusion(gs_engine* eng, int on);
/* Fold the norms of a four-slot next step into the previous write (two pairing coordinates per thread) on or off
   (default on; off: such a step takes its own norm sweep). */
int gs_set_fused4(gs_engine* eng, int on);
/* Cap one sweep at 2^lg blocks (1..14; default 9, plain sweeps of one-slot steps 12, or 14 from 2^30 cosets on); tiles
   beyond the cap loop inside each block.  Sets the one-slot caps to lg as well; lg = 0 restores all three defaults. */
int gs_set_grid_log2(gs_engine* eng, int lg);
/* Cap the plain sweeps of one-slot steps alone at 2^lg blocks (1..14; default 12, or 14 from 2^30 cosets on); lg = 0
   restores those two defaults. */
int gs_set_grid_log2_one_slot(gs_engine* eng, int lg);
/* log2 run length of the fused layout (complex64: 1 default or 2; complex128: 0; -1 restores the default). */
int gs_set_fused_run(gs_engine* eng, int lr);
/* Tile order of the sweeps: 0 forward (default), 1 reverse on odd steps, 2 reverse on every step. */
int gs_set_walk(gs_engine* eng, int mode);
/* Each XCD takes a contiguous eighth of every sweep's block coordinates (on = 1, default) or blocks go round-robin
   (0). */
int gs_set_xcd_blocks(gs_engine* eng, int on);
/* Plain sweeps of steps with at most ntq_max slots per coset (0..2; default 0, off) take two runs per thread, the
   second at the top coordinate (far = 1) or at the next unit bit (far = 0). */
int gs_set_plain_pairs(gs_engine* eng, int ntq_max, int far);
/* Runs of at most max_steps (2..8; 0: off; default 8) consecutive diagonal steps -- one slot, one output, slot offset
   0, the same live table -- of at least 2^min_nk cosets (10..40; default 20): one pass sums |v|^2 per pattern of the
   run's signs, one block makes the run's decisions, one pass multiplies every entry by its pattern's factor. */
int gs_set_diag_runs(gs_engine* eng, int max_steps, int min_nk);
/* The run of diagonal steps holding step k, out[5]: first step, steps, log2 blocks, log2 tiles per block, and 1 when
   its product pass can also sum the next (two-slot) step's Gram matrix (all -1: none). */
int gs_step_run(const gs_engine* eng, int k, int* out);
/* A sweep whose blocks leave their partial sums unreduced has them reduced by every block of the next sweep, which
   then needs no separate reduction launch (on = 1; on >= 2: only when the next sweep's step has at most 2^on live
   slots), or by a one-block kernel between the two (0, default); the results are bitwise the same. */
int gs_set_fold_reduce(gs_engine* eng, int on);
/* Runs of two or more steps with at most 2^nk cosets go to one single-block launch (nk -1: never; default 10). */
int gs_set_small_steps(gs_engine* eng, int nk);
/* Device race check of step k (debug mode; overwrites the arena, used as an owner map of the slots): layout 0 the
   plain sweeps' coset coordinates, 1 the fused sweep's, 2 the plan's own, 3 the run of diagonal steps starting at k.
   counts[4]: slots read by two cosets, slots written twice (for a run: visited twice), writes onto a slot another
   coset reads (for a run: slots outside the step's table), slots checked -- the reference's in-place asserts
   (inplace.py:126-137) on the layouts the kernels actually walk. */
int gs_race_check(gs_engine* eng, int k, int layout, int64_t* counts);
/* One plain device copy of `bytes` (read + write), *ms its duration: calibrates hardware byte counters. */
int gs_copy_probe(int device, int64_t bytes, double* ms);

/* Message of the last failure on the calling thread ("" if none). */
const char* gs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GSAMP_AMD_H */
