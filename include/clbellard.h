/*
 * clbellard.h -- C ABI of the MI355X-native Bellard-formula hex digit extraction library
 * (libclbellard.so).  Plain C types only; no exceptions cross this boundary.
 *
 * Every entry point returns CLB_OK (0) or a negative CLB_E* code; clb_last_error() gives the
 * message of the calling thread's last failure.  Outputs are caller-owned.  A clb_ctx owns the
 * device buffers and the HIP stream of one GPU; calls on one context must not overlap, calls on
 * different contexts may run from different threads.
 *
 * 192-bit values are uint64_t[3] in the reference's bigSum order: w[0] is the most significant
 * word (Bellard.cl:1082-1084 "ulsum0"), the value is read as the binary fraction w / 2^192.
 *
 * Reference surface replaced by this header (toropippi/CLBellard):
 *   Bellard_192.hsp      the host app: normalisation, normal mode, work mode, fsave
 *   KernelFuncs.as       BatchCalc / AddAns / Set0mem* launch orchestration
 *   HSPCL64.as + .dll    the HSP->OpenCL plugin calls on this path (HCLinit, HCLSetDev,
 *                        HCLCreateBuffer, HCLSetKrns, HCLDoKrn1, HCLReadBuffer, ...)
 *   Bellard.cl,
 *   Bellard_div2by1.cl   the device kernels (Sglobal*_192*, Sglobal_after_192, Ssum_192, Sum7_192)
 *   hexsademical.as      hexadecimal_i32 / hexadecimal_dd output formatting
 *   workmodeSum0_6.hsp   the work-file summer
 */
#ifndef CLBELLARD_H
#define CLBELLARD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLB_OK 0
#define CLB_EINVAL -1  /* bad argument (digits out of range, bad formula, null pointer ...) */
#define CLB_EDEVICE -2 /* HIP runtime / device failure */
#define CLB_ERANGE -3  /* a modulus is beyond what the selected kernel supports */
#define CLB_EIO -4     /* file could not be read or written */
#define CLB_ENOMEM -5  /* host or device allocation failed */
#define CLB_EFORMAT -6 /* malformed worklist or work file */

/* Precision modes of the reference GUI, Bellard_192.hsp:64 ("106bit", "192bit", "192bit(beta)").
 * All three evaluate every term exactly as floor(2^192 frac(+-2^(10(d-k))/m)), the per-term rule
 * of the reference's all-integer kernel (Bellard_div2by1.cl:327-347); they differ only in how the
 * result is written: CLB_PREC_DD106 emits the 28-hex-digit double-double layout of
 * hexadecimal_dd, the two 192-bit modes emit all 48 digits.  With clb_set_replay they compute what
 * the reference's kernels of that precision compute instead, rounding included. */
#define CLB_PREC_DD106 0
#define CLB_PREC_U192 1
#define CLB_PREC_U192_BETA 2

/* Term kernels.  Both produce bit-identical sums. */
#define CLB_KERNEL_AUTO 0  /* SPLIT where every modulus allows it, else PLAIN */
#define CLB_KERNEL_PLAIN 1 /* all-integer: 64-bit Montgomery on 32-bit limb products */
#define CLB_KERNEL_SPLIT 2 /* FP64 error-free modexp + FP64-reciprocal quotient digits */

typedef struct clb_ctx clb_ctx;

/* One term-range job: head terms k in [k_begin, min(k_end, d)) of formula j at loop count d,
 * plus the tail terms k = d..d+19 iff the range reaches d (KernelFuncs.as:56-59). */
typedef struct {
  uint64_t d;
  uint64_t k_begin;
  uint64_t k_end;
  int32_t formula;
  int32_t reserved;
} clb_job;

/* Message of this thread's last failed call ("" if none). */
const char *clb_last_error(void);
/* Library version string. */
const char *clb_version(void);

/* HCLinit + HCLGetDevCount (HSPCL64.as:11, :29; Bellard_192.hsp:16-17). */
int clb_get_dev_count(int *count);
/* HCLGetDeviceInfo(CL_DEVICE_NAME) (HSPCL64.as:21; Bellard_192.hsp:36-44, :528). */
int clb_get_device_name(int device, char *name, size_t cap);

/* HCLSetDev + LoadOpenCLFile + GlobalMemSet (Bellard_192.hsp:250-255, :376-420): selects the GPU,
 * allocates the partial-sum and answer buffers, zeroes the answer. */
int clb_ctx_create(int device, int precision, clb_ctx **out);
int clb_ctx_destroy(clb_ctx *ctx);
/* Kernel choice (CLB_KERNEL_*) for the following calls; default CLB_KERNEL_AUTO. */
int clb_set_kernel(clb_ctx *ctx, int kernel);
/* Launch shape: persistent workgroups per compute unit (0 = library default). */
int clb_set_launch(clb_ctx *ctx, int blocks_per_cu);
/* Tile length (no reference counterpart; the reference's work-items take GWS-strided terms,
 * KernelFuncs.as:38-52): terms per lane per tile, 16, 4 or 1; or 0 (default) to choose per
 * launch the longest whose tiles give every resident wave at least min_tiles_per_wave tiles
 * (0, the default: one tile to every 8th wave, so 16 except for the smallest launches).  Results
 * are identical for every choice. */
int clb_set_tiling(clb_ctx *ctx, int terms_per_lane, int min_tiles_per_wave);
/* Tile schedule of the term kernel (no reference counterpart; the reference's launches have a
 * fixed shape, 2^24 work-items in workgroups of 256, Bellard_192.hsp:5-7).  waves_per_group: the
 * workgroup size in waves, 1, 2, 4 or 8 (0: 4, i.e. 256 work-items); min_chunk: tiles per grab of
 * the tile queue, at most (0: 256); chunks_per_wave: > 0 for a guided queue whose chunks halve from
 * phase to phase, that many per wave per phase, down to min_chunk (0 or -1: fixed chunks, the
 * default); max_replicas: the most copies of an accumulator slot, a power of 2 (0: 64).  Results
 * are identical for every choice. */
int clb_set_schedule(clb_ctx *ctx, int waves_per_group, int min_chunk, int chunks_per_wave,
                     int max_replicas);
/* Autotuner (no reference counterpart): times the tile lengths 16, 4, 1 against 1, 3/4 and 1/2
 * of the resident workgroups per CU on a sample of the normal-mode computation at position n
 * (the seven formulas' head terms around d/2, sized to spend about budget_ms; best of 3 runs
 * each), then, one at a time on that choice, the workgroup size (64, 128, 512 work-items at the
 * same waves per CU), the tile queue (fixed chunks of 64 or 1024 tiles, a guided queue) and at
 * most 4 slot copies; keeps the fastest on the context (as clb_set_tiling / clb_set_launch /
 * clb_set_schedule would; a candidate must beat the best so far, starting from the defaults, by
 * 2%) and returns the tile length and workgroups per CU (outputs may be NULL; clb_get_schedule
 * reads the rest).
 * Every candidate's sum is compared; a difference returns CLB_EDEVICE.  The exact kernels are
 * timed even on a context in replay mode (clb_set_replay), whose setting is kept. */
int clb_autotune(clb_ctx *ctx, uint64_t n, double budget_ms, int *terms_per_lane,
                 int *blocks_per_cu);
/* The context's launch settings (no reference counterpart): terms per lane (0: chosen per
 * launch), workgroups per CU (0: full residency), then clb_set_schedule's four values.  Outputs
 * may be NULL. */
int clb_get_schedule(clb_ctx *ctx, int *terms_per_lane, int *blocks_per_cu, int *waves_per_group,
                     int *min_chunk, int *chunks_per_wave, int *max_replicas);
/* A context's launch settings as a file, so that a later run reuses a tuning (no reference
 * counterpart: the reference's launch shape is compile-time constants, Bellard_192.hsp:5-7).
 * Text, one key=value per line, in any order, each key once: device (clb_get_device_name's
 * string), terms_per_lane, min_tiles_per_wave (clb_set_tiling), blocks_per_cu (clb_set_launch),
 * waves_per_group, min_chunk, chunks_per_wave, max_replicas (clb_set_schedule); integers in
 * decimal.  Blank lines are skipped.
 * clb_read_schedule / clb_write_schedule are host only: CLB_EIO if the file cannot be read or
 * written, CLB_EFORMAT for an unknown, repeated or missing key, a value that is not an integer
 * or a file above 4096 bytes, CLB_EINVAL for a value outside its setter's range or a device name
 * that is empty, too long or holds a control character.
 * clb_save_schedule writes the context's settings with its device's name; clb_load_schedule
 * reads a file and applies it as the setters would, or returns CLB_EINVAL when the file names
 * another device (nothing is applied on any error). */
typedef struct {
  char device[256];
  int terms_per_lane, min_tiles_per_wave, blocks_per_cu;
  int waves_per_group, min_chunk, chunks_per_wave, max_replicas;
} clb_schedule;
int clb_read_schedule(const char *path, clb_schedule *out);
int clb_write_schedule(const char *path, const clb_schedule *s);
int clb_save_schedule(clb_ctx *ctx, const char *path);
int clb_load_schedule(clb_ctx *ctx, const char *path);
/* Launch mode (no reference counterpart; the reference enqueues every kernel from the HSP host,
 * KernelFuncs.as:25-74): 1 = run each call's kernel sequence (head and tail terms, combine) as a
 * HIP graph, captured on the first call with a given buffer layout and kernel choice and replayed
 * by later calls (the call's host-to-device copy, which also clears the accumulators, stays
 * outside the graph); 0 (default) = launch the kernels one by one.  Results are identical. */
int clb_set_graph(clb_ctx *ctx, int enable);
/* Reference replay.  By default every term is evaluated exactly (floor(r' 2^192 / m), the rule of
 * the reference's Sglobal64mtg_192_Refresh, Bellard_div2by1.cl:313-359), so results do not depend
 * on launch shapes.  global_work_size > 0 instead reproduces, bit for bit, the results of the
 * reference's own kernels launched with that many work-items, their double-double rounding
 * included.  The reference launches CLB_REF_GLOBAL_WORK_SIZE work-items (Bellard_192.hsp:5-7);
 * 0 turns replay off.
 *   Precision 1 and 2: Sglobal32mtg_192 (Bellard.cl:771-823), Sglobal64mtg_192 (:992-1084,
 *   precision 1) and Sglobal_after_192 (:1153-1253), Refresh kept where the reference uses it.
 *   Precision 0 (the 106-bit mode): Sglobal32mtg / Sglobal64mtg (:729-766, :914-985) with their
 *   256-lane trees and the 65536 per-group accumulators, Sglobal_after (:1118-1150), Ssum and
 *   Sum7 (:1274-1356).  Every result (frac, partials, dev_ans, work files) is then a
 *   double-double, returned in the 3-word outputs as the bit patterns of hi (word 0) and lo
 *   (word 1) with word 2 = 0, and printed with hexadecimal_dd.  The operators act on the
 *   context's own bigSumMem of 65536 double-doubles (cleared by clb_set0mem), as the reference's
 *   kernels do: clb_sglobal is one launch of the given global work size over [offset, k_max)
 *   (Sglobal32mtg when m(offset) < 2^32, else Sglobal64mtg; CLB_EINVAL for a range that crosses
 *   loopmax32, which the reference's 32-bit kernel would truncate), clb_sglobal_after adds the
 *   tail into bigSum[0], clb_ssum runs Ssum's two passes (bigSum[0] = the total), clb_batch_calc
 *   all three in BatchCalc's order, and clb_add_ans Sum7; the partial (clb_read_partial) is
 *   bigSum[0] after each call.
 *   Loop counts with 10 d + 9 >= 2^53 replay the reference's -D OVERDOUBLE build
 *   (Bellard_192.hsp:472-474): from m = 2^53 on, numerator and modulus split into double-doubles
 *   and divided with dd_div (Bellard.cl:937-962, :1017-1046); moduli must stay below 2^63
 *   (CLB_ERANGE otherwise).
 * CLB_EINVAL for a size that is not a multiple of 256 up to 2^24.  Replayed calls never use graph
 * mode. */
#define CLB_REF_GLOBAL_WORK_SIZE 16777216ull
int clb_set_replay(clb_ctx *ctx, uint64_t global_work_size);

/* BuildOptions core (Bellard_192.hsp:449-450, :468-474): N -> odd loop count d and N' (a multiple
 * of 5, the 1-based hex position of the first output digit); overdouble = (10d + 9 >= 2^53).
 * CLB_EINVAL for N <= 4 or N > 999999999999999999. */
int clb_normalize(uint64_t n_in, uint64_t *n_out, uint64_t *d_out, int *overdouble);
/* numer_denom_init (KernelFuncs.as:3-13): c_j, sign flag (1 = minus), den0, den1 for j = 0..6. */
int clb_formula(int j, uint64_t *numer, int64_t *numesign, uint64_t *den0, uint64_t *den1);

/* Set0memUlong3 / Set0memdd (KernelFuncs.as:15-23): clear the formula partial (bigSumMem). */
int clb_set0mem(clb_ctx *ctx);
/* BatchCalc k_start,k_end,formulano (KernelFuncs.as:25-66): adds formula j's terms of
 * [k_start, min(k_end, d)), plus the tail iff the range reaches d, into the partial. */
int clb_batch_calc(clb_ctx *ctx, uint64_t d, uint64_t k_start, uint64_t k_end, int j);
/* AddAns formulano (KernelFuncs.as:68-74 -> Sum7_192, Bellard.cl:1359-1398): partial <<= log2 c_j
 * (mod 2^192, written back to the partial) and answer += partial. */
int clb_add_ans(clb_ctx *ctx, int j);
/* HCLReadBuffer of bigSumMem / dev_ans (Bellard_192.hsp:287-288, :292). */
int clb_read_partial(clb_ctx *ctx, uint64_t w[3]);
int clb_read_ans(clb_ctx *ctx, uint64_t w[3]);
/* GlobalMemSet's answer clear (Bellard_192.hsp:413-419). */
int clb_clear_ans(clb_ctx *ctx);

/* Sglobal*mtg_192* kernel operator (Bellard.cl:771, :992; Bellard_div2by1.cl:313): adds the head
 * terms k in [offset, k_max) of the formula (numesign, den0, den1) into the partial. */
int clb_sglobal(clb_ctx *ctx, uint64_t offset, uint64_t d, uint64_t k_max, int64_t numesign,
                uint64_t den0, uint64_t den1);
/* Sglobal_after_192 kernel operator (Bellard.cl:1153-1253): adds the tail terms k = d..d+19. */
int clb_sglobal_after(clb_ctx *ctx, uint64_t d, int64_t numesign, uint64_t den0, uint64_t den1);
/* Ssum kernel operator, both passes (KernelFuncs.as:62-65, Bellard.cl:1274-1297 / :1300-1321):
 * reduces the partial's per-group accumulators into the partial.  Only the 106-bit-mode replay
 * keeps per-group accumulators (clb_set_replay); everywhere else the operators add straight into
 * the partial with exact 192-bit sums, whose order does not change the result, and this call
 * does nothing. */
int clb_ssum(clb_ctx *ctx);

/* Allloop (Bellard_192.hsp:264-296): the full normal-mode computation at hex position N.
 * frac = frac(16^(N'-1) pi) (192 bits), hex = the display string (precision dependent),
 * n_out = N' (may be NULL). */
int clb_normal(clb_ctx *ctx, uint64_t n, uint64_t frac[3], char hex[49], uint64_t *n_out);

/* Many digit positions in one persistent launch (no reference counterpart; the reference
 * computes one position per run).  frac[3*i..] = frac(16^(N'_i - 1) pi).  At most 2^21
 * positions per call (CLB_EINVAL beyond). */
int clb_batch_positions(clb_ctx *ctx, const uint64_t *n, size_t count, uint64_t *frac);
/* Many term-range jobs in one persistent launch (work mode's jobs batched, Bellard_192.hsp:
 * 298-340).  partial[3*i..] = c_j * (job i's sum) mod 1, the value fsave writes.  At most 2^24
 * jobs per call (CLB_EINVAL beyond). */
int clb_run_jobs(clb_ctx *ctx, const clb_job *jobs, size_t count, uint64_t *partial);

/* clb_run_jobs without any allocation (no counterpart in the reference; for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same jobs, same results, bit for bit.  The call's tables, accumulator and results live in a
 * fixed array of CLB_NOALLOC_BYTES in each GPU's memory (and a pinned host staging buffer of the
 * same size), set up once per device by the first clb_ctx_create on it and held until the process
 * ends; the call itself allocates nothing.  At most CLB_NOALLOC_MAX_JOBS jobs per call
 * (CLB_EINVAL beyond), not in replay mode (CLB_EINVAL), never as a captured graph.  The array
 * keeps the last call's contents; every call clears what must start at zero.  Calls on contexts
 * of the same GPU share the array: within one process the library runs them one at a time (a
 * lock held for the call); unsynchronised access to it from outside the library would race. */
#define CLB_NOALLOC_BYTES 1048576u
#define CLB_NOALLOC_MAX_JOBS 4096u
int clb_run_jobs_noalloc(clb_ctx *ctx, const clb_job *jobs, size_t count, uint64_t *partial);

/* hexadecimal_i32 with the host word swap (Bellard_192.hsp:354-359, hexsademical.as:104-127). */
int clb_hex48(const uint64_t w[3], char hex[49]);
/* hexadecimal_dd (hexsademical.as:21-101): double-double -> 28 hex digits + 20 zeros.  CLB_EINVAL
 * for a non-finite part or one outside [-1, 1] (the reference's values lie in [-1, 1)). */
int clb_hex_dd(const double dd[2], char hex[49]);
/* The double-double (hi, lo) nearest to the 192-bit fraction w (round to nearest, 106 bits). */
int clb_u192_to_dd(const uint64_t w[3], double dd[2]);

/* Work mode (Bellard_192.hsp:88-340, :514-533).  Reads the worklist (digits=, precision=, j_b
 * lines), runs every job whose range starts at or below d and whose file
 * <work_dir>/<j>/<j>_<b>.txt does not exist yet, in launches bounded by clb_set_worklist_launch,
 * and writes each launch's files before the next launch starts.  batch_size 0 selects the
 * reference's 2^37 terms per batch.  device_label is written on the file's second line (NULL:
 * "<device>\t<device name>").  *jobs_run receives the number of files written, also when a later
 * launch fails (may be NULL). */
int clb_run_worklist(clb_ctx *ctx, const char *worklist_path, const char *work_dir,
                     uint64_t batch_size, const char *device_label, int *jobs_run);
/* Work-mode launch bounds (the reference launches and saves one job at a time,
 * Bellard_192.hsp:309-333; here consecutive pending jobs share one persistent launch).  A launch
 * takes jobs in worklist order while their estimated time stays within max_seconds (0: no time
 * bound; default 60) and their count within max_jobs (0: no count bound, the default); it holds at
 * least one job.  Each launch's files are written before the next launch starts, so an
 * interrupted run loses at most one launch, and a file's third line is its job's share of its
 * launch's time (by estimated cost).  max_launches > 0 makes clb_run_worklist return after that
 * many launches (the files written so far stay; a later call resumes), 0 (default) runs the whole
 * list.  CLB_EINVAL for a negative or non-finite max_seconds. */
int clb_set_worklist_launch(clb_ctx *ctx, uint32_t max_jobs, double max_seconds,
                            uint32_t max_launches);
/* *chktxtstr + *workStart2 (Bellard_192.hsp:176-247): parse a worklist without running it.
 * digits, precision: the values of lines 0 and 1 (digits clamped to [0, 999999999999999999]);
 * up to cap jobs (formula, batch index) from the following lines into jobs[2*i], jobs[2*i+1];
 * *njobs = number of job lines (may exceed cap).  Blank lines are skipped. */
int clb_parse_worklist(const char *worklist_path, uint64_t *digits, int *precision,
                       int64_t *jobs, size_t cap, size_t *njobs);
/* fsave (Bellard_192.hsp:514-533): writes <work_dir>/<j>/<j>_<b>.txt (creating the directories):
 * the value w formatted for the precision, then device_label, then "<ms>ms", CR LF separated. */
int clb_write_work_file(const char *work_dir, int j, int64_t b, const uint64_t w[3],
                        int precision, const char *device_label, int64_t ms);
/* workmodeSum0_6 (workmodeSum0_6.hsp:1-49): for j = 0..6 and b = 0, 1, ... until the first missing
 * file, adds line 0 of <work_dir>/<j>/<j>_<b>.txt mod 16^48.  counts[j] = files used. */
int clb_work_sum(const char *work_dir, char hex[49], int counts[7]);

/* Issue-rate microbenchmark (no reference counterpart; measurement for the roofline).  Runs one
 * VALU instruction (CLB_PEAK_FMA_F64 ... CLB_PEAK_FMA_F32) or one engine's whole modular square
 * (CLB_PEAK_FSQ: the FP64 engine's six-instruction square; CLB_PEAK_ISQ: the integer engine's
 * 64-bit Montgomery square) on 8 independent chains per lane, 8 waves per SIMD on every CU, for
 * about target_ms.  out[0] = lane-operations (or modular squares) per second, out[1] = the shader
 * clock measured during the run (GHz: the shader-clock and 100 MHz real-time counters read by the
 * first wave of the first 8 workgroups, one per XCD, averaged),
 * out[2] = the run's ms, out[3] = compute units.  Allocates its own small buffers. */
#define CLB_PEAK_FMA_F64 0
#define CLB_PEAK_MUL_F64 1
#define CLB_PEAK_ADD_F64 2
#define CLB_PEAK_RNDNE_F64 3
#define CLB_PEAK_LDEXP_F64 4
#define CLB_PEAK_MAD_U64_U32 5
#define CLB_PEAK_MUL_HI_U32 6
#define CLB_PEAK_MUL_LO_U32 7
#define CLB_PEAK_ADD_U32 8
#define CLB_PEAK_FMA_F32 9
#define CLB_PEAK_FSQ 10
#define CLB_PEAK_ISQ 11
#define CLB_PEAK_COUNT 12
int clb_issue_rate(int device, int op, double target_ms, double out[4]);

/* Timing of the last clb_normal / clb_batch_positions / clb_run_jobs / clb_batch_calc call on
 * the context's stream (HIP events): the head-term kernel (in graph mode: the whole replayed
 * graph) and the whole call, in ms. */
int clb_last_timing(clb_ctx *ctx, double *head_kernel_ms, double *total_ms);

/* Shader clock of the head-term kernel of the last clb_normal / clb_batch_positions /
 * clb_run_jobs / clb_batch_calc / clb_sglobal / clb_sglobal_after call (the calls
 * clb_last_timing times; no counterpart in the reference: a measurement for the roofline report,
 * whose peak is quoted at a nominal clock while GPUs run at 2.1-2.4 GHz).  The first wave of
 * workgroup x < CLB_CLOCK_XCDS, one per XCD, reads the shader-clock and 100 MHz real-time
 * counters at the kernel's start and when it leaves the tile loop; ghz[x] (optional) receives
 * each one's mean clock in GHz and *mean_ghz (optional) their mean.  CLB_EINVAL when that call
 * ran no head-term kernel (the 106-bit-mode replay, clb_run_jobs_noalloc, or no call yet).
 * Synchronises the context's stream. */
#define CLB_CLOCK_XCDS 8
int clb_last_clock(clb_ctx *ctx, double ghz[CLB_CLOCK_XCDS], double *mean_ghz);

#ifdef __cplusplus
}
#endif
#endif
