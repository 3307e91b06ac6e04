/*
 * ref_runner.c -- runs the reference's own OpenCL kernels (Bellard.cl, Bellard_div2by1.cl,
 * compiled from the reference sources into oracle/_ref/ by oracle/Makefile) through
 * the OpenCL runtime, replaying the reference host's launch sequence.
 *
 * TEST INFRASTRUCTURE ONLY (the reference's own kernels for parity tests, and their GPU rate beside ours).  Only the
 * GPU tests load it; the product never does.
 *
 * Replayed host code (toropippi/CLBellard):
 *   Bellard_192.hsp:5-7        BIGSUM_SIZE 65536, LOCAL_WORK_SIZE 256, GWS = 65536 * 256
 *   Bellard_192.hsp:388-405    kernels chosen per precision (0: dd, 1: 192bit, 2: 192bit beta)
 *   Bellard_192.hsp:408-420    buffers bigSum, smallSum, dev_ans; dev_ans cleared
 *   Bellard_192.hsp:264-296    Allloop: per formula Set0mem, BatchCalc(0, d), AddAns
 *   KernelFuncs.as:15-23       Set0memUlong3 / Set0memdd (FillBuffer8)
 *   KernelFuncs.as:25-66       BatchCalc: GWS*16-term launches below loopmax32, GWS*32 above,
 *                              the tail iff offset == d, Ssum 65536 -> 256 -> 1
 *   KernelFuncs.as:68-74       AddAns -> Sum7
 * The programs are the default build (no -cl-denorms-are-zero, no -cl-fast-relaxed-math); as the
 * reference host does (Bellard_192.hsp:472-474), a call whose loop count has 10 d + 9 >= 2^53
 * runs the kernels of Bellard.cl built with -D OVERDOUBLE (Bellard_od.gfx950.co).
 */
#define CL_TARGET_OPENCL_VERSION 120
#include <CL/cl.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define BIGSUM_SIZE 65536u
#define LWS 256u
#define GWS (BIGSUM_SIZE * LWS)

static const uint64_t kNumer[7] = {32, 1, 256, 64, 4, 4, 1};
static const int64_t kNumeSign[7] = {1, 1, 0, 1, 1, 1, 0};
static const uint64_t kDen0[7] = {4, 4, 10, 10, 10, 10, 10};
static const uint64_t kDen1[7] = {1, 3, 1, 3, 5, 7, 9};

static cl_context g_ctx;
static cl_command_queue g_q;
static cl_program g_prg, g_prg2, g_prg_od;
/* [build][precision]: build 0 the default program, 1 Bellard.cl with -D OVERDOUBLE */
static cl_kernel g_k32[2][3], g_k64[2][3], g_kafter[2][3], g_ksum[2][3], g_ksum7[2][3], g_kfill;
static int g_od; /* the build of the current call */
static cl_mem g_big[3], g_small[3], g_ans[3];
static char g_err[512];
static double g_last_ms;

const char *ref_last_error(void) { return g_err; }

/* Work-items per term-kernel launch: the reference's GWS unless a test sets a smaller one (a
 * multiple of LWS).  The kernels stride by get_global_size(0), so a smaller launch gives each
 * work-item its 16 or 32 terms on a small range.  Returns -1 for a bad size. */
static uint64_t g_gws = GWS;
int ref_set_gws(uint64_t gws) {
  if (gws == 0 || gws % LWS || gws > GWS) return -1;
  g_gws = gws;
  return 0;
}
double ref_last_ms(void) { return g_last_ms; }

static int failf(const char *what, cl_int e) {
  snprintf(g_err, sizeof g_err, "%s failed (%d)", what, (int)e);
  return -1;
}

static unsigned char *read_file(const char *path, size_t *n) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  unsigned char *buf = (unsigned char *)malloc((size_t)sz);
  if (buf && fread(buf, 1, (size_t)sz, f) != (size_t)sz) {
    free(buf);
    buf = NULL;
  }
  fclose(f);
  *n = (size_t)sz;
  return buf;
}

static cl_program load_program(cl_device_id dev, const char *path, cl_int *e) {
  size_t n = 0;
  unsigned char *bin = read_file(path, &n);
  if (!bin) {
    *e = CL_INVALID_BINARY;
    snprintf(g_err, sizeof g_err, "cannot read %s", path);
    return NULL;
  }
  const unsigned char *bins[1] = {bin};
  cl_int st = 0;
  cl_program p = clCreateProgramWithBinary(g_ctx, 1, &dev, &n, bins, &st, e);
  free(bin);
  if (*e != CL_SUCCESS) return NULL;
  *e = clBuildProgram(p, 1, &dev, "", NULL, NULL);
  return p;
}

/* Creates the context on the first GPU and loads both programs from co_dir. */
int ref_init(const char *co_dir) {
  cl_platform_id plat;
  cl_uint np = 0;
  cl_int e = clGetPlatformIDs(1, &plat, &np);
  if (e != CL_SUCCESS || np == 0) return failf("clGetPlatformIDs", e);
  cl_device_id dev;
  cl_uint nd = 0;
  e = clGetDeviceIDs(plat, CL_DEVICE_TYPE_GPU, 1, &dev, &nd);
  if (e != CL_SUCCESS || nd == 0) return failf("clGetDeviceIDs(GPU)", e);
  g_ctx = clCreateContext(NULL, 1, &dev, NULL, NULL, &e);
  if (e != CL_SUCCESS) return failf("clCreateContext", e);
  g_q = clCreateCommandQueue(g_ctx, dev, 0, &e);
  if (e != CL_SUCCESS) return failf("clCreateCommandQueue", e);
  char path[1024];
  snprintf(path, sizeof path, "%s/Bellard.gfx950.co", co_dir);
  g_prg = load_program(dev, path, &e);
  if (e != CL_SUCCESS) return failf("Bellard program", e);
  snprintf(path, sizeof path, "%s/Bellard_div2by1.gfx950.co", co_dir);
  g_prg2 = load_program(dev, path, &e);
  if (e != CL_SUCCESS) return failf("Bellard_div2by1 program", e);
  snprintf(path, sizeof path, "%s/Bellard_od.gfx950.co", co_dir);
  g_prg_od = load_program(dev, path, &e);
  if (e != CL_SUCCESS) return failf("Bellard -D OVERDOUBLE program", e);
  /* Bellard_192.hsp:388-405 */
  const char *k32[3] = {"Sglobal32mtg", "Sglobal32mtg_192", "Sglobal32mtg_192"};
  const char *k64[3] = {"Sglobal64mtg", "Sglobal64mtg_192", "Sglobal64mtg_192_Refresh"};
  const char *kaf[3] = {"Sglobal_after", "Sglobal_after_192", "Sglobal_after_192"};
  const char *ksm[3] = {"Ssum", "Ssum_192", "Ssum_192"};
  const char *ks7[3] = {"Sum7", "Sum7_192", "Sum7_192"};
  for (int p = 0; p < 3; ++p) {
    for (int b = 0; b < 2; ++b) {
      cl_program prg = b ? g_prg_od : g_prg;
      g_k32[b][p] = clCreateKernel(prg, k32[p], &e);
      if (e != CL_SUCCESS) return failf(k32[p], e);
      g_k64[b][p] = clCreateKernel(p == 2 ? g_prg2 : prg, k64[p], &e);
      if (e != CL_SUCCESS) return failf(k64[p], e);
      g_kafter[b][p] = clCreateKernel(prg, kaf[p], &e);
      if (e != CL_SUCCESS) return failf(kaf[p], e);
      g_ksum[b][p] = clCreateKernel(prg, ksm[p], &e);
      if (e != CL_SUCCESS) return failf(ksm[p], e);
      g_ksum7[b][p] = clCreateKernel(prg, ks7[p], &e);
      if (e != CL_SUCCESS) return failf(ks7[p], e);
    }
    /* Bellard_192.hsp:409-411: 8 * (2 + (precision != 0)) bytes per slot */
    const size_t w = 8u * (p == 0 ? 2u : 3u);
    g_big[p] = clCreateBuffer(g_ctx, CL_MEM_READ_WRITE, BIGSUM_SIZE * w, NULL, &e);
    if (e != CL_SUCCESS) return failf("bigSum", e);
    g_small[p] = clCreateBuffer(g_ctx, CL_MEM_READ_WRITE, BIGSUM_SIZE / 256 * w, NULL, &e);
    if (e != CL_SUCCESS) return failf("smallSum", e);
    g_ans[p] = clCreateBuffer(g_ctx, CL_MEM_READ_WRITE, w, NULL, &e);
    if (e != CL_SUCCESS) return failf("dev_ans", e);
  }
  g_kfill = clCreateKernel(g_prg, "FillBuffer8", &e);
  if (e != CL_SUCCESS) return failf("FillBuffer8", e);
  return 0;
}

static int launch(cl_kernel k, size_t global, size_t local) {
  cl_int e = clEnqueueNDRangeKernel(g_q, k, 1, NULL, &global, &local, 0, NULL, NULL);
  return e == CL_SUCCESS ? 0 : failf("clEnqueueNDRangeKernel", e);
}

/* FillBuffer8(A, 0, 0) over n words (KernelFuncs.as:15-23, Bellard_192.hsp:414-418). */
static int fill0(cl_mem buf, size_t words, size_t local) {
  const cl_ulong zero = 0;
  const cl_uint off = 0;
  clSetKernelArg(g_kfill, 0, sizeof(cl_mem), &buf);
  clSetKernelArg(g_kfill, 1, sizeof(cl_ulong), &zero);
  clSetKernelArg(g_kfill, 2, sizeof(cl_uint), &off);
  return launch(g_kfill, words, local);
}

static int set_sglobal(cl_kernel k, int p, cl_ulong offset, cl_ulong d, cl_ulong kmax, int j) {
  const cl_long sign = kNumeSign[j];
  const cl_ulong den0 = kDen0[j], den1 = kDen1[j];
  cl_int e = clSetKernelArg(k, 0, sizeof(cl_mem), &g_big[p]);
  e |= clSetKernelArg(k, 1, sizeof(cl_ulong), &offset);
  e |= clSetKernelArg(k, 2, sizeof(cl_ulong), &d);
  e |= clSetKernelArg(k, 3, sizeof(cl_ulong), &kmax);
  e |= clSetKernelArg(k, 4, sizeof(cl_long), &sign);
  e |= clSetKernelArg(k, 5, sizeof(cl_ulong), &den0);
  e |= clSetKernelArg(k, 6, sizeof(cl_ulong), &den1);
  return e == CL_SUCCESS ? 0 : failf("clSetKernelArg", e);
}

static int set0mem(int p) { return fill0(g_big[p], BIGSUM_SIZE * (p == 0 ? 2u : 3u), 64); }

/* Sglobal_after* on one work-item (KernelFuncs.as:56-59). */
static int after(int p, uint64_t d, int j) {
  const cl_long sign = kNumeSign[j];
  const cl_ulong dd = d, den0 = kDen0[j], den1 = kDen1[j];
  cl_kernel k = g_kafter[g_od][p];
  clSetKernelArg(k, 0, sizeof(cl_mem), &g_big[p]);
  clSetKernelArg(k, 1, sizeof(cl_ulong), &dd);
  clSetKernelArg(k, 2, sizeof(cl_long), &sign);
  clSetKernelArg(k, 3, sizeof(cl_ulong), &den0);
  clSetKernelArg(k, 4, sizeof(cl_ulong), &den1);
  return launch(k, 1, 1);
}

/* Ssum 65536 -> 256, then 256 -> 1 (KernelFuncs.as:61-65). */
static int ssum(int p) {
  cl_kernel s = g_ksum[g_od][p];
  clSetKernelArg(s, 0, sizeof(cl_mem), &g_big[p]);
  clSetKernelArg(s, 1, sizeof(cl_mem), &g_small[p]);
  if (launch(s, BIGSUM_SIZE, LWS)) return -1;
  clSetKernelArg(s, 0, sizeof(cl_mem), &g_small[p]);
  clSetKernelArg(s, 1, sizeof(cl_mem), &g_big[p]);
  return launch(s, BIGSUM_SIZE / 256, LWS);
}

/* KernelFuncs.as:25-66 */
static int batch_calc(int p, uint64_t d, uint64_t k_start, uint64_t k_end, int j) {
  uint64_t k_max = k_start, offset = k_start;
  uint64_t loopmax32 = ((1ull << 32) - kDen1[j] - 1) / kDen0[j] + 1;
  if (loopmax32 > d) loopmax32 = d;
  if (loopmax32 > k_end) loopmax32 = k_end;
  while (offset < loopmax32) {
    k_max = k_max + g_gws * 16 < loopmax32 ? k_max + g_gws * 16 : loopmax32;
    cl_kernel k = g_k32[g_od][p];
    if (set_sglobal(k, p, offset, d, k_max, j) || launch(k, g_gws, LWS)) return -1;
    offset = k_max;
  }
  const uint64_t loopmax64 = k_end < d ? k_end : d;
  while (offset < loopmax64) {
    k_max = k_max + g_gws * 32 < loopmax64 ? k_max + g_gws * 32 : loopmax64;
    cl_kernel k = g_k64[g_od][p];
    if (set_sglobal(k, p, offset, d, k_max, j) || launch(k, g_gws, LWS)) return -1;
    offset = k_max;
  }
  if (offset == d && after(p, d, j)) return -1;
  return ssum(p);
}

/* KernelFuncs.as:68-74 */
static int add_ans(int p, int j) {
  const cl_ulong numer = kNumer[j];
  cl_kernel k = g_ksum7[g_od][p];
  clSetKernelArg(k, 0, sizeof(cl_mem), &g_ans[p]);
  clSetKernelArg(k, 1, sizeof(cl_mem), &g_big[p]);
  clSetKernelArg(k, 2, sizeof(cl_ulong), &numer);
  return launch(k, 1, 1);
}

static int read_words(cl_mem buf, int p, uint64_t *out) {
  cl_int e = clEnqueueReadBuffer(g_q, buf, CL_TRUE, 0, 8u * (p == 0 ? 2u : 3u), out, 0, NULL, NULL);
  return e == CL_SUCCESS ? 0 : failf("clEnqueueReadBuffer", e);
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

/* Picks the program build for loop count d (Bellard_192.hsp:472-474); moduli up to 10 d + 9 must
 * stay below 2^63 (the -D OVERDOUBLE kernels' condition, Bellard.cl:1041). */
static int check_d(uint64_t d) {
  if (d > (9223372036854775807ull - 9) / 10) {
    snprintf(g_err, sizeof g_err, "d = %llu: moduli reach 2^63", (unsigned long long)d);
    return -1;
  }
  g_od = d * 10 + 9 >= 9007199254740992ull;
  return 0;
}

/* Allloop, Bellard_192.hsp:264-296.  ans: 2 words (the double-double bits) for precision 0,
 * 3 words (bigSum order, most significant first) otherwise. */
int ref_normal(uint64_t d, int precision, uint64_t *ans) {
  if (precision < 0 || precision > 2 || check_d(d)) return -1;
  const double t0 = now_ms();
  if (fill0(g_ans[precision], precision == 0 ? 2 : 3, precision == 0 ? 2 : 3)) return -1;
  for (int j = 0; j < 7; ++j) {
    if (set0mem(precision) || batch_calc(precision, d, 0, d, j) || add_ans(precision, j))
      return -1;
  }
  if (read_words(g_ans[precision], precision, ans)) return -1;
  g_last_ms = now_ms() - t0;
  return 0;
}

/* Set0mem + BatchCalc(k_start, k_end, j) + AddAns into a cleared answer: the value fsave writes
 * for a work-mode job (Bellard_192.hsp:309-333).  part: bigSum[0..] after Sum7. */
int ref_job(uint64_t d, uint64_t k_start, uint64_t k_end, int j, int precision, uint64_t *part) {
  if (precision < 0 || precision > 2 || j < 0 || j > 6 || check_d(d)) return -1;
  const double t0 = now_ms();
  if (fill0(g_ans[precision], precision == 0 ? 2 : 3, precision == 0 ? 2 : 3) ||
      set0mem(precision) || batch_calc(precision, d, k_start, k_end, j) || add_ans(precision, j))
    return -1;
  if (read_words(g_big[precision], precision, part)) return -1;
  g_last_ms = now_ms() - t0;
  return 0;
}

/* The reference's kernels one at a time on its bigSumMem, for the operator-level tests:
 * op 0 Set0mem, 1 Sglobal32mtg*(offset = a, d = b, k_max = c), 2 Sglobal64mtg*(a, b, c),
 * 3 Sglobal_after*(d = b), 4 Ssum's two passes, then (every op) bigSum[0..] into out (2 words
 * for precision 0, 3 otherwise).  Launches use the work-item count of ref_set_gws. */
int ref_op(int op, int precision, uint64_t a, uint64_t b, uint64_t c, int j, uint64_t *out) {
  const int p = precision;
  if (p < 0 || p > 2 || j < 0 || j > 6 || op < 0 || op > 4) {
    snprintf(g_err, sizeof g_err, "ref_op: bad argument");
    return -1;
  }
  if ((op == 1 || op == 2 || op == 3) && check_d(b)) return -1;
  int rc = 0;
  switch (op) {
    case 0: rc = set0mem(p); break;
    case 1:
    case 2: {
      cl_kernel k = op == 1 ? g_k32[g_od][p] : g_k64[g_od][p];
      rc = set_sglobal(k, p, a, b, c, j) || launch(k, g_gws, LWS);
      break;
    }
    case 3: rc = after(p, b, j); break;
    case 4: rc = ssum(p); break;
  }
  if (rc) return -1;
  return read_words(g_big[p], p, out);
}

void ref_shutdown(void) {
  if (g_q) clFinish(g_q);
  for (int p = 0; p < 3; ++p) {
    for (int b = 0; b < 2; ++b) {
      if (g_k32[b][p]) clReleaseKernel(g_k32[b][p]);
      if (g_k64[b][p]) clReleaseKernel(g_k64[b][p]);
      if (g_kafter[b][p]) clReleaseKernel(g_kafter[b][p]);
      if (g_ksum[b][p]) clReleaseKernel(g_ksum[b][p]);
      if (g_ksum7[b][p]) clReleaseKernel(g_ksum7[b][p]);
    }
    if (g_big[p]) clReleaseMemObject(g_big[p]);
    if (g_small[p]) clReleaseMemObject(g_small[p]);
    if (g_ans[p]) clReleaseMemObject(g_ans[p]);
  }
  if (g_kfill) clReleaseKernel(g_kfill);
  if (g_prg) clReleaseProgram(g_prg);
  if (g_prg2) clReleaseProgram(g_prg2);
  if (g_prg_od) clReleaseProgram(g_prg_od);
  if (g_q) clReleaseCommandQueue(g_q);
  if (g_ctx) clReleaseContext(g_ctx);
  memset(g_k32, 0, sizeof g_k32);
  memset(g_k64, 0, sizeof g_k64);
  memset(g_kafter, 0, sizeof g_kafter);
  memset(g_ksum, 0, sizeof g_ksum);
  memset(g_ksum7, 0, sizeof g_ksum7);
  memset(g_big, 0, sizeof g_big);
  memset(g_small, 0, sizeof g_small);
  memset(g_ans, 0, sizeof g_ans);
  g_kfill = NULL;
  g_prg = g_prg2 = g_prg_od = NULL;
  g_q = NULL;
  g_ctx = NULL;
}
