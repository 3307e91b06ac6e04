// clb_hostlib.hpp -- host-only parts of libclbellard.so shared by the GPU host code (clb_host.cpp)
// and the formatting / file code (clb_format.cpp): the formula table, the error message of the
// calling thread, digit normalisation, output formats, worklists and work files.  No HIP: the
// file compiles with any C++17 compiler (the sanitizer build of the tests uses g++).
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

namespace clb {

// KernelFuncs.as:9-12: numerator, sign flag (1 = minus), denominator den0*k + den1.
inline constexpr uint64_t kNumer[7] = {32, 1, 256, 64, 4, 4, 1};
inline constexpr int64_t kNumeSign[7] = {1, 1, 0, 1, 1, 1, 0};
inline constexpr uint64_t kDen0[7] = {4, 4, 10, 10, 10, 10, 10};
inline constexpr uint64_t kDen1[7] = {1, 3, 1, 3, 5, 7, 9};
// Sum7_192 (Bellard.cl:1364-1368): multiplication by c_j is a left shift by log2(c_j).
inline constexpr uint8_t kShift[7] = {5, 0, 8, 6, 2, 2, 0};

inline constexpr uint64_t kMaxDigits = 999999999999999999ull;  // Bellard_192.hsp:450
inline constexpr uint64_t kBatchSize = 137438953472ull;        // Bellard_192.hsp:19, 2^37 terms

// Records the calling thread's error message (clb_last_error) and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// BuildOptions core (Bellard_192.hsp:468-474); CLB_EINVAL outside [5, kMaxDigits].
int normalize(uint64_t n_in, uint64_t* n_out, uint64_t* d_out, int* overdouble);
// hexadecimal_i32 with the host word swap: 48 lowercase hex digits, w[0] first.
void hex48(const uint64_t w[3], char out[49]);
// hexadecimal_dd (hexsademical.as:21-101) for a finite double-double with |hi| <= 1.
void hex_dd(const double dd[2], char out[49]);
// The double-double nearest to the 192-bit fraction w.
int u192_to_dd(const uint64_t w[3], double dd[2]);

// Internal format code: the 3 words hold a double-double's bit patterns (hi, lo, 0), as the
// 106-bit-mode replay returns its results.
inline constexpr int kDDWords = 3;
// The display string of a result in `precision` (0, 1, 2, or kDDWords).
void format_result(int precision, const uint64_t w[3], char hex[49]);

bool file_exists(const std::string& p);
// fsave (Bellard_192.hsp:514-533): <work_dir>/<j>/<j>_<b>.txt.
int write_work_file(const std::string& work_dir, int j, int64_t b, const uint64_t w[3],
                    int precision, const std::string& label, long long ms);

// *chktxtstr (Bellard_192.hsp:176-207), *workStart (:221-227), *workStart2 (:229-247).
struct Worklist {
  int64_t digits = 0;
  int precision = -1;
  std::vector<std::pair<int, int64_t>> jobs;
};
int parse_worklist(const char* path, Worklist* wl);

}  // namespace clb
