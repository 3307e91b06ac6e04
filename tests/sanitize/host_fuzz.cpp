// Sanitizer harness for the host-only entry points of libclbellard.so (clb_format.cpp), built with
// g++ -fsanitize=address,undefined by tests/test_sanitizers.py and linked with that source file
// directly (no HIP).  It feeds the entry points that parse or format external data -- worklists,
// work files, launch-settings files, digit positions, 192-bit values, double-doubles -- with
// random, malformed and
// boundary inputs; any memory error or undefined behaviour aborts the run.  Results are checked
// where they are known (round trips, status codes).
#include <sys/stat.h>
#include <unistd.h>

#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/clbellard.h"

static int failures = 0;
#define CHECK(cond)                                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond);         \
      ++failures;                                                                \
    }                                                                            \
  } while (0)

static void write_file(const std::string& path, const std::string& body) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) abort();
  fwrite(body.data(), 1, body.size(), f);
  fclose(f);
}

static std::string random_bytes(std::mt19937_64& rng, size_t n, const char* alphabet) {
  std::string s;
  const size_t a = alphabet ? strlen(alphabet) : 0;
  for (size_t i = 0; i < n; ++i) s.push_back(alphabet ? alphabet[rng() % a] : (char)(rng() & 255));
  return s;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: host_fuzz <scratch dir> [iterations]\n");
    return 2;
  }
  const std::string dir = argv[1];
  const int iters = argc > 2 ? atoi(argv[2]) : 2000;
  std::mt19937_64 rng(20261204);

  // normalisation and the formula table
  const uint64_t edges[] = {0, 1, 4, 5, 6, 9, 10, 11, 999999999999999999ull,
                            1000000000000000000ull, UINT64_MAX};
  for (uint64_t n : edges) {
    uint64_t a, b;
    int od;
    const int rc = clb_normalize(n, &a, &b, &od);
    CHECK((rc == CLB_OK) == (n >= 5 && n <= 999999999999999999ull));
    if (rc == CLB_OK) CHECK(a % 5 == 0 && b % 2 == 1);
  }
  for (int i = 0; i < iters; ++i) {
    uint64_t a, b;
    int od;
    (void)clb_normalize(rng(), &a, &b, &od);
    (void)clb_normalize(rng() % 1000000000000000000ull, nullptr, nullptr, nullptr);
  }
  for (int j = -3; j < 10; ++j) {
    uint64_t c, d0, d1;
    int64_t s;
    CHECK((clb_formula(j, &c, &s, &d0, &d1) == CLB_OK) == (j >= 0 && j <= 6));
  }

  // formats
  for (int i = 0; i < iters; ++i) {
    uint64_t w[3] = {rng(), rng(), rng()};
    char hex[49];
    CHECK(clb_hex48(w, hex) == CLB_OK && strlen(hex) == 48);
    double dd[2];
    CHECK(clb_u192_to_dd(w, dd) == CLB_OK);
    CHECK(clb_hex_dd(dd, hex) == CLB_OK && strlen(hex) == 48);
    // arbitrary bit patterns: NaN, infinities, huge values are refused, the rest formatted
    double raw[2];
    uint64_t bits[2] = {rng(), rng()};
    memcpy(raw, bits, sizeof raw);
    const int rc = clb_hex_dd(raw, hex);
    CHECK(rc == CLB_OK || rc == CLB_EINVAL);
    // in-range values with parts of any sign and size, including improper double-doubles
    double v[2] = {std::ldexp((double)(rng() >> 11), -53) * ((rng() & 1) ? 1 : -1),
                   std::ldexp((double)(rng() >> 11), -53 - (int)(rng() % 80)) * ((rng() & 1) ? 1 : -1)};
    CHECK(clb_hex_dd(v, hex) == CLB_OK);
  }
  const double specials[][2] = {{1.0, 0.0}, {-1.0, 0.0}, {1.0, 1.0}, {-1.0, -1.0}, {0.5, -0x1p-60},
                                {0.5, -0x1p-80}, {0.0, -0.0}, {1.0 - 0x1p-53, 0x1p-54}, {0x1p-1074, 0}};
  for (const auto& s : specials) {
    char hex[49];
    CHECK(clb_hex_dd(s, hex) == CLB_OK);
  }
  CHECK(clb_hex48(nullptr, nullptr) == CLB_EINVAL);

  // worklists: random text, long lines, numbers beyond 64 bits, CR/LF mixtures
  const std::string wl = dir + "/worklist.txt";
  const char* alpha = "0123456789_=digitsprecsion \r\n\t-+abcxyz";
  for (int i = 0; i < iters; ++i) {
    std::string body;
    switch (i % 5) {
      case 0: body = random_bytes(rng, rng() % 300, alpha); break;
      case 1: body = random_bytes(rng, rng() % 300, nullptr); break;
      case 2:
        body = "digits=" + std::to_string(rng()) + std::string(rng() % 40, '9') + "\r\nprecision=" +
               std::to_string(rng() % 5) + "\r\n";
        for (int k = 0; k < (int)(rng() % 20); ++k)
          body += std::to_string(rng() % 9) + "_" + std::to_string(rng() % 1000000) + "\r\n";
        break;
      case 3: body = "digits=1000000\nprecision=1\n" + std::string(5000 + rng() % 5000, '7') + "\n"; break;
      default: body = std::string(rng() % 3, '\n'); break;
    }
    write_file(wl, body);
    uint64_t digits = 0;
    int prec = 0;
    int64_t jobs[16];
    size_t n = 0;
    const int rc = clb_parse_worklist(wl.c_str(), &digits, &prec, jobs, 8, &n);
    CHECK(rc == CLB_OK || rc == CLB_EFORMAT);
    if (rc == CLB_OK) CHECK(digits <= 999999999999999999ull && prec >= -1 && prec <= 2);
  }
  CHECK(clb_parse_worklist((dir + "/missing.txt").c_str(), nullptr, nullptr, nullptr, 0, nullptr) ==
        CLB_EIO);

  // work files: written and summed, then corrupted
  const std::string work = dir + "/work";
  uint64_t total[3] = {0, 0, 0};
  for (int b = 0; b < 20; ++b) {
    uint64_t w[3] = {rng(), rng(), rng()};
    CHECK(clb_write_work_file(work.c_str(), 3, b, w, 1, "0\tfuzz", (int64_t)rng()) == CLB_OK);
    // total += w mod 2^192 (w[0] most significant)
    unsigned __int128 c = (unsigned __int128)total[2] + w[2];
    total[2] = (uint64_t)c;
    c = (c >> 64) + total[1] + w[1];
    total[1] = (uint64_t)c;
    total[0] = total[0] + w[0] + (uint64_t)(c >> 64);
  }
  char hex[49], want[49];
  int counts[7];
  CHECK(clb_work_sum(work.c_str(), hex, counts) == CLB_OK && counts[3] == 20);
  CHECK(clb_hex48(total, want) == CLB_OK && strcmp(hex, want) == 0);
  CHECK(clb_write_work_file(work.c_str(), 7, 0, total, 1, "x", 1) == CLB_EINVAL);
  CHECK(clb_write_work_file(work.c_str(), 0, -1, total, 1, "x", 1) == CLB_EINVAL);
  for (int i = 0; i < iters / 4; ++i) {
    const std::string p = work + "/3/3_" + std::to_string(rng() % 20) + ".txt";
    std::string body;
    switch (i % 4) {
      case 0: body = random_bytes(rng, rng() % 100, "0123456789abcdefABCDEF\r\n"); break;
      case 1: body = random_bytes(rng, rng() % 100, nullptr); break;
      case 2: body = ""; break;
      default: body = std::string(48, 'f') + "\r\n" + random_bytes(rng, rng() % 50, nullptr); break;
    }
    write_file(p, body);
    const int rc = clb_work_sum(work.c_str(), hex, counts);
    CHECK(rc == CLB_OK || rc == CLB_EFORMAT);
  }
  CHECK(clb_work_sum(nullptr, hex, counts) == CLB_EINVAL);

  // launch-settings files: a round trip, then random and mutated text
  clb_schedule s0;
  memset(&s0, 0, sizeof s0);
  snprintf(s0.device, sizeof s0.device, "%s", "AMD Instinct MI355X");
  s0.terms_per_lane = 16;
  s0.blocks_per_cu = 8;
  s0.waves_per_group = 4;
  s0.min_chunk = 256;
  s0.max_replicas = 64;
  const std::string sp = dir + "/sched.txt";
  CHECK(clb_write_schedule(sp.c_str(), &s0) == CLB_OK);
  clb_schedule s1;
  CHECK(clb_read_schedule(sp.c_str(), &s1) == CLB_OK && memcmp(&s0, &s1, sizeof s0) == 0);
  const std::string good = "device=g\nterms_per_lane=4\nmin_tiles_per_wave=0\nblocks_per_cu=0\n"
                           "waves_per_group=2\nmin_chunk=64\nchunks_per_wave=-1\nmax_replicas=8\n";
  for (int i = 0; i < iters; ++i) {
    std::string body;
    switch (i % 4) {
      case 0: body = random_bytes(rng, rng() % 300, "=\r\n-0123456789abcdeimnprstuvwx_"); break;
      case 1: body = random_bytes(rng, rng() % 300, nullptr); break;
      case 2: {
        body = good;
        for (int k = 0; k < 3; ++k) body[rng() % body.size()] = (char)(rng() & 255);
        break;
      }
      default: body = good + random_bytes(rng, rng() % 5000, "=\n9"); break;
    }
    write_file(sp, body);
    clb_schedule s;
    const int rc = clb_read_schedule(sp.c_str(), &s);
    CHECK(rc == CLB_OK || rc == CLB_EFORMAT || rc == CLB_EINVAL);
    if (rc == CLB_OK) CHECK(strlen(s.device) > 0 && strlen(s.device) < sizeof s.device);
  }
  CHECK(clb_read_schedule(nullptr, &s1) == CLB_EINVAL);
  CHECK(clb_write_schedule(sp.c_str(), nullptr) == CLB_EINVAL);
  printf("host_fuzz: %d iterations, %d failed checks, last error '%s'\n", iters, failures,
         clb_last_error());
  return failures ? 1 : 0;
}
