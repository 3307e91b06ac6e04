// clb_cli.cpp -- command-line front end over libclbellard.so, replacing the reference's GUI
// (Bellard_192.hsp normal mode and work mode) and workmodeSum0_6.exe.
//
//   clbellard normal <digits> [--precision 0|1|2] [--device i] [--kernel auto|plain|split]
//                    [--replay | --replay-gws GWS] [--autotune MS] [--tune-cache FILE]
//   clbellard work <worklist.txt> [--workdir work] [--batch-size B] [--device i]
//                  [--launch-seconds S] [--launch-jobs J] [--replay | --replay-gws GWS]
//                  [--autotune MS] [--tune-cache FILE]
//   clbellard sum [--workdir work]
//   clbellard batch <positions.txt | -> [--precision 0|1|2] [--device i] [--graph]
//                   [--autotune MS] [--tune-cache FILE] [--replay | --replay-gws GWS]
// work: the pending jobs run in launches of at most S estimated seconds (default 60) and J jobs
// (default: any number), each launch's files written before the next (clb_set_worklist_launch).
// --replay: the reference kernels' own output bits (clb_set_replay) with the reference's
// 16777216 work-items per launch; --replay-gws: with GWS work-items per launch.
// batch: one position per line (file, or standard input for "-"), all of them in one persistent
// launch (clb_batch_positions); prints "N' digits" per position in input order.  --graph replays
// the captured launch sequence (clb_set_graph).
// --autotune MS first tunes the launch settings for about MS milliseconds (clb_autotune) on the
// position (normal), the worklist's position (work) or the largest position (batch);
// --tune-cache FILE loads them from FILE when it exists (clb_load_schedule; it must have been
// written for this device), and otherwise writes the tuned settings there (clb_save_schedule).
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clbellard.h"

static int usage() {
  fprintf(stderr,
          "usage:\n"
          "  clbellard normal <digits> [--precision 0|1|2] [--device i] [--kernel auto|plain|split]\n"
          "                   [--replay | --replay-gws GWS] [--autotune MS] [--tune-cache FILE]\n"
          "  clbellard work <worklist.txt> [--workdir DIR] [--batch-size B] [--device i]\n"
          "                 [--launch-seconds S] [--launch-jobs J] [--replay | --replay-gws GWS]\n"
          "                 [--autotune MS] [--tune-cache FILE]\n"
          "  clbellard sum [--workdir DIR]\n"
          "  clbellard batch <positions.txt | -> [--precision 0|1|2] [--device i] [--graph]\n"
          "                  [--autotune MS] [--tune-cache FILE] [--replay | --replay-gws GWS]\n");
  return 2;
}

static int die(int rc) {
  fprintf(stderr, "error %d: %s\n", rc, clb_last_error());
  return 1;
}

// A decimal number: digits only (no sign, no blanks), at most 2^64 - 1.
static bool parse_u64(const char* s, size_t len, unsigned long long* out) {
  if (len == 0) return false;
  unsigned long long v = 0;
  for (size_t i = 0; i < len; ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    const unsigned d = (unsigned)(s[i] - '0');
    if (v > (ULLONG_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

static bool parse_u64(const char* s, unsigned long long* out) {
  return s && parse_u64(s, strlen(s), out);
}

// Positions, one per line (blanks around the number allowed); blank lines skipped.  Returns
// false on an unreadable file or a line that is not a decimal number.
static bool read_positions(const std::string& path, std::vector<uint64_t>* out) {
  FILE* f = path == "-" ? stdin : fopen(path.c_str(), "r");
  if (!f) return false;
  char line[4096];
  bool ok = true;
  while (ok && fgets(line, sizeof line, f)) {
    const size_t n = strlen(line);
    if (n + 1 == sizeof line && line[n - 1] != '\n') {  // longer than any number
      ok = false;
      break;
    }
    size_t a = 0, b = n;
    while (a < b && (line[a] == ' ' || line[a] == '\t')) ++a;
    while (b > a && (line[b - 1] == ' ' || line[b - 1] == '\t' || line[b - 1] == '\r' ||
                     line[b - 1] == '\n'))
      --b;
    if (a == b) continue;
    unsigned long long v = 0;
    ok = parse_u64(line + a, b - a, &v);
    out->push_back(v);
  }
  if (f != stdin) fclose(f);
  return ok;
}

// The display string of a result: the 106-bit mode prints a double-double (hexadecimal_dd), which
// the replay of that mode returns as its bit patterns and the exact mode as 192 bits.
static int format(const uint64_t w[3], int precision, bool replay, char hex[49]) {
  if (precision != 0) return clb_hex48(w, hex);
  double dd[2];
  if (replay) {
    memcpy(&dd[0], &w[0], 8);
    memcpy(&dd[1], &w[1], 8);
  } else if (int rc = clb_u192_to_dd(w, dd)) {
    return rc;
  }
  return clb_hex_dd(dd, hex);
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  const std::string cmd = argv[1];
  int precision = 1, device = 0, kernel = CLB_KERNEL_AUTO;
  std::string workdir = "work", positional;
  unsigned long long batch = 0, replay = 0, launch_jobs = 0;
  double launch_seconds = 60.0;
  bool graph = false;
  double tune_ms = 0.0;
  std::string tune_cache;
  // Every option is checked here, before any GPU call: a malformed one prints the usage.
  for (int i = 2; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    unsigned long long u = 0;
    if (a == "--precision" && parse_u64(v = next(), &u) && u <= 2) {
      precision = (int)u;
    } else if (a == "--device" && parse_u64(v = next(), &u) && u <= 1024) {
      device = (int)u;
    } else if (a == "--workdir" && (v = next())) {
      workdir = v;
    } else if (a == "--batch-size" && parse_u64(v = next(), &u)) {
      batch = u;
    } else if (a == "--kernel" && (v = next()) &&
               (!strcmp(v, "auto") || !strcmp(v, "plain") || !strcmp(v, "split"))) {
      kernel = !strcmp(v, "plain") ? CLB_KERNEL_PLAIN : !strcmp(v, "split") ? CLB_KERNEL_SPLIT
                                                                            : CLB_KERNEL_AUTO;
    } else if (a == "--launch-jobs" && parse_u64(v = next(), &u) && u <= 0xffffffffull) {
      launch_jobs = u;
    } else if (a == "--launch-seconds" && (v = next())) {
      char* end = nullptr;
      launch_seconds = strtod(v, &end);
      if (end == v || *end != 0 || !(launch_seconds >= 0.0 && launch_seconds <= 1e9)) return usage();
    } else if (a == "--graph") {
      graph = true;
    } else if (a == "--autotune" && (v = next())) {
      char* end = nullptr;
      tune_ms = strtod(v, &end);
      if (end == v || *end != 0 || !(tune_ms > 0.0 && tune_ms < 1e7)) return usage();
    } else if (a == "--tune-cache" && (v = next()) && *v) {
      tune_cache = v;
    } else if (a == "--replay") {
      replay = CLB_REF_GLOBAL_WORK_SIZE;
    } else if (a == "--replay-gws" && parse_u64(v = next(), &u) && u > 0) {
      replay = u;
    } else if (positional.empty() && (a[0] != '-' || a == "-")) {
      positional = a;
    } else {
      return usage();
    }
  }
  // The command, the work-item count of the replay and every digit position are checked here too,
  // still before the first GPU call (clb_normalize is host-only).
  if (cmd != "normal" && cmd != "work" && cmd != "sum" && cmd != "batch") return usage();
  if (replay && (replay % 256 != 0 || replay > CLB_REF_GLOBAL_WORK_SIZE)) return usage();
  unsigned long long digits = 0;
  if (cmd == "normal" && (!parse_u64(positional.c_str(), &digits) ||
                          clb_normalize(digits, nullptr, nullptr, nullptr) != CLB_OK))
    return usage();
  if (cmd == "sum") {
    char hex[49];
    int counts[7];
    if (int rc = clb_work_sum(workdir.c_str(), hex, counts)) return die(rc);
    printf("file num\n");
    for (int j = 0; j < 7; ++j) printf("%d:%d\n", j, counts[j]);
    printf("\n%s\n", hex);
    return 0;
  }
  if (positional.empty()) return usage();
  std::vector<uint64_t> pos;
  if (cmd == "batch" && !read_positions(positional, &pos)) {
    fprintf(stderr, "cannot read positions from %s (one decimal number per line)\n",
            positional.c_str());
    return 1;
  }
  for (uint64_t n : pos)
    if (clb_normalize(n, nullptr, nullptr, nullptr) != CLB_OK) {
      fprintf(stderr, "position %llu: %s\n", (unsigned long long)n, clb_last_error());
      return usage();
    }
  // The position the launch settings are tuned on, and a cache file that is there already (read
  // and checked here; a missing one is written after tuning).
  uint64_t tune_n = digits;
  for (uint64_t n : pos) tune_n = n > tune_n ? n : tune_n;
  if (cmd == "work" && tune_ms > 0.0) {
    uint64_t wl_digits = 0;
    size_t njobs = 0;
    if (clb_parse_worklist(positional.c_str(), &wl_digits, nullptr, nullptr, 0, &njobs) != CLB_OK ||
        clb_normalize(wl_digits, nullptr, nullptr, nullptr) != CLB_OK) {
      fprintf(stderr, "%s: %s\n", positional.c_str(), clb_last_error());
      return usage();
    }
    tune_n = wl_digits;
  }
  bool cached = false;
  if (!tune_cache.empty()) {
    clb_schedule s;
    const int r = clb_read_schedule(tune_cache.c_str(), &s);
    if (r != CLB_OK && r != CLB_EIO) {
      fprintf(stderr, "%s\n", clb_last_error());
      return usage();
    }
    cached = r == CLB_OK;
  }
  clb_ctx* ctx = nullptr;
  if (int rc = clb_ctx_create(device, precision, &ctx)) return die(rc);
  int rc = clb_set_kernel(ctx, kernel);
  if (!rc && replay) rc = clb_set_replay(ctx, replay);
  if (!rc && cached) {
    rc = clb_load_schedule(ctx, tune_cache.c_str());
    if (!rc) fprintf(stderr, "launch settings from %s\n", tune_cache.c_str());
  } else if (!rc && tune_ms > 0.0 && tune_n > 0) {
    rc = clb_autotune(ctx, tune_n, tune_ms, nullptr, nullptr);
    if (!rc && !tune_cache.empty()) {
      rc = clb_save_schedule(ctx, tune_cache.c_str());
      if (!rc) fprintf(stderr, "launch settings written to %s\n", tune_cache.c_str());
    }
  }
  if (!rc && cmd == "normal") {
    uint64_t w[3], n_out = 0;
    char hex[49];
    rc = clb_normal(ctx, digits, w, hex, &n_out);
    if (!rc) {
      double head, total;
      clb_last_timing(ctx, &head, &total);
      printf("digits=%llu\n%s\nAll calc time %.0fms\n", (unsigned long long)n_out, hex, total);
    }
  } else if (!rc && cmd == "batch") {
    if (graph) rc = clb_set_graph(ctx, 1);
    std::vector<uint64_t> frac(3 * pos.size());
    if (!rc && !pos.empty()) rc = clb_batch_positions(ctx, pos.data(), pos.size(), frac.data());
    for (size_t i = 0; !rc && i < pos.size(); ++i) {
      uint64_t n_out = 0;
      char hex[49];
      rc = clb_normalize(pos[i], &n_out, nullptr, nullptr);
      if (!rc) rc = format(&frac[3 * i], precision, replay != 0, hex);
      if (!rc) printf("%llu %s\n", (unsigned long long)n_out, hex);
    }
    if (!rc) {
      double head, total;
      clb_last_timing(ctx, &head, &total);
      fprintf(stderr, "%zu position(s), All calc time %.0fms\n", pos.size(), total);
    }
  } else if (!rc && cmd == "work") {
    int ran = 0;
    rc = clb_set_worklist_launch(ctx, (uint32_t)launch_jobs, launch_seconds, 0);
    if (!rc) rc = clb_run_worklist(ctx, positional.c_str(), workdir.c_str(), batch, nullptr, &ran);
    printf("%d job file(s) written to %s\n", ran, workdir.c_str());  // also after a failed launch
  }
  clb_ctx_destroy(ctx);
  return rc ? die(rc) : 0;
}
