// clb_cli.cpp -- command-line front end over libclbellard.so, replacing the reference's GUI
// (Bellard_192.hsp normal mode and work mode) and workmodeSum0_6.exe.
//
//   clbellard normal <digits> [--precision 0|1|2] [--device i] [--kernel auto|plain|split]
//                    [--replay | --replay-gws GWS]
//   clbellard work <worklist.txt> [--workdir work] [--batch-size B] [--device i]
//                  [--replay | --replay-gws GWS]
//   clbellard sum [--workdir work]
// --replay: the reference kernels' own output bits (clb_set_replay) with the reference's
// 16777216 work-items per launch; --replay-gws: with GWS work-items per launch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/clbellard.h"

static int usage() {
  fprintf(stderr,
          "usage:\n"
          "  clbellard normal <digits> [--precision 0|1|2] [--device i] [--kernel auto|plain|split]\n"
          "                   [--replay | --replay-gws GWS]\n"
          "  clbellard work <worklist.txt> [--workdir DIR] [--batch-size B] [--device i]\n"
          "                 [--replay | --replay-gws GWS]\n"
          "  clbellard sum [--workdir DIR]\n");
  return 2;
}

static int die(int rc) {
  fprintf(stderr, "error %d: %s\n", rc, clb_last_error());
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  const std::string cmd = argv[1];
  int precision = 1, device = 0, kernel = CLB_KERNEL_AUTO;
  std::string workdir = "work", positional;
  unsigned long long batch = 0, replay = 0;
  for (int i = 2; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (a == "--precision" && (v = next())) {
      precision = atoi(v);
    } else if (a == "--device" && (v = next())) {
      device = atoi(v);
    } else if (a == "--workdir" && (v = next())) {
      workdir = v;
    } else if (a == "--batch-size" && (v = next())) {
      batch = strtoull(v, nullptr, 10);
    } else if (a == "--kernel" && (v = next())) {
      kernel = !strcmp(v, "plain") ? CLB_KERNEL_PLAIN : !strcmp(v, "split") ? CLB_KERNEL_SPLIT
                                                                            : CLB_KERNEL_AUTO;
    } else if (a == "--replay") {
      replay = CLB_REF_GLOBAL_WORK_SIZE;
    } else if (a == "--replay-gws" && (v = next())) {
      replay = strtoull(v, nullptr, 10);
    } else if (positional.empty() && a[0] != '-') {
      positional = a;
    } else {
      return usage();
    }
  }
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
  clb_ctx* ctx = nullptr;
  if (int rc = clb_ctx_create(device, precision, &ctx)) return die(rc);
  int rc = clb_set_kernel(ctx, kernel);
  if (!rc && replay) rc = clb_set_replay(ctx, replay);
  if (!rc && cmd == "normal") {
    uint64_t w[3], n_out = 0;
    char hex[49];
    rc = clb_normal(ctx, strtoull(positional.c_str(), nullptr, 10), w, hex, &n_out);
    if (!rc) {
      double head, total;
      clb_last_timing(ctx, &head, &total);
      printf("digits=%llu\n%s\nAll calc time %.0fms\n", (unsigned long long)n_out, hex, total);
    }
  } else if (!rc && cmd == "work") {
    int ran = 0;
    rc = clb_run_worklist(ctx, positional.c_str(), workdir.c_str(), batch, nullptr, &ran);
    if (!rc) printf("%d job file(s) written to %s\n", ran, workdir.c_str());
  } else if (!rc) {
    clb_ctx_destroy(ctx);
    return usage();
  }
  clb_ctx_destroy(ctx);
  return rc ? die(rc) : 0;
}
