// clb_format.cpp -- host-only entry points of libclbellard.so: digit normalisation, the formula
// table, the output formats of hexsademical.as, worklists, work files and the work-file summer
// (workmodeSum0_6.hsp), and the calling thread's error message.  No HIP: this file also builds
// with g++ under AddressSanitizer and UndefinedBehaviorSanitizer for the tests.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clbellard.h"
#include "clb_hostlib.hpp"

namespace clb {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int normalize(uint64_t n_in, uint64_t* n_out, uint64_t* d_out, int* overdouble) {
  if (n_in <= 4 || n_in > kMaxDigits)
    return fail(CLB_EINVAL, "invalid digits %llu (need 5 <= N <= %llu)",
                (unsigned long long)n_in, (unsigned long long)kMaxDigits);
  uint64_t d = ((n_in - 1) * 2 - 3) / 5;  // Bellard_192.hsp:468
  if (d % 2 == 0) d--;                    // :469
  if (n_out) *n_out = (d * 5 + 3) / 2 + 1;  // :470, a multiple of 5
  if (d_out) *d_out = d;
  if (overdouble) *overdouble = d * 10 + 9 >= 9007199254740992ull;  // :472
  return CLB_OK;
}

void hex48(const uint64_t w[3], char out[49]) {
  static const char hx[] = "0123456789abcdef";
  for (int i = 0; i < 3; ++i)
    for (int b = 0; b < 16; ++b) out[16 * i + b] = hx[(w[i] >> (60 - 4 * b)) & 15];
  out[48] = 0;
}

// HSP's int() of a double: truncation toward zero, as the x86 conversion the HSP runtime uses
// (cvttsd2si), which gives INT_MIN for NaN and values outside the int range (hexadecimal_dd
// reaches them after its borrow quirk, hexsademical.as:73-97); a C cast would be undefined there.
static int hsp_int_of(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
  return (int)v;
}

static void twosum(double a, double b, double* x, double* y) {
  double s = a + b;
  double t = s - a;
  *y = (a - (s - t)) + (b - t);
  *x = s;
}

// addcab of hexsademical.as:12-19.
static void dd_add(double* h, double* l, double bh, double bl) {
  double x, y;
  twosum(*h, bh, &x, &y);
  y = y + *l + bl;
  twosum(x, y, h, l);
}

// hexadecimal_dd, hexsademical.as:21-101: 14 base-256 digits of a double-double in [-1, 1),
// with carries and borrows pushed into the already written (more significant) bytes of a
// 24-byte little-endian buffer, printed as 48 hex digits (the last 20 are zero).
void hex_dd(const double dd[2], char out[49]) {
  double hh = dd[0], ll = dd[1];
  uint32_t words[6] = {0, 0, 0, 0, 0, 0};
  uint8_t* bytes = reinterpret_cast<uint8_t*>(words);
  if (hh < 0.0) dd_add(&hh, &ll, 1.0, 0.0);
  auto carry = [&](int idx, int dir) {
    if (idx >= 24) return;
    int b = idx / 4;
    uint32_t add = 1u << ((idx % 4) * 8);
    words[b] += dir > 0 ? add : (uint32_t)0 - add;
    for (int rep = 5 - b; rep > 0; --rep) {
      const int32_t v = (int32_t)words[b];
      const bool more = dir > 0 ? (v >= 0 && v < (int32_t)add) : (v < 0 && v >= -(int32_t)add);
      if (!more) break;
      ++b;
      add = 1;
      words[b] += dir > 0 ? 1u : 0xffffffffu;
    }
  };
  for (int cnt = 0; cnt < 14; ++cnt) {
    hh *= 256.0;
    ll *= 256.0;
    twosum(hh, ll, &hh, &ll);
    int digit = hsp_int_of(hh);
    if (digit < 256 && hh >= 0.0) {
      bytes[23 - cnt] = (uint8_t)digit;
      hh -= digit;
      continue;
    }
    if (digit == 256) {
      carry(23 - cnt + 1, +1);
      bytes[23 - cnt] = 0;
      hh -= 256.0;
    }
    if (hh < 0.0) {
      digit = hsp_int_of(hh + 256.0);
      carry(23 - cnt + 1, -1);
      bytes[23 - cnt] = (uint8_t)digit;
      dd_add(&hh, &ll, 256.0, 0.0);
    }
  }
  static const char hx[] = "0123456789abcdef";
  for (int i = 0; i < 24; ++i) {
    out[2 * i] = hx[bytes[23 - i] >> 4];
    out[2 * i + 1] = hx[bytes[23 - i] & 15];
  }
  out[48] = 0;
}

// Round a non-negative integer given as up to 3 words (most significant first) times 2^-192 to
// the nearest double (ties to even).
static double round_u192(const uint64_t w[3], bool* exact) {
  int lead = -1;
  for (int i = 0; i < 3 && lead < 0; ++i)
    if (w[i]) lead = (2 - i) * 64 + (63 - __builtin_clzll(w[i]));
  *exact = true;
  if (lead < 0) return 0.0;
  auto bit = [&](int p) -> int { return p < 0 ? 0 : (int)((w[2 - p / 64] >> (p % 64)) & 1); };
  uint64_t mant = 0;
  const int low = lead - 52;  // position of the last kept bit
  for (int p = lead; p >= low; --p) mant = (mant << 1) | (uint64_t)bit(p);
  bool sticky = false;
  for (int p = low - 2; p >= 0; --p)
    if (bit(p)) {
      sticky = true;
      break;
    }
  const int rbit = low - 1 >= 0 ? bit(low - 1) : 0;
  if (low - 1 >= 0 && (rbit || sticky)) *exact = false;
  if (rbit && (sticky || (mant & 1))) mant++;
  return std::ldexp((double)mant, low - 192);
}

int u192_to_dd(const uint64_t w[3], double dd[2]) {
  bool exact;
  const double hi = round_u192(w, &exact);
  // remainder w - hi * 2^192 as a signed 192-bit integer
  uint64_t h[3] = {0, 0, 0};
  int ex;
  const double fr = std::frexp(hi, &ex);  // hi = fr * 2^ex, fr in [0.5, 1)
  if (hi != 0.0) {
    const uint64_t mant = (uint64_t)std::ldexp(fr, 53);
    const int pos = ex - 53 + 192;  // bit position of mant's lsb
    for (int b = 0; b < 54; ++b)
      if ((mant >> b) & 1) {
        const int p = pos + b;
        if (p >= 0 && p < 192) h[2 - p / 64] |= 1ull << (p % 64);
      }
  }
  uint64_t r[3];
  uint64_t borrow = 0;
  for (int i = 2; i >= 0; --i) {
    const uint64_t a = w[i], b = h[i];
    const uint64_t d1 = a - b;
    const uint64_t b1 = a < b;
    const uint64_t d2 = d1 - borrow;
    const uint64_t b2 = d1 < borrow;
    r[i] = d2;
    borrow = b1 | b2;
  }
  const bool neg = (r[0] >> 63) != 0;
  if (neg) {
    uint64_t c = 1;
    for (int i = 2; i >= 0; --i) {
      r[i] = ~r[i] + c;
      c = (c && r[i] == 0) ? 1 : 0;
    }
  }
  bool ex2;
  const double lo = round_u192(r, &ex2);
  dd[0] = hi;
  dd[1] = neg ? -lo : lo;
  return CLB_OK;
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

static int read_text(const std::string& path, std::string* out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return fail(CLB_EIO, "cannot open %s: %s", path.c_str(), strerror(errno));
  out->clear();
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, n);
  fclose(f);
  return CLB_OK;
}

static std::vector<std::string> split_lines(const std::string& text) {
  std::vector<std::string> lines;
  std::string cur;
  for (char c : text) {
    if (c == '\n') {
      if (!cur.empty() && cur.back() == '\r') cur.pop_back();
      lines.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty() && cur.back() == '\r') cur.pop_back();
  if (!cur.empty()) lines.push_back(cur);
  return lines;
}

static void strrep(std::string* s, const std::string& from) {
  size_t p;
  while (!from.empty() && (p = s->find(from)) != std::string::npos) s->erase(p, from.size());
}

// HSP int() / int64() of a string: optional sign, then leading decimal digits (0 if none).
static int64_t hsp_int(const std::string& s) {
  size_t i = 0;
  bool neg = false;
  if (i < s.size() && (s[i] == '-' || s[i] == '+')) neg = s[i++] == '-';
  int64_t v = 0;
  for (; i < s.size() && s[i] >= '0' && s[i] <= '9'; ++i) {
    if (v > (INT64_MAX - 9) / 10) break;
    v = v * 10 + (s[i] - '0');
  }
  return neg ? -v : v;
}

static int make_dir(const std::string& p) {
  if (mkdir(p.c_str(), 0755) == 0 || errno == EEXIST) return CLB_OK;
  return fail(CLB_EIO, "cannot create %s: %s", p.c_str(), strerror(errno));
}

void format_result(int precision, const uint64_t w[3], char hex[49]) {
  if (precision == kDDWords) {
    double dd[2];
    memcpy(dd, w, sizeof dd);
    hex_dd(dd, hex);
  } else if (precision == CLB_PREC_DD106) {
    double dd[2];
    u192_to_dd(w, dd);
    hex_dd(dd, hex);
  } else {
    hex48(w, hex);
  }
}

// fsave, Bellard_192.hsp:514-533: hex line, device line, "<ms>ms"; HSP's "\n" is CR LF.
int write_work_file(const std::string& work_dir, int j, int64_t b, const uint64_t w[3],
                    int precision, const std::string& label, long long ms) {
  if (int rc = make_dir(work_dir)) return rc;
  const std::string dir = work_dir + "/" + std::to_string(j);
  if (int rc = make_dir(dir)) return rc;
  const std::string path = dir + "/" + std::to_string(j) + "_" + std::to_string(b) + ".txt";
  char hex[49];
  format_result(precision, w, hex);
  const std::string body = std::string(hex) + "\r\n" + label + "\r\n" + std::to_string(ms) + "ms";
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return fail(CLB_EIO, "cannot write %s: %s", path.c_str(), strerror(errno));
  const bool ok = fwrite(body.data(), 1, body.size(), f) == body.size();
  if (fclose(f) != 0 || !ok) return fail(CLB_EIO, "write failed: %s", path.c_str());
  return CLB_OK;
}

int parse_worklist(const char* path, Worklist* wl) {
  std::string text;
  if (int rc = read_text(path, &text)) return rc;
  const std::vector<std::string> lines = split_lines(text);
  if (lines.size() < 2) return fail(CLB_EFORMAT, "worklist needs digits= and precision= lines");
  std::string digits = lines[0];
  strrep(&digits, "digits");
  strrep(&digits, "=");
  strrep(&digits, " ");
  int64_t nd = hsp_int(digits);
  if (nd < 0) nd = 0;
  if ((uint64_t)nd > kMaxDigits) nd = (int64_t)kMaxDigits;
  std::string prec = lines[1];
  strrep(&prec, "precision");
  strrep(&prec, "=");
  strrep(&prec, " ");
  wl->digits = nd;
  wl->precision = (!prec.empty() && prec[0] >= '0' && prec[0] <= '2') ? prec[0] - '0' : -1;
  wl->jobs.clear();
  for (size_t li = 2; li < lines.size(); ++li) {
    const std::string& t = lines[li];
    if (t.empty()) continue;  // a blank line is not a job
    const int j = (int)hsp_int(t.substr(0, 1));
    const int64_t b = hsp_int(t.size() > 2 ? t.substr(2) : std::string());
    if (j < 0 || j > 6 || b < 0) return fail(CLB_EFORMAT, "bad job line '%s'", t.c_str());
    wl->jobs.push_back({j, b});
  }
  return CLB_OK;
}

// The launch-settings file: keys, and the ranges their setters accept.
struct SchedKey {
  const char* name;
  int clb_schedule::*field;
  int lo, hi;
};
static const SchedKey kSchedKeys[] = {
    {"terms_per_lane", &clb_schedule::terms_per_lane, 0, 16},
    {"min_tiles_per_wave", &clb_schedule::min_tiles_per_wave, 0, 1024},
    {"blocks_per_cu", &clb_schedule::blocks_per_cu, 0, 64},
    {"waves_per_group", &clb_schedule::waves_per_group, 0, 8},
    {"min_chunk", &clb_schedule::min_chunk, 0, 65536},
    {"chunks_per_wave", &clb_schedule::chunks_per_wave, -1, 64},
    {"max_replicas", &clb_schedule::max_replicas, 0, 64},
};

static bool is_pow2_or_zero(int v) { return v >= 0 && (v & (v - 1)) == 0; }

// Values the setters accept (clb_set_tiling, clb_set_launch, clb_set_schedule).
static int check_schedule(const clb_schedule& s) {
  const size_t len = strnlen(s.device, sizeof s.device);
  if (len == 0 || len == sizeof s.device) return fail(CLB_EINVAL, "device name empty or too long");
  for (size_t i = 0; i < len; ++i)
    if ((unsigned char)s.device[i] < 0x20 || s.device[i] == 0x7f)
      return fail(CLB_EINVAL, "device name holds a control character");
  for (const SchedKey& k : kSchedKeys)
    if (s.*k.field < k.lo || s.*k.field > k.hi)
      return fail(CLB_EINVAL, "%s = %d outside [%d, %d]", k.name, s.*k.field, k.lo, k.hi);
  const int t = s.terms_per_lane;
  if (t != 0 && t != 1 && t != 4 && t != 16)
    return fail(CLB_EINVAL, "terms_per_lane must be 0, 1, 4 or 16");
  const int w = s.waves_per_group;
  if (w != 0 && w != 1 && w != 2 && w != 4 && w != 8)
    return fail(CLB_EINVAL, "waves_per_group must be 0, 1, 2, 4 or 8");
  if (!is_pow2_or_zero(s.max_replicas))
    return fail(CLB_EINVAL, "max_replicas must be 0 or a power of 2");
  return CLB_OK;
}

// A decimal int with an optional minus sign, nothing else.
static bool parse_int(const std::string& v, int* out) {
  size_t i = v.size() && v[0] == '-' ? 1 : 0;
  if (i == v.size() || v.size() - i > 9) return false;
  long long x = 0;
  for (; i < v.size(); ++i) {
    if (v[i] < '0' || v[i] > '9') return false;
    x = x * 10 + (v[i] - '0');
  }
  *out = (int)(v[0] == '-' ? -x : x);
  return true;
}

// At most cap bytes of a file; CLB_EFORMAT if it holds more.
static int read_text_capped(const char* path, size_t cap, std::string* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return fail(CLB_EIO, "cannot open %s: %s", path, strerror(errno));
  out->assign(cap + 1, '\0');
  const size_t n = fread(&(*out)[0], 1, cap + 1, f);
  fclose(f);
  if (n > cap) return fail(CLB_EFORMAT, "%s: above %zu bytes", path, cap);
  out->resize(n);
  return CLB_OK;
}

}  // namespace clb

using namespace clb;

extern "C" {

const char* clb_last_error(void) { return g_last_error.c_str(); }

const char* clb_version(void) { return "clbellard-mi355x 0.1 (gfx950)"; }

int clb_normalize(uint64_t n_in, uint64_t* n_out, uint64_t* d_out, int* overdouble) {
  return normalize(n_in, n_out, d_out, overdouble);
}

int clb_formula(int j, uint64_t* numer, int64_t* numesign, uint64_t* den0, uint64_t* den1) {
  if (j < 0 || j > 6) return fail(CLB_EINVAL, "formula %d (need 0..6)", j);
  if (numer) *numer = kNumer[j];
  if (numesign) *numesign = kNumeSign[j];
  if (den0) *den0 = kDen0[j];
  if (den1) *den1 = kDen1[j];
  return CLB_OK;
}

int clb_hex48(const uint64_t w[3], char hex[49]) {
  if (!w || !hex) return fail(CLB_EINVAL, "null pointer");
  hex48(w, hex);
  return CLB_OK;
}

int clb_hex_dd(const double dd[2], char hex[49]) {
  if (!dd || !hex) return fail(CLB_EINVAL, "null pointer");
  if (!std::isfinite(dd[0]) || !std::isfinite(dd[1]) || std::fabs(dd[0]) > 1.0 ||
      std::fabs(dd[1]) > 1.0)
    return fail(CLB_EINVAL, "double-double (%g, %g) outside [-1, 1]", dd[0], dd[1]);
  hex_dd(dd, hex);
  return CLB_OK;
}

int clb_u192_to_dd(const uint64_t w[3], double dd[2]) {
  if (!w || !dd) return fail(CLB_EINVAL, "null pointer");
  return u192_to_dd(w, dd);
}

int clb_parse_worklist(const char* worklist_path, uint64_t* digits, int* precision,
                       int64_t* jobs, size_t cap, size_t* njobs) {
  if (!worklist_path) return fail(CLB_EINVAL, "null path");
  Worklist wl;
  if (int rc = parse_worklist(worklist_path, &wl)) return rc;
  if (digits) *digits = (uint64_t)wl.digits;
  if (precision) *precision = wl.precision;
  if (njobs) *njobs = wl.jobs.size();
  for (size_t i = 0; jobs && i < wl.jobs.size() && i < cap; ++i) {
    jobs[2 * i] = wl.jobs[i].first;
    jobs[2 * i + 1] = wl.jobs[i].second;
  }
  return CLB_OK;
}

int clb_write_work_file(const char* work_dir, int j, int64_t b, const uint64_t w[3],
                        int precision, const char* device_label, int64_t ms) {
  if (!work_dir || !w) return fail(CLB_EINVAL, "null pointer");
  if (j < 0 || j > 6 || b < 0) return fail(CLB_EINVAL, "bad job %d_%lld", j, (long long)b);
  if (precision < 0 || precision > 2) return fail(CLB_EINVAL, "precision must be 0, 1 or 2");
  return write_work_file(work_dir, j, b, w, precision, device_label ? device_label : "", ms);
}


int clb_read_schedule(const char* path, clb_schedule* out) {
  if (!path || !out) return fail(CLB_EINVAL, "null pointer");
  std::string text;
  if (int rc = read_text_capped(path, 4096, &text)) return rc;
  clb_schedule s;
  memset(&s, 0, sizeof s);
  bool seen_device = false;
  bool seen[sizeof kSchedKeys / sizeof kSchedKeys[0]] = {};
  for (const std::string& line : split_lines(text)) {
    if (line.empty()) continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) return fail(CLB_EFORMAT, "%s: line without '='", path);
    const std::string key = line.substr(0, eq), val = line.substr(eq + 1);
    if (key == "device") {
      if (seen_device) return fail(CLB_EFORMAT, "%s: device given twice", path);
      if (val.size() >= sizeof s.device) return fail(CLB_EINVAL, "%s: device name too long", path);
      memcpy(s.device, val.data(), val.size());
      s.device[val.size()] = 0;
      seen_device = true;
      continue;
    }
    size_t k = 0;
    while (k < sizeof kSchedKeys / sizeof kSchedKeys[0] && key != kSchedKeys[k].name) ++k;
    if (k == sizeof kSchedKeys / sizeof kSchedKeys[0])
      return fail(CLB_EFORMAT, "%s: unknown key '%.64s'", path, key.c_str());
    if (seen[k]) return fail(CLB_EFORMAT, "%s: %s given twice", path, kSchedKeys[k].name);
    if (!parse_int(val, &(s.*kSchedKeys[k].field)))
      return fail(CLB_EFORMAT, "%s: %s is not an integer", path, kSchedKeys[k].name);
    seen[k] = true;
  }
  if (!seen_device) return fail(CLB_EFORMAT, "%s: no device", path);
  for (size_t k = 0; k < sizeof kSchedKeys / sizeof kSchedKeys[0]; ++k)
    if (!seen[k]) return fail(CLB_EFORMAT, "%s: no %s", path, kSchedKeys[k].name);
  if (int rc = check_schedule(s)) return rc;
  *out = s;
  return CLB_OK;
}

int clb_write_schedule(const char* path, const clb_schedule* s) {
  if (!path || !s) return fail(CLB_EINVAL, "null pointer");
  if (int rc = check_schedule(*s)) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(CLB_EIO, "cannot create %s: %s", path, strerror(errno));
  int bad = fprintf(f, "device=%s\n", s->device) < 0;
  for (const SchedKey& k : kSchedKeys) bad |= fprintf(f, "%s=%d\n", k.name, s->*k.field) < 0;
  bad |= fclose(f) != 0;
  if (bad) return fail(CLB_EIO, "cannot write %s", path);
  return CLB_OK;
}

int clb_work_sum(const char* work_dir, char hex[49], int counts[7]) {
  if (!work_dir || !hex) return fail(CLB_EINVAL, "null pointer");
  int sum[48] = {0};
  for (int j = 0; j < 7; ++j) {
    int used = 0;
    for (long long b = 0;; ++b) {
      char name[64];
      snprintf(name, sizeof name, "/%d/%d_%lld.txt", j, j, b);
      const std::string path = std::string(work_dir) + name;
      if (!file_exists(path)) break;  // workmodeSum0_6.hsp:14-16: stop at the first gap
      std::string text;
      if (int rc = read_text(path, &text)) return rc;
      const std::vector<std::string> lines = split_lines(text);
      if (lines.empty() || lines[0].size() < 48)
        return fail(CLB_EFORMAT, "%s: first line is not 48 hex digits", path.c_str());
      const std::string& line = lines[0];
      for (int c = 0; c < 48; ++c) {  // :21-28, least significant digit first
        const char ch = line[47 - c];
        int v;
        if (ch >= '0' && ch <= '9')
          v = ch - '0';
        else if (ch >= 'a' && ch <= 'f')
          v = ch - 'a' + 10;
        else
          return fail(CLB_EFORMAT, "%s: '%c' is not a lowercase hex digit", path.c_str(), ch);
        sum[47 - c] += v;
        if (sum[47 - c] >= 16) {
          sum[47 - c] -= 16;
          if (c != 47) sum[46 - c]++;
        }
      }
      ++used;
    }
    if (counts) counts[j] = used;
  }
  static const char hx[] = "0123456789abcdef";
  for (int i = 0; i < 48; ++i) hex[i] = hx[sum[i]];
  hex[48] = 0;
  return CLB_OK;
}

}  // extern "C"
