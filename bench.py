#!/usr/bin/env python3
"""Benchmark: Bellard terms/s at hex position N = 10^12 on one MI355X.

One step = the complete normal-mode computation at N = 10^12 (d = 399,999,999,999): the seven
formulas' head terms k in [0, d) in one persistent launch of the precision-split (fp64 exact)
kernel, the 7 x 20 tail terms, and the combine.  value = d / (seconds per step), one "term" being
one k over all seven formulas (seven modular exponentiations).

Also reported, outside the timed steps:
  * the plain all-integer kernel and the split kernel on the same seeded windows (64 windows of
    2^26 consecutive k per formula, starts from SplitMix64 seed 2), bit-compared;
  * the same windows summed by the CPU oracle on a bounded sample, bit-compared, and its rate
    (cpu_baseline, kind "port");
  * roofline of the dominant kernel (the head-term kernel) from its HIP-event launch time.

    python bench.py [--gpus 1] [--steps 2] [--warmup 1]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Bellard terms/sec (7 modexps each) at hex pos 10^12; frac err vs exact oracle"
POSITION = 10**12
PEAK_FP64_TFLOPS = 78.6  # MI355X published FP64 peak (one FMA = 2 flops per lane per cycle)
# VALU issue: one wave64 instruction per CU per clock = 64 lane-operations; 256 CUs at 2.4 GHz
PEAK_VALU_LANE_SLOTS_PER_S = 256 * 64 * 2.4e9
# 32x32->64-bit multiplies (v_mad_u64_u32, v_mul_hi_u32) at full rate: 256 CUs x 64 lanes x 2.4 GHz
PEAK_INT32_MUL_PER_S = 256 * 64 * 2.4e9
NUMER_SUM = 32 + 1 + 256 + 64 + 4 + 4 + 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def splitmix64(seed):
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def sum_bitlen_10x(d):
    """sum over x = 1..d of bitlen(10 x): the exponents 10 (d - k) of the head terms."""
    total, b = 0, 1
    while True:
        lo = max(1, -(-(1 << (b - 1)) // 10))
        hi = min(d, ((1 << b) - 1) // 10)
        if lo > d:
            break
        if hi >= lo:
            total += b * (hi - lo + 1)
        b += 1
    return total


def algorithmic_fp64_slots(d):
    """FP64 issue slots one normal-mode launch needs at minimum: per head term and formula,
    6 per modular square (exact product hi/lo, quotient estimate, rounding, residue, sum) for
    bitlen(e) - 1 squares, plus 16 for the four 48-bit quotient digits."""
    squares = sum_bitlen_10x(d) - d
    return 7 * (6 * squares + 16 * d)


def algorithmic_fp64_flops(d):
    """FP64 flops in those slots (fma = 2, mul and add = 1, rounding = 0): 7 per modular square
    (mul, fma, mul, rint, fma, add), 20 per term's four digits (two fmas each, and the sums)."""
    squares = sum_bitlen_10x(d) - d
    return 7 * (7 * squares + 20 * d)


def int32_word_products(d):
    """32x32-bit word products a 2-limb (64-bit) Montgomery engine needs for one normal-mode
    launch: per head term and formula, bitlen(e) - 1 squares and 3 quotient-word steps of
    2 s^2 + s = 10 word products each (CIOS, s = 2 limbs)."""
    squares = sum_bitlen_10x(d) - d
    return 7 * 10 * (squares + 3 * d)


# 32x32-bit multiply instructions (v_mad_u64_u32, v_mul_lo_u32, v_mul_hi_u32) of the plain
# kernel, counted in its gfx950 ISA (k_head<3, 16> and its device functions compiled alone): 31
# per 3-bit window of three Montgomery squares (10 per plain square, 11 with the window digit's
# shift), 46 per term besides (the first square 11, -m^-1 10, the three exact-division quotient
# words 24, the carried 2^64 mod m 1).
PLAIN_MULS_PER_WINDOW = 31
PLAIN_MULS_PER_TERM = 46


def plain_multiplies(d, k0, k1):
    """Multiply instructions the plain kernel issues for the head terms k in [k0, k1) of one
    formula: its Montgomery exponent is e + 192 = 10 (d - k) + 192, taken in 3-bit windows from
    the top (one square for the top window, three for each further one)."""
    total = 0
    lo_x, hi_x = d - k1 + 1, d - k0  # x = d - k
    bl = 1
    while True:
        # x with bitlen(10 x + 192) == bl: 10 x + 192 in [2^(bl-1), 2^bl - 1]
        a = max(lo_x, -(-((1 << (bl - 1)) - 192) // 10))
        b = min(hi_x, ((1 << bl) - 1 - 192) // 10)
        if a > hi_x:
            break
        if b >= a:
            windows = (bl + 2) // 3
            total += (b - a + 1) * (PLAIN_MULS_PER_WINDOW * (windows - 1) + PLAIN_MULS_PER_TERM)
        bl += 1
    return total


def cpu_model() -> str:
    """The host CPU's model name and logical CPU count, as /proc/cpuinfo gives them."""
    try:
        with open("/proc/cpuinfo") as f:
            names = [ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")]
        if names:
            return f"{names[0]} ({len(names)} logical CPUs)"
    except OSError:
        pass
    return "unknown CPU"


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get("head_kernel_hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def reference_digits_check(clb, hexs, d):
    """The exact mode's N = 10^12 output against the reference kernels' own outputs there
    (tests/golden/ref_positions.json, measured on MI355X): leading hex digits in common and the
    distance, which must stay within the double-double kernels' bound (2 units of 2^-192 per term
    and c_j, plus the tail kernel; the 106-bit mode: 2^-104 per addition)."""
    path = os.path.join(ROOT, "tests", "golden", "ref_positions.json")
    with open(path) as f:
        rows = json.load(f)["positions"]
    ours = int(hexs, 16)
    out = {}
    for row in rows:
        if row["n"] != 10**12:
            continue
        p = row["precision"]
        if p == 0:
            hi, lo = float.fromhex(row["hi"]), float.fromhex(row["lo"])
            from fractions import Fraction
            theirs = (Fraction(hi) + Fraction(lo)) * (1 << 192)
            diff = abs(theirs - ours)
            dist = float(diff) if diff else 0.0
            bound = 7 * (d + 20) << (192 - 104)
            common = 0
            ref_hex = clb.hex_dd(hi, lo)
            ours_hex = clb.hex_dd(*clb.u192_to_dd(clb.int_to_u192(ours)))
        else:
            theirs = int(row["hex"], 16)
            dist = abs(theirs - ours)
            bound = NUMER_SUM * (2 * (d + 20) + (1 << 16))
            ref_hex, ours_hex = row["hex"], hexs
        common = next((i for i, (a, b) in enumerate(zip(ref_hex, ours_hex)) if a != b),
                      len(ours_hex))
        import math
        out[f"precision_{p}"] = {"hex_digits_equal": common,
                                 "distance_log2": round(math.log2(dist) - 192, 2) if dist else None,
                                 "bound_log2": round(math.log2(bound) - 192, 2)}
        if dist > bound or common < (20 if p == 0 else 36):
            raise SystemExit(f"exact mode {hexs} vs reference precision {p} {ref_hex}: "
                             f"{common} digits equal, distance beyond the bound")
    return out


def measure_frac_err(clb, orc, ctx, d, starts, span, threads):
    """Fraction error of each mode against the 256-bit oracle (orc_job256: every term to 2^-256,
    the tail to k = d + 25) on sub-windows of `span` consecutive k at the seeded window starts,
    all seven formulas.  Modes: the exact kernels (split and plain, bit-identical), and the replay
    of the reference's precision 2, 1 and 0 kernels launched with span / 32 work-items, so that
    every work-item sums 32 terms as in the reference's full launches (Bellard.cl:992-1084,
    KernelFuncs.as:45-53).  Returns per mode the largest window error, the mean error per term
    and the whole-position error extrapolated from the sample (the windows are uniform in
    [0, d)), as log2 of the fraction, and the sampled sums."""
    import math
    import struct
    from fractions import Fraction
    from clb_testlib import frac_error_256
    jobs = [(d, st, st + span, j) for j in range(7) for st in starts]
    t0 = time.perf_counter()
    exact256 = [orc.job256(dd, k0, k1, j, threads=threads) for dd, k0, k1, j in jobs]
    cpu_s = time.perf_counter() - t0
    modes = {}
    ctx.set_kernel(clb.KERNEL_SPLIT)
    modes["exact"] = [clb.u192_to_int(w) for w in ctx.run_jobs(jobs)]
    ctx.set_kernel(clb.KERNEL_PLAIN)
    plain = [clb.u192_to_int(w) for w in ctx.run_jobs(jobs)]
    if plain != modes["exact"]:
        raise SystemExit("plain and split kernels disagree on the error windows")
    ctx.set_kernel(clb.KERNEL_AUTO)
    gws = max(256, (span // 32) // 256 * 256)
    for prec in (clb.PREC_U192_BETA, clb.PREC_U192, clb.PREC_DD106):
        with clb.Context(0, prec) as rc:
            rc.set_replay(gws)
            res = rc.run_jobs(jobs)
        if prec == clb.PREC_DD106:  # double-double bit patterns (hi, lo) -> 256-bit integers
            vals = []
            for w in res:
                hi, lo = (struct.unpack("<d", struct.pack("<Q", x))[0] for x in w[:2])
                v = (Fraction(hi) + Fraction(lo)) % 1
                vals.append(int(v * (1 << 256)))
            modes["replay_p0"] = [("dd", v) for v in vals]
        else:
            modes[f"replay_p{prec}"] = [clb.u192_to_int(w) for w in res]
    terms = 7 * span * len(starts)
    out = {"what": "signed error of each mode's sum against the 256-bit oracle, log2 of |error| "
                   "as a fraction: the largest of the 448 sampled windows, the sum over all of "
                   "them, the mean per term and formula, and the whole-position error "
                   "extrapolated linearly from the sample (an upper estimate where rounding "
                   "errors of both signs cancel, as in the replay modes)"}
    for name, vals in modes.items():
        errs = []
        for v, x in zip(vals, exact256):
            if isinstance(v, tuple):
                e = (v[1] - x) & ((1 << 256) - 1)
                e = e - (1 << 256) if e > (1 << 255) else e
            else:
                e = frac_error_256(v, x)
            errs.append(e)
        worst = max(abs(e) for e in errs)
        total = sum(errs)
        full = total * (d / (span * len(starts)))

        def lg(u):
            return round(math.log2(abs(u)) - 256, 2) if u else None
        out[name] = {"max_window_log2": lg(worst), "sample_sum_log2": lg(total),
                     "mean_per_term_formula_log2": lg(total / terms),
                     "full_position_estimate_log2": lg(full)}
    out["plain"] = "bit-identical to exact"
    return out, cpu_s, terms


def bench_batched(clb, args):
    """Many digit positions (log-uniform in [10^6, 10^11], SplitMix64 seed 0, multiples of 5) in
    one persistent launch; a sample is re-run one position at a time and compared bit for bit."""
    rng = splitmix64(0)
    pos = []
    for _ in range(args.positions):
        u = (next(rng) >> 11) * 2.0**-53
        pos.append(5 * round(10 ** (6 + 5 * u) / 5))
    terms = sum(clb.normalize(p)[1] for p in pos)
    ctx = clb.Context(0, clb.PREC_U192)
    # tuned where the work is: the position below which half of the terms lie (the count median
    # of a log-uniform draw holds a negligible share of them)
    ranked = sorted(pos)
    half, acc, median = terms / 2, 0, ranked[-1]
    for p in ranked:
        acc += clb.normalize(p)[1]
        if acc >= half:
            median = p
            break
    tpl, bpc = ctx.autotune(median, 3000.0)
    sched = ctx.get_schedule()
    log(f"autotune at N={median} (term-weighted median): {tpl} terms per lane, {bpc} workgroups "
        f"per CU, {sched}")
    ctx.set_graph(not args.no_graph)
    log(f"batched: {len(pos)} positions, {terms:.4g} terms, "
        f"{'eager launches' if args.no_graph else 'captured HIP graph'}")
    fracs, dt, head_ms = [], 0.0, 0.0
    per = -(-len(pos) // args.batches)
    for b in range(0, len(pos), per):
        t0 = time.perf_counter()
        fracs += ctx.batch_positions(pos[b:b + per])
        dt += time.perf_counter() - t0
        head_ms += ctx.last_timing()[0]
        log(f"launch of positions {b}..{min(b + per, len(pos)) - 1}: {dt:.2f}s so far")
    pick = sorted(range(len(pos)), key=lambda i: pos[i])
    sample = [pick[int(i * (len(pick) - 1) / max(1, args.check - 1))] for i in range(args.check)]
    for i in sample:
        if ctx.normal(pos[i])[0] != fracs[i]:
            raise SystemExit(f"position {pos[i]}: batched result differs from the single run")
    log(f"{len(sample)} positions re-run alone: bit-identical")
    print(json.dumps({
        "metric": "Bellard terms/sec, batched digit positions", "value": terms / dt,
        "unit": "terms/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(pos)} positions log-uniform in [1e6, 1e11], seed 0, "
                               f"{args.batches} persistent launch(es)", "positions": len(pos),
                   "terms": terms,
                   "graph": not args.no_graph, "terms_per_lane": tpl,
                   "workgroups_per_cu": bpc, "schedule": sched, "tuned_at": median,
                   "head_kernel_ms": head_ms,
                   "checked_single": len(sample)}}), flush=True)
    ctx.close()
    return 0


def bench_small(clb, args):
    """Call latency of small positions: back-to-back clb_normal calls at N = 10^5 (d = 39999),
    10^6 and 10^7, with the tile length chosen per launch (auto) or forced to 16, 4 or 1 terms
    per lane, and auto with the captured-graph replay; every mode's digits compared."""
    modes = [("auto", 0, False), ("tpl16", 16, False), ("tpl4", 4, False), ("tpl1", 1, False),
             ("auto+graph", 0, True)]
    table = {}
    for n in (100000, 1000000, 10000000):
        calls = max(5, args.calls * 100000 // n)
        ref = None
        row = {}
        for name, tpl, graph in modes:
            ctx = clb.Context(0, clb.PREC_U192)
            ctx.set_tiling(tpl)
            ctx.set_graph(graph)
            ctx.normal(n)  # warm-up (and capture)
            t0 = time.perf_counter()
            for _ in range(calls):
                frac = ctx.normal(n)[0]
            dt = time.perf_counter() - t0
            ctx.close()
            if ref is None:
                ref = frac
            elif frac != ref:
                raise SystemExit(f"N={n} {name}: digits differ from the first mode")
            row[name] = round(1e6 * dt / calls, 1)
        table[str(n)] = row
        log(f"N={n}: " + ", ".join(f"{k} {v} us" for k, v in row.items()))
    best = table["100000"]["auto"]
    print(json.dumps({
        "metric": "clb_normal call latency at N=1e5", "value": best, "unit": "us/call",
        "n_gpus": 1, "steps": args.calls, "warmup": 1, "ms_per_step": best / 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "back-to-back normal-mode calls at N=1e5, 1e6, 1e7 per tile "
                               "length and launch mode (us per call)", "table": table}}),
          flush=True)
    return 0


def bench_worklist(clb, args):
    """64 uneven term-range jobs (lengths U[1e8, 1e11], SplitMix64 seed 1, rescaled to tile
    [0, d) at N = 2.5e12) x 7 formulas in one launch, written as work files, summed like
    workmodeSum0_6, and compared with the single normal-mode run."""
    import tempfile
    n_pos = args.position or 2500000000000
    n_pos, d, _ = clb.normalize(n_pos)
    rng = splitmix64(1)
    lens = [10**8 + next(rng) % (10**11 - 10**8 + 1) for _ in range(64)]
    total, acc, cuts = sum(lens), 0, [0]
    for ln in lens:
        acc += ln
        cuts.append((d * acc + total // 2) // total)
    jobs = [(d, cuts[i], cuts[i + 1] if i < 63 else d + 1, j) for j in range(7) for i in range(64)]
    ctx = clb.Context(0, clb.PREC_U192)
    log(f"worklist: N={n_pos} d={d}, {len(jobs)} jobs")
    t0 = time.perf_counter()
    parts = ctx.run_jobs(jobs)
    dt = time.perf_counter() - t0
    head_ms = ctx.last_timing()[0]
    log(f"jobs launch: {dt:.2f}s")
    with tempfile.TemporaryDirectory() as work:
        for idx, (jd, k0, k1, j) in enumerate(jobs):
            clb.write_work_file(work, j, idx % 64, parts[idx], clb.PREC_U192, "0\tMI355X",
                                int(1e3 * dt))
        summed, counts = clb.work_sum(work)
    t1 = time.perf_counter()
    single, single_hex, _ = ctx.normal(n_pos)
    dt_single = time.perf_counter() - t1
    if summed != single_hex or counts != [64] * 7:
        raise SystemExit(f"work-file sum {summed} != single run {single_hex}")
    log(f"summed work files == single run ({dt_single:.2f}s): {summed}")
    print(json.dumps({
        "metric": "Bellard terms/sec, uneven worklist jobs", "value": d / dt, "unit": "terms/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * dt, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"64 uneven jobs x 7 formulas tiling [0, d) at N={n_pos}, one "
                               "persistent launch, work files summed", "position": n_pos, "d": d,
                   "head_kernel_ms": head_ms, "single_run_s": dt_single, "digits": summed}}),
          flush=True)
    ctx.close()
    return 0


def bench_peaks(clb, args):
    """Issue rates of the instructions the term kernels are made of, and of the two engines' whole
    modular squares (clb_issue_rate: 8 independent chains per lane, 8 waves per SIMD on every CU),
    per CU per clock both at the nominal 2.4 GHz and at the shader clock measured during each
    run.  The VALU can issue one wave64 instruction per CU per clock, i.e. 64 lane-operations."""
    rows = {}
    for op, name in enumerate(clb.PEAK_OPS):
        r = clb.issue_rate(op, args.peak_ms)
        cus = r["cus"]
        rows[name] = {
            "per_s": r["per_s"], "sclk_ghz": round(r["sclk_ghz"], 3), "ms": round(r["ms"], 1),
            "per_cu_per_clk_nominal": round(r["per_s"] / (cus * 2.4e9), 2),
            "per_cu_per_clk_measured": round(r["per_s"] / (cus * r["sclk_ghz"] * 1e9), 2)
            if r["sclk_ghz"] > 0 else None}
        log(f"{name:24s} {r['per_s']:.4e}/s  sclk {r['sclk_ghz']:.3f} GHz  "
            f"{rows[name]['per_cu_per_clk_nominal']:6.2f} per CU per clock at 2.4 GHz, "
            f"{rows[name]['per_cu_per_clk_measured']} at the measured clock")
    fma = rows["v_fma_f64"]
    print(json.dumps({
        "metric": "VALU issue rate, v_fma_f64 lanes per CU per clock (measured clock)",
        "value": fma["per_cu_per_clk_measured"], "unit": "lanes/CU/clk", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": fma["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "issue-rate microbenchmark: 8 independent chains per lane, 8 "
                               "waves per SIMD, every CU", "rates": rows}}), flush=True)
    return 0


def bench_schedule(clb, args):
    """Sweep of the term kernel's launch knobs (clb_set_schedule, clb_set_tiling, clb_set_launch)
    on the normal-mode computation at --position (default 10^11 here, 1.8 s per run) and on small
    positions' call latency; every setting's digits must equal the first's.  Knobs: workgroup size
    (1, 2, 4, 8 waves, each at full occupancy, i.e. 8 waves per SIMD), the tile queue (fixed
    chunks of 256 or 64 tiles, or the guided schedule with 1, 2 or 4 chunks per wave per phase),
    the slot copies (1, 4, 64 at most) and resident workgroups per CU."""
    n = args.position or 10**11
    n, d, _ = clb.normalize(n)
    ctx = clb.Context(0, clb.PREC_U192)
    ref = ctx.normal(n)[0]

    clocks = {}

    def timed(label, setup, pos=n, reps=2):
        setup()
        best = None
        for _ in range(reps):
            frac = ctx.normal(pos)[0]
            h, t = ctx.last_timing()
            if pos == n and frac != ref:
                raise SystemExit(f"{label}: digits differ")
            if best is None or h < best:
                best, ghz = h, ctx.last_clock()[0]
        ctx.set_schedule()
        ctx.set_launch(0)
        ctx.set_tiling(0)
        clocks[label] = {"sclk_ghz": round(ghz, 3), "ms_x_ghz": round(best * ghz, 1)}
        log(f"{label:40s} head kernel {best:10.3f} ms at {ghz:.3f} GHz ({best * ghz:.1f} ms*GHz)")
        return round(best, 3)

    rows = {}
    if args.knobs == "all":
        for w in (1, 2, 4, 8):
            rows[f"workgroup {64 * w}"] = timed(f"workgroup {64 * w}",
                                                lambda w=w: ctx.set_schedule(waves_per_group=w))
    queues = (("queue fixed 256 (default)", 256, -1), ("queue fixed 64", 64, -1),
              ("queue guided 1/wave, min 64", 64, 1), ("queue guided 2/wave, min 64", 64, 2),
              ("queue guided 4/wave, min 64", 64, 4), ("queue guided 2/wave, min 16", 16, 2))
    if args.knobs == "queue":  # the larger chunks, and guided queues ending in 256-tile chunks
        queues = (("queue fixed 256 (default)", 256, -1), ("queue fixed 1024", 1024, -1),
                  ("queue fixed 4096", 4096, -1), ("queue guided 4/wave, min 256", 256, 4),
                  ("queue guided 8/wave, min 256", 256, 8))
    for label, mc, cpw in queues:
        rows[label] = timed(label, lambda mc=mc, cpw=cpw: ctx.set_schedule(0, mc, cpw, 0))
    if args.knobs == "all":
        for bpc in (4, 6):
            rows[f"{bpc} workgroups per CU"] = timed(f"{bpc} workgroups per CU",
                                                     lambda bpc=bpc: ctx.set_launch(bpc))
    small = {}
    for pos in ((100000, 1000000, 10000000) if args.knobs == "all" else ()):
        ref_small = ctx.normal(pos)[0]
        for r in (1, 4, 64):
            ctx.set_schedule(max_replicas=r)
            best = None
            for _ in range(args.calls // 10 + 3):
                if ctx.normal(pos)[0] != ref_small:
                    raise SystemExit("slot copies change the digits")
                t = ctx.last_timing()[1]
                best = t if best is None else min(best, t)
            small[f"N={pos} max_replicas={r}"] = round(1e3 * best, 1)
            log(f"N={pos} max_replicas {r}: call {1e3 * best:.1f} us")
        for w in (1, 2, 4, 8):
            ctx.set_schedule(waves_per_group=w)
            best = None
            for _ in range(args.calls // 10 + 3):
                if ctx.normal(pos)[0] != ref_small:
                    raise SystemExit("workgroup size changes the digits")
                t = ctx.last_timing()[1]
                best = t if best is None else min(best, t)
            small[f"N={pos} workgroup {64 * w}"] = round(1e3 * best, 1)
            log(f"N={pos} workgroup {64 * w}: call {1e3 * best:.1f} us")
        ctx.set_schedule()
    ctx.close()
    base = rows["queue fixed 256 (default)"]
    print(json.dumps({
        "metric": f"term-kernel time at N={n} per launch setting (ms)", "value": base,
        "unit": "ms", "n_gpus": 1, "steps": 2, "warmup": 1, "ms_per_step": base,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"launch-knob sweep ({args.knobs}): normal mode at N={n} (best of "
                               "2) and call latency at N=1e5, 1e6, 1e7 (us)",
                   "head_kernel_ms": rows, "head_kernel_clock": clocks,
                   "small_call_us": small}}), flush=True)
    return 0


def bench_replay(clb, args):
    """Normal mode at N (default 10^12) replaying the reference's kernels of precision 2, 1 and 0
    (clb_set_replay with the reference's 2^24 work-items per launch): the reference's own output
    bits, its double-double rounding included.  The leading 24 hex digits (16 in the 106-bit mode)
    are compared with the exact mode's."""
    n_pos, d, _ = clb.normalize(args.position or POSITION)
    exact = None
    rows = {}
    for prec in (clb.PREC_U192_BETA, clb.PREC_U192, clb.PREC_DD106):
        ctx = clb.Context(0, prec)
        if exact is None:
            exact = ctx.normal(n_pos)[1]
            log(f"exact mode: {exact}")
        ctx.set_replay(clb.REF_GLOBAL_WORK_SIZE)
        t0 = time.perf_counter()
        _, hexs, _ = ctx.normal(n_pos)
        dt = time.perf_counter() - t0
        ctx.close()
        agree = 16 if prec == clb.PREC_DD106 else 24
        if hexs[:agree] != exact[:agree]:
            raise SystemExit(f"precision {prec} replay {hexs} disagrees with exact {exact}")
        rows[str(prec)] = {"seconds": round(dt, 3), "digits": hexs}
        log(f"replay precision {prec}: {dt:.2f}s {hexs}")
    print(json.dumps({
        "metric": "Bellard terms/sec, reference replay (precision 1)",
        "value": d / rows["1"]["seconds"],
        "unit": "terms/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * rows["1"]["seconds"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"normal mode at N={n_pos} replaying the reference's kernels of "
                               "precision 2, 1 and 0 (2^24 work-items per launch)",
                   "position": n_pos,
                   "d": d, "precision": rows}}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--position", type=int, default=None,
                    help="digit position (default per workload: normal and replay 10^12, "
                         "worklist 2.5x10^12, schedule 10^11)")
    ap.add_argument("--windows", type=int, default=64)
    ap.add_argument("--window-log2", type=int, default=26)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--err-window-log2", type=int, default=18,
                    help="k per seeded window in the fraction-error measurement")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--peak-ms", type=float, default=200.0,
                    help="peaks: target duration of each microbenchmark run")
    ap.add_argument("--workload", choices=["normal", "batched", "worklist", "small", "replay",
                                           "peaks", "schedule"],
                    default="normal",
                    help="normal: the metric (N = 10^12); batched: many positions in one launch; "
                         "worklist: uneven term-range jobs written as work files and summed; "
                         "small: call latency of small positions, eager vs captured graph; "
                         "replay: the reference's precision-1/2 output bits at N; "
                         "peaks: VALU issue-rate microbenchmarks; "
                         "schedule: sweep of the launch knobs")
    ap.add_argument("--batches", type=int, default=8,
                    help="batched: split the positions over this many launches (each launch "
                         "prints when it ends; 4096 positions in one launch run ~10 minutes "
                         "without a word)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched: launch the kernels one by one instead of a captured graph")
    ap.add_argument("--calls", type=int, default=200, help="small: calls per mode")
    ap.add_argument("--knobs", choices=["all", "queue"], default="all",
                    help="schedule: every knob, or only the tile queue (fixed 256, 1024, 4096 "
                         "tiles per chunk and guided queues ending in 256-tile chunks)")
    ap.add_argument("--positions", type=int, default=4096, help="batched: number of positions")
    ap.add_argument("--check", type=int, default=8,
                    help="batched: positions re-run alone and compared bit for bit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        a = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} bytes")
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); across GPUs the "
              "digit extraction would shard by term ranges (work-mode jobs), one process per GPU, "
              "with one 24-byte partial per job summed on the host", file=sys.stderr)
        return 2

    import clbellard_amd as clb

    if args.workload == "batched":
        return bench_batched(clb, args)
    if args.workload == "worklist":
        return bench_worklist(clb, args)
    if args.workload == "small":
        return bench_small(clb, args)
    if args.workload == "replay":
        return bench_replay(clb, args)
    if args.workload == "peaks":
        return bench_peaks(clb, args)
    if args.workload == "schedule":
        return bench_schedule(clb, args)

    n_pos, d, _ = clb.normalize(args.position or POSITION)
    ctx = clb.Context(0, clb.PREC_U192, kernel=clb.KERNEL_SPLIT)
    log(f"device {clb.device_name(0)}; N={n_pos} d={d}")

    for i in range(args.warmup):
        t0 = time.perf_counter()
        ctx.normal(n_pos)
        log(f"warmup {i}: {time.perf_counter() - t0:.2f}s")

    head_ms, step_s, sclk, result = [], [], [], None
    for i in range(args.steps):
        t0 = time.perf_counter()
        frac, hexs, _ = ctx.normal(n_pos)
        dt = time.perf_counter() - t0
        h, tot = ctx.last_timing()
        ghz = ctx.last_clock()[0]  # after the timed call: a 256-byte read-back
        head_ms.append(h)
        step_s.append(dt)
        sclk.append(ghz)
        if result is not None and result != frac:
            raise SystemExit("nondeterministic result across steps")
        result = frac
        log(f"step {i}: {dt:.3f}s (head kernel {h / 1e3:.3f}s, {ghz:.3f} GHz) digits {hexs}")
    total_s = sum(step_s)
    ms_per_step = 1e3 * total_s / len(step_s)
    value = d * len(step_s) / total_s
    head_avg_s = sum(head_ms) / len(head_ms) / 1e3
    sclk_ghz = sum(sclk) / len(sclk)
    slots = algorithmic_fp64_slots(d)
    achieved_slots = slots / head_avg_s  # FP64 VALU lane-slots per second
    flops = algorithmic_fp64_flops(d)
    int_products = int32_word_products(d)
    ref_check = reference_digits_check(clb, clb.hex48(result), d) if n_pos == 10**12 else None

    # seeded windows: plain vs split, GPU vs oracle
    rng = splitmix64(2)
    span = 1 << args.window_log2
    windows = [next(rng) % (d - span) for _ in range(args.windows)]
    jobs = [(d, s, s + span, j) for j in range(7) for s in windows]
    win_terms = span * len(windows)
    ctx.set_kernel(clb.KERNEL_SPLIT)
    split_parts = ctx.run_jobs(jobs)
    t0 = time.perf_counter()
    split_parts = ctx.run_jobs(jobs)
    split_win_s = ctx.last_timing()[1] / 1e3
    ctx.set_kernel(clb.KERNEL_PLAIN)
    plain_parts = ctx.run_jobs(jobs)
    plain_head_s, plain_win_s = (x / 1e3 for x in ctx.last_timing())
    plain_muls = 7 * sum(plain_multiplies(d, s, s + span) for s in windows)
    if plain_parts != split_parts:
        raise SystemExit("plain and split kernels disagree on the seeded windows")
    log(f"windows: split {win_terms / split_win_s:.3e} terms/s, plain "
        f"{win_terms / plain_win_s:.3e} terms/s, bit-identical")

    cpu, frac_err = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from clb_testlib import load_oracle, w2i, SHIFT  # the oracle, timed as the baseline
        orc = load_oracle()
        threads = min(16, os.cpu_count() or 1)
        # sample: the first window of each formula, shortened to fit the CPU time budget
        probe = 1 << 14
        t0 = time.perf_counter()
        orc.range_sum(d, windows[0], windows[0] + probe, 0, threads=threads)
        per_term = (time.perf_counter() - t0) / probe
        sample = max(probe, min(span, int(args.cpu_seconds / 7 / max(per_term, 1e-9))))
        sample_jobs = [(d, windows[0], windows[0] + sample, j) for j in range(7)]
        ctx.set_kernel(clb.KERNEL_SPLIT)
        gpu_sample = ctx.run_jobs(sample_jobs)
        t0 = time.perf_counter()
        ok = True
        for j, g in enumerate(gpu_sample):
            s = orc.range_sum(d, windows[0], windows[0] + sample, j, threads=threads)
            ok = ok and ((s << SHIFT[j]) & ((1 << 192) - 1)) == w2i(g)
        cpu_s = time.perf_counter() - t0
        if not ok:
            raise SystemExit("GPU and CPU oracle disagree on the sample window")
        cpu = {"value": sample / cpu_s, "unit": "terms/s", "cores": threads, "kind": "port",
               "sample": f"k in [{windows[0]}, {windows[0] + sample}) of N=10^12, all 7 formulas, "
                         f"bit-identical to the GPU sum; {threads} threads on {cpu_model()}"}
        log(f"cpu oracle: {sample / cpu_s:.3e} terms/s on {threads} threads")
        frac_err, err_cpu_s, err_terms = measure_frac_err(
            clb, orc, ctx, d, windows, 1 << args.err_window_log2, threads)
        frac_err["sample"] = (f"{len(windows)} seeded windows x 2^{args.err_window_log2} k x 7 "
                              f"formulas at N={n_pos} ({err_terms:.3e} term-formulas; the 256-bit "
                              f"oracle took {err_cpu_s:.1f} s on {threads} threads)")
        frac_err["bound_full_position_log2"] = \
            round(-192 + math.log2(NUMER_SUM * (d + 20)), 2)
        log(f"frac err: " + ", ".join(f"{k} {v['full_position_estimate_log2']}"
                                      for k, v in frac_err.items() if isinstance(v, dict)))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"normal-mode hex digit extraction at N={n_pos}: 7 formulas x {d} head "
                        "terms + 140 tail terms + combine, one persistent launch",
            "position": n_pos, "d": d, "kernel": "split (fp64 exact, 3-bit windows)",
            "digits": clb.hex48(result),
            "frac_err_bound_log2": round(-192 + math.log2(NUMER_SUM * (d + 20)), 2),
            "reference_kernels_output": ref_check,
        },
        "frac_err": frac_err,
        "roofline": {
            "bound": "valu-fp64-issue",
            "what": "algorithmic FP64 VALU lane-slots per second of the head kernel (6 per modular "
                    "square for bitlen(e)-1 squares, 16 per term's quotient digits) against the "
                    "VALU issue peak of one wave64 instruction per CU per clock (256 CUs x 64 "
                    "lanes x 2.4 GHz); no MFMA is used",
            "achieved": achieved_slots / 1e12, "peak": PEAK_VALU_LANE_SLOTS_PER_S / 1e12,
            "unit": "Tslot/s", "frac": achieved_slots / PEAK_VALU_LANE_SLOTS_PER_S,
            "sclk_ghz_measured": sclk_ghz,
            "frac_at_measured_clock": achieved_slots / (256 * 64 * sclk_ghz * 1e9),
            "clock_what": "the head kernel's shader clock during the timed steps (the first wave "
                          "on each XCD reads the shader-clock and 100 MHz real-time counters at "
                          "its start and end; mean over XCDs and steps), and the issue fraction "
                          "at that clock; the peak and frac above use the nominal 2.4 GHz.  It "
                          "can exceed 1: the algorithmic count takes bitlen(e) - 1 squares from "
                          "x = 2, and the kernel starts from 2^t mod m (t up to 49 + log2 m) in "
                          "one step, about six squares fewer, while it also issues rounding, "
                          "ldexp and integer instructions that the count leaves out",
            "traffic": load_traffic(),
            "head_kernel_ms": 1e3 * head_avg_s, "algorithmic_fp64_slots": slots,
            "fp64_flops": {"achieved": flops / head_avg_s / 1e12, "peak": PEAK_FP64_TFLOPS,
                           "unit": "TFLOP/s",
                           "frac": flops / head_avg_s / 1e12 / PEAK_FP64_TFLOPS,
                           "what": "fma = 2, mul and add = 1, rounding = 0"},
        },
        "int_mul_view": {
            "what": "32x32-bit word products/s that a 2-limb Montgomery engine would need for the "
                    "same terms (bitlen(e)-1 squares + 3 quotient words, 10 products each), against "
                    "the full-rate integer-multiply issue peak; the split kernel does this work on "
                    "the FP64 pipe, so the fraction can exceed 1",
            "achieved": int_products / head_avg_s, "peak": PEAK_INT32_MUL_PER_S,
            "unit": "products/s", "frac": int_products / head_avg_s / PEAK_INT32_MUL_PER_S,
            "modular_squares_per_s": 7 * (sum_bitlen_10x(d) - d) / head_avg_s,
        },
        "plain_kernel": {"value": win_terms / plain_win_s, "unit": "terms/s",
                         "windows": f"{args.windows} x 2^{args.window_log2} k per formula, seed 2",
                         "int_mul_issue": {
                             "what": "32x32-bit multiply instructions per second of the plain "
                                     "kernel on the windows (31 per 3-bit window of Montgomery "
                                     "squares, 46 per term besides, counted in its ISA) against "
                                     "the full-rate multiply issue peak (256 CUs x 64 lanes x "
                                     "2.4 GHz)",
                             "achieved": plain_muls / plain_head_s,
                             "peak": PEAK_INT32_MUL_PER_S, "unit": "multiplies/s",
                             "frac": plain_muls / plain_head_s / PEAK_INT32_MUL_PER_S}},
        "split_kernel_windows": {"value": win_terms / split_win_s, "unit": "terms/s"},
        "split_vs_plain": plain_win_s / split_win_s,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
