"""GPU parity tests: the HIP kernels of libclbellard.so (through the C ABI) against the CPU
oracle, the golden pi digits, and the size-independent properties of the exact-integer sums
(partition invariance, batched == single, plain == split)."""
import ctypes
import os
import random
import sys

import pytest

import clbellard_amd as clb
from clb_testlib import (DEN0, DEN1, M192, NUMESIGN, ROOT, SHIFT, bound_ulps, i2w, py_normalize,
                         w2i)

pytestmark = pytest.mark.gpu
R = random.Random(424242)


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = clb.Context(0, clb.PREC_U192)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_plain():
    c = clb.Context(0, clb.PREC_U192, kernel=clb.KERNEL_PLAIN)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx_split():
    c = clb.Context(0, clb.PREC_U192, kernel=clb.KERNEL_SPLIT)
    yield c
    c.close()


SMALL = [5, 10, 15, 20, 25, 30, 50, 100, 260, 500, 1000, 1005, 4100, 9995, 10000, 65540, 100000,
         123455, 250000]


@pytest.mark.parametrize("n", SMALL)
def test_normal_bit_exact_vs_oracle(ctx_plain, ctx_split, oracle, pi_fixtures, n):
    note(f"normal N={n}")
    want = oracle.normal(n)
    for c in (ctx_split, ctx_plain):
        frac, hexs, nn = c.normal(n)
        assert nn == n
        assert w2i(frac) == want, (n, c)
        assert hexs == f"{want:048x}"
    _, d = py_normalize(n)
    ref = int(pi_fixtures["digits"][str(n)], 16)
    assert (ref - want) % (1 << 192) <= bound_ulps(d)


@pytest.mark.parametrize("n", [500000, 777775, 1000000])
def test_normal_vs_pi_fixture(ctx_plain, ctx_split, pi_fixtures, n):
    note(f"normal N={n}")
    a = w2i(ctx_split.normal(n)[0])
    b = w2i(ctx_plain.normal(n)[0])
    assert a == b
    _, d = py_normalize(n)
    ref = int(pi_fixtures["digits"][str(n)], 16)
    assert (ref - a) % (1 << 192) <= bound_ulps(d)


@pytest.mark.parametrize("n", [10**7, 10**8, 10**9, 10**10])
def test_published_digits(ctx, ctx_plain, pi_fixtures, n):
    note(f"published N={n}")
    frac, hexs, _ = ctx.normal(n)
    assert hexs[:14] == pi_fixtures["published"][str(n)]
    if n <= 10**9:
        assert ctx_plain.normal(n)[0] == frac


def _ranges():
    cases = []
    # (d, k0, k1, j) across every kernel variant: m < 2^42 (W3), < 2^46 (W2), < 2^48 (W1),
    # < 2^55 (plain, lazy Montgomery), < 2^63 (plain, reduced residues)
    for d in (1, 3, 7, 999, 1000001, 399999999999, 2 * 10**12 + 1, 3 * 10**12 + 1,
              5 * 10**13 + 1, 10**15 + 1, 3 * 10**17 + 1):
        for _ in range(3):
            j = R.randrange(7)
            k0 = R.randrange(0, d)
            k1 = min(d, k0 + R.randrange(1, 3000))
            cases.append((d, k0, k1, j))
        cases.append((d, max(0, d - 2500), d, R.randrange(7)))       # reaches d: tail included
        cases.append((d, d, d + 10, R.randrange(7)))                # tail only
        cases.append((d, max(0, d - 100), d + 10**6, R.randrange(7)))  # k_end past d
    cases.append((1001, 1002, 2000, 0))  # beyond d: nothing
    cases.append((1001, 5, 5, 3))        # empty
    return cases


@pytest.mark.parametrize("kernel", [clb.KERNEL_AUTO, clb.KERNEL_PLAIN])
def test_batch_calc_vs_oracle(oracle, kernel):
    with clb.Context(0, clb.PREC_U192, kernel=kernel) as c:
        for d, k0, k1, j in _ranges():
            note(f"batch_calc kernel={kernel} d={d} k=[{k0},{k1}) j={j}")
            c.set0mem()
            c.batch_calc(d, k0, k1, j)
            assert w2i(c.read_partial()) == oracle.batch_calc(d, k0, k1, j), (d, k0, k1, j)


def test_split_variants_vs_oracle(ctx_split, oracle):
    # force each fp64 window width at the top of its modulus range
    for d, span in ((4 * 10**11 + 1, 4000), (2 * 10**12 + 1, 4000), (2 * 10**13 + 1, 4000)):
        for j in range(7):
            k0 = R.randrange(0, d - span)
            note(f"split d={d} k0={k0} j={j}")
            ctx_split.set0mem()
            ctx_split.batch_calc(d, k0, k0 + span, j)
            assert w2i(ctx_split.read_partial()) == oracle.batch_calc(d, k0, k0 + span, j)


KIND_LIMITS = (1 << 42, 1 << 46, 1 << 48, 1 << 55, 1 << 63)


@pytest.mark.parametrize("limit", KIND_LIMITS, ids=lambda t: f"2^{t.bit_length() - 1}")
def test_kernel_kind_limits_vs_oracle(ctx, ctx_plain, oracle, limit):
    """Head ranges whose largest modulus sits just below each limit at which the host changes the
    engine (fp64 3-, 2- and 1-bit windows below 2^42, 2^46 and 2^48; lazy Montgomery residues below
    2^55; reduced residues below 2^63), so that each engine runs at the top of its residue bounds,
    and ranges that straddle the limit, so that the next engine takes moduli just above it.  4096
    terms, so that full tiles take the fast term loop; d = 3 k, so that the exponents are long."""
    for j in (0, 2, 6):
        k_top = (limit - 1 - DEN1[j]) // DEN0[j]  # largest k whose modulus is below the limit
        d = min(999999999999999999, 3 * k_top)
        if d <= k_top:
            continue  # 4k + 1 reaches 2^63 only beyond the largest position
        for k0, k1 in ((k_top - 4095, k_top + 1), (k_top - 2047, k_top + 2049)):
            if DEN0[j] * (k1 - 1) + DEN1[j] >= 1 << 63:
                continue
            want = oracle.batch_calc(d, k0, k1, j)
            for name, c in (("auto", ctx), ("plain", ctx_plain)):
                note(f"limit 2^{limit.bit_length() - 1}: d={d} k=[{k0},{k1}) j={j} {name}")
                c.set0mem()
                c.batch_calc(d, k0, k1, j)
                assert w2i(c.read_partial()) == want, (limit, d, k0, k1, j)


def test_split_rejects_large_moduli(ctx_split):
    d = 10**15 + 1  # moduli 10k + 1 near k = d exceed 2^49
    with pytest.raises(clb.ClbError) as e:
        ctx_split.batch_calc(d, d - 10, d, 2)
    assert e.value.code == -3


def test_accumulates_across_calls(ctx, oracle):
    d, j = 10**9 + 1, 4
    ctx.set0mem()
    total = 0
    for k0 in (0, 12345, 9999999):
        ctx.batch_calc(d, k0, k0 + 5000, j)
        total += oracle.batch_calc(d, k0, k0 + 5000, j)
    assert w2i(ctx.read_partial()) == total & M192


def test_kernel_operators_generic_formula(ctx, oracle):
    for _ in range(6):
        den0 = 2 * R.randrange(1, 20)
        den1 = 2 * R.randrange(0, 20) + 1
        numesign = R.randrange(2)
        d = R.choice([10**6 + 1, 10**11 + 1])
        k0 = R.randrange(0, d - 600)
        k1 = k0 + R.randrange(1, 600)
        ctx.set0mem()
        ctx.sglobal(k0, d, k1, numesign, den0, den1)
        want = sum(oracle.term192(d, k, numesign, den0, den1) for k in range(k0, k1)) & M192
        assert w2i(ctx.read_partial()) == want
        ctx.set0mem()
        ctx.sglobal_after(d, numesign, den0, den1)
        assert w2i(ctx.read_partial()) == oracle.tail192(d, numesign, den0, den1)
    with pytest.raises(clb.ClbError):
        ctx.sglobal(0, 100, 101, 0, 4, 1)  # k_max beyond d
    with pytest.raises(clb.ClbError):
        ctx.sglobal(0, 100, 10, 0, 3, 1)   # even moduli


def test_allloop_through_operators_equals_normal(ctx, oracle):
    n = 123455
    _, d = py_normalize(n)
    ctx.clear_ans()
    for j in range(7):
        ctx.set0mem()
        ctx.batch_calc(d, 0, d, j)
        ctx.add_ans(j)
        part = w2i(ctx.read_partial())
        assert part == (oracle.batch_calc(d, 0, d, j) << SHIFT[j]) & M192
    assert w2i(ctx.read_ans()) == oracle.normal(n)
    assert w2i(ctx.normal(n)[0]) == oracle.normal(n)


def test_batch_positions_equal_single_runs(ctx):
    rng = random.Random(0)
    pos = [5 * round(10 ** (3 + 4 * rng.random()) / 5) for _ in range(48)] + [5, 10, 1000000]
    note(f"batch of {len(pos)} positions")
    batched = ctx.batch_positions(pos)
    for p, b in zip(pos, batched):
        assert ctx.normal(p)[0] == b, p


def test_run_jobs_tiling_sums_to_normal(ctx):
    n = 10**7
    _, d = py_normalize(n)
    cuts = sorted(R.sample(range(1, d), 9))
    bounds = [0] + cuts + [d]
    jobs = [(d, a, b, j) for j in range(7) for a, b in zip(bounds, bounds[1:])]
    parts = ctx.run_jobs(jobs)
    total = sum(w2i(p) for p in parts) & M192
    assert total == w2i(ctx.normal(n)[0])
    # one job per formula covering everything is the same as the jobs sum per formula
    whole = ctx.run_jobs([(d, 0, d + 5, j) for j in range(7)])
    assert sum(w2i(p) for p in whole) & M192 == total


def test_launch_shape_does_not_change_bits(oracle):
    n = 65540
    want = oracle.normal(n)
    with clb.Context(0) as c:
        for bpc in (1, 2, 5, 8, 16):
            c.set_launch(bpc)
            assert w2i(c.normal(n)[0]) == want, bpc


def _worklist(path, digits, precision, jobs):
    with open(path, "w", newline="") as f:
        f.write(f"digits={digits}\r\nprecision={precision} (0=106bit , 1=192bit)\r\n")
        f.write("\r\n".join(jobs) + "\r\n")


@pytest.mark.parametrize("precision", [1, 0])
def test_worklist_end_to_end(tmp_path, precision):
    n = 1000000
    _, d = py_normalize(n)
    batch = 100000
    nb = d // batch + 1
    jobs = [f"{j}_{b}" for j in range(7) for b in range(nb)] + ["3_99"]  # 3_99 starts past d
    wl = os.path.join(str(tmp_path), "worklist.txt")
    _worklist(wl, n, precision, jobs)
    work = os.path.join(str(tmp_path), "work")
    with clb.Context(0, precision) as c:
        ran = c.run_worklist(wl, work, batch_size=batch, device_label="0\ttest-device")
        assert ran == 7 * nb
        assert c.run_worklist(wl, work, batch_size=batch) == 0  # resume: all files exist
        os.remove(os.path.join(work, "2", "2_1.txt"))
        assert c.run_worklist(wl, work, batch_size=batch) == 1
        frac, hexs, _ = c.normal(n)
    body = open(os.path.join(work, "0", "0_0.txt"), "rb").read().decode()
    line0, line1, line2 = body.split("\r\n")
    assert len(line0) == 48 and line1 == "0\ttest-device" and line2.endswith("ms")
    if precision == 1:
        total, counts = clb.work_sum(work)
        assert counts == [nb] * 7
        assert total == hexs
    else:
        assert line0.endswith("0" * 20)


def test_worklist_writes_files_launch_by_launch(tmp_path):
    """Work mode keeps the reference's per-job checkpoint (Bellard_192.hsp:307-333) at launch
    granularity: with at most 5 jobs per launch and one launch per call, each call writes exactly
    the next 5 pending files of the list, a rerun skips what exists, and the finished files sum to
    the normal-mode digits.  A time bound below any job's cost gives one job per launch.  Each
    file's third line is its job's share of its launch's time."""
    n = 1000000
    _, d = py_normalize(n)
    batch = 50000
    nb = d // batch + 1
    order = [(j, b) for j in range(7) for b in range(nb)]
    wl = os.path.join(str(tmp_path), "worklist.txt")
    _worklist(wl, n, 1, [f"{j}_{b}" for j, b in order])
    work = os.path.join(str(tmp_path), "work")

    def done():
        return [(j, b) for j, b in order
                if os.path.exists(os.path.join(work, str(j), f"{j}_{b}.txt"))]

    with clb.Context(0, clb.PREC_U192) as c:
        c.set_worklist_launch(max_jobs=5, max_seconds=0, max_launches=1)
        for call in range(3):
            note(f"worklist call {call}: one launch of at most 5 jobs")
            assert c.run_worklist(wl, work, batch_size=batch) == 5
            assert done() == order[:5 * (call + 1)]
        c.set_worklist_launch(max_jobs=0, max_seconds=1e-9, max_launches=4)
        assert c.run_worklist(wl, work, batch_size=batch) == 4  # one job per launch
        assert done() == order[:19]
        c.set_worklist_launch()  # defaults: 60 s launches, until the list is done
        assert c.run_worklist(wl, work, batch_size=batch) == len(order) - 19
        assert c.run_worklist(wl, work, batch_size=batch) == 0
        hexs = c.normal(n)[1]
        with pytest.raises(clb.ClbError):
            c.set_worklist_launch(max_seconds=-1.0)
    total, counts = clb.work_sum(work)
    assert counts == [nb] * 7 and total == hexs
    for j, b in order:
        body = open(os.path.join(work, str(j), f"{j}_{b}.txt"), "rb").read().decode()
        ms = body.split("\r\n")[2]
        assert ms.endswith("ms") and int(ms[:-2]) >= 0


def test_graph_mode_replays_match_eager(ctx, oracle):
    """Graph mode (clb_set_graph): every replay of a captured sequence, with eager work on another
    context and synchronisations in between, equals the eager result; calls of other sizes reuse
    or recapture the graph as their buffers require."""
    g = clb.Context(0, clb.PREC_U192)
    g.set_graph(True)
    try:
        rng = random.Random(7)
        for rep in range(3):
            for n in (5, 100000, 65540, 10**7):
                note(f"graph replay {rep}: normal N={n}")
                want = ctx.normal(n)
                assert g.normal(n) == want, (rep, n)
            pos = [5 * round(10 ** (3 + 4 * rng.random()) / 5) for _ in range(12 + 20 * rep)]
            note(f"graph replay {rep}: batch of {len(pos)} positions")
            assert g.batch_positions(pos) == ctx.batch_positions(pos), rep
            _, d = py_normalize(10**8)
            jobs = [(d, 1000 * rep, d // 3 + rep, j) for j in range(7)]
            assert g.run_jobs(jobs) == ctx.run_jobs(jobs), rep
            # the accumulating path (BatchCalc into the partial, AddAns)
            g.set0mem()
            g.clear_ans()
            ctx.set0mem()
            ctx.clear_ans()
            for j in range(7):
                g.set0mem()
                g.batch_calc(d, 0, d, j)
                g.add_ans(j)
                ctx.set0mem()
                ctx.batch_calc(d, 0, d, j)
                ctx.add_ans(j)
            assert g.read_ans() == ctx.read_ans(), rep
        assert w2i(g.normal(250000)[0]) == oracle.normal(250000)
    finally:
        g.close()


@pytest.mark.parametrize("tpl", [16, 4, 1])
def test_tile_lengths_give_identical_bits(oracle, tpl):
    """Every terms-per-lane choice, both engines, against the oracle and the automatic choice."""
    for kernel in (clb.KERNEL_SPLIT, clb.KERNEL_PLAIN):
        with clb.Context(0, clb.PREC_U192, kernel=kernel) as c:
            c.set_tiling(tpl)
            for n in (5, 1005, 65540, 250000):
                note(f"tiles of {tpl} terms per lane, kernel {kernel}: N={n}")
                assert w2i(c.normal(n)[0]) == oracle.normal(n), (tpl, kernel, n)
            _, d = py_normalize(10**9)
            jobs = [(d, 12345 + 777 * j, 12345 + 777 * j + 3 * (1 << 20) + j, j) for j in range(7)]
            jobs += [(d, d - 5000 - j, d + 3, j) for j in range(7)]
            got = c.run_jobs(jobs)
            c.set_tiling(0)
            assert got == c.run_jobs(jobs), (tpl, kernel)


def test_autotune_keeps_bits(oracle):
    """clb_autotune picks a valid tile length and grid, and results stay bit-identical."""
    with clb.Context(0, clb.PREC_U192) as c:
        for n, budget in ((250000, 50.0), (10**9, 300.0)):
            tpl, bpc = c.autotune(n, budget)
            sched = c.get_schedule()
            note(f"autotune N={n}: {tpl} terms per lane, {bpc} workgroups per CU, {sched}")
            assert tpl in (16, 4, 1) and bpc >= 1 and sched["terms_per_lane"] == tpl
            assert sched["waves_per_group"] in (1, 2, 4, 8)
            assert sched["min_chunk"] in (64, 256, 1024) and sched["max_replicas"] in (4, 64)
        assert w2i(c.normal(250000)[0]) == oracle.normal(250000)
        with clb.Context(0, clb.PREC_U192) as ref:
            assert c.normal(10**8) == ref.normal(10**8)


def test_schedule_file_carries_a_tuning(tmp_path):
    """clb_save_schedule writes the context's launch settings under its device's name;
    clb_load_schedule on a new context restores them (same bits), and refuses a file written for
    another device without changing anything."""
    p = str(tmp_path / "tune.txt")
    with clb.Context(0, clb.PREC_U192) as a:
        a.set_tiling(4, 3)
        a.set_launch(6)
        a.set_schedule(2, 1024, 0, 4)
        a.save_schedule(p)
        want_sched, want = a.get_schedule(), a.normal(10**7)
    saved = clb.read_schedule(p)
    assert saved["device"] == clb.device_name(0)
    assert saved["min_tiles_per_wave"] == 3 and saved["terms_per_lane"] == 4
    with clb.Context(0, clb.PREC_U192) as b:
        default = b.get_schedule()
        b.load_schedule(p)
        assert b.get_schedule() == want_sched != default
        assert b.normal(10**7) == want
        other = str(tmp_path / "other.txt")
        clb.write_schedule(other, dict(saved, device="another GPU", blocks_per_cu=2))
        with pytest.raises(clb.ClbError) as e:
            b.load_schedule(other)
        assert e.value.code == clb.EINVAL and b.get_schedule() == want_sched


def test_cli_tune_cache(tmp_path):
    """`clbellard normal N --autotune MS --tune-cache FILE` tunes and writes FILE; the next run
    loads it instead of tuning, with the same digits."""
    import subprocess
    cli = os.path.join(ROOT, "clbellard_amd", "clbellard")
    p = str(tmp_path / "tune.txt")
    args = [cli, "normal", "100000000", "--autotune", "200", "--tune-cache", p]
    first = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert first.returncode == 0 and "written to" in first.stderr, first.stderr
    assert clb.read_schedule(p)["device"] == clb.device_name(0)
    second = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert second.returncode == 0 and "launch settings from" in second.stderr, second.stderr
    assert first.stdout.splitlines()[:2] == second.stdout.splitlines()[:2]


def test_out_of_range_moduli_are_refused(ctx):
    """Loop counts whose moduli leave the supported range fail before any launch (CLB_ERANGE)."""
    with pytest.raises(clb.ClbError) as e:
        ctx.run_jobs([(10**18 - 1, 10**18 - 10, 10**18, 2)])  # tail and head moduli near 10^19
    assert e.value.code == clb.ERANGE
    assert w2i(ctx.normal(1000)[0]) != 0  # the context stays usable


def test_call_size_limits(ctx):
    """More positions or jobs than one call's block can hold fail with CLB_EINVAL before anything
    is read or allocated (a one-element array with a larger count is never read past its end)."""
    one = (ctypes.c_uint64 * 1)(10**6)
    out = (ctypes.c_uint64 * 3)()
    assert clb.lib.clb_batch_positions(ctx._ctx, one, (1 << 21) + 1, out) == clb.EINVAL
    job = (clb._Job * 1)()
    assert clb.lib.clb_run_jobs(ctx._ctx, job, (1 << 24) + 1, out) == clb.EINVAL
    assert w2i(ctx.batch_positions([10**6])[0]) == w2i(ctx.normal(10**6)[0])


def test_contexts_on_threads(oracle):
    """Calls on different contexts may run from different threads (include/clbellard.h): four
    threads, each with its own context and stream, compute different positions at the same time,
    three times each; every result equals the oracle's, and each thread's error message (an
    invalid position of its own) is its own."""
    import threading
    ns = [65540, 100000, 250000, 777775]
    want = {n: oracle.normal(n) for n in ns}
    got, errors, failures = {}, {}, []
    start = threading.Barrier(len(ns))

    def work(n):
        try:
            with clb.Context(0, clb.PREC_U192) as c:
                start.wait()
                got[n] = [w2i(c.normal(n)[0]) for _ in range(3)]
                with pytest.raises(clb.ClbError) as e:
                    c.batch_calc(n, 0, n, 7 + ns.index(n))  # formula ids past 6, one per thread
                errors[n] = str(e.value)
        except BaseException as e:  # reported by the main thread
            failures.append((n, repr(e)))

    note(f"threads: positions {ns}")
    threads = [threading.Thread(target=work, args=(n,)) for n in ns]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not failures, failures
    for n in ns:
        assert got[n] == [want[n]] * 3, n
        assert f"formula {7 + ns.index(n)} (" in errors[n], errors[n]


def test_issue_rate_microbenchmark():
    """clb_issue_rate runs, measures a plausible shader clock, and the FP64 FMA issues within the
    VALU's one wave64 instruction per CU per clock."""
    for op in (0, 5, 10):
        r = clb.issue_rate(op, 20.0)
        note(f"{r['op']}: {r['per_s']:.3e}/s at {r['sclk_ghz']:.3f} GHz")
        assert r["per_s"] > 0 and 0.5 < r["sclk_ghz"] < 3.5 and r["cus"] > 0
        if op == 0:
            # runs read 65.2-65.5 lanes per CU per clock: the clock of the eight sampled waves
            # differs from the chip's average by a few percent; a miscount (two lane-operations
            # per instruction, a wrong CU count) would read 128 or more
            assert r["per_s"] / (r["cus"] * r["sclk_ghz"] * 1e9) <= 64 * 1.2


def test_head_kernel_clock_samples():
    """clb_last_clock: after a head-kernel launch every XCD's sample gives a plausible shader
    clock; before any call, and after calls that run no head-term kernel with samples (the
    106-bit-mode replay, the no-allocation entry point), it refuses."""
    _, d = py_normalize(10**10)
    with clb.Context(0, clb.PREC_U192) as c:
        with pytest.raises(clb.ClbError):
            c.last_clock()
        c.normal(10**10)  # ~0.16 s: long enough for the counters to mean something
        mean, per = c.last_clock()
        note(f"head kernel clock {mean:.3f} GHz, per XCD {[round(g, 3) for g in per]}")
        assert len(per) == clb.CLOCK_XCDS and all(0.5 < g < 3.5 for g in per)
        assert abs(mean - sum(per) / len(per)) < 1e-9
        c.set_graph(True)  # the captured graph writes them too
        c.normal(10**9)
        assert 0.5 < c.last_clock()[0] < 3.5
        c.set_graph(False)
        c.run_jobs_noalloc([(d, 0, 1000, 0)])
        with pytest.raises(clb.ClbError):
            c.last_clock()
    with clb.Context(0, clb.PREC_DD106) as c:
        c.set_replay(clb.REF_GLOBAL_WORK_SIZE)
        c.normal(10**6)
        with pytest.raises(clb.ClbError):
            c.last_clock()


def test_schedule_knobs_give_identical_bits(oracle):
    """Every workgroup size, tile-queue schedule and slot-copy limit (clb_set_schedule) gives the
    oracle's bits, on small launches (static tiles) and large ones (the guided queue); bad values
    are refused."""
    with clb.Context(0, clb.PREC_U192) as c:
        _, d = py_normalize(10**9)
        jobs = [(d, 0, d + 1, j) for j in range(7)]
        want = c.run_jobs(jobs)
        for w, mc, cpw, r in ((1, 0, 0, 0), (2, 0, 0, 0), (8, 0, 0, 0), (0, 256, -1, 0),
                              (0, 1, -1, 0), (0, 16, 4, 0), (0, 64, 1, 1), (8, 7, 3, 4)):
            note(f"schedule waves={w} min_chunk={mc} chunks_per_wave={cpw} replicas={r}")
            c.set_schedule(w, mc, cpw, r)
            for n in (5, 65540, 250000):
                assert w2i(c.normal(n)[0]) == oracle.normal(n), (w, mc, cpw, r, n)
            assert c.run_jobs(jobs) == want, (w, mc, cpw, r)
        for bad in ((3, 0, 0, 0), (0, -1, 0, 0), (0, 0, -2, 0), (0, 0, 0, 3), (16, 0, 0, 0)):
            with pytest.raises(clb.ClbError):
                c.set_schedule(*bad)
