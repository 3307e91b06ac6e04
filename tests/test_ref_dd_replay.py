"""The double-double replay of the reference's precision-1/2 kernels against those kernels.

The reference's 192-bit kernels round: each work-item keeps two double-double sums over its
strided terms (Bellard.cl:771-823, :992-1084) and converts them once (dd2_to_ulong3,
:506-578); the tail kernel is double-double throughout (:1153-1253).  orc_ref_batch192 restates
that arithmetic operation by operation, and the library's replay mode (clb_set_replay,
clb_replay.hip) reproduces it on the GPU.  Both are held, bit for bit, to the reference's own
kernels (oracle/_ref, GPU) on every case of ref_dd_cases() and on whole positions.
"""
import os
import struct
import sys

import pytest

import clbellard_amd as clb
from clb_testlib import (M192, REF_LIB, SHIFT, RefKernels, load_oracle, py_normalize,
                         ref_dd106_cases, ref_dd_cases, ref_od_cases, w2i)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(REF_LIB),
                                 reason="oracle/_ref not built (needs the reference sources)")]

CASES = ref_dd_cases()


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def ref():
    note("reference kernels: loading oracle/_ref")
    r = RefKernels()
    yield r
    r.set_gws()
    r.close()


@pytest.fixture(scope="module")
def orc():
    return load_oracle()


@pytest.fixture(scope="module")
def ctxs():
    c = {p: clb.Context(0, p) for p in (clb.PREC_U192, clb.PREC_U192_BETA)}
    c.update({(p, "plain"): clb.Context(0, p, kernel=clb.KERNEL_PLAIN)
              for p in (clb.PREC_U192, clb.PREC_U192_BETA)})
    yield c
    for x in c.values():
        x.close()


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d{}_k{}-{}_j{}_p{}_g{}".format(*c))
def test_oracle_replay_equals_reference_kernels(ref, orc, case):
    d, k0, k1, j, p, gws = case
    note(f"reference kernels d={d} k=[{k0},{k1}) j={j} precision={p} gws={gws}")
    ref.set_gws(gws)
    got = w2i(ref.job(d, k0, k1, j, p))
    want = (orc.ref_batch192(d, k0, k1, j, p, gws) << SHIFT[j]) & M192
    assert got == want, f"{got:048x} != {want:048x}"


@pytest.mark.parametrize("case", ref_dd106_cases(), ids=lambda c: "d{}_k{}-{}_j{}_g{}".format(*c))
def test_oracle_106bit_replay_equals_reference_kernels(ref, orc, case):
    d, k0, k1, j, gws = case
    note(f"reference kernels, 106-bit mode, d={d} k=[{k0},{k1}) j={j} gws={gws}")
    ref.set_gws(gws)
    assert ref.job(d, k0, k1, j, 0) == orc.ref_job106(d, k0, k1, j, gws)


@pytest.mark.parametrize("n", [10**6, 10**7])
def test_oracle_106bit_positions_equal_reference_kernels(ref, orc, n):
    ref.set_gws()
    assert ref.normal(py_normalize(n)[1], 0) == orc.ref_normal106(n)


@pytest.mark.parametrize("engine", ["auto", "plain"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "d{}_k{}-{}_j{}_p{}_g{}".format(*c))
def test_hip_replay_equals_reference_kernels(ref, ctxs, case, engine):
    """Residues from the fp64 engine (auto: moduli below 2^48) and from the integer engine."""
    d, k0, k1, j, p, gws = case
    note(f"replay ({engine}) d={d} k=[{k0},{k1}) j={j} precision={p} gws={gws}")
    ctx = ctxs[p] if engine == "auto" else ctxs[(p, "plain")]
    ctx.set_replay(gws)
    ref.set_gws(gws)
    got = w2i(ctx.run_jobs([(d, k0, k1, j)])[0])
    want = w2i(ref.job(d, k0, k1, j, p))
    assert got == want, f"{got:048x} != {want:048x}"


@pytest.mark.parametrize("n", [10**6, 10**7, 10**8, 10**9, 2 * 10**9, 10**10])
@pytest.mark.parametrize("p", [1, 2])
def test_hip_replay_positions_equal_reference_kernels(ref, ctxs, n, p):
    """Normal mode at the reference's own launch size: all 48 hex digits equal."""
    nn, d = py_normalize(n)
    note(f"replay position N={nn} precision {p}")
    ctx = ctxs[p]
    ctx.set_replay(clb.REF_GLOBAL_WORK_SIZE)
    ref.set_gws()
    frac, hexs, _ = ctx.normal(n)
    want = ref.normal(d, p)
    note(f"  ours {hexs}\n  ref  {clb.hex48(want)}")
    assert w2i(frac) == w2i(want)
    ctx.set_replay(0)
    assert ctx.normal(n)[1][:24] == hexs[:24]  # the exact mode agrees on the leading digits


def test_replay_arguments(ctxs):
    ctx = ctxs[clb.PREC_U192]
    for bad in (100, clb.REF_GLOBAL_WORK_SIZE * 2):
        with pytest.raises(clb.ClbError) as e:
            ctx.set_replay(bad)
        assert e.value.code == clb.EINVAL
    ctx.set_replay(clb.REF_GLOBAL_WORK_SIZE)
    try:
        d_over = ((1 << 63) - 9) // 10 + 1  # moduli reach 2^63 (Bellard.cl:1041 needs less)
        with pytest.raises(clb.ClbError) as e:
            ctx.run_jobs([(d_over, d_over - 100, d_over, 0)])
        assert e.value.code == clb.ERANGE
    finally:
        ctx.set_replay(0)


@pytest.mark.parametrize("case", ref_od_cases(), ids=lambda c: "d{}_k{}-{}_j{}_p{}_g{}".format(*c))
def test_hip_overdouble_replay_equals_reference_kernels(ref, ctxs, case):
    """Loop counts with 10 d + 9 >= 2^53: the reference's kernels built with -D OVERDOUBLE
    (dd_div of the split numerator and modulus from 2^53 on), all three precisions, residues
    from the integer engine (3-bit windows below 2^55, the wide engine above)."""
    d, k0, k1, j, p, gws = case
    note(f"OVERDOUBLE replay d={d} k=[{k0},{k1}) j={j} precision={p} gws={gws}")
    ref.set_gws(gws)
    if p == 0:
        with clb.Context(0, clb.PREC_DD106) as c:
            c.set_replay(gws)
            assert dd_of_words(c.run_jobs([(d, k0, k1, j)])[0]) == ref.job(d, k0, k1, j, 0)
        return
    ctx = ctxs[(p, "plain")]
    ctx.set_replay(gws)
    try:
        got = w2i(ctx.run_jobs([(d, k0, k1, j)])[0])
    finally:
        ctx.set_replay(0)
    want = w2i(ref.job(d, k0, k1, j, p))
    assert got == want, f"{got:048x} != {want:048x}"


@pytest.mark.parametrize("p", [1, 2])
def test_replay_work_files_equal_reference_jobs(ref, ctxs, tmp_path, p):
    """Work mode (Batchloop + fsave, Bellard_192.hsp:298-333) replayed: every work file equals the
    reference kernels' job (Set0mem, BatchCalc, AddAns) bit for bit, double-double ranges and
    tails included, so the files sum (workmodeSum) to what the reference's own work files sum to.
    That is not the normal-mode result: each job restarts the launch chunking at its k_start, so
    above loopmax32 its double-double grouping differs from one run from 0 (the replayed normal
    mode equals the reference kernels' normal mode, checked separately)."""
    n = 25_000_000_000
    _, d = py_normalize(n)
    batch = 1 << 30
    nb = d // batch + 1
    wl = tmp_path / "worklist.txt"
    lines = [f"digits={n}", f"precision={p}"] + [f"{j}_{b}" for j in range(7) for b in range(nb)]
    wl.write_bytes(("\r\n".join(lines) + "\r\n").encode())
    work = str(tmp_path / "work")
    ctx = ctxs[p]
    ctx.set_replay(clb.REF_GLOBAL_WORK_SIZE)
    ref.set_gws()
    try:
        assert ctx.run_worklist(str(wl), work, batch_size=batch) == 7 * nb
        frac, digits, _ = ctx.normal(n)
    finally:
        ctx.set_replay(0)
    theirs_total = 0
    for j in range(7):
        for b in range(nb):
            k0, k1 = b * batch, (b + 1) * batch
            line0 = open(os.path.join(work, str(j), f"{j}_{b}.txt"), "rb").read().split(b"\r\n")[0]
            theirs = w2i(ref.job(d, k0, k1, j, p))
            assert int(line0, 16) == theirs, (j, b)
            theirs_total = (theirs_total + theirs) % (1 << 192)
    assert w2i(frac) == w2i(ref.normal(d, p))
    assert clb.work_sum(work) == (f"{theirs_total:048x}", [nb] * 7)
    if p == 2:  # the double-double range (below loopmax32) lies in batch 0 either way
        assert f"{theirs_total:048x}" == digits
    note(f"precision {p}: {7 * nb} replayed work files equal the reference kernels' jobs")


# ----------------------------------------------------------------- the 106-bit mode


@pytest.fixture(scope="module")
def ctx106():
    c = clb.Context(0, clb.PREC_DD106)
    yield c
    c.close()


def dd_of_words(w):
    """(hi, lo) from the 3-word result of the 106-bit-mode replay."""
    return tuple(struct.unpack("<d", struct.pack("<Q", x))[0] for x in w[:2])


@pytest.mark.parametrize("case", ref_dd106_cases(), ids=lambda c: "d{}_k{}-{}_j{}_g{}".format(*c))
def test_hip_106bit_replay_equals_reference_kernels(ref, ctx106, case):
    d, k0, k1, j, gws = case
    note(f"106-bit replay d={d} k=[{k0},{k1}) j={j} gws={gws}")
    ctx106.set_replay(gws)
    ref.set_gws(gws)
    assert dd_of_words(ctx106.run_jobs([(d, k0, k1, j)])[0]) == ref.job(d, k0, k1, j, 0)


@pytest.mark.parametrize("n", [10**6, 10**7, 10**8, 10**9])
def test_hip_106bit_replay_positions(ref, ctx106, n):
    """Allloop in the 106-bit mode: dev_ans and the printed 28 hex digits equal the reference's."""
    nn, d = py_normalize(n)
    note(f"106-bit replay position N={nn}")
    ctx106.set_replay(clb.REF_GLOBAL_WORK_SIZE)
    ref.set_gws()
    frac, hexs, _ = ctx106.normal(n)
    want = ref.normal(d, 0)
    note(f"  ours {hexs}\n  ref  {clb.hex_dd(*want)}")
    assert dd_of_words(frac) == want
    assert hexs == clb.hex_dd(*want)


def test_hip_106bit_replay_operators(ref, ctx106):
    """Set0mem / BatchCalc / AddAns by hand (Allloop, Bellard_192.hsp:264-296) on the context's
    own bigSumMem, with a BatchCalc split in two calls (the group accumulators carry over, as the
    reference's do), against the reference kernels' normal mode."""
    n = 10**7
    _, d = py_normalize(n)
    ctx106.set_replay(clb.REF_GLOBAL_WORK_SIZE)
    ref.set_gws()
    ctx106.clear_ans()
    for j in range(7):
        ctx106.set0mem()
        ctx106.batch_calc(d, 0, d, j)
        ctx106.add_ans(j)
    assert dd_of_words(ctx106.read_ans()) == ref.normal(d, 0)


def loopmax32(j):
    _, _, den0, den1 = clb.formula(j)
    return ((1 << 32) - den1 - 1) // den0 + 1


class OpPair:
    """The library's operators and the reference's kernels side by side on their own bigSumMem:
    every call is made on both and bigSum[0] (the library's partial) compared bit for bit."""

    def __init__(self, ctx, ref, j):
        self.ctx, self.ref, self.j = ctx, ref, j
        _, self.sign, self.den0, self.den1 = clb.formula(j)
        self.calls = 0

    def check(self, want, what):
        got = dd_of_words(self.ctx.read_partial())
        assert got == want, (what, self.j, got, want)
        self.calls += 1

    def set0mem(self):
        self.ctx.set0mem()
        self.check(self.ref.op(0, 0), "Set0memdd")

    def sglobal(self, offset, d, k_max):
        self.ctx.sglobal(offset, d, k_max, self.sign, self.den0, self.den1)
        op = 1 if offset < loopmax32(self.j) else 2
        self.check(self.ref.op(op, 0, offset, d, k_max, self.j), ("Sglobal", op, offset, k_max))

    def after(self, d):
        self.ctx.sglobal_after(d, self.sign, self.den0, self.den1)
        self.check(self.ref.op(3, 0, 0, d, 0, self.j), ("Sglobal_after", d))

    def ssum(self):
        self.ctx.ssum()
        self.check(self.ref.op(4, 0), "Ssum")

    def batch_calc(self, d, k_start, k_end, gws):
        """BatchCalc (KernelFuncs.as:25-66) spelled out as single kernel calls."""
        offset = k_max = k_start
        lm32 = min(loopmax32(self.j), d, k_end)
        while offset < lm32:
            k_max = min(k_max + gws * 16, lm32)
            self.sglobal(offset, d, k_max)
            offset = k_max
        lm64 = min(k_end, d)
        while offset < lm64:
            k_max = min(k_max + gws * 32, lm64)
            self.sglobal(offset, d, k_max)
            offset = k_max
        if offset == d:
            self.after(d)
        self.ssum()


def op_cases():
    x = loopmax32(2)
    g = 4096
    return [  # d, k_start, k_end, j, work-items per launch
        (1234567, 1000000, 1234567, 0, g),               # several launches, then the tail
        (999999999, x - 3 * g * 16 - 5, x + 2 * g * 32 + 7, 2, g),  # across loopmax32
        (500001, 0, 500001, 6, clb.REF_GLOBAL_WORK_SIZE),  # Allloop's single launch
    ]


@pytest.mark.parametrize("case", op_cases(), ids=lambda c: "d{}_k{}-{}_j{}_g{}".format(*c))
def test_hip_106bit_replay_kernel_operators(ref, ctx106, case):
    """Sglobal32mtg / Sglobal64mtg / Sglobal_after / Ssum one call at a time on the context's
    65536 group accumulators (clb_sglobal, clb_sglobal_after, clb_ssum) equal the reference's
    kernels after every call, and their BatchCalc sequence equals clb_batch_calc."""
    d, k0, k1, j, gws = case
    note(f"106-bit kernel operators d={d} k=[{k0},{k1}) j={j} gws={gws}")
    ctx106.set_replay(gws)
    ref.set_gws(gws)
    pair = OpPair(ctx106, ref, j)
    pair.set0mem()
    pair.batch_calc(d, k0, k1, gws)
    by_ops = ctx106.read_partial()
    ctx106.set0mem()
    ctx106.batch_calc(d, k0, k1, j)
    assert ctx106.read_partial() == by_ops
    note(f"  {pair.calls} calls equal the reference kernels; BatchCalc equal")


def test_hip_106bit_replay_operator_orders(ref, ctx106):
    """Orders BatchCalc never makes, still the reference kernels' results after every call:
    tails before head terms and twice in a row, an empty launch (its work-group 0 still applies
    sumfix2 to bigSum[0]), Ssum twice, launches after Ssum on the accumulators it leaves, and two
    BatchCalcs without Set0mem between them."""
    gws = 8192
    ctx106.set_replay(gws)
    ref.set_gws(gws)
    d, j = 2000001, 3
    p = OpPair(ctx106, ref, j)
    p.set0mem()
    p.after(d)
    p.after(d)
    p.sglobal(500, d, 500)
    p.sglobal(0, d, 300000)
    p.after(d)
    p.sglobal(10, d, 5)
    p.ssum()
    p.ssum()
    p.sglobal(300000, d, 700000)
    p.ssum()
    p.batch_calc(d, 700000, d, gws)
    p.batch_calc(d, 0, 123457, gws)
    x = loopmax32(j)
    with pytest.raises(clb.ClbError) as e:  # one 32-bit launch past loopmax32
        ctx106.sglobal(x - 10, 10**9 + 1, x + 10, p.sign, p.den0, p.den1)
    assert e.value.code == clb.EINVAL
    note(f"106-bit operator orders: {p.calls} calls equal the reference kernels")


def test_hip_106bit_replay_work_files(ref, ctx106, tmp_path):
    """Work mode in the 106-bit mode replayed: each file's digits are hexadecimal_dd of the
    reference kernels' job (Set0memdd, BatchCalc, AddAns, bigSum[0] as fsave writes it)."""
    n = 2_500_000_000
    _, d = py_normalize(n)
    batch = 1 << 28
    nb = d // batch + 1
    wl = tmp_path / "worklist.txt"
    lines = [f"digits={n}", "precision=0"] + [f"{j}_{b}" for j in range(7) for b in range(nb)]
    wl.write_bytes(("\r\n".join(lines) + "\r\n").encode())
    work = str(tmp_path / "work")
    ctx106.set_replay(clb.REF_GLOBAL_WORK_SIZE)
    ref.set_gws()
    assert ctx106.run_worklist(str(wl), work, batch_size=batch) == 7 * nb
    for j in range(7):
        for b in range(nb):
            line0 = open(os.path.join(work, str(j), f"{j}_{b}.txt"), "rb").read().split(b"\r\n")[0]
            want = clb.hex_dd(*ref.job(d, b * batch, (b + 1) * batch, j, 0))
            assert line0.decode() == want, (j, b)
