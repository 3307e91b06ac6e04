"""GPU: the two seeded workloads at their full sizes.

* Uneven worklist jobs at N = 2.5e12 (d = 999,999,999,999): 64 term ranges whose lengths are
  drawn from [10^8, 10^11] with SplitMix64 seed 1 and rescaled to tile [0, d) exactly, times the
  seven formulas: 448 jobs in one persistent launch (clb_run_jobs), each written as a
  work/j/j_i.txt file (fsave's format, Bellard_192.hsp:514-533) and the files summed as
  workmodeSum0_6 does (workmodeSum0_6.hsp:1-49).  The sum must equal the position's single
  normal-mode result bit for bit; that result is the fixture tests/golden/position_2500000000000.json,
  computed by the all-integer and the precision-split kernels with equal digits.  About 50 s.
* Batched positions: 4096 positions log-uniform in [10^6, 10^11] (SplitMix64 seed 0, multiples of
  5), 1.35e13 terms.  The launch shape is autotuned once (clb_autotune, at the draw's
  term-weighted median position) and every launch replays a captured HIP graph (clb_set_graph).  The positions go in eight persistent launches of 512
  consecutive draws (each launch about 75 s, one test item each, so that progress is reported
  between them); a size-spread sample is then recomputed one position at a time on an untuned,
  eager context and must be bit-identical, and the smallest positions also equal the oracle.
  About 10 minutes.  The 64 smallest positions are also batched in the 106-bit mode with the
  reference's double-double kernels replayed, against single runs.
"""
import json
import os
import sys

import pytest

import clbellard_amd as clb
from clb_testlib import (GOLDEN, M192, batched_positions, py_normalize, uneven_jobs, w2i)

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, file=sys.stderr, flush=True)


def test_uneven_jobs_work_files_sum_to_single_run(tmp_path):
    with open(os.path.join(GOLDEN, "position_2500000000000.json")) as f:
        single = json.load(f)
    n, d = py_normalize(2500000000000)
    assert n == single["n"] and d == 999999999999
    cuts = uneven_jobs(d, 64, seed=1)
    assert cuts[0] == 0 and cuts[-1] == d and all(a < b for a, b in zip(cuts, cuts[1:]))
    # the last job reaches d, so it also carries the tail (KernelFuncs.as:56-59)
    jobs = [(d, cuts[i], cuts[i + 1], j) for j in range(7) for i in range(64)]
    note(f"uneven jobs: {len(jobs)} jobs at N={n} in one launch")
    with clb.Context(0, clb.PREC_U192) as ctx:
        parts = ctx.run_jobs(jobs)
        head_ms, total_ms = ctx.last_timing()
    note(f"  {total_ms / 1e3:.1f} s ({d / (total_ms / 1e3):.3e} terms/s)")
    work = str(tmp_path / "work")
    for idx, (_, k0, _, j) in enumerate(jobs):
        clb.write_work_file(work, j, idx % 64, parts[idx], clb.PREC_U192, "0\tMI355X",
                            int(total_ms / len(jobs)))
    summed, counts = clb.work_sum(work)
    assert counts == [64] * 7
    assert summed == single["hex"]
    assert f"{sum(w2i(p) for p in parts) & M192:048x}" == single["hex"]


POSITIONS = batched_positions(4096, seed=0)
PER_LAUNCH = 512
RESULTS = {}


def term_weighted_median(positions):
    """The position below which half of the draw's terms lie: where the work is (the count median,
    N = 2.9e8 for this draw, holds a negligible share of the 1.35e13 terms)."""
    ranked = sorted(positions)
    weights = [py_normalize(p)[1] for p in ranked]
    half, acc = sum(weights) / 2, 0
    for p, w in zip(ranked, weights):
        acc += w
        if acc >= half:
            return p
    return ranked[-1]


@pytest.fixture(scope="module")
def tuned():
    ctx = clb.Context(0, clb.PREC_U192)
    median = term_weighted_median(POSITIONS)
    tpl, bpc = ctx.autotune(median, 3000.0)
    note(f"batched positions: autotuned at N={median} (term-weighted median): {tpl} terms per "
         f"lane, {bpc} workgroups per CU, {ctx.get_schedule()}; captured HIP graph")
    ctx.set_graph(True)
    yield ctx
    ctx.close()


@pytest.mark.parametrize("part", range(len(POSITIONS) // PER_LAUNCH))
def test_batched_positions_launch(tuned, part):
    idx = list(range(part * PER_LAUNCH, (part + 1) * PER_LAUNCH))
    pos = [POSITIONS[i] for i in idx]
    terms = sum(py_normalize(p)[1] for p in pos)
    note(f"batched positions {idx[0]}..{idx[-1]}: {terms:.3e} terms in one launch")
    fracs = tuned.batch_positions(pos)
    head_ms, total_ms = tuned.last_timing()
    note(f"  {total_ms / 1e3:.1f} s ({terms / (total_ms / 1e3):.3e} terms/s)")
    assert len(fracs) == len(pos)
    RESULTS.update(zip(idx, fracs))


def test_batched_positions_equal_single_runs(oracle):
    if len(RESULTS) != len(POSITIONS):
        pytest.fail("the batched launches did not all run")
    by_size = sorted(range(len(POSITIONS)), key=lambda i: POSITIONS[i])
    sample = sorted({by_size[round(s * (len(by_size) - 1) / 15)] for s in range(16)})
    with clb.Context(0, clb.PREC_U192) as single:
        for i in sample:
            note(f"single run of position {i}: N={POSITIONS[i]}")
            assert single.normal(POSITIONS[i])[0] == RESULTS[i], (i, POSITIONS[i])
    for i in by_size[:2]:
        assert w2i(RESULTS[i]) == oracle.normal(POSITIONS[i]), (i, POSITIONS[i])


def test_batched_positions_106bit_replay_equal_single_runs():
    """The same draw's 64 smallest positions in the 106-bit mode, the reference's double-double
    kernels replayed (each position's seven BatchCalcs on their own 65536 group accumulators, then
    Sum7): the batched double-doubles equal single runs bit for bit."""
    pos = sorted(POSITIONS)[:64]
    note(f"106-bit replay: {len(pos)} positions {pos[0]}..{pos[-1]} batched, then one by one")
    with clb.Context(0, clb.PREC_DD106) as c:
        c.set_replay(clb.REF_GLOBAL_WORK_SIZE)
        batched = c.batch_positions(pos)
        for n, b in zip(pos, batched):
            assert c.normal(n)[0] == b, n
