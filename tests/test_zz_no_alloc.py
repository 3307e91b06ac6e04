"""GPU: clb_run_jobs_noalloc (the entry point that allocates nothing) against clb_run_jobs, the
entry point it mirrors, bit for bit, through the C ABI: two calls in a row, the second smaller (the
shared device array keeps the first call's contents, which must not leak into the second), edge
jobs (empty ranges, ranges past d, tail-only jobs), the plain kernel, the largest call, and the
size and mode errors."""
import random
import sys

import pytest

import clbellard_amd as clb
from clb_testlib import py_normalize, w2i

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, file=sys.stderr, flush=True)


def _jobs(rng, n, count):
    _, d = py_normalize(n)
    cuts = sorted(rng.sample(range(1, d), count - 1))
    bounds = [0] + cuts + [d]
    return [(d, a, b, j) for j in range(7) for a, b in zip(bounds, bounds[1:])]


def test_noalloc_equals_run_jobs_twice_second_smaller(oracle):
    rng = random.Random(20261204)
    with clb.Context(0, clb.PREC_U192) as ctx:
        first = _jobs(rng, 10**9, 9)
        first += [(1001, 5, 5, 3), (1001, 1002, 2000, 0), (399999, 399999, 400010, 6),
                  (399999, 399990, 500000, 2)]
        note(f"no-alloc call 1: {len(first)} jobs")
        a = ctx.run_jobs_noalloc(first)
        assert a == ctx.run_jobs(first)
        for job, part in zip(first[-4:], a[-4:]):
            d, k0, k1, j = job
            want = (oracle.batch_calc(d, k0, k1, j) << clb_shift(j)) % (1 << 192)
            assert w2i(part) == want, job
        second = _jobs(rng, 10**6, 2)[:3]
        note(f"no-alloc call 2: {len(second)} jobs")
        b = ctx.run_jobs_noalloc(second)
        assert b == ctx.run_jobs(second)
        assert len(b) == 3
        # the whole of a position through the no-alloc entry point equals the normal run
        n = 10**7
        whole = ctx.run_jobs_noalloc(_jobs(rng, n, 3))
        assert sum(w2i(p) for p in whole) % (1 << 192) == w2i(ctx.normal(n)[0])


def clb_shift(j):
    return (5, 0, 8, 6, 2, 2, 0)[j]


def test_noalloc_plain_kernel_and_largest_call():
    rng = random.Random(7)
    _, d = py_normalize(10**8)
    jobs = []
    for _ in range(clb.NOALLOC_MAX_JOBS):
        k0 = rng.randrange(0, d)
        jobs.append((d, k0, k0 + rng.randrange(0, 5000), rng.randrange(7)))
    jobs[-1] = (d, d - 100, d + 1, 4)  # one job with the tail
    with clb.Context(0, clb.PREC_U192, kernel=clb.KERNEL_PLAIN) as ctx:
        note(f"no-alloc: {len(jobs)} jobs, plain kernel")
        got = ctx.run_jobs_noalloc(jobs)
        assert got == ctx.run_jobs(jobs)
        small = jobs[:5]
        assert ctx.run_jobs_noalloc(small) == ctx.run_jobs(small)
        with pytest.raises(clb.ClbError) as e:
            ctx.run_jobs_noalloc(jobs + [jobs[0]])
        assert e.value.code == clb.EINVAL
        ctx.set_replay(clb.REF_GLOBAL_WORK_SIZE)
        with pytest.raises(clb.ClbError) as e:
            ctx.run_jobs_noalloc(small)
        assert e.value.code == clb.EINVAL
        ctx.set_replay(0)
        assert ctx.run_jobs_noalloc(small) == ctx.run_jobs(small)


def test_noalloc_contexts_share_the_device_array():
    """Two contexts on the same GPU alternate calls of different sizes; each result equals
    clb_run_jobs's."""
    rng = random.Random(11)
    with clb.Context(0, clb.PREC_U192) as c1, clb.Context(0, clb.PREC_U192) as c2:
        for rep in range(3):
            j1 = _jobs(rng, 10**8, 4 + rep)
            j2 = _jobs(rng, 10**6, 2)
            r1, r2 = c1.run_jobs_noalloc(j1), c2.run_jobs_noalloc(j2)
            assert r1 == c2.run_jobs(j1) and r2 == c1.run_jobs(j2), rep
