"""GPU tests of the job server (crispy_server_*): a job's result in a mix of
concurrent jobs is bit-identical to its result alone and to the oracle, and
admission keeps the charged device bytes within the HBM budget."""
import sys

import numpy as np
import pytest

import crispy_amd as ca
from crispy_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_job(n, seed, K=6, band=10, law=("lognormal", 300, 0.35, 150, 550)):
    seqs, _ = synth.generate(n=n, seed=seed, species=max(5, n // 100), div=(0.03, 0.10), law=law,
                             p_sub=0.01, p_indel=0.005)
    reads = ca.Reads.from_list(seqs)
    return dict(reads=reads, K=K, band=band, kthr=ca.kmer_default_threshold(reads.n),
                nthr=ca.nw_default_threshold(reads.n))


def alone(job):
    with ca.Context(0) as c:
        c.upload_reads(ca.KMER, job["reads"])
        c.upload_reads(ca.NW, job["reads"])
        c.kmer_run(job["K"], job["kthr"])
        c.nw_run_on_kmer(job["band"], job["nthr"])
        return c.nw_fetch()


def submit(srv, job):
    return srv.submit(job["reads"], job["K"], job["kthr"], job["band"], job["nthr"])


def same(a, b):
    return a[0].shape == b[0].shape and np.array_equal(a[0], b[0]) and \
        np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


JOBS = [(300, 101, 6, 10), (2500, 102, 6, 10), (800, 103, 8, 10), (4000, 104, 6, 5),
        (1200, 105, 7, 2), (6000, 106, 6, 10), (150, 107, 6, 1), (3000, 108, 8, 10)]


@pytest.fixture(scope="module")
def jobs():
    return [make_job(n, s, K, b) for n, s, K, b in JOBS]


@pytest.fixture(scope="module")
def solo(jobs):
    return [alone(j) for j in jobs]


def test_large_served_job_matches_oracle(jobs):
    """A 20,000-read job of the served workload's length law (the top of its
    500-20,000 read range), served while 5 other jobs run, against the oracle
    on 128 seeded random rows of its k-mer matrix plus its first and last rows:
    its NW pairs in those rows, and their distances, are the oracle's."""
    big = make_job(20000, 1999)
    r = big["reads"]
    with ca.Server(0, 4, 64 << 30) as srv:
        others = [submit(srv, j) for j in jobs[1:6]]
        jid = submit(srv, big)
        p, d = srv.fetch(jid)
        for o in others:
            srv.fetch(o)
    ro = O.Reads(r.bases, r.offsets)
    rng = np.random.default_rng(77)
    rows = np.unique(np.concatenate([rng.choice(r.n - 1, 128, replace=False), [0, r.n - 2]]))
    kp, _ = O.kmer_rows(ro, big["K"], big["kthr"], rows)
    od, _, _ = O.nw_pairs(ro, kp, big["band"])
    keep = np.array([O.passes(float(x), big["nthr"]) for x in od], bool)
    want_p = np.asarray(kp, np.int32).reshape(-1, 2)[keep]
    want_d = np.asarray(od, np.float32)[keep]
    order = np.lexsort((want_p[:, 1], want_p[:, 0]))
    sel = np.isin(p[:, 0], rows)
    say(f"served 20k job: {len(d)} NW pairs, {int(sel.sum())} in {len(rows)} sampled rows")
    assert int(sel.sum()) > 0
    assert same((want_p[order], want_d[order]), (p[sel], d[sel]))


def test_small_job_matches_oracle(jobs, solo):
    job, (p, d) = jobs[0], solo[0]
    r = job["reads"]
    ro = O.Reads(r.bases, r.offsets)
    kp, _ = O.kmer_all(ro, job["K"], job["kthr"], nthreads=O.num_threads())
    od, _, _ = O.nw_pairs(ro, kp, job["band"])
    keep = np.array([O.passes(x, job["nthr"]) for x in od], bool)
    want_p = np.asarray(kp, np.int32).reshape(-1, 2)[keep]
    want_d = np.asarray(od, np.float32)[keep]
    # the threaded oracle emits pairs in thread order; the result is in (i, j) order
    order = np.lexsort((want_p[:, 1], want_p[:, 0]))
    say(f"oracle job: n={r.n} kmer {len(kp)} nw {len(want_d)}")
    assert len(want_d) > 0
    assert same((want_p[order], want_d[order]), (p, d))


@pytest.mark.parametrize("slots", [1, 3, 8])
def test_mix_matches_alone(jobs, solo, slots):
    """All jobs submitted at once (twice over); every result equals the job alone."""
    with ca.Server(0, slots, 32 << 30) as srv:
        ids = [submit(srv, j) for j in jobs + jobs]
        res = [srv.fetch(i) for i in ids]
    for k, r in enumerate(res):
        say(f"slots={slots} job {k}: {len(r[1])} pairs")
        assert same(r, solo[k % len(jobs)]), f"job {k} differs in the mix"


def test_jobs_overlap_with_free_slots(jobs):
    with ca.Server(0, 4, 32 << 30) as srv:
        ids = [submit(srv, jobs[5]) for _ in range(8)]
        info = [srv.wait(i) for i in ids]
        for i in ids:
            srv.fetch(i)
    iv = sorted((x["start_ms"], x["end_ms"]) for x in info)
    overlap = any(iv[k + 1][0] < iv[k][1] for k in range(len(iv) - 1))
    say("intervals", [(round(a, 2), round(b, 2)) for a, b in iv])
    assert overlap, "four free slots never ran two jobs at once"


def test_budget_admits_one_at_a_time(jobs):
    """A budget that fits the largest job but not two: jobs never overlap and
    the charged bytes stay within the budget."""
    est = [ca.job_bytes(j["reads"], j["K"], j["band"]) for j in jobs]
    budget = int(max(est) * 1.05)
    with ca.Server(0, 4, budget) as srv:
        big = int(np.argmax(est))
        ids = [submit(srv, jobs[big]) for _ in range(4)]
        info = [srv.wait(i) for i in ids]
        st = srv.status()
        for i in ids:
            srv.fetch(i)
    iv = sorted((x["start_ms"], x["end_ms"]) for x in info)
    say("est", max(est), "held", [x["held_bytes"] for x in info], "peak", st["peak_reserved_bytes"])
    for x in info:
        assert x["held_bytes"] <= x["est_bytes"], "job held more than its admission estimate"
    assert all(iv[k + 1][0] >= iv[k][1] for k in range(len(iv) - 1)), "jobs overlapped"
    assert st["peak_reserved_bytes"] <= budget


def test_estimate_covers_holdings(jobs):
    with ca.Server(0, 2, 32 << 30) as srv:
        ids = [submit(srv, j) for j in jobs]
        info = [srv.wait(i) for i in ids]
        for i in ids:
            srv.fetch(i)
    # a slot's buffers only grow, so after the k-th job (FIFO starts) it holds
    # at most the largest estimate among jobs 0..k
    top = 0
    for j, x in zip(jobs, info):
        top = max(top, x["est_bytes"])
        say(f"n={j['reads'].n} K={j['K']} band={j['band']} est {x['est_bytes']} held {x['held_bytes']}")
        assert x["held_bytes"] <= top


def test_errors():
    big = make_job(2000, 109)
    est = ca.job_bytes(big["reads"], 6, 10)
    with ca.Server(0, 2, est // 2) as srv:
        with pytest.raises(ca.CrispyError) as e:
            submit(srv, big)
        assert e.value.code == -3
        with pytest.raises(ca.CrispyError) as e:
            srv.submit(big["reads"], 9, 0.3, 10, 0.3)
        assert e.value.code == -1
        # a read longer than the reference's limit fails the job, not the server
        bad = ca.Reads.from_list([b"ACGT" * 600, b"ACGT" * 10])
        jid = srv.submit(bad, 6, 0.3, 10, 0.3)
        with pytest.raises(ca.CrispyError) as e:
            srv.fetch(jid)
        assert e.value.code == -1
        # the server keeps working
        small = make_job(200, 110)
        p, d = srv.fetch(submit(srv, small))
        assert same((p, d), alone(small))
        # empty and single-read jobs
        for reads in (ca.Reads(b"", np.zeros(1, np.int64)), ca.Reads.from_list([b"ACGTACGTAA"])):
            p, d = srv.fetch(srv.submit(reads, 6, 0.3, 10, 0.3))
            assert len(d) == 0 and p.shape == (0, 2)


def test_csr_views(jobs, solo):
    """The device CSR of a result equals the (i, j) pairs grouped by row, from
    a context and from the server."""
    job, (p, d) = jobs[1], solo[1]
    n = job["reads"].n
    want_rp = np.searchsorted(p[:, 0], np.arange(n + 1), side="left").astype(np.int64)
    with ca.Context(0) as c:
        c.upload_reads(ca.KMER, job["reads"])
        c.upload_reads(ca.NW, job["reads"])
        c.kmer_run(job["K"], job["kthr"])
        c.nw_run_on_kmer(job["band"], job["nthr"])
        rp, cols, vals = c.nw_csr(n)
        assert np.array_equal(rp, want_rp) and np.array_equal(cols, p[:, 1])
        assert np.array_equal(vals.view(np.uint32), d.view(np.uint32))
        # a row count below the largest row, and pairs not in (i, j) order, are refused
        with pytest.raises(ca.CrispyError):
            c.nw_csr(int(p[:, 0].max()))
        c.nw_run(p[::-1].copy(), job["band"], job["nthr"])
        with pytest.raises(ca.CrispyError):
            c.nw_csr(n)
    with ca.Server(0, 2, 32 << 30) as srv:
        jid = submit(srv, job)
        # row_ptr sized for another read count: refused, the job is kept
        with pytest.raises(ca.CrispyError):
            srv.fetch_csr(jid, n - 1)
        rp2, cols2, vals2 = srv.fetch_csr(jid, n)
        # a job is fetched once
        with pytest.raises(ca.CrispyError):
            srv.fetch(jid)
    assert np.array_equal(rp2, want_rp) and np.array_equal(cols2, p[:, 1])
    assert np.array_equal(vals2.view(np.uint32), d.view(np.uint32))


def near_copies(n, L, seed):
    """n reads that are one L-bp read with 0-2 substitutions each: every pair
    passes both filters, far more pairs per read than a job estimate allows."""
    rng = np.random.default_rng(seed)
    alph = np.frombuffer(b"ACGT", np.uint8)
    base = alph[rng.integers(0, 4, L)]
    seqs = set()
    while len(seqs) < n:
        s = base.copy()
        for _ in range(int(rng.integers(0, 3))):
            s[int(rng.integers(0, L))] = alph[int(rng.integers(0, 4))]
        seqs.add(s.tobytes())
    return ca.Reads.from_list(sorted(seqs))


def test_context_memory_limit():
    """A context whose buffers would grow past its limit fails the call with
    CRISPY_ERR_NOMEM and keeps what it held; with the limit lifted the same
    run gives the same result as a context that never had one."""
    reads = near_copies(1500, 300, 5)
    job = dict(reads=reads, K=6, band=10, kthr=ca.kmer_default_threshold(reads.n),
               nthr=ca.nw_default_threshold(reads.n))
    want = alone(job)
    with ca.Context(0) as c:
        c.upload_reads(ca.KMER, reads)
        c.upload_reads(ca.NW, reads)
        held = c.device_bytes()
        c.mem_limit(held + (1 << 20))
        with pytest.raises(ca.CrispyError) as e:
            c.kmer_run(job["K"], job["kthr"])
            c.nw_run_on_kmer(job["band"], job["nthr"])
        assert e.value.code == -3
        assert c.device_bytes() <= held + (1 << 20)
        c.mem_limit(0)
        c.kmer_run(job["K"], job["kthr"])
        c.nw_run_on_kmer(job["band"], job["nthr"])
        assert same(c.nw_fetch(), want)


def test_job_outgrowing_its_charge_runs_alone(jobs):
    """A job whose buffers outgrow its admission estimate (3,000 near-copies of
    one read: every pair passes, ~1,500 per read against the estimate's 256)
    is held to its charge while other jobs run: it stops at the cap, runs again
    alone charged the whole budget, and gives the result it gives alone; the
    charged bytes never exceed the budget and the other jobs are unaffected."""
    reads = near_copies(3000, 300, 6)
    big = dict(reads=reads, K=6, band=10, kthr=ca.kmer_default_threshold(reads.n),
               nthr=ca.nw_default_threshold(reads.n))
    want = alone(big)
    say(f"near copies: {len(want[1])} NW pairs, estimate {ca.job_bytes(reads, 6, 10)} bytes")
    budget = 4 << 30
    with ca.Server(0, 3, budget) as srv:
        before = [submit(srv, j) for j in jobs[1:4]]
        jid = submit(srv, big)
        after = [submit(srv, j) for j in jobs[4:6]]
        info = srv.wait(jid)
        p, d = srv.fetch(jid)
        others = [srv.fetch(i) for i in before + after]
        st = srv.status()
    say("big job:", info, "peak", st["peak_reserved_bytes"])
    assert info["ran_alone_again"]
    assert same((p, d), want)
    assert st["peak_reserved_bytes"] <= budget
    for k, r in zip([1, 2, 3, 4, 5], others):
        assert same(r, alone(jobs[k]))
