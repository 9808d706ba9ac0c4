"""Served form of the kmerDist -> genDist pipeline under a Poisson job stream.

The reference runs one sample per process (crispy.sh:19-20).  Here a stream
of jobs (each a read set through the k-mer filter and banded NW) arrives at a
`crispy_server` (include/crispy_hip.h): `slots` contexts behind a FIFO queue
with HBM-budget admission.  Two policies are compared on the same arrivals:

  serial     slots=1, one job at a time (a process per sample, minus start-up)
  batched    slots=S, jobs admitted while others run share the device

Jobs: a seeded pool of synthetic read sets (crispy_amd.synth), sizes
uniform in {n_lo..n_hi} reads, log-normal lengths (median 300, sigma 0.35,
clipped to 150-550), k=6, band 10, default thresholds.  Arrival rates
are multiples of the serial service rate (1 / mean serial job time).  Every
job's result in the mix is compared bit for bit with its result alone.

    python -m crispy_amd.serve --slots 4 --jobs 48 --rates 0.5,1,2 --feeders 4
"""
import argparse
import hashlib
import json
import threading
import time

import numpy as np

from . import api, synth


def make_pool(count, seed, n_lo=500, n_hi=20000, K=6, band=10):
    """Seeded job pool: job k has n ~ U{n_lo..n_hi} reads (drawn from `seed`)
    generated with seed 1000 + k: log-normal lengths (median 300, sigma
    0.35, clipped to 150-550), n/200 species at 3-10 %, 1 % substitutions,
    0.5 % indels."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        species = max(5, n // 200)
        seqs, _ = synth.generate(n=n, seed=1000 + k, species=species,
                                 div=(0.03, 0.10), law=("lognormal", 300, 0.35, 150, 550),
                                 p_sub=0.01, p_indel=0.005)
        reads = api.Reads.from_list(seqs)
        pool.append(dict(reads=reads, K=K, band=band,
                         kthr=api.kmer_default_threshold(reads.n),
                         nthr=api.nw_default_threshold(reads.n)))
    return pool


def _digest(pairs, dists):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(pairs).tobytes())
    h.update(np.ascontiguousarray(dists).tobytes())
    return h.hexdigest()


def _submit(srv, job):
    return srv.submit(job["reads"], job["K"], job["kthr"], job["band"], job["nthr"])


def serial_service(srv, pool, repeats=1):
    """Runs every pool job alone (submit, wait) `repeats` times; returns the
    per-job mean service time (s) and each job's result digest."""
    times = np.zeros(len(pool))
    digests = []
    for r in range(repeats):
        for k, job in enumerate(pool):
            t0 = time.perf_counter()
            jid = _submit(srv, job)
            p, d = srv.fetch(jid)
            times[k] += time.perf_counter() - t0
            if r == 0:
                digests.append((_digest(p, d), len(d)))
    return times / repeats, digests


def poisson_run(srv, pool, arrivals, digests, feeders=4):
    """Submits pool[k % len(pool)] at arrivals[k] seconds from `feeders`
    threads (job k from thread k % feeders; a submit copies the job's reads,
    and ctypes releases the GIL meanwhile, so one thread alone caps the
    achievable arrival rate); returns per-job latency (s), the makespan (s),
    the achieved arrival rate (jobs/s over the submission span) and the
    number of results that differ from the job's result alone."""
    n = len(arrivals)
    ids = [None] * n
    t_sub = np.zeros(n)
    start = time.perf_counter()

    def feeder(f):
        for k in range(f, n, feeders):
            delay = start + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            t_sub[k] = time.perf_counter()
            ids[k] = _submit(srv, pool[k % len(pool)])

    threads = [threading.Thread(target=feeder, args=(f,)) for f in range(feeders)]
    for th in threads:
        th.start()
    lat = np.zeros(n)
    mismatches = 0
    last_end = start
    for k in range(n):
        while ids[k] is None:
            time.sleep(0.0005)
        info = srv.wait(ids[k])
        lat[k] = (info["end_ms"] - info["submit_ms"]) / 1e3
        p, d = srv.fetch(ids[k])
        last_end = max(last_end, time.perf_counter())
        if (_digest(p, d), len(d)) != digests[k % len(pool)]:
            mismatches += 1
    for th in threads:
        th.join()
    span = t_sub.max() - t_sub.min()
    return lat, last_end - start, (n - 1) / span if span > 0 else float("nan"), mismatches


def run(slots=4, jobs=48, rates=(0.5, 1.0, 2.0), pool_size=16, seed=7, budget_gib=64.0,
        n_lo=500, n_hi=20000, device=0, log=print, feeders=4):
    pool = make_pool(pool_size, seed, n_lo, n_hi)
    sizes = [j["reads"].n for j in pool]
    budget = int(budget_gib * (1 << 30))
    log(f"pool: {len(pool)} jobs, reads {min(sizes)}-{max(sizes)} (mean {np.mean(sizes):.0f})")
    servers = {"serial": api.Server(device, 1, budget), "batched": api.Server(device, slots, budget)}
    try:
        # warm both servers (code objects, first buffers), then time the serial policy
        for name, srv in servers.items():
            for jid in [_submit(srv, job) for job in pool[: max(1, slots)]]:
                srv.fetch(jid)
        svc, digests = serial_service(servers["serial"], pool, repeats=2)
        mean_svc = float(np.mean(svc))
        mu = 1.0 / mean_svc
        log(f"serial service: mean {mean_svc * 1e3:.2f} ms/job -> {mu:.1f} jobs/s")
        rng = np.random.default_rng(seed + 1)
        results = []
        for f in rates:
            lam = f * mu
            arrivals = np.cumsum(rng.exponential(1.0 / lam, jobs))
            arrivals -= arrivals[0]
            row = dict(rate_factor=f, arrival_rate=lam)
            for name, srv in servers.items():
                lat, span, offered, bad = poisson_run(srv, pool, arrivals, digests, feeders)
                st = srv.status()
                row[name] = dict(jobs_per_s=jobs / span, offered_per_s=offered,
                                 p50_ms=float(np.percentile(lat, 50) * 1e3),
                                 p99_ms=float(np.percentile(lat, 99) * 1e3), mismatches=bad,
                                 peak_reserved_gib=st["peak_reserved_bytes"] / (1 << 30))
                log(f"rate {f:.2f}x ({lam:.1f}/s) {name:8s}: offered {offered:7.1f}/s  {jobs / span:7.1f} jobs/s  "
                    f"p50 {row[name]['p50_ms']:8.2f} ms  p99 {row[name]['p99_ms']:8.2f} ms  "
                    f"mismatches {bad}")
            results.append(row)
        return dict(slots=slots, jobs=jobs, pool=len(pool), reads_min=min(sizes),
                    reads_max=max(sizes), reads_mean=float(np.mean(sizes)),
                    budget_gib=budget_gib, serial_ms_per_job=mean_svc * 1e3, rates=results)
    finally:
        for srv in servers.values():
            srv.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--jobs", type=int, default=48, help="jobs per arrival rate")
    ap.add_argument("--rates", default="0.5,1,2", help="multiples of the serial service rate")
    ap.add_argument("--pool", type=int, default=16, help="distinct read sets")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--n-lo", type=int, default=500)
    ap.add_argument("--n-hi", type=int, default=20000)
    ap.add_argument("--feeders", type=int, default=4, help="submitting threads")
    ap.add_argument("--json", default=None, help="write the summary here")
    a = ap.parse_args(argv)
    out = run(a.slots, a.jobs, tuple(float(x) for x in a.rates.split(",")), a.pool, a.seed,
              a.budget_gib, a.n_lo, a.n_hi, log=lambda s: print(s, flush=True), feeders=a.feeders)
    print(json.dumps(out), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
