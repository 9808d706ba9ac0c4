"""Benchmark of the kmerDist -> genDist hot path on one MI355X.

One step = the whole distance pipeline over a workload (default: the
250,000-read large-amplicon workload, 400-900 bp, 2,000 species, k = 6,
band 5) with the reads already resident in HBM:
  k-mer profiles -> all n(n-1)/2 k-mer distances + filter + (i, j) sort ->
  banded NW on every passing pair + filter -> the output side: the
  reference's file split and each file's (dist, i, j) sort on the device,
  copied to the host (what genDist writes as .npair_f / .ndist_f).
`value` is read pairs per second through that step (n(n-1)/2 / step time).
The JSON line also carries the k-mer pairs/s and the NW GCUPS (in-band
cells), the roofline of the dominant kernel (timed with HIP events on the
stream it runs on), the CPU oracle timed on a bounded sample (all cores and
the reference's shipped thread settings), the served form under a Poisson
job stream (jobs/s, p50/p99 latency) and the 100,000-read k = 6 step against
its own CPU sample.  Before the timed steps the library's autotuners pick
the k-mer exact-count form and the NW slice width on the workload
(crispy_kmer_autotune, crispy_nw_autotune); the line reports the choice and
the default settings' step time beside the tuned one (--no-autotune times
the defaults only).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload reads250k]
"""
import argparse
import json
import shutil
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-mer pairs/s + banded-NW GCUPS, 1 GPU; served jobs/s with p50/p99 latency"
VALU_PEAK_TOPS = 39.3    # 256 CU x 64 lanes x 2.4 GHz int32 lane-ops/s
MFMA_FP4_PEAK_TOPS = 10100.0  # MI355X dense FP4 matrix ops/s (10.1 PFLOP/s)
NW_OPS_PER_CELL = 13     # nominal int ops per in-band DP cell
KMER_KERNELS = ("kmer_profile", "kmer_tile", "kmer_cand_pop", "kmer_cand_sort", "kmer_cand", "kmer_sort")
NW_KERNELS = ("nw_bucket", "nw_band_reg", "nw_band_sliced", "nw_generic", "nw_filter")
# the NW launches of one step run concurrently on side streams; this event
# pair on the main stream brackets all of them (fork to join)
NW_PHASE = "nw_dp_phase"

def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(seqs, kmer_pairs, K, band, kthr, budget_s=24.0, kmer_threads=None,
                 nw_threads=None, nw_cpus=None):
    """The oracle (the reference's CPU loops restated, OpenMP) on a bounded
    sample of the same workload: k-mer rows [0, R) against all j > i with
    the reference's per-pair critical section, and NW on the first k-mer
    pairs.  Rates are extrapolated to the whole workload.  Threads default
    to every core; with nw_cpus the NW threads are confined to that many
    CPUs (genDist's shipped pinning)."""
    from oracle import oracle as O
    reads = O.Reads.from_list(seqs)
    n = reads.n
    procs = O.num_threads()
    kt = kmer_threads or procs
    nt = nw_threads or procs
    # k-mer: grow the row range until ~half the budget is spent
    rows, t_k, pairs_k = 0, 0.0, 0
    step = 64
    while t_k < budget_s / 2 and rows < n:
        r1 = min(n, rows + step)
        t0 = time.perf_counter()
        O.kmer_all(reads, K, kthr, nthreads=kt, rows=(rows, r1), cap=1 << 22)
        t_k += time.perf_counter() - t0
        pairs_k += sum(n - 1 - i for i in range(rows, r1))
        rows = r1
        step *= 2
    kmer_rate = pairs_k / t_k
    # NW: the first k-mer-passing pairs until the rest of the budget is spent
    done, t_n, chunk = 0, 0.0, 2000
    cells = 0
    lens = reads.lengths()
    while t_n < budget_s / 2 and done < len(kmer_pairs):
        sub = kmer_pairs[done:done + chunk]
        t0 = time.perf_counter()
        if nw_cpus:
            O.nw_pairs_as_shipped(reads, sub, band, nt, nw_cpus)
        else:
            O.nw_pairs(reads, sub, band, nthreads=nt)
        t_n += time.perf_counter() - t0
        l1 = lens[sub[:, 0]].astype(np.int64)
        l2 = lens[sub[:, 1]].astype(np.int64)
        cells += band_cells_np(l1, l2, (l1 + l2) // (2 * band))
        done += len(sub)
        chunk *= 2
    nw_rate = done / t_n if done else float("inf")
    total_pairs = n * (n - 1) // 2
    t_all = total_pairs / kmer_rate + len(kmer_pairs) / nw_rate
    nw_desc = (f"NW {nt} threads confined to {nw_cpus} CPUs" if nw_cpus else f"NW {nt} threads")
    return {
        "value": total_pairs / t_all, "unit": "pairs/s", "cores": max(kt, nt) if not nw_cpus else kt,
        "kind": "port",
        "threads": f"k-mer {kt} threads; {nw_desc}; {procs} CPUs available",
        "sample": (f"oracle (OpenMP restatement of kmerCPU.cpp:11-76 and genCPU.cpp:290-526) on "
                   f"k-mer rows [0,{rows}) = {pairs_k} pairs in {t_k:.1f} s and NW on the first "
                   f"{done} k-mer-passing pairs in {t_n:.1f} s; whole job extrapolated as "
                   f"{total_pairs} pairs / kmer rate + {len(kmer_pairs)} NW pairs / NW rate"),
        "kmer_pairs_per_s": kmer_rate, "nw_pairs_per_s": nw_rate,
        "nw_gcups_in_band": cells / t_n / 1e9 if t_n else None,
    }


def band_cells_np(L1, L2, kb):
    """sum over pairs of #{1<=i<=L2, 1<=j<=L1 : |i-j| <= kb} (vectorised)."""
    total = 0
    for i in range(1, int(L2.max()) + 1):
        hi = np.minimum(L1, i + kb)
        lo = np.maximum(1, i - kb)
        total += int(np.where(i <= L2, np.clip(hi - lo + 1, 0, None), 0).sum())
    return total


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` ("kmer_tile", or "nw_dp_phase": the
    fused family launches of one step) from the committed PMC summary of this
    workload (two rocprofv3 --pmc runs, scaled by a copy kernel of known size),
    or None when there is none."""
    import re
    names = {"kmer_tile": f"kmer_tile_{workload}_counters.md",
             "nw_dp_phase": f"kmer_nw_step_{workload}_traffic.md"}
    path = os.path.join(ROOT, "profiles", names[kernel])
    if not os.path.exists(path):
        return None, None
    m = re.search(kernel + r" traffic bytes per (?:step|launch): (\d+)", open(path).read())
    return (int(m.group(1)) if m else None), os.path.relpath(path, ROOT)


def describe_law(law):
    if law[0] == "uniform":
        return f"{law[1]}-{law[2]} bp"
    return f"log-normal median {law[1]} bp sigma {law[2]} clipped {law[3]}-{law[4]} bp"


def input_side(args):
    """Raw input of the workload: its reads drawn with replacement (about a
    third duplicates), shuffled, labelled, 80-column FASTA.  Times the device
    merge (sort + duplicate runs, HIP events) and, on the host, the reader,
    the writer, and the oracle's whole preprocess (std qsort restatement)."""
    import tempfile
    import crispy_amd as ca
    from crispy_amd import synth
    from oracle import oracle as O
    seqs, _, cfg = synth.make(args.workload)
    rng = np.random.default_rng(99)
    n = int(len(seqs) * 1.5)
    pick = rng.integers(0, len(seqs), n)
    tmp = tempfile.mkdtemp(prefix="crispy_input_")
    raw = os.path.join(tmp, "raw.fa")
    with open(raw, "wb") as f:
        for k, p in enumerate(pick):
            s = seqs[p]
            f.write(b">read_%d sample\n" % k + b"\n".join(s[i:i + 80] for i in range(0, len(s), 80)) + b"\n")
    t0 = time.perf_counter()
    rec = ca.preprocess_read(raw)
    t_read = time.perf_counter() - t0
    ctx = ca.Context(0)
    for _ in range(args.warmup):
        ctx.preprocess_order(rec.reads, rec.freq)
    ctx.timing(True)
    ctx.timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        order, freq = ctx.preprocess_order(rec.reads, rec.freq)
    dt = (time.perf_counter() - t0) / args.steps
    sort_ms = ctx.timing_query("pre_sort")[0] / args.steps
    merge_ms = ctx.timing_query("pre_merge")[0] / args.steps
    t0 = time.perf_counter()
    ca.preprocess_write(os.path.join(tmp, "raw.fa_Clean"), rec, order, freq)
    t_write = time.perf_counter() - t0
    t0 = time.perf_counter()
    m = O.preprocess_file(raw, os.path.join(tmp, "oracle_Clean"))
    t_oracle = time.perf_counter() - t0
    same = (open(os.path.join(tmp, "raw.fa_Clean"), "rb").read() ==
            open(os.path.join(tmp, "oracle_Clean"), "rb").read())
    ctx.close()
    print(json.dumps({"metric": "preprocess-equivalent merge, records/s", "value": rec.n / dt,
                      "unit": "records/s", "records": rec.n, "unique": len(order), "oracle_unique": m,
                      "identical_to_oracle": same, "ms_per_merge": dt * 1e3, "device_sort_ms": sort_ms,
                      "device_runs_ms": merge_ms, "host_read_s": t_read, "host_write_s": t_write,
                      "oracle_whole_preprocess_s": t_oracle, "workload": args.workload,
                      "steps": args.steps, "warmup": args.warmup}))
    shutil.rmtree(tmp)


def plain_vs_fused(args):
    """The one-launch-per-stage plain GPU form (the reference's two executables
    chained through files, kmerDist.cpp -> .kpair -> genDist: the k-mer pairs
    come back to the host, then NW runs over 589,824-pair chunks of them with
    a fetch after each, genCUDA.cu:156-226; NW DP kernels launched one by
    one) against the fused form (crispy_nw_run_on_kmer: the k-mer result stays
    on the device, the DP phase replays its captured graph) on the same reads;
    both results must be identical."""
    import crispy_amd as ca
    from crispy_amd import synth
    seqs, _, cfg = synth.make(args.workload)
    reads = ca.Reads.from_list(seqs)
    n = reads.n
    K, band = (args.k or cfg["K"]), cfg["band"]
    kthr, nthr = ca.kmer_default_threshold(n), ca.nw_default_threshold(n)
    ctx = ca.Context(0)
    ctx.upload_reads(ca.KMER, reads)
    ctx.upload_reads(ca.NW, reads)
    chunk = 589824

    def plain():
        ctx.graphs(False)
        ctx.kmer_run(K, kthr)
        kp, _ = ctx.kmer_fetch(dists=False)
        outs = []
        for c0 in range(0, len(kp), chunk):
            ctx.nw_run(kp[c0:c0 + chunk], band, nthr)
            outs.append(ctx.nw_fetch())
        return (np.concatenate([o[0] for o in outs]), np.concatenate([o[1] for o in outs])) if outs else None

    def fused():
        ctx.graphs(True)
        ctx.kmer_run(K, kthr)
        ctx.nw_run_on_kmer(band, nthr)
        return ctx.nw_fetch()

    res = {}
    for name, fn in (("plain", plain), ("fused", fused)):
        for _ in range(args.warmup):
            out = fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        ctx.sync()
        res[name] = ((time.perf_counter() - t0) / args.steps * 1e3, out)
        log(f"{name}: {res[name][0]:.2f} ms/step (incl. fetching the NW result)")
    same = (np.array_equal(res["plain"][1][0], res["fused"][1][0]) and
            np.array_equal(res["plain"][1][1].view(np.uint32), res["fused"][1][1].view(np.uint32)))
    ctx.close()
    pairs = n * (n - 1) // 2
    print(json.dumps({"metric": "k-mer pairs/s, fused vs one-launch-per-stage form",
                      "unit": "pairs/s", "workload": args.workload, "n_reads": n, "k": K, "band": band,
                      "plain_ms_per_step": res["plain"][0], "fused_ms_per_step": res["fused"][0],
                      "plain_pairs_per_s": pairs / (res["plain"][0] * 1e-3),
                      "fused_pairs_per_s": pairs / (res["fused"][0] * 1e-3),
                      "speedup": res["plain"][0] / res["fused"][0], "identical_results": same,
                      "nw_passing_pairs": int(len(res["fused"][1][1])), "steps": args.steps,
                      "warmup": args.warmup}))


def served_leg(seconds=5.0, slots=4, rate_factor=2.0, pool_size=16, seed=7):
    """The served form on the C4 job stream: a seeded pool of jobs of
    500-20,000 reads (log-normal 150-550 bp, k = 6, band 10), Poisson arrivals
    at `rate_factor` x the one-at-a-time service rate for about `seconds`,
    against `slots` slots and against one slot (one job at a time) on the same
    arrivals; every result is compared with the job's result alone."""
    from crispy_amd import api, serve
    pool = serve.make_pool(pool_size, seed)
    budget = 64 << 30
    one, many = api.Server(0, 1, budget), api.Server(0, slots, budget)
    try:
        for srv in (one, many):
            for jid in [serve._submit(srv, job) for job in pool[:slots]]:
                srv.fetch(jid)
        svc, digests = serve.serial_service(one, pool, repeats=2)
        mu = 1.0 / float(np.mean(svc))
        lam = rate_factor * mu
        jobs = max(64, int(lam * seconds))
        rng = np.random.default_rng(seed + 1)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, jobs))
        arrivals -= arrivals[0]
        out = {}
        for name, srv in (("slots", many), ("one_slot", one)):
            lat, span, offered, bad = serve.poisson_run(srv, pool, arrivals, digests)
            out[name] = dict(jobs_per_s=jobs / span, offered_per_s=offered,
                             p50_ms=float(np.percentile(lat, 50) * 1e3),
                             p99_ms=float(np.percentile(lat, 99) * 1e3), mismatches=bad)
            log(f"served {name}: {jobs / span:.1f} jobs/s, p50 {out[name]['p50_ms']:.1f} ms, "
                f"p99 {out[name]['p99_ms']:.1f} ms, {bad} results differ from alone")
        sizes = [j["reads"].n for j in pool]
        return {"jobs_per_s": out["slots"]["jobs_per_s"], "p50_ms": out["slots"]["p50_ms"],
                "p99_ms": out["slots"]["p99_ms"], "slots": slots,
                "one_slot_jobs_per_s": out["one_slot"]["jobs_per_s"],
                "one_slot_p50_ms": out["one_slot"]["p50_ms"], "one_slot_p99_ms": out["one_slot"]["p99_ms"],
                "offered_jobs_per_s": out["slots"]["offered_per_s"], "jobs": jobs,
                "arrivals": f"Poisson at {rate_factor}x the one-at-a-time rate ({mu:.1f} jobs/s)",
                "results_differing_from_alone": out["slots"]["mismatches"] + out["one_slot"]["mismatches"],
                "job_reads": f"{min(sizes)}-{max(sizes)} (mean {np.mean(sizes):.0f}), 150-550 bp, k=6, band 10"}
    finally:
        one.close()
        many.close()


KMER_FORMS = {0: "exact counts in the tile", 1: "candidates (hashed for k >= 6) + exact pass",
              2: "candidates on the exact bitmaps + exact pass"}


def run_workload(ca, workload, K_override, steps, warmup, graphs, output_side=True, timers=True,
                 autotune=True):
    """Times `steps` pipeline steps of `workload` after `warmup` untimed ones
    (and, with timers, the same steps again with the per-kernel HIP-event
    timers on).  With autotune, the default settings' step time is taken
    first (as many steps), then the k-mer exact-count form and the NW slice
    width are tuned on the workload and the timed steps run tuned.  Returns
    a dict of the measurements plus the context."""
    from crispy_amd import synth
    from crispy_amd.pipeline import OutputBuffers, fused_output_files
    t0 = time.perf_counter()
    seqs, _, cfg = synth.make(workload)
    reads = ca.Reads.from_list(seqs)
    n = reads.n
    K, band = (K_override or cfg["K"]), cfg["band"]
    kthr, nthr = ca.kmer_default_threshold(n), ca.nw_default_threshold(n)
    log(f"workload {workload}: n={n} K={K} band={band} kthr={kthr:.4f} nthr={nthr:.4f} "
        f"(generated in {time.perf_counter() - t0:.1f}s)")
    ctx = ca.Context(0)
    ctx.upload_reads(ca.KMER, reads)
    ctx.upload_reads(ca.NW, reads)
    ctx.graphs(graphs)

    host_out = OutputBuffers()  # the output files' host arrays, reused from step to step

    def step(out=True):
        kc = ctx.kmer_run(K, kthr)
        nc = ctx.nw_run_on_kmer(band, nthr)
        files = fused_output_files(ctx, host_out) if out else None
        return kc, nc, files

    for _ in range(warmup):
        kc, nc, files = step()
    ctx.sync()
    tune = None
    if autotune:
        t0 = time.perf_counter()
        for _ in range(steps):
            step(output_side)
        ctx.sync()
        default_ms = (time.perf_counter() - t0) / steps * 1e3
        default_bits = ctx.kmer_tile_bits()
        t0 = time.perf_counter()
        kmode = ctx.kmer_autotune(K, kthr, repeats=2)
        width = ctx.nw_autotune(band, repeats=2)
        # the defaults: exact-count form by pair count (-1), slice width by the cost model (0)
        default_mode = 1 if K > 6 or n * (n - 1) // 2 >= 10 ** 9 else 0
        tuned_bits = ctx.kmer_tile_bits()
        tune = {"kmer_exact_pass": kmode, "kmer_form": KMER_FORMS[kmode], "kmer_tile_bits": tuned_bits,
                "nw_slice_width": width,
                "nw_slice_policy": "cost model" if width == 0 else f"{width} cells per lane",
                "chosen_is_default": kmode == default_mode and width == 0 and tuned_bits == default_bits,
                "default_ms_per_step": default_ms, "tuning_s": time.perf_counter() - t0}
        for _ in range(warmup):
            step()
        ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        kc, nc, files = step(output_side)
    ctx.sync()
    dt = time.perf_counter() - t0
    res = dict(ctx=ctx, seqs=seqs, cfg=cfg, n=n, K=K, band=band, kthr=kthr, nthr=nthr, kc=kc, nc=nc,
               tile_bits=ctx.kmer_tile_bits(),
               ms_step=dt / steps * 1e3, nfiles=len(files) if files is not None else None)
    if tune is not None:
        tune["tuned_ms_per_step"] = res["ms_step"]
        res["autotune"] = tune
        log(f"autotune: k-mer form {kmode} ({KMER_FORMS[kmode]}), NW slice width {tune['nw_slice_policy']}; "
            f"{default_ms:.2f} ms/step default -> {res['ms_step']:.2f} ms/step tuned")
    if output_side:
        # the same steps without the output side (file split, sort, copy out)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(False)
        ctx.sync()
        res["ms_step_no_output"] = (time.perf_counter() - t0) / steps * 1e3
    if timers:
        ctx.timing(True)
        ctx.timing_reset()
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(steps):
            step(False)
        ctx.sync()
        res["ms_step_timers"] = (time.perf_counter() - t1) / steps * 1e3
        kern = {}
        for name in KMER_KERNELS + NW_KERNELS + (NW_PHASE,):
            ms, launches, units = ctx.timing_query(name)
            kern[name] = {"total_ms": ms, "launches": launches, "units": units,
                          "avg_ms": ms / launches if launches else 0.0}
        ctx.timing(False)
        res["kern"] = kern
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="reads250k")
    ap.add_argument("--k", type=int, default=None, help="override the workload's k")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baselines")
    ap.add_argument("--no-served", action="store_true", help="skip the served-form leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the 100,000-read k=6 leg")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel chain kernel by kernel instead of replaying captured graphs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the default settings only (no k-mer form / NW slice-width tuning)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds per CPU baseline leg")
    ap.add_argument("--served-seconds", type=float, default=5.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB read + 1 GiB write elementwise kernel (3x), to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--plain-vs-fused", action="store_true",
                    help="time the one-launch-per-stage form (k-mer pairs to the host, NW in "
                         "589,824-pair chunks) against the fused form on the workload")
    ap.add_argument("--input-side", action="store_true",
                    help="time the preprocess-equivalent merge (device) against the oracle's "
                         "preprocess on a raw FASTA with duplicates, instead of the step")
    args = ap.parse_args()
    if args.input_side:
        return input_side(args)
    if args.plain_vs_fused:
        return plain_vs_fused(args)
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.add(a, 1, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 1 << 30, "bytes_written": 1 << 30,
                          "dispatches": 3}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the path shards by k-mer tile range with no "
              f"collective (see crispy_kmer_run's tile range), and this run uses exactly one "
              f"device; got --gpus {args.gpus}, WORLD_SIZE {world}.", file=sys.stderr)
        sys.exit(2)

    import crispy_amd as ca

    r = run_workload(ca, args.workload, args.k, args.steps, args.warmup, not args.no_graphs,
                     autotune=not args.no_autotune)
    ctx, n, K, band, kern = r["ctx"], r["n"], r["K"], r["band"], r["kern"]
    steps = args.steps
    total_pairs = n * (n - 1) // 2
    k_ms = sum(kern[k]["total_ms"] for k in KMER_KERNELS) / steps
    phase_ms = kern[NW_PHASE]["total_ms"] / steps
    cells = kern[NW_PHASE]["units"] / steps
    log(f"step {r['ms_step']:.2f} ms ({r['ms_step_no_output']:.2f} ms without the output side, "
        f"{r['nfiles']} file(s)): kmer kernels {k_ms:.2f} ms ({r['kc']} pairs pass), NW DP phase "
        f"{phase_ms:.2f} ms ({r['nc']} pass), cells/step {cells:.3e}")
    for k, v in kern.items():
        log(f"  {k:14s} {v['total_ms'] / steps:9.3f} ms/step  launches {v['launches']:5d}  "
            f"avg {v['avg_ms']:.4f} ms")

    kt = kern["kmer_tile"]
    # the tile is an FP4 matrix product over the presence vectors (the k <= 6
    # codes, or their 1,280, 1,536 or 2,048 hash buckets for the k = 6 candidate
    # filter): 2 ops x kdim per pair
    kdim = r["tile_bits"]
    kmer_tile_tops = kt["units"] * 2 * kdim / (kt["total_ms"] * 1e-3) / 1e12 if kt["total_ms"] else 0
    dp_launches = sum(kern[k]["launches"] for k in ("nw_band_reg", "nw_band_sliced", "nw_generic"))
    if phase_ms >= kt["total_ms"] / steps:
        # One DP phase per step: the fused register / sliced family launches run
        # concurrently on the worker streams; the phase is timed fork to join.
        ach = cells * NW_OPS_PER_CELL / (phase_ms * 1e-3) / 1e12
        traffic, src = measured_traffic(args.workload, "nw_dp_phase")
        roofline = {"kernel": "nw_dp_phase", "bound": "valu", "achieved": ach,
                    "peak": VALU_PEAK_TOPS, "unit": "Tops/s (int32 lane-ops)",
                    "frac": ach / VALU_PEAK_TOPS,
                    "work": f"{NW_OPS_PER_CELL} int ops x in-band DP cells ({cells:.4e}/step)",
                    "phase_ms": phase_ms, "avg_launch_ms": phase_ms,
                    "dp_form": ("kernel launches" if args.no_graphs else
                                "captured HIP graph: plan kernel + one launch per kernel family"),
                    "family_launches_per_step": dp_launches / steps if args.no_graphs else None,
                    "traffic": traffic, "traffic_source": src}
    else:
        traffic, src = measured_traffic(args.workload, "kmer_tile")
        roofline = {"kernel": "kmer_tile", "bound": "mfma", "achieved": kmer_tile_tops,
                    "peak": MFMA_FP4_PEAK_TOPS, "unit": "TFLOP/s (FP4 MFMA)",
                    "frac": kmer_tile_tops / MFMA_FP4_PEAK_TOPS,
                    "work": (f"2 x {kdim} FP4 ops (presence-vector dot product"
                             f"{f', 4,096 6-mer codes hashed into {kdim:,} buckets' if kdim in (1280, 1536, 2048) else ''}) per read pair, "
                             f"{int(kt['units'] / max(1, kt['launches']))} pairs per launch"),
                    "avg_launch_ms": kt["avg_ms"], "traffic": traffic, "traffic_source": src}

    out = {
        "metric": METRIC, "value": total_pairs / (r["ms_step"] * 1e-3), "unit": "pairs/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": r["ms_step"],
        "ms_per_step_without_output_side": r["ms_step_no_output"],
        "ms_per_step_with_kernel_timers": r["ms_step_timers"],
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} synthetic 16S-like reads "
                               f"{describe_law(r['cfg']['law'])}, {r['cfg']['species']} species, "
                               f"k={K}, band {band}, one kmerDist+genDist pass incl. the output "
                               f"side ({r['nfiles']} file(s) split, sorted by (dist, i, j) on the "
                               f"device, copied to the host)",
                   "n_reads": n, "k": K, "band": band, "parallelism": "single GPU"},
        "kmer_pairs_per_s": total_pairs / (k_ms * 1e-3),
        "nw_gcups": cells / (phase_ms * 1e-3) / 1e9,
        "kmer_tile_mfma_frac": kmer_tile_tops / MFMA_FP4_PEAK_TOPS,
        "kmer_passing_pairs": r["kc"], "nw_passing_pairs": r["nc"],
        "kernels_ms_per_step": {k: v["total_ms"] / steps for k, v in kern.items()},
        "roofline": roofline,
    }
    if "autotune" in r:
        out["autotune"] = r["autotune"]
    kp = None
    if not args.no_cpu:
        kp, _ = ctx.kmer_fetch(dists=False)
    ctx.close()
    if not args.no_cpu:
        from oracle import oracle as O
        procs = O.num_threads()
        out["cpu_baseline"] = cpu_baseline(r["seqs"], kp, K, band, r["kthr"], args.cpu_budget)
        out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        # kmerDist's default threads (numProcs / 2, kmerMain.cpp:232-236); genDist's
        # numProcs threads pinned to 8 CPUs (genMain.cpp:49-73, -DSET_AFFINITY)
        out["cpu_baseline_as_shipped"] = cpu_baseline(
            r["seqs"], kp, K, band, r["kthr"], args.cpu_budget,
            kmer_threads=procs // 2 if procs > 2 else 1, nw_threads=procs, nw_cpus=min(8, procs))
        out["vs_cpu_as_shipped"] = out["value"] / out["cpu_baseline_as_shipped"]["value"]
        log(f"CPU oracle: {out['cpu_baseline']['value']:.3e} pairs/s all cores, "
            f"{out['cpu_baseline_as_shipped']['value']:.3e} pairs/s as shipped")
    if not args.no_served:
        out["served"] = served_leg(args.served_seconds)
    if not args.no_c3 and args.workload != "reads100k":
        c3 = run_workload(ca, "reads100k", 6, steps, args.warmup, not args.no_graphs, timers=False,
                          autotune=not args.no_autotune)
        c3n = c3["n"] * (c3["n"] - 1) // 2
        entry = {"workload": "reads100k, k=6, band 10 (the same step incl. the output side)",
                 "ms_per_step": c3["ms_step"], "pairs_per_s": c3n / (c3["ms_step"] * 1e-3),
                 "nw_passing_pairs": c3["nc"]}
        if "autotune" in c3:
            entry["autotune"] = c3["autotune"]
        c3kp = None
        if not args.no_cpu:
            c3kp, _ = c3["ctx"].kmer_fetch(dists=False)
        c3["ctx"].close()
        if not args.no_cpu:
            cb = cpu_baseline(c3["seqs"], c3kp, 6, c3["band"], c3["kthr"], args.cpu_budget)
            entry["cpu_pairs_per_s"] = cb["value"]
            entry["cpu_sample"] = cb["sample"]
            entry["cpu_cores"] = cb["cores"]
            entry["vs_cpu"] = entry["pairs_per_s"] / cb["value"]
        log(f"reads100k k=6: {c3['ms_step']:.2f} ms/step, {entry['pairs_per_s']:.3e} pairs/s"
            + (f", {entry['vs_cpu']:.0f}x the CPU oracle" if "vs_cpu" in entry else ""))
        out["reads100k_k6"] = entry
    print(json.dumps(out))


if __name__ == "__main__":
    main()
