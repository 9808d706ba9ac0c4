"""Benchmark of the kmerDist -> genDist hot path on one MI355X.

One step = the whole distance pipeline over the 20,000-read workload
(synthetic 16S-like reads, 200-300 bp, 100 species, k = 6, band 10) with the
reads already resident in HBM:
  k-mer profiles -> all n(n-1)/2 k-mer distances + filter + (i, j) sort ->
  banded NW on every passing pair + filter.
`value` is read pairs per second through that step (n(n-1)/2 / step time).
The JSON line also carries the k-mer pairs/s and the NW GCUPS (in-band
cells), the roofline of the dominant kernel (timed with HIP events on the
stream it runs on) and the CPU oracle timed on a bounded sample.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload reads20k]
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
KMER_KERNELS = ("kmer_profile", "kmer_tile", "kmer_cand_sort", "kmer_cand", "kmer_sort")
NW_KERNELS = ("nw_bucket", "nw_band_reg", "nw_band_sliced", "nw_generic", "nw_filter")
# the NW launches of one step run concurrently on side streams; this event
# pair on the main stream brackets all of them (fork to join)
NW_PHASE = "nw_dp_phase"

def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(seqs, kmer_pairs, K, band, kthr, budget_s=24.0):
    """The oracle (the reference's CPU loops restated, OpenMP) on a bounded
    sample of the same workload: k-mer rows [0, R) against all j > i with
    the reference's per-pair critical section, and NW on the first k-mer
    pairs.  Rates are extrapolated to the whole workload."""
    from oracle import oracle as O
    reads = O.Reads.from_list(seqs)
    n = reads.n
    threads = O.num_threads()
    # k-mer: grow the row range until ~half the budget is spent
    rows, t_k, pairs_k = 0, 0.0, 0
    step = 64
    while t_k < budget_s / 2 and rows < n:
        r1 = min(n, rows + step)
        t0 = time.perf_counter()
        O.kmer_all(reads, K, kthr, nthreads=threads, rows=(rows, r1), cap=1 << 22)
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
        O.nw_pairs(reads, sub, band, nthreads=threads)
        t_n += time.perf_counter() - t0
        l1 = lens[sub[:, 0]].astype(np.int64)
        l2 = lens[sub[:, 1]].astype(np.int64)
        cells += band_cells_np(l1, l2, (l1 + l2) // (2 * band))
        done += len(sub)
        chunk *= 2
    nw_rate = done / t_n
    total_pairs = n * (n - 1) // 2
    t_all = total_pairs / kmer_rate + len(kmer_pairs) / nw_rate
    return {
        "value": total_pairs / t_all, "unit": "pairs/s", "cores": threads, "kind": "port",
        "sample": (f"oracle (OpenMP restatement of kmerCPU.cpp:11-76 and genCPU.cpp:290-526) on "
                   f"k-mer rows [0,{rows}) = {pairs_k} pairs in {t_k:.1f} s and NW on the first "
                   f"{done} k-mer-passing pairs in {t_n:.1f} s; whole job extrapolated as "
                   f"{total_pairs} pairs / kmer rate + {len(kmer_pairs)} NW pairs / NW rate"),
        "kmer_pairs_per_s": kmer_rate, "nw_pairs_per_s": nw_rate,
        "nw_gcups_in_band": cells / t_n / 1e9,
    }


def band_cells_np(L1, L2, kb):
    """sum over pairs of #{1<=i<=L2, 1<=j<=L1 : |i-j| <= kb} (vectorised)."""
    total = 0
    for i in range(1, int(L2.max()) + 1):
        hi = np.minimum(L1, i + kb)
        lo = np.maximum(1, i - kb)
        total += int(np.where(i <= L2, np.clip(hi - lo + 1, 0, None), 0).sum())
    return total


def measured_traffic(workload):
    """HBM bytes per NW DP phase (the fused family launches of one step) from
    the committed PMC summary of this workload (two rocprofv3 --pmc runs,
    scaled by a copy kernel of known size), or None when there is none."""
    import re
    path = os.path.join(ROOT, "profiles", f"kmer_nw_step_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"nw_dp_phase traffic bytes per step: (\d+)", open(path).read())
    return int(m.group(1)) if m else None


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="reads20k")
    ap.add_argument("--k", type=int, default=None, help="override the workload's k")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel chain kernel by kernel instead of replaying captured graphs")
    ap.add_argument("--cpu-budget", type=float, default=24.0)
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
    from crispy_amd import synth

    t0 = time.perf_counter()
    seqs, _, cfg = synth.make(args.workload)
    reads = ca.Reads.from_list(seqs)
    n = reads.n
    K, band = (args.k or cfg["K"]), cfg["band"]
    kthr, nthr = ca.kmer_default_threshold(n), ca.nw_default_threshold(n)
    log(f"workload {args.workload}: n={n} K={K} band={band} kthr={kthr:.4f} nthr={nthr:.4f} "
        f"(generated in {time.perf_counter() - t0:.1f}s)")

    ctx = ca.Context(0)
    ctx.upload_reads(ca.KMER, reads)
    ctx.upload_reads(ca.NW, reads)
    ctx.graphs(not args.no_graphs)

    def step():
        kc = ctx.kmer_run(K, kthr)
        nc = ctx.nw_run_on_kmer(band, nthr)
        return kc, nc

    for _ in range(args.warmup):
        kc, nc = step()
    ctx.sync()
    # the timed steps run without the per-kernel event timers; the same
    # number of steps then runs again with them for the kernel breakdown
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kc, nc = step()
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.timing(True)
    ctx.timing_reset()
    ctx.sync()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    dt_timed = time.perf_counter() - t1
    kern = {}
    for name in KMER_KERNELS + NW_KERNELS + (NW_PHASE,):
        ms, launches, units = ctx.timing_query(name)
        kern[name] = {"total_ms": ms, "launches": launches, "units": units,
                      "avg_ms": ms / launches if launches else 0.0}
    ctx.timing(False)
    total_pairs = n * (n - 1) // 2
    ms_step = dt / args.steps * 1e3
    k_ms = sum(kern[k]["total_ms"] for k in KMER_KERNELS) / args.steps
    phase_ms = kern[NW_PHASE]["total_ms"] / args.steps
    cells = kern[NW_PHASE]["units"] / args.steps
    log(f"step {ms_step:.2f} ms: kmer kernels {k_ms:.2f} ms ({kc} pairs pass), NW DP phase "
        f"{phase_ms:.2f} ms ({nc} pass), cells/step {cells:.3e}")
    for k, v in kern.items():
        log(f"  {k:14s} {v['total_ms'] / args.steps:9.3f} ms/step  launches {v['launches']:5d}  "
            f"avg {v['avg_ms']:.4f} ms")

    kt = kern["kmer_tile"]
    # the tile is an FP4 matrix product over the k<=6 presence vectors: 2 ops x kdim per pair
    kdim = 32 * max(16, 4 ** min(K, 6) // 32)
    kmer_tile_tops = kt["units"] * 2 * kdim / (kt["total_ms"] * 1e-3) / 1e12 if kt["total_ms"] else 0
    dp_launches = sum(kern[k]["launches"] for k in ("nw_band_reg", "nw_band_sliced", "nw_generic"))
    if phase_ms >= kt["total_ms"] / args.steps:
        # One DP phase per step: the fused register / sliced family launches run
        # concurrently on the worker streams; the phase is timed fork to join.
        ach = cells * NW_OPS_PER_CELL / (phase_ms * 1e-3) / 1e12
        roofline = {"kernel": "nw_dp_phase", "bound": "valu", "achieved": ach,
                    "peak": VALU_PEAK_TOPS, "unit": "Tops/s (int32 lane-ops)",
                    "frac": ach / VALU_PEAK_TOPS,
                    "work": f"{NW_OPS_PER_CELL} int ops x in-band DP cells ({cells:.4e}/step)",
                    "phase_ms": phase_ms, "avg_launch_ms": phase_ms,
                    "dp_form": ("kernel launches" if args.no_graphs else
                                "captured HIP graph: plan kernel + one launch per kernel family"),
                    "family_launches_per_step": dp_launches / args.steps if args.no_graphs else None,
                    "traffic": measured_traffic(args.workload),
                    "traffic_source": f"profiles/kmer_nw_step_{args.workload}_traffic.md"}
    else:
        roofline = {"kernel": "kmer_tile", "bound": "mfma", "achieved": kmer_tile_tops,
                    "peak": MFMA_FP4_PEAK_TOPS, "unit": "TFLOP/s (FP4 MFMA)",
                    "frac": kmer_tile_tops / MFMA_FP4_PEAK_TOPS,
                    "work": f"2 x {kdim} FP4 ops (presence-vector dot product) per read pair",
                    "avg_launch_ms": kt["avg_ms"], "traffic": None}

    out = {
        "metric": METRIC, "value": total_pairs / (dt / args.steps), "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "ms_per_step_with_kernel_timers": dt_timed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} synthetic 16S-like reads "
                               f"{describe_law(cfg['law'])}, {cfg['species']} species, "
                               f"k={K}, band {band}, one kmerDist+genDist pass",
                   "n_reads": n, "k": K, "band": band, "parallelism": "single GPU"},
        "kmer_pairs_per_s": total_pairs / (k_ms * 1e-3),
        "nw_gcups": cells / (phase_ms * 1e-3) / 1e9,
        "kmer_tile_mfma_frac": kmer_tile_tops / MFMA_FP4_PEAK_TOPS,
        "kmer_passing_pairs": kc, "nw_passing_pairs": nc,
        "kernels_ms_per_step": {k: v["total_ms"] / args.steps for k, v in kern.items()},
        "roofline": roofline,
    }
    if not args.no_cpu:
        kp, _ = ctx.kmer_fetch(dists=False)
        out["cpu_baseline"] = cpu_baseline(seqs, kp, K, band, kthr, args.cpu_budget)
        out["vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
