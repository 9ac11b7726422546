"""Benchmarks of the HMM scoring path on one MI355X.

Default (--workload classify), the headline:

Workload (one step): classify 100,000 simulated reads (3 channels, 12
timesteps) against 1,000 dye-seqs (lengths 6-30, 0-5 dyes per channel) with a
5-sigma pruning cutoff -- 10^8 read x dye-seq pairs -- on the FP32 rescaled
path, inputs resident in HBM. Prints one JSON line.

Also measured (untimed for `value`): top-1 agreement with the FP64 GPU path on
the same reads, per-kernel HIP-event durations, algorithmic FLOPs of the
pruned work, and the compiled reference's multithreaded CPU classify on a
bounded sample of the same reads.

--workload hybrid: classify hybrid (-k 100 -s 0.5 -H 1000 -p 5) of 10^5 reads
against 20,000 dye-seqs (C=3, T=12, lengths 6-30, 0-5 dyes per channel) with
the kd-tree over their simulated, deduplicated dye-tracks (100 per dye-seq by
default); one step is one full classify call from host arrays (kNN search on
the GPU, vote on the host, HMM rescoring of the 1,000 candidates on the GPU).
CPU baseline: the reference's HybridClassifier on a bounded sample.

--workload shortlist: the learned shortlist (MLP 36 -> 512 x 3 -> 20,000,
BF16 on the matrix cores) trained on 10^7 simulated reads of the 20,000
dye-seqs above, then 10^5 reads classified by FP32 HMM rescoring of its
top-256 candidates (full-HMM fallback when the shortlist holds under half the
posterior mass), against the plain
FP64 GPU HMM over all 20,000 dye-seqs (-p 5 both).

--workload simulate: the GPU simulator (simulate dt): 20,000 dye-seqs x 1,000
peptides each (2x10^7 simulated dye tracks), deduplicated on the device by a
radix sort of the packed tracks; one step is one call. CPU baseline: the host
simulator (same streams) on 2x10^5 peptides.

--workload fit: Baum-Welch fitting of one dye-seq ("..0..1..2..0..1..2..",
2 dyes in each of 3 channels, T = 30) to 10^6 simulated reads; one step is one
EM iteration (GPU E-step over every read + host M-step). CPU baseline: the
reference's serial E-step loop on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
# HBM bytes of one classify step of the default workload (prep, every bin's kernels and the FP64
# decision), from two rocprofv3 --pmc runs of this bench (FETCH_SIZE, then WRITE_SIZE, per
# dispatch, summed over the step), each scaled by what the same counter reads for a 512 MiB
# device copy in the same run; refreshed whenever the classify kernels change.
CLASSIFY_HBM_BYTES = 13_452_000_000
CLASSIFY_HBM_SOURCE = ("rocprofv3 --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per step "
                       "(profiles/classify_fp32_100k_x_1k_cut5.md)")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(model, strings, rads, cutoff, sample_reads):
    """The reference's own HMMClassifier (oracle/_ref, compiled from its
    sources) on a bounded sample of the same reads, OpenMP over reads; falls
    back to the oracle's C restatement when the reference library is absent."""
    import numpy as np
    from tests._oracle import Oracle, Reference
    o = Oracle()
    threads = o.L.wpo_num_threads()
    sample = np.ascontiguousarray(rads[:sample_reads])
    counts = np.ones(len(strings), dtype=np.int32)
    ids = np.arange(len(strings), dtype=np.int32)
    ref = Reference.load()
    kind = "reference" if ref is not None else "port"
    impl = ref if ref is not None else o
    t0 = time.perf_counter()
    impl.classify(model, strings, counts, ids, sample, cutoff)
    dt = time.perf_counter() - t0
    pairs = sample.shape[0] * len(strings)
    return {"value": pairs / dt, "unit": "pairs/s", "cores": int(threads), "kind": kind,
            "sample": f"{sample.shape[0]} reads x {len(strings)} dye-seqs of the same workload, "
                      f"{dt:.1f} s wall",
            "host": host_cpu(),
            "build": ("the reference's hmm/classifier sources compiled by oracle/ref.mk with its release flags "
                      "(g++ -O2 -flto, OpenMP over reads as HMMClassifier::classify does)" if ref is not None else
                      "the oracle's C restatement (gcc -O2, OpenMP over reads)")}


def host_cpu():
    """CPU model, logical CPUs this process may use, and the machine's count."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--reads", type=int, default=100000)
    ap.add_argument("--dye-seqs", type=int, default=1000)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--cpu-sample-reads", type=int, default=3000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--no-north-star", action="store_true",
                    help="leave out the learned-shortlist leg (10^5 reads x 20,000 dye-seqs) of the default line")
    ap.add_argument("--no-fit-leg", action="store_true",
                    help="leave out the Baum-Welch E-step leg (10^6 reads, T = 30) of the default line")
    ap.add_argument("--shortlist-reads", type=int, default=100000)
    ap.add_argument("--labels", choices=["hmm", "truth"], default="hmm",
                    help="shortlist training labels: the HMM's pick among same-dye-count dye-seqs, or the source")
    ap.add_argument("--workload", choices=["classify", "fit", "hybrid", "shortlist", "simulate"], default="classify")
    ap.add_argument("--peptides-per-dye-seq", type=int, default=1000)
    ap.add_argument("--train-reads", type=int, default=10000000)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--tracks-per-dye-seq", type=int, default=1000)
    ap.add_argument("--serial-bins", action="store_true",
                    help="run each call's bins in order on one stream (profiling: each kernel's duration is its own)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(reads would shard across GPUs by read index with no exchange between them)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "fit":
        return bench_fit(args)
    if args.workload == "hybrid":
        return bench_hybrid(args)
    if args.workload == "shortlist":
        return bench_shortlist(args)
    if args.workload == "simulate":
        return bench_simulate(args)

    import numpy as np
    import torch

    import whatprot_amd as w

    T, C, S, N, cutoff = 12, 3, args.dye_seqs, args.reads, 5.0
    model = w.SequencingModel.uniform(C)  # seqparams_atto647n_x3 priors with mu = 1
    log(f"generating {S} dye-seqs and {N} reads")
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1001)
    rads, true_idx = w.simulate_radiometries(model, T, strings, N, 2001, gpu=True)
    entries = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    prec = w.WP_PRECISION_FP32 if args.precision == "fp32" else w.WP_PRECISION_FP64
    clf = w.HMMClassifier(T, C, model, w.SequencingSettings(cutoff), entries, prec)
    if args.serial_bins:
        clf.set_branches(False)

    dev = torch.device("cuda:0")
    d_rads = torch.from_numpy(rads).to(dev)
    d_id = torch.empty(N, dtype=torch.int32, device=dev)
    d_adj = torch.empty(N, dtype=torch.float64, device=dev)
    d_ls = torch.empty(N, dtype=torch.float64, device=dev)
    d_lt = torch.empty(N, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    if args.copy_mb:
        src = torch.empty(args.copy_mb * (1 << 20) // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_mb} MiB")
        del src, dst

    def step_of(c, out_id):
        def step():
            c.classify_device(N, d_rads.data_ptr(), out_id.data_ptr(), d_adj.data_ptr(), d_ls.data_ptr(),
                              d_lt.data_ptr(), stream.cuda_stream)
        return step

    def event_steps(step, warmup, steps):
        """W untimed steps, then K steps timed with HIP events on the launch
        stream, bracketed by torch.cuda.synchronize(): (ms per step, wall s)."""
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps, time.perf_counter() - t0

    step = step_of(clf, d_id)
    log(f"warmup {args.warmup}, timing {args.steps} steps")
    ms_per_step, wall = event_steps(step, args.warmup, args.steps)
    ids = d_id.cpu().numpy()

    # per-kernel durations over the same steps, HIP events on the same stream
    clf.set_timing(True)
    prep_ms, cls_ms = [], []
    for _ in range(args.steps):
        step()
        a, b = clf.last_timing()
        prep_ms.append(a)
        cls_ms.append(b)
    clf.set_timing(False)
    torch.cuda.synchronize()
    classify_ms = float(np.mean(cls_ms))
    log(f"ms/step {ms_per_step:.3f}  prep {np.mean(prep_ms):.3f} ms  classify {classify_ms:.3f} ms")

    variants, tune_ms = clf.variants()
    pairs_nonempty, states, flops = clf.work(rads)
    achieved = flops / (classify_ms * 1e-3) / 1e12

    result = {
        "metric": "HMM read x dye-seq pairs/s on 1 MI355X (FP32 rescaled, 5-sigma pruning)",
        "value": N * S / (ms_per_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == w.WP_PRECISION_FP32 else "fp64",
        "data": "synthetic (seeded simulator, seqparams_atto647n_x3 priors, mu = 1)",
        "config": {"workload": "classify hmm: 100k reads x 1k dye-seqs, C=3, T=12, -p 5",
                   "reads": N, "dye_seqs": S, "channels": C, "timesteps": T, "dist_cutoff": cutoff,
                   "precision": args.precision},
        "reads_per_s": N / (ms_per_step * 1e-3),
        "wall_s": wall,
        "launch": {"graph_replay": True, "per_bin_read_lists": True,
                   "quad_rows_variant_per_bin": [int(v) for v in variants],
                   "autotune_ms_per_bin": [[round(float(x), 3) for x in row] for row in tune_ms]},
        "nonempty_pairs": pairs_nonempty,
        "state_timesteps": states,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": CLASSIFY_HBM_BYTES if (N, S, prec) == (100000, 1000, w.WP_PRECISION_FP32) else None,
                     "traffic_source": CLASSIFY_HBM_SOURCE,
                     "kernel": "the classify launches of one call (every bin's kernels and wide passes, and the "
                               "FP64 decision of near-ties)",
                     "kernel_ms": classify_ms, "flops_per_launch": flops},
    }
    if prec == w.WP_PRECISION_FP32:
        # the FP64 decision of near-ties: what it rescored, what it costs, and
        # what the FP32 argmax alone would give
        n_ref, n_all = clf.refine_stats(N)
        clf.set_refine(False)
        d_id_raw = torch.empty(N, dtype=torch.int32, device=dev)
        ms_raw, _ = event_steps(step_of(clf, d_id_raw), 1, args.steps)
        clf.set_refine(True)
        ids_raw = d_id_raw.cpu().numpy()
        result["fp64_decision"] = {"reads_rescored": n_ref, "rescored_over_every_candidate": n_all,
                                   "ms_per_step_without": ms_raw, "ms_added": ms_per_step - ms_raw}
    if not args.no_fp64_check and prec == w.WP_PRECISION_FP32:
        log("FP64 path on the same reads, same step protocol")
        clf64 = w.HMMClassifier(T, C, model, w.SequencingSettings(cutoff), entries, w.WP_PRECISION_FP64)
        d_id64 = torch.empty(N, dtype=torch.int32, device=dev)
        ms64, _ = event_steps(step_of(clf64, d_id64), 1, max(2, args.steps // 2))
        ids64 = d_id64.cpu().numpy()
        result["top1_vs_fp64"] = float((ids == ids64).mean())
        result["fp64_decision"]["fp32_argmax_top1_vs_fp64"] = float((ids_raw == ids64).mean())
        result["fp64_gpu_ms_per_step"] = ms64
        result["fp64_gpu_pairs_per_s"] = N * S / (ms64 * 1e-3)
        result["speedup_vs_fp64_gpu"] = ms64 / ms_per_step
        result["top1_vs_truth"] = float((ids == true_idx).mean())
    if not args.no_cpu_baseline:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(model, strings, rads, cutoff, args.cpu_sample_reads)
    if not args.no_north_star:
        log("north star: learned shortlist at 10^5 reads x 20,000 dye-seqs")
        result["north_star"] = shortlist_leg(args, steps=3)
    if not args.no_fit_leg:
        log("fit: Baum-Welch E-step of 10^6 reads at T = 30")
        fa = argparse.Namespace(**vars(args))
        fa.reads, fa.precision, fa.steps, fa.warmup, fa.no_cpu_baseline = 100000, "fp32", 2, 1, True
        r = fit_result(fa)
        result["fit"] = {"workload": r["config"]["workload"], "reads_per_s": r["value"],
                         "ms_per_em_iteration": r["ms_per_step"], "roofline": r["roofline"],
                         "fp32_vs_fp64_max_rel_stat": r.get("fp32_vs_fp64_max_rel_stat"),
                         "timing": "1 untimed EM iteration, then 2 (GPU E-step over every read, device "
                                   "reduction, host M-step) between synchronizations"}
    print(json.dumps(result), flush=True)


def timed_steps(step, warmup, steps):
    """W untimed steps, then K steps bracketed by torch.cuda.synchronize();
    returns seconds per step (wall clock: the steps are synchronous calls or
    end in a synchronize)."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def shortlist_leg(args, steps, warmup=1, fp64_steps=2):
    """The learned-shortlist workload: MLP 36 -> 512 x 3 -> 20,000 trained on
    simulated reads of 20,000 dye-seqs, then N reads classified by FP32 HMM
    rescoring of each read's top-256 (full-HMM fallback below shortlist
    posterior mass 0.5), against the plain FP64 GPU HMM over every dye-seq.
    Both sides are timed the same way: radiometries resident in device
    memory, outputs to device memory, W untimed steps, then K steps between
    synchronizations."""
    import numpy as np
    import torch

    import whatprot_amd as w

    T, C, S, N, H, cutoff = 12, 3, 20000, args.shortlist_reads, 256, 5.0
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1002)
    entries = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    log(f"shortlist: simulating {args.train_reads} training reads and {N} query reads")
    t0 = time.perf_counter()
    tx, ty = w.simulate_radiometries(model, T, strings, args.train_reads, 4003, gpu=True)
    t_sim = time.perf_counter() - t0
    qx, qy = w.simulate_radiometries(model, T, strings, N, 3002, gpu=True)
    clf64 = w.HMMClassifier(T, C, model, w.SequencingSettings(cutoff), entries, w.WP_PRECISION_FP64)
    clf32 = w.HMMClassifier(T, C, model, w.SequencingSettings(cutoff), entries, w.WP_PRECISION_FP32)
    t_label = 0.0
    if args.labels == "hmm":
        log("shortlist: HMM labels (the FP32 HMM's pick among each read's same-dye-count dye-seqs)")
        t0 = time.perf_counter()
        ty_h = w.hmm_labels(clf32, tx, ty)
        t_label = time.perf_counter() - t0
        log(f"shortlist: {np.mean(ty_h != ty):.4f} of the labels moved ({t_label:.1f} s)")
        ty = ty_h
    tx = tx.reshape(len(tx), -1).astype(np.float32)
    net = w.ShortlistMLP(T * C, 512, 3, S, seed=1)
    log(f"shortlist: training {args.epochs} epochs ({t_sim:.0f} s simulating)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = net.train(tx, ty, batch=4096, epochs=args.epochs, lr=1e-3, seed=9)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del tx, ty
    flops_step = 2 * 4096 * (64 * 512 + 2 * 512 * 512 + 512 * S) * 3
    train_tflops = flops_step * (args.train_reads // 4096) * args.epochs / t_train / 1e12
    log("shortlist: plain FP64 GPU HMM over every dye-seq")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    d_qx = torch.from_numpy(np.ascontiguousarray(qx)).to(dev)
    d_id = torch.empty(N, dtype=torch.int32, device=dev)
    d_adj = torch.empty(N, dtype=torch.float64, device=dev)

    def step64():
        clf64.classify_device(N, d_qx.data_ptr(), d_id.data_ptr(), d_adj.data_ptr(), 0, 0, stream.cuda_stream)

    t64 = timed_steps(step64, warmup, fp64_steps)
    ids64 = d_id.cpu().numpy()
    log(f"shortlist: FP64 {t64 * 1e3:.1f} ms/step; shortlist steps")
    idx, post = net.shortlist(qx, H)
    recall = float(np.mean([ids64[i] in idx[i] for i in range(N)]))
    times, out = {}, {}
    for key, (thr, crit) in {"mass": (0.5, "mass"), "max": (0.5, "max"), "none": (0.0, "max")}.items():
        res = {}

        def step_sl():
            res["nf"] = net.classify_device(clf32, N, d_qx.data_ptr(), H, d_id.data_ptr(), d_adj.data_ptr(), thr,
                                            crit, stream.cuda_stream)

        times[key] = timed_steps(step_sl, warmup, steps)
        out[key] = (d_id.cpu().numpy(), d_adj.cpu().numpy(), res["nf"])
    ids, adj, nf = out["mass"]
    return {
        "workload": f"shortlist: MLP 36->512x3->{S} (BF16 MFMA) trained on {args.train_reads} simulated reads "
                    f"({args.epochs} epochs), {N} reads, top-{H} rescored by FP32 HMM, full-HMM fallback below "
                    f"shortlist posterior mass 0.5, -p 5",
        "reads": N, "dye_seqs": S,
        "reads_per_s": N / times["mass"],
        "ms_per_step": times["mass"] * 1e3,
        "plain_fp64_ms_per_step": t64 * 1e3,
        "plain_fp64_reads_per_s": N / t64,
        "speedup_vs_plain_fp64": t64 / times["mass"],
        "target_speedup": 50.0,
        "recall_fp64_top1_at_256": recall,
        "top1_vs_fp64": float(np.mean(ids == ids64)),
        "fallback_fraction": nf / N,
        "timing": f"radiometries resident in device memory and device outputs on both sides, {warmup} untimed "
                  f"step(s), then {steps} (shortlist) and {fp64_steps} (FP64) steps between synchronizations",
        "labels": args.labels,
        "label_s": t_label,
        "train_s": t_train,
        "train_tflops_bf16": train_tflops,
        "epoch_loss": [float(x) for x in losses],
        "max_posterior_criterion": {"reads_per_s": N / times["max"], "speedup_vs_plain_fp64": t64 / times["max"],
                                    "fallback_fraction": out["max"][2] / N,
                                    "top1_vs_fp64": float(np.mean(out["max"][0] == ids64))},
        "no_fallback": {"reads_per_s": N / times["none"], "speedup_vs_plain_fp64": t64 / times["none"],
                        "top1_vs_fp64": float(np.mean(out["none"][0] == ids64))},
        "top1_vs_truth": float(np.mean(ids == qy)),
    }


def bench_hybrid(args):
    import numpy as np
    import torch

    import whatprot_amd as w

    T, C, S, N = 12, 3, 20000, args.reads
    k, sig, H, cutoff = 100, 0.5, 1000, 5.0
    model = w.SequencingModel.uniform(C)
    log(f"generating {S} dye-seqs, {args.tracks_per_dye_seq} tracks each, {N} reads")
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1002)
    seqs = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    t0 = time.perf_counter()
    tracks = w.simulate_dye_tracks(model, T, seqs, args.tracks_per_dye_seq, 2002, gpu=True)
    t_sim = time.perf_counter() - t0
    rads, true_idx = w.simulate_radiometries(model, T, strings, N, 3002, gpu=True)
    log(f"{len(tracks)} distinct dye-tracks ({t_sim:.1f} s); building the hybrid classifier (host kd-tree)")
    prec = w.WP_PRECISION_FP32 if args.precision == "fp32" else w.WP_PRECISION_FP64
    t0 = time.perf_counter()
    hc = w.HybridClassifier(T, C, model, w.SequencingSettings(cutoff), k, sig, tracks, H, seqs, prec)
    t_build = time.perf_counter() - t0
    log(f"built in {t_build:.1f} s; building the kNN classifier for the search-only timing")
    nn = w.NNClassifier(T, C, model, k, sig, tracks)
    log("warmup")
    for _ in range(args.warmup):
        hc.classify_arrays(rads[: min(N, 2000)])
    torch.cuda.synchronize()
    log("kNN search alone")
    t0 = time.perf_counter()
    cnt, _, _ = nn.search(rads)
    t_knn = time.perf_counter() - t0
    log(f"timing {args.steps} hybrid classify calls")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ids, adj = hc.classify_arrays(rads)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    result = {
        "metric": "hybrid classify reads/s on 1 MI355X (kNN k=100 + HMM over top 1000, -p 5)",
        "value": N / (ms * 1e-3),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded simulator, seqparams_atto647n_x3 priors with mu = 1)",
        "config": {"workload": f"classify hybrid: {N} reads, {S} dye-seqs, {args.tracks_per_dye_seq} simulated "
                               f"dye-tracks per dye-seq ({len(tracks)} distinct), C=3, T=12, k=100, sigma=0.5, "
                               f"H=1000, -p 5", "reads": N, "dye_seqs": S, "dye_tracks": len(tracks),
                   "precision": args.precision},
        "knn_search_ms": t_knn * 1e3,
        "knn_queries_per_s": N / t_knn,
        "knn_mean_neighbours": float(np.mean(cnt)),
        "hmm_pairs_per_step": int(N * H),
        "build_s": t_build,
        "top1_vs_truth": float((ids == true_idx).mean()),
    }
    if not args.no_cpu_baseline and len(tracks) > 2_000_000:
        # the reference builds its kd-tree and precomputations single-threaded
        # in one call that runs for minutes at this size
        result["cpu_baseline"] = None
        result["cpu_baseline_note"] = (f"not run at {len(tracks)} tracks (the reference's HybridClassifier "
                                       "construction alone runs for minutes); see --tracks-per-dye-seq 100")
    elif not args.no_cpu_baseline:
        from tests._oracle import Oracle, Reference
        ref = Reference.load()
        if ref is not None:
            n = min(N, 200)
            log(f"CPU baseline: the reference's HybridClassifier on {n} reads (its kd-tree build first)")
            # construction alone (kd-tree build, precomputations), then with n reads
            t1 = time.perf_counter()
            ref.hybrid_classify(model, tracks, k, sig, H, strings, [1] * S, list(range(S)), rads[:0], cutoff)
            d0 = time.perf_counter() - t1
            t1 = time.perf_counter()
            rid, _, _ = ref.hybrid_classify(model, tracks, k, sig, H, strings, [1] * S, list(range(S)), rads[:n],
                                            cutoff)
            d1 = time.perf_counter() - t1 - d0
            log(f"reference: construction {d0:.1f} s, {n} reads {d1:.1f} s")
            result["cpu_baseline"] = {"value": n / d1, "unit": "reads/s", "cores": int(Oracle().L.wpo_num_threads()),
                                      "kind": "reference",
                                      "sample": f"{n} reads of the same workload through HybridClassifier::classify "
                                                f"(OpenMP over reads), construction ({d0:.1f} s) excluded, "
                                                f"{d1:.1f} s wall"}
            result["cpu_agreement_ids"] = float((rid == ids[:n]).mean())
            if prec == w.WP_PRECISION_FP32:
                hc64 = w.HybridClassifier(T, C, model, w.SequencingSettings(cutoff), k, sig, tracks, H, seqs,
                                          w.WP_PRECISION_FP64)
                ids64, _ = hc64.classify_arrays(rads[:n])
                result["cpu_agreement_ids_fp64"] = float((rid == ids64).mean())
    print(json.dumps(result), flush=True)


def bench_shortlist(args):
    r = shortlist_leg(args, args.steps)
    result = {
        "metric": "learned-shortlist classify reads/s on 1 MI355X (MLP top-256 + FP32 HMM rescoring, -p 5)",
        "value": r["reads_per_s"],
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (MLP) + fp32 (HMM)",
        "data": "synthetic (seeded simulator, seqparams_atto647n_x3 priors with mu = 1)",
        "config": {"workload": r["workload"], "reads": r["reads"], "dye_seqs": r["dye_seqs"],
                   "train_reads": args.train_reads, "epochs": args.epochs},
    }
    result.update({k: v for k, v in r.items() if k not in ("workload", "reads", "dye_seqs", "ms_per_step")})
    print(json.dumps(result), flush=True)


def bench_simulate(args):
    """GPU dye-track simulation with deduplication (the input stage of
    classify nn / hybrid): 20,000 dye-seqs x --peptides-per-dye-seq peptides,
    C = 3, T = 12. One step = one wp_simulate_dye_tracks_gpu call (simulate,
    sort-based dedup, copy of the distinct tracks to the host)."""
    import numpy as np
    import torch

    import whatprot_amd as w

    T, C, S, per = 12, 3, 20000, args.peptides_per_dye_seq
    model = w.SequencingModel.uniform(C)
    strings = w.generate_dye_seqs(C, S, 6, 30, 0, 5, 1002)
    seqs = [w.DyeSeqEntry(s, 1, i) for i, s in enumerate(strings)]
    peptides = S * per
    cap = min(peptides, 12_000_000)
    log(f"warmup ({peptides} peptides per step)")
    for _ in range(args.warmup):
        tr = w.simulate_dye_tracks(model, T, seqs, per, 7, gpu=True, max_tracks=cap)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        tr = w.simulate_dye_tracks(model, T, seqs, per, 100 + k, gpu=True, max_tracks=cap)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ntr = len(tr)
    log(f"{ntr} distinct tracks, {dt * 1e3:.1f} ms/step")
    # algorithmic bytes: the T*C one-byte counts of every peptide written once
    # and read once by the key packing, plus the distinct tracks (T*C shorts)
    bytes_step = peptides * T * C * 2 + ntr * T * C * 2
    achieved = bytes_step / dt / 1e9
    result = {
        "metric": "simulated dye-track peptides/s with dedup on 1 MI355X",
        "value": peptides / dt,
        "unit": "peptides/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 counts + fp64 rng",
        "data": "synthetic (seeded simulator, seqparams_atto647n_x3 priors)",
        "config": {"workload": f"simulate dt: {S} dye-seqs x {per} peptides, C=3, T=12, dedup",
                   "dye_seqs": S, "peptides": peptides, "distinct_tracks": ntr},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None,
                     "kernel": "whole call (simulate + radix-sort dedup + host copy)"},
    }
    if not args.no_cpu_baseline:
        n = 10
        t0 = time.perf_counter()
        w.simulate_dye_tracks(model, T, seqs, n, 100)
        d1 = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": S * n / d1, "unit": "peptides/s", "cores": 1, "kind": "port",
                                  "sample": f"{S * n} peptides of the same workload through the host simulator "
                                            f"(same streams, std::map dedup), {d1:.1f} s wall"}
    print(json.dumps(result), flush=True)


def fit_flops_per_read(T, D):
    """Algorithmic FLOPs of one read's E-step: the forward pass by the
    per-state counts of the classify roofline over full (uncut) boxes, the
    backward pass (the same operators transposed) and the statistics (taken
    as one more pass of the same size)."""
    nb = 1
    for d in D:
        nb *= d + 1
    binom = sum(2 * (d + 2) for d in D)
    fwd = nb * (3 + 2 + binom)  # t = 0: initial block, duds, emission
    for t in range(1, T):
        fwd += (t + 1) * nb * (3 + 4 + 8 + 2 + binom)
    return 3 * fwd


def bench_fit(args):
    print(json.dumps(fit_result(args)), flush=True)


def fit_result(args):
    import numpy as np
    import torch

    import whatprot_amd as w

    T, C, ds = 30, 3, "..0..1..2..0..1..2.."
    N = args.reads if args.reads != 100000 else 1000000
    D = [ds.count(str(c)) for c in range(C)]
    truth = w.SequencingModel.uniform(C, p_edman_failure=0.07, p_bleach=0.04, p_dud=0.08, p_detach=0.04)
    start = w.SequencingModel.uniform(C)
    log(f"simulating {N} reads of {ds}")
    rads, _ = w.simulate_radiometries(truth, T, [ds], N, 3004, gpu=True)
    prec = w.WP_PRECISION_FP32 if args.precision == "fp32" else w.WP_PRECISION_FP64
    f = w.HMMFitter(T, C, ds, rads, prec)
    log("warmup")
    m = start
    for _ in range(args.warmup):
        st, _ = f.estep(m)
        m, _ = w.fit_mstep(T, C, ds, m, N, st)
    torch.cuda.synchronize()
    log(f"timing {args.steps} EM iterations")
    t0 = time.perf_counter()
    m = start
    for _ in range(args.steps):
        st, _ = f.estep(m)
        m, _ = w.fit_mstep(T, C, ds, m, N, st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.steps
    flops = fit_flops_per_read(T, D) * N
    achieved = flops / (ms * 1e-3) / 1e12
    peak = PEAK_FP32_TFLOPS if prec == w.WP_PRECISION_FP32 else 78.6
    result = {
        "metric": "Baum-Welch E-step reads/s on 1 MI355X (fit, uncut)",
        "value": N / (ms * 1e-3),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded simulator, seqparams_atto647n_x3 priors with mu = 1)",
        "config": {"workload": "fit: 1 dye-seq L=20, 2 dyes x 3 channels, T=30, 1e6 reads, one EM iteration per step",
                   "reads": N, "timesteps": T, "channels": C, "dye_seq": ds, "precision": args.precision},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel": "fit_rows_kernel",
                     "flops_per_launch": flops},
    }
    if prec == w.WP_PRECISION_FP32:
        f64 = w.HMMFitter(T, C, ds, rads[:20000], w.WP_PRECISION_FP64)
        f32 = w.HMMFitter(T, C, ds, rads[:20000], w.WP_PRECISION_FP32)
        a, la = f64.estep(start)
        b, lb = f32.estep(start)
        result["fp32_vs_fp64_max_rel_stat"] = float(np.max(np.abs(b - a) / np.maximum(np.abs(a), 1e-300)))
        result["fp32_vs_fp64_rel_logL"] = float(abs(lb - la) / abs(la))
    if not args.no_cpu_baseline:
        from tests._oracle import Reference
        ref = Reference.load()
        if ref is not None:
            n = min(N, 1000)
            t1 = time.perf_counter()
            ref.fit_estep(start, ds, rads[:n])
            d1 = time.perf_counter() - t1
            result["cpu_baseline"] = {"value": n / d1, "unit": "reads/s", "cores": 1, "kind": "reference",
                                      "sample": f"{n} reads of the same workload, one serial E-step "
                                                f"(HMMFitter::fit's loop), {d1:.1f} s wall"}
    return result


if __name__ == "__main__":
    main()
