"""Throughput of the `aln` search path on one MI355X.

Workload (default): 200,000 synthetic 150-nt reads, translated search
(aln -q d -t p) against 1,000,000 synthetic proteins (~330M aa, one chunk
resident in HBM), BLOSUM62 11/1, -b 10 -v 1, SEG on.  A step is one full
search of all reads: query lists, seed search / distance / ungapped
extension on the device, chain filter, gapped extension on the device,
AddResults, BuildResults and tabular output.

Three forms run on the same inputs: the plain GPU form (stages one after the
other, jobs in hit order), the batched form (query blocks pipelined so that
the host stages of one block overlap the device stages of the next) and the
batched learned-scheduled form (the same, with gapped extensions ordered on
the device by the MLP's predicted cost; the MLP is trained on the GPU from
the plain run's extension traces during set-up).  Their outputs must be
byte-identical.  `value` is the batched learned-scheduled form's query
residues per second (query residues = amino acids of the six translated
frames); the other forms' rates are reported beside it.

The CPU baseline is the oracle (a single-threaded restatement of the
reference's CPU aligner) timed on a bounded sample of the same reads.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("ghostz-gpu_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM bytes per work unit of the seed kernels at this workload, measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and calibrated against a copy kernel
TRAFFIC_PROFILE = "profiles/hbm_traffic_translated_reads_200k_vs_db_1m.txt"


def traffic_per_unit(kernel):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), TRAFFIC_PROFILE)
    try:
        for line in open(path):
            f = line.split()
            if f and f[0] == "gz::" + kernel and len(f) >= 7:
                return float(f[6])
    except OSError:
        pass
    return None


def query_residues(view):
    c = view["codes"]
    return int(np.count_nonzero(c != 25))


def train_cost_model(gz, db, db_codes, rng, n_reads, epochs, seed, max_traces=2_000_000):
    """Logs extension traces with the plain schedule on training reads, then
    trains the MLP on the GPU.  Returns (flat weights, report dict)."""
    import synthetic
    names, res, offs = synthetic.translated_reads(rng, db_codes, n_reads)
    q = gz.Queries(names, res, offs, translated=True)
    al = gz.Aligner(db, record_traces=1)
    al.align(q)
    tr = al.traces()
    if len(tr) > max_traces:  # a uniform sample keeps training to seconds
        tr = tr[np.sort(rng.choice(len(tr), max_traces, replace=False))]
    concat = al.last_concatenated_queries()
    chunk_seq = {c: db.chunk(c)["sequence"] for c in np.unique(tr[:, 2])} if len(tr) else {}
    feats = np.zeros((len(tr), 128), np.uint8)
    for c, seqc in chunk_seq.items():
        for rev in (0, 1):
            sel = np.nonzero((tr[:, 2] == c) & (tr[:, 3] == rev))[0]
            if len(sel):
                feats[sel] = gz.cost_features(concat, seqc, tr[sel, 0], tr[sel, 1], rev)
    target = np.log2(np.maximum(tr[:, 4], 1)).astype(np.float32)
    perm = rng.permutation(len(tr))
    n_test = max(1, len(tr) // 10)
    test, train = perm[:n_test], perm[n_test:]
    mlp = gz.Mlp(gz.Mlp.init_weights(seed))
    batch = 1024
    t0 = time.perf_counter()
    steps = 0
    for ep in range(epochs):
        order = rng.permutation(train)
        for b in range(0, len(order) - batch + 1, batch):
            idx = order[b:b + batch]
            mlp.train_step(feats[idx], target[idx], 1e-3)
            steps += 1
    train_s = time.perf_counter() - t0
    pred = mlp.predict(feats[test])
    mae = float(np.abs(pred - target[test]).mean())
    base = float(np.abs(target[test] - target[train].mean()).mean())
    # inference throughput (kernel time only, all traces)
    mlp.predict(feats)
    seeds_per_s = len(feats) / (mlp.last_kernel_ms / 1e3) if mlp.last_kernel_ms > 0 else None
    report = {"traces": int(len(tr)), "held_out": int(n_test), "epochs": epochs,
              "train_steps": steps, "train_seconds": round(train_s, 3),
              "mae_log2_cells": round(mae, 4), "mae_constant_predictor": round(base, 4),
              "inference_seeds_per_s": seeds_per_s}
    return mlp.weights(), report


def cpu_baseline(gz, oracle_py, db, names, res, offs, m, nl, budget_s):
    """Oracle (single thread) on the first reads of the workload."""
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    n = 128
    while True:
        qs = gz.Queries(names[:n], res[:int(offs[n])], offs[:n + 1], translated=True)
        view = qs.view()
        t0 = time.perf_counter()
        oracle_py.align(view, chunks, info, m, nl)
        dt = time.perf_counter() - t0
        log(f"cpu baseline: {n} reads in {dt:.2f} s")
        if dt >= budget_s / 3 or n >= len(names):
            return {"value": query_residues(view) / dt, "unit": "query aa/s", "cores": 1,
                    "kind": "port", "sample": f"first {n} reads of the workload vs the full "
                    f"database ({dt:.1f} s, oracle, one thread)"}
        n = min(len(names), int(n * max(2.0, budget_s / 2 / max(dt, 1e-3))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the workload (reads and database proteins)")
    ap.add_argument("--train-reads", type=int, default=40000)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--pipeline-blocks", type=int, default=0,
                    help="query blocks per chunk in the batched forms (0: library default)")
    ap.add_argument("--no-learned", action="store_true",
                    help="plain schedule only (for profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also launch a 1 GiB streaming copy kernel (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the search "
              "path has no multi-GPU form in this build", file=sys.stderr)
        sys.exit(2)

    import gz_native as gz
    import synthetic
    gz.check(gz.lib().gz_set_device(0))
    m, nl = gz.blosum62()
    log(f"host: os.cpu_count {os.cpu_count()}, affinity {len(os.sched_getaffinity(0))}, "
        f"OMP_NUM_THREADS {os.environ.get('OMP_NUM_THREADS')}")

    n_db = max(1000, int(1_000_000 * args.scale))
    n_reads = max(1000, int(200_000 * args.scale))
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    db_codes = synthetic.protein_database(rng, n_db)
    names, res, offs = synthetic.translated_reads(rng, db_codes, n_reads)
    dn, dr, do = synthetic._pack(db_codes, "subject_")
    log(f"workload generated in {time.perf_counter() - t0:.1f} s: {n_db} proteins "
        f"({len(dr)} aa), {n_reads} reads")
    t0 = time.perf_counter()
    db = gz.Database(dn, dr, do)
    del dr
    queries = gz.Queries(names, res, offs, translated=True)
    qv = queries.view()
    q_aa = query_residues(qv)
    log(f"host preparation {time.perf_counter() - t0:.1f} s: {q_aa} query aa")

    if args.calibrate_copy:
        nbytes = 1 << 30
        src, dst = gz.DeviceBuffer(nbytes), gz.DeviceBuffer(nbytes)
        for _ in range(3):
            gz.check(gz.lib().gz_copy_kernel_async(dst.ptr, src.ptr, nbytes, None))
        gz.check(gz.lib().gz_stream_synchronize(None))
        src.free()
        dst.free()
        log("copy kernel: 3 x 1 GiB")

    t0 = time.perf_counter()
    weights, learn = None, None
    if not args.no_learned:
        train_reads = max(1000, int(args.train_reads * min(1.0, args.scale * 10)))
        weights, learn = train_cost_model(gz, db, db_codes, np.random.default_rng(args.seed + 1),
                                          train_reads, args.epochs, args.seed)
    log(f"cost model trained in {time.perf_counter() - t0:.1f} s: {learn}")

    plain = gz.Aligner(db, schedule=gz.SCHEDULE_PLAIN)
    forms = [("plain", plain)]
    if not args.no_learned:
        pb = {"pipeline_blocks": args.pipeline_blocks} if args.pipeline_blocks else {}
        batched = gz.Aligner(db, schedule=gz.SCHEDULE_PLAIN, batched=1, **pb)
        learned = gz.Aligner(db, schedule=gz.SCHEDULE_LEARNED, batched=1, **pb)
        learned.set_cost_model(weights)
        forms += [("batched", batched), ("learned", learned)]

    def run(al, k, label):
        outs, stats, times = [], [], []
        for i in range(k):
            # the step is the C call: queries in host memory to the output
            # text (the copy into a Python bytes object is not part of it)
            text, st = al.align(queries)
            times.append(al.last_call_seconds)
            outs.append(text)
            stats.append(st)
            nl_count = text.count(b"\n")
            log(f"{label} step {i}: {times[-1]:.3f} s, {nl_count} lines, "
                f"seed {st['seconds_seed']:.3f} s, gapped {st['seconds_gapped']:.3f} s "
                f"(kernel {st['gapped_kernel_ms']:.1f} ms)")
        return outs, stats, times

    results = {}
    for name, al in forms:
        run(al, args.warmup, name + " warmup")
        results[name] = run(al, args.steps, name)
    p_out, p_st, p_t = results["plain"]
    l_out, l_st, l_t = results["learned" if "learned" in results else "plain"]
    identical = all(o == p_out[0] for r in results.values() for o in r[0])
    if not identical:
        log("ERROR: the forms' outputs differ")

    p_time = sum(p_t)
    l_time = sum(l_t)
    value = q_aa * args.steps / l_time
    plain_value = q_aa * args.steps / p_time
    form_values = {name: q_aa * args.steps / sum(r[2]) for name, r in results.items()}

    # kernel times of the learned form (its two streams overlap, so these
    # include sharing the GPU) and of the plain form (one kernel at a time)
    def kernel_ms(st):
        return {"seed_rep": st["seed_rep_kernel_ms"], "seed_value": st["seed_value_kernel_ms"],
                "gapped": st["gapped_kernel_ms"], "schedule": st["schedule_kernel_ms"],
                "traceback": st["traceback_kernel_ms"]}
    learned_kernels = kernel_ms(l_st[-1])
    plain_kernels = kernel_ms(p_st[-1])
    # roofline of the dominant kernel, from the HIP-event times of the plain
    # form's timed steps, where each kernel has the GPU to itself
    st = p_st[-1]
    # the seed kernels are the largest single kernels of the step (the gapped
    # time is spread over four kernels: two sides, two ring widths)
    kern = {k: plain_kernels[k] for k in ("seed_rep", "seed_value")}
    dom = max(kern, key=kern.get)
    k = 0 if dom == "seed_rep" else 1
    units = st["units_rep"] if k == 0 else st["units_value"]
    launches = max(1, st["seed_rep_launches"] if k == 0 else st["seed_value_launches"])
    ms = st["seed_rep_kernel_ms"] if k == 0 else st["seed_value_kernel_ms"]
    avg_ms = ms / launches
    # algorithmic bytes per unit: 4 (entry position) + 4 (database position)
    # + 2 x 8 (two packed database words around the seed)
    per_unit = 24.0
    achieved = per_unit * units / launches / (avg_ms / 1e3) / 1e9
    name = "SeedKernel<%s>" % ("false" if k == 0 else "true")
    measured = traffic_per_unit(name) if args.scale == 1.0 else None
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured * units / launches if measured else None,
            "kernel": name, "avg_launch_ms": avg_ms, "units_per_launch": units / launches,
            "algorithmic_bytes_per_unit": per_unit,
            "traffic_bytes_per_unit": measured,
            "traffic_source": TRAFFIC_PROFILE if measured else None}

    cpu = None
    if not args.no_cpu:
        import oracle_py
        cpu = cpu_baseline(gz, oracle_py, db, names, res, offs, m, nl, args.cpu_seconds)

    line = {
        "metric": "query aa/s vs synthetic DB: learned-scheduled vs plain GPU, bit-exact hits",
        "value": value, "unit": "query aa/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * l_time / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "translated 150-nt reads vs synthetic protein database "
                   "(aln -q d -t p, BLOSUM62 11/1, -b 10 -v 1, SEG)",
                   "reads": n_reads, "db_proteins": n_db, "db_residues": int(db.info().total_length),
                   "query_aa": q_aa, "schedule": "learned"},
        "plain_value": plain_value, "plain_ms_per_step": 1e3 * p_time / args.steps,
        "form_values": form_values,
        "learned_over_plain": value / plain_value, "outputs_identical": identical,
        "kernel_ms_per_step": learned_kernels,
        "plain_kernel_ms_per_step": plain_kernels,
        "search_stats": {k: st[k] for k in ("lists", "units_rep", "units_value", "seed_hits",
                                             "chained_hits", "gapped_extensions", "results",
                                             "sequential_build_chunks")},
        "stage_seconds": {k: round(v, 4) for k, v in st.items()
                                                       if k.startswith("seconds_")},
        "cost_model": learn,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if not identical:
        sys.exit(1)


if __name__ == "__main__":
    main()
