"""Throughput of the `aln` search path on one MI355X.

Workload (default, `--workload reads1m_db4m`): 1,000,000 synthetic 150-nt
reads, translated search (aln -q d -t p) against 4,000,000 synthetic proteins
(~1.32G aa in four database chunks of ~330M aa, all resident in HBM),
BLOSUM62 11/1, -b 10 -v 1, SEG on.  The reads form three query chunks of
333,334 / 333,333 / 333,333 reads.  A step is one query chunk searched
against all four database chunks, exactly as the reference's per-query-chunk
loop does it: query lists, seed search / distance / ungapped extension on the
device, chain filter, gapped extension on the device, AddResults over the
four database chunks, BuildResults and tabular output.  Steps cycle through
the three query chunks.  `--workload reads200k_db1m` is the smaller
translated workload (200,000 reads against 1,000,000 proteins, one chunk
each), where a step is the whole search.

Three forms run on the same inputs: the plain GPU form (stages one after the
other, jobs in hit order), the batched form (query blocks pipelined so that
the host stages of one block overlap the device stages of the next) and the
batched learned-scheduled form (the same, with gapped extensions ordered on
the device by the MLP's predicted cost; the MLP is trained on the GPU from
the plain form's extension traces during set-up).  Their outputs must be
byte-identical for every query chunk.  `value` is the batched
learned-scheduled form's query residues per second over K timed steps
(query residues = amino acids of the six translated frames); the plain and
batched forms run one timed step per query chunk (after their own warm-up)
and their rates are reported beside it.

A deterministic subset of the reads (16 sub-chunks of 256 reads, each its
own query chunk) is searched by the GPU (learned form) and by the oracle, a
restatement of the reference's CPU aligner, on all host cores (one sub-chunk
per thread); the outputs must be byte-identical, and the oracle's rate is the
CPU baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--scale S]
"""
import argparse
import hashlib
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("ghostz-gpu_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# int32 VALU: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 39.3
GAPPED_OPS_PER_CELL = 10
MLP_FLOP_PER_SEED = 2 * (1024 * 256 + 256 * 64 + 64)

WORKLOADS = {
    "reads1m_db4m": {"kind": "translated", "reads": 1_000_000, "db_proteins": 4_000_000,
                     "db_chunks": 4, "query_chunks": 3, "train_reads": 10_000},
    "reads200k_db1m": {"kind": "translated", "reads": 200_000, "db_proteins": 1_000_000,
                       "db_chunks": 1, "query_chunks": 1, "train_reads": 40_000},
    # protein queries of Pareto-tailed lengths (30 to 35,000 aa, median 250,
    # 20% with three planted homologs) against 1,000,000 proteins
    "pareto50k_db1m": {"kind": "uneven_protein", "reads": 50_000, "db_proteins": 1_000_000,
                       "db_chunks": 1, "query_chunks": 2, "train_reads": 4_000},
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM bytes per work unit of the seed kernels, measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE and calibrated against a copy kernel
TRAFFIC_PROFILES = {
    "reads1m_db4m": "profiles/hbm_traffic_translated_reads_1m_vs_db_4m.txt",
}


def traffic_per_unit(path, kernel):
    path = os.path.join(ROOT, path)
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


def training_queries(gz, kind, rng, db_codes, n):
    """Training queries drawn like the workload's: translated reads, or
    homologs of random database proteins (30-90% identity) for protein
    queries."""
    import synthetic
    if kind == "translated":
        names, res, offs = synthetic.translated_reads(rng, db_codes, n)
        return gz.Queries(names, res, offs, translated=True)
    qs = [synthetic.homolog(rng, db_codes[int(rng.integers(len(db_codes)))],
                            rng.uniform(0.3, 0.9)) for _ in range(n)]
    names, res, offs = synthetic._pack(qs, "train_")
    return gz.Queries(names, res, offs)


def train_cost_model(gz, db, db_codes, rng, n_reads, epochs, seed, max_traces=0,
                     kind="translated", batch=4096, lr=1e-3):
    """Logs extension traces with the plain schedule on training queries, then
    trains the MLP on the GPU (every logged trace unless max_traces caps them;
    10% held out; the training set resident on the device).  Returns (flat
    weights, report dict)."""
    q = training_queries(gz, kind, rng, db_codes, n_reads)
    al = gz.Aligner(db, record_traces=1)
    _, st = al.align(q)
    tr = al.traces()
    all_traces = len(tr)
    # the plain-order gapped kernels' VALU roofline (each kernel alone on the
    # GPU): 10 int32 operations per DP cell
    cells = int(st["gapped_cells"])
    gk_ms = float(st["gapped_kernel_ms"])
    gapped_roof = None
    if cells and gk_ms > 0:
        tops = GAPPED_OPS_PER_CELL * cells / (gk_ms / 1e3) / 1e12
        gapped_roof = {"bound": "valu_int32", "achieved": tops, "peak": VALU_PEAK_TOPS,
                       "unit": "Tops/s", "frac": tops / VALU_PEAK_TOPS, "cells": cells,
                       "extensions": int(st["gapped_extensions"]), "kernel_ms": gk_ms,
                       "launches": int(st["gapped_launches"]),
                       "workload": f"{n_reads} training reads, plain order"}
    if max_traces and len(tr) > max_traces:  # a uniform sample
        tr = tr[np.sort(rng.choice(len(tr), max_traces, replace=False))]
    concat = al.last_concatenated_queries()
    chunk_seq = {c: db.chunk_sequence(c) for c in np.unique(tr[:, 2])} if len(tr) else {}
    feats = np.zeros((len(tr), 128), np.uint8)
    t_feat = time.perf_counter()
    jobs = []
    for c, seqc in chunk_seq.items():
        for rev in (0, 1):
            sel = np.nonzero((tr[:, 2] == c) & (tr[:, 3] == rev))[0]
            for part in np.array_split(sel, max(1, len(sel) // 250_000)):
                if len(part):
                    jobs.append((c, rev, part))

    def featurize(job):
        c, rev, part = job
        feats[part] = gz.cost_features(concat, chunk_seq[c], tr[part, 0], tr[part, 1], rev)

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(host_threads()) as pool:
        list(pool.map(featurize, jobs))
    feat_s = time.perf_counter() - t_feat
    del chunk_seq
    target = np.log2(np.maximum(tr[:, 4], 1)).astype(np.float32)
    perm = rng.permutation(len(tr))
    n_test = max(1, len(tr) // 10)
    test, train = perm[:n_test], perm[n_test:]
    mlp = gz.Mlp(gz.Mlp.init_weights(seed))
    # the held-in rows, resident on the device; each epoch one permutation
    train_feats, train_target = feats[train], target[train]
    order = np.concatenate([rng.permutation(len(train)) for _ in range(epochs)])
    steps = len(order) // batch
    t0 = time.perf_counter()
    mean_loss, train_ms = mlp.train_epochs(train_feats, train_target, order, batch, lr)
    train_s = time.perf_counter() - t0
    del train_feats, train_target, order
    pred = mlp.predict(feats[test])
    mae = float(np.abs(pred - target[test]).mean())
    base = float(np.abs(target[test] - target[train].mean()).mean())
    # inference throughput (kernel time only, all sampled traces)
    mlp.predict(feats)
    infer_s = mlp.last_kernel_ms / 1e3
    seeds_per_s = len(feats) / infer_s if infer_s > 0 else None
    infer_roof = None
    if seeds_per_s:
        tf = seeds_per_s * MLP_FLOP_PER_SEED / 1e12
        infer_roof = {"bound": "mfma_bf16", "achieved": tf, "peak": BF16_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": tf / BF16_PEAK_TFLOPS}
    # forward + two backward products per seed
    train_flop = 3 * MLP_FLOP_PER_SEED * steps * batch
    report = {"traces_logged": int(all_traces), "traces": int(len(tr)), "held_out": int(n_test),
              "epochs": epochs, "batch": batch, "lr": lr, "train_steps": steps,
              "feature_seconds": round(feat_s, 3),
              "train_seconds": round(train_s, 3), "train_device_ms": round(train_ms, 1),
              "train_tflops": round(train_flop / (train_ms / 1e3) / 1e12, 3)
              if train_ms > 0 else None,
              "train_mfma_frac": round(train_flop / (train_ms / 1e3) / 1e12 / BF16_PEAK_TFLOPS, 4)
              if train_ms > 0 else None,
              "mean_train_loss": round(mean_loss, 4),
              "mae_log2_cells": round(mae, 4), "mae_constant_predictor": round(base, 4),
              "inference_seeds_per_s": seeds_per_s, "inference_roofline": infer_roof}
    return mlp.weights(), report, gapped_roof


def oracle_subset(gz, oracle_py, db, al, names, res, offs, m, nl, n_sub, sub_reads,
                  translated=True):
    """A deterministic subset of the reads, cut into n_sub sub-chunks of
    sub_reads reads, each searched as its own query chunk by the GPU (al)
    and by the oracle (one sub-chunk per host thread).  Returns (identical,
    cpu baseline dict)."""
    info = db.info()
    chunks = [db.chunk(c) for c in range(info.number_chunks)]
    stride = max(1, len(names) // (n_sub * sub_reads))
    sub_q, views = [], []
    for k in range(n_sub):
        idx = [(k * sub_reads + j) * stride for j in range(sub_reads)]
        idx = [i for i in idx if i < len(names)]
        nm = [names[i] for i in idx]
        rs = b"".join(res[int(offs[i]):int(offs[i + 1])] for i in idx)
        ln = np.array([int(offs[i + 1] - offs[i]) for i in idx], np.uint64)
        of = np.zeros(len(idx) + 1, np.uint64)
        np.cumsum(ln, out=of[1:])
        q = gz.Queries(nm, rs, of, translated=translated)
        sub_q.append(q)
        views.append(q.view())
    gpu = [al.align(q)[0] for q in sub_q]
    cpu = [None] * n_sub
    err = []

    def work(k):
        try:
            cpu[k] = oracle_py.align(views[k], chunks, info, m, nl)[0]
        except Exception as e:  # reported below
            err.append(e)
    threads = [threading.Thread(target=work, args=(k,)) for k in range(n_sub)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    dt = time.perf_counter() - t0
    if err:
        raise err[0]
    identical = all(g == c for g, c in zip(gpu, cpu))
    mism = [k for k in range(n_sub) if gpu[k] != cpu[k]]
    if mism:
        log(f"ERROR: GPU output differs from the oracle on sub-chunks {mism}")
    aa = sum(query_residues(v) for v in views)
    nreads = sum(int(v["number_queries"]) for v in views)
    base = {"value": aa / dt, "unit": "query aa/s", "cores": n_sub, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{nreads} reads (every {stride}th read of the workload, {n_sub} "
                      f"sub-chunks of {sub_reads} reads, each its own query chunk) against "
                      f"the full database; oracle, one sub-chunk per thread on {n_sub} "
                      f"threads, {dt:.1f} s wall",
            "reads_per_s": nreads / dt, "lines": int(sum(c.count(b"\n") for c in cpu))}
    return identical, base


def tuner_report(al):
    """The batched form's autotuner after the run: per query-length bucket,
    the configuration in use and the measured kernel ms per job of each."""
    chosen, ms, trials = al.tuner_state()
    names = ["ring32/queue", "ring16/queue", "ring32/grid", "ring16/grid"]
    out = {}
    for b, label in enumerate(["<=128", "<=512", "<=2048", ">2048"]):
        if chosen[b] < 0:
            continue
        out[label] = {"chosen": names[chosen[b]],
                      "us_per_job": {names[c]: round(1e3 * ms[b, c], 5)
                                     for c in range(4) if trials[b, c]}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="reads1m_db4m", choices=sorted(WORKLOADS))
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the workload (reads and database proteins)")
    ap.add_argument("--train-reads", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--train-traces", type=int, default=0,
                    help="cap on the logged traces the cost model trains on (0: all)")
    ap.add_argument("--train-batch", type=int, default=4096)
    ap.add_argument("--train-lr", type=float, default=1e-3)
    ap.add_argument("--train-only", action="store_true",
                    help="stop after the cost model's training report")
    ap.add_argument("--compare-warmup", type=int, default=2,
                    help="warm-up steps of the plain and batched forms")
    ap.add_argument("--subset-reads", type=int, default=256,
                    help="reads per oracle sub-chunk (16 sub-chunks)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--pipeline-blocks", type=int, default=0,
                    help="query blocks per chunk in the batched forms (0: library default)")
    ap.add_argument("--profile-form", choices=["plain", "batched", "length_sorted", "learned"],
                    default=None,
                    help="run only this form (W warm-up and K timed steps) and print its "
                         "rate; for profiler runs, no result line")
    ap.add_argument("--no-learned", action="store_true",
                    help="plain form only (for profiling runs)")
    ap.add_argument("--sorted-form", action="store_true",
                    help="also run the batched form in query-side-length order")
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
    log(f"host: os.cpu_count {os.cpu_count()}, threads {host_threads()}, "
        f"OMP_NUM_THREADS {os.environ.get('OMP_NUM_THREADS')}, cpu {cpu_model()}")

    wl = WORKLOADS[args.workload]
    n_db = max(1000, int(wl["db_proteins"] * args.scale))
    n_reads = max(1000, int(wl["reads"] * args.scale))
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    translated = wl["kind"] == "translated"
    db_codes = synthetic.protein_database(rng, n_db)
    if translated:
        names, res, offs = synthetic.translated_reads(rng, db_codes, n_reads)
    else:
        lens = synthetic.pareto_lengths(rng, n_reads)
        qcodes = [synthetic.random_codes(rng, int(L)) for L in lens]
        for qc in qcodes:
            if rng.random() < 0.2:
                synthetic.plant(rng, db_codes, qc, 3)
        names, res, offs = synthetic._pack(qcodes, "query_")
        del qcodes
    dn, dr, do = synthetic._pack(db_codes, "subject_")
    total_aa = int(do[-1])
    log(f"workload generated in {time.perf_counter() - t0:.1f} s: {n_db} proteins "
        f"({total_aa} aa), {n_reads} reads")
    t0 = time.perf_counter()
    # database chunks of equal residue budgets (db -l): subjects are added to
    # a chunk while its residues stay below the budget
    chunk_size = total_aa // wl["db_chunks"] + 4096 if wl["db_chunks"] > 1 else 0
    db = gz.Database(dn, dr, do, chunk_size=chunk_size)
    del dr, dn
    n_chunks = int(db.info().number_chunks)
    t_db = time.perf_counter() - t0
    # query chunks: equal thirds of the reads, each one query chunk (every
    # third stays below the default -l of 2^27 framed codes)
    nq = wl["query_chunks"]
    if translated:
        bounds = [n_reads * i // nq for i in range(nq + 1)]
    else:  # equal residue shares
        cum = np.asarray(offs, np.int64)
        bounds = [0] + [int(np.searchsorted(cum, cum[-1] * i // nq)) for i in range(1, nq)] + \
            [n_reads]
    qchunks, q_aa = [], []
    t0 = time.perf_counter()
    for i in range(nq):
        a, b = bounds[i], bounds[i + 1]
        base = int(offs[a])
        q = gz.Queries(names[a:b], res[base:int(offs[b])],
                       np.asarray(offs[a:b + 1], np.uint64) - np.uint64(base),
                       translated=translated)
        qchunks.append(q)
        q_aa.append(query_residues(q.view()))
    log(f"database built in {t_db:.1f} s ({n_chunks} chunks); queries prepared in "
        f"{time.perf_counter() - t0:.1f} s: {nq} query chunks, {sum(q_aa)} query aa")

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
    weights, learn, gapped_roof = None, None, None
    if not args.no_learned:
        tr_reads = args.train_reads or wl["train_reads"]
        tr_reads = max(1000, int(tr_reads * min(1.0, args.scale * 10)))
        weights, learn, gapped_roof = train_cost_model(
            gz, db, db_codes, np.random.default_rng(args.seed + 1), tr_reads, args.epochs,
            args.seed, max_traces=args.train_traces, kind=wl["kind"], batch=args.train_batch,
            lr=args.train_lr)
    del db_codes
    log(f"cost model trained in {time.perf_counter() - t0:.1f} s: {learn}")
    if args.train_only:
        print(json.dumps({"workload": args.workload, "cost_model": learn,
                          "gapped_plain_order": gapped_roof}), flush=True)
        return

    pb = {"pipeline_blocks": args.pipeline_blocks} if args.pipeline_blocks else {}
    plain = gz.Aligner(db, schedule=gz.SCHEDULE_PLAIN)
    forms = [("plain", plain)]
    if not args.no_learned:
        batched = gz.Aligner(db, schedule=gz.SCHEDULE_PLAIN, batched=1, **pb)
        learned = gz.Aligner(db, schedule=gz.SCHEDULE_LEARNED, batched=1, **pb)
        learned.set_cost_model(weights)
        forms += [("batched", batched)]
        if args.sorted_form:
            forms += [("length_sorted", gz.Aligner(db, schedule=gz.SCHEDULE_LENGTH_SORTED,
                                                   batched=1, **pb))]
        forms += [("learned", learned)]

    if args.profile_form:
        al = {"plain": plain}.get(args.profile_form)
        if al is None:
            for name, a in forms:
                if name == args.profile_form:
                    al = a
        if al is None and args.profile_form == "length_sorted":
            al = gz.Aligner(db, schedule=gz.SCHEDULE_LENGTH_SORTED, batched=1, **pb)
        forms = [(args.profile_form, al)]

    digests = {}  # query chunk -> set of output digests over all forms and runs

    def run(al, first, k, label):
        stats, times, aa = [], [], 0
        for i in range(k):
            c = (first + i) % nq
            # the step is the C call: queries in host memory to the output
            # text (the copy into a Python bytes object is not part of it)
            text, st = al.align(qchunks[c])
            times.append(al.last_call_seconds)
            digests.setdefault(c, set()).add(hashlib.sha256(text).hexdigest())
            stats.append(st)
            aa += q_aa[c]
            lines = text.count(b"\n")
            log(f"{label} step {i} (query chunk {c}): {times[-1]:.3f} s, "
                f"{lines} lines, seed {st['seconds_seed']:.3f} s, chain "
                f"{st['seconds_chain']:.3f} s, build {st['seconds_build']:.3f} s, gapped "
                f"{st['seconds_gapped']:.3f} s (kernel {st['gapped_kernel_ms']:.1f} ms, "
                f"schedule {st['schedule_kernel_ms']:.1f} ms, predicted jobs "
                f"{st['gated_jobs']} of {st['gapped_extensions']}; seed units rep "
                f"{st['units_rep']} value {st['units_value']})")
            del text
        return stats, times, aa

    results = {}
    for name, al in forms:
        flagship = name == forms[-1][0]
        w = args.warmup if flagship else min(args.warmup, args.compare_warmup)
        k = args.steps if flagship else nq
        run(al, 0, w, name + " warmup")
        results[name] = run(al, w, k, name)
    flag = forms[-1][0]
    if args.profile_form:
        r = results[flag]
        log(f"{flag}: {r[2] / sum(r[1]):.4g} query aa/s over {len(r[1])} steps")
        return
    identical = all(len(v) == 1 for v in digests.values())
    if not identical:
        log("ERROR: the forms' outputs differ")
    form_values = {name: r[2] / sum(r[1]) for name, r in results.items()}
    form_ms = {name: 1e3 * sum(r[1]) / len(r[1]) for name, r in results.items()}
    l_st, l_t, l_aa = results[flag]
    p_st, p_t, p_aa = results["plain"]
    value = form_values[flag]
    plain_value = form_values["plain"]

    def total(sts, key):
        return sum(s[key] for s in sts)

    def kernel_ms(sts):
        n = len(sts)
        return {"seed_rep": total(sts, "seed_rep_kernel_ms") / n,
                "seed_value": total(sts, "seed_value_kernel_ms") / n,
                "gapped": total(sts, "gapped_kernel_ms") / n,
                "schedule": total(sts, "schedule_kernel_ms") / n,
                "traceback": total(sts, "traceback_kernel_ms") / n}
    learned_kernels = kernel_ms(l_st)
    plain_kernels = kernel_ms(p_st)
    # roofline of the dominant kernel, from the HIP-event times of the plain
    # form's timed steps, where each kernel has the GPU to itself: the seed
    # kernels are the largest single kernels of the step (the gapped time is
    # spread over four kernels: two sides, two ring widths)
    kern = {k: plain_kernels[k] for k in ("seed_rep", "seed_value")}
    dom = max(kern, key=kern.get)
    k = 0 if dom == "seed_rep" else 1
    units = total(p_st, "units_rep" if k == 0 else "units_value")
    launches = max(1, total(p_st, "seed_rep_launches" if k == 0 else "seed_value_launches"))
    ms = total(p_st, "seed_rep_kernel_ms" if k == 0 else "seed_value_kernel_ms")
    avg_ms = ms / launches
    # algorithmic bytes per unit: 4 (entry position) + 4 (database position)
    # + 2 x 8 (two packed database words around the seed)
    per_unit = 24.0
    achieved = per_unit * units / launches / (avg_ms / 1e3) / 1e9
    name = "SeedKernel<%s>" % ("false" if k == 0 else "true")
    tp = TRAFFIC_PROFILES.get(args.workload)
    measured = traffic_per_unit(tp, name) if tp and args.scale == 1.0 else None
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured * units / launches if measured else None,
            "kernel": name, "avg_launch_ms": avg_ms, "units_per_launch": units / launches,
            "algorithmic_bytes_per_unit": per_unit,
            "traffic_bytes_per_unit": measured,
            "traffic_source": tp if measured else None}
    stage_roofs = {"gapped_plain_order": gapped_roof,
                   "mlp_inference": learn["inference_roofline"] if learn else None}

    cpu, subset_ok = None, None
    if not args.no_cpu:
        import oracle_py
        subset_ok, cpu = oracle_subset(gz, oracle_py, db, forms[-1][1], names, res, offs, m,
                                       nl, host_threads(),
                                       args.subset_reads if translated else 16,
                                       translated=translated)

    sched_ms = learned_kernels["schedule"]
    line = {
        "metric": "query aa/s vs synthetic DB: learned-scheduled vs plain GPU, bit-exact hits",
        "value": value, "unit": "query aa/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sum(l_t) / len(l_t),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": ("translated 150-nt reads" if translated else
                                "protein queries of Pareto-tailed lengths") +
                   " vs synthetic protein database (aln -q " + ("d" if translated else "p") +
                   " -t p, BLOSUM62 11/1, -b 10 -v 1, SEG); step = one query chunk against "
                   "every database chunk",
                   "name": args.workload, "reads": n_reads, "db_proteins": n_db,
                   "db_residues": int(db.info().total_length), "db_chunks": n_chunks,
                   "query_chunks": nq, "query_aa": int(sum(q_aa)),
                   "schedule": flag},
        "plain_value": plain_value, "plain_ms_per_step": form_ms["plain"],
        "form_values": form_values, "form_ms_per_step": form_ms,
        "form_steps": {name: len(r[1]) for name, r in results.items()},
        "learned_over_plain": value / plain_value,
        "learned_over_batched": (value / form_values["batched"]) if "batched" in form_values
        else None,
        "outputs_identical": identical, "oracle_subset_identical": subset_ok,
        "kernel_ms_per_step": learned_kernels,
        "plain_kernel_ms_per_step": plain_kernels,
        "schedule_ms_vs_gapped_ms_saved": None if args.no_learned else {
            "schedule_ms": sched_ms,
            "gapped_ms_batched": kernel_ms(results["batched"][0])["gapped"],
            "gapped_ms_learned": learned_kernels["gapped"],
            "predicted_jobs": total(l_st, "gated_jobs") // len(l_st),
            "gapped_jobs": total(l_st, "gapped_extensions") // len(l_st)},
        "search_stats": {k: total(p_st, k) // len(p_st) for k in (
            "lists", "units_rep", "units_value", "seed_hits", "chained_hits",
            "gapped_extensions", "results", "sequential_build_chunks",
            "traceback_host_jobs")},
        "stage_seconds": {k: round(total(p_st, k) / len(p_st), 4) for k in p_st[0]
                          if k.startswith("seconds_")},
        "cost_model": learn,
        "autotuner": {name: tuner_report(al) for name, al in forms if name != "plain"},
        "graphs": {name: dict(zip(("instantiated", "updated"), al.graph_counts()))
                   for name, al in forms if name != "plain"},
        "roofline": roof,
        "stage_rooflines": stage_roofs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if not identical or subset_ok is False:
        sys.exit(1)


if __name__ == "__main__":
    main()
