#!/usr/bin/env python3
"""Benchmark of the MI355X propagator-completion engine.

Headline workload (one step = propagating one batch of new search nodes to
their fixpoints):
  one seeded 50,000-variable / 200,000-propagator ternary network from the
  planted generator (every propagator holds under a hidden assignment;
  lb ~ U[0, 5e5], ub = lb + U[1, 5e5] around it, 10% constants, 25% Boolean
  variables, 30% of the propagators in chains of 200), told and run to its
  root fixpoint; 4,096 search-node subproblems, each the root fixpoint with
  1-20 random splits (x <= v or x > v inside the current bounds) and a
  random subset of 2,000-200,000 active propagators; int32 interval stores.
A step loads a new set of 4,096 stores from device memory (one of four sets
staged in HBM before the timed region, each drawn with its own seed) and
runs the batched engine to every instance's fixpoint: the load (stores,
emptiness, base store and the values that differ from it) is inside the
timing.  The engine's default first sweep evaluates only the propagators
adjacent to a value that differs from the base store and those the base
store does not satisfy (the others compute in each instance what they
compute on the base, where they tighten nothing); the same steps with a full
first sweep (lockstep) are timed beside it and must reach the same stores.
Beside them: the same batch replayed from its own reset image (no new
stores), the plain engine (one sweep launch per instance and sweep), a
search step on the device, and a bounded sample of instances on the CPU
oracle (the reference's Gauss-Seidel algorithm, restated) on the host
cores, whose results are checked bit-for-bit against the GPU's.  The
roofline is the dominant kernel's of the headline step (the one with the
most device time per step: the load, the full sweep or the marked phase).

The default line also carries the other workloads, each measured the same
way (HBM-resident inputs, HIP-event kernel timing, CPU oracle beside it):
  single:  one 1,000,000-variable / 4,000,000-propagator fixpoint (plain
           engine), ms per fixpoint;
  bitset:  1,024 bit-removal subproblems of one 200,000-word /
           800,000-propagator bitset network (values in [0, 61]);
  serving: Poisson arrivals at 2,000 requests/s over 24 networks of 1k-1M
           propagators, 1-10 splits per request, replayed in real time through
           the continuous-batching server: p50 / p99 latency, beside the
           same trace served one request at a time.
--config 1|2|3|4 prints the line of one workload alone.

Prints one JSON line (the driver's contract).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lala-pc_amd"))

import numpy as np  # noqa: E402

import pcx  # noqa: E402

METRIC = "fixpoints/sec batched vs plain; propagator evals/s; %HBM roofline; p99 latency"
HBM_PEAK_GBS = 8000.0
BYTES_READ_PER_EVAL = 40     # 16 B propagator + 3 x 8 B operand intervals
BYTES_PER_TIGHTENING = 4     # one int32 bound written (a bitset word change writes 8 B)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


BOX_CPU_SHARE = 16   # host cores a one-GPU box grants a job (worker pools are sized to it)


def host_cpus():
    """(threads the CPU baseline uses, host facts): the affinity count capped
    at the box's CPU share of 16, beside the affinity and os.cpu_count()."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    used = max(1, min(aff, BOX_CPU_SHARE))
    return used, {"affinity_cpus": aff, "os_cpu_count": os.cpu_count(), "cpu_share": BOX_CPU_SHARE}


def host_threads():
    return host_cpus()[0]


def cpu_baseline(value, threads, kind, sample):
    t, info = host_cpus()
    return dict({"value": value, "unit": "fixpoints/s", "cores": threads, "kind": kind,
                 "sample": sample}, **info)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated), or
    None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return int(float(line.split("=")[1]))
    return None


def roofline(evals, tightened, kernel_ms, launches, kernel, workload, tight_bytes=BYTES_PER_TIGHTENING,
             algo_bytes=None):
    """Algorithmic bytes per launch / average launch duration (HIP events on
    the engine's stream around every launch of the dominant kernel).  Sweep
    kernels: 40 B per evaluation and the bytes of each tightening; a load
    passes its own algorithmic bytes."""
    if algo_bytes is None:
        algo_bytes = evals * BYTES_READ_PER_EVAL + tightened * tight_bytes
    launches = max(1, launches)
    achieved = (algo_bytes / launches) / (kernel_ms * 1e-3 / launches) / 1e9 if kernel_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel, workload),
            "kernel": kernel, "launches_per_step": launches,
            "algorithmic_bytes_per_launch": int(algo_bytes / launches),
            "avg_launch_us": round(kernel_ms * 1e3 / launches, 2)}


class StatSum:
    """Sums of pcx.Stats over the timed steps, split into full sweeps (every
    active propagator) and marked-only sweeps."""

    FIELDS = ("kernel_ms", "kernel_launches", "evals", "tightened", "full_ms", "full_launches",
              "full_evals", "full_tightened", "load_ms")

    def __init__(self):
        self.v = {k: 0 for k in self.FIELDS}
        self.lockstep = 0

    def add(self, st):
        for k in self.FIELDS:
            self.v[k] += getattr(st, k)
        self.lockstep = int(st.lockstep)

    def roofline(self, steps, full_kernel, marked_kernel, workload, tight_bytes=BYTES_PER_TIGHTENING,
                 load=None):
        """Roofline of the dominant kernel (the class with the most kernel
        time per step): the full sweeps, the marked sweeps, or -- load =
        (kernel, algorithmic bytes per load) for steps that load new stores
        -- the device load; plus the per-step figures of every class."""
        v = {k: x / steps for k, x in self.v.items()}
        full = (v["full_evals"], v["full_tightened"], v["full_ms"], v["full_launches"])
        marked = (v["evals"] - v["full_evals"], v["tightened"] - v["full_tightened"],
                  v["kernel_ms"] - v["full_ms"], v["kernel_launches"] - v["full_launches"])
        classes = {"full": (full, full_kernel, None), "marked": (marked, marked_kernel, None)}
        if load is not None and v["load_ms"] > 0:
            classes["load"] = ((0, 0, v["load_ms"], 1), load[0], load[1])
        name = max(classes, key=lambda c: classes[c][0][2])
        dom, kernel, ab = classes[name]
        rf = roofline(*dom, kernel, workload, tight_bytes, algo_bytes=ab)
        rf["kernel_classes_per_step"] = {
            c: {"kernel": k, "launches": d[3], "kernel_ms": round(d[2], 3), "evals": int(d[0]),
                "algorithmic_bytes": int(ab if ab is not None else d[0] * BYTES_READ_PER_EVAL + d[1] * tight_bytes)}
            for c, (d, k, ab) in classes.items()}
        return rf


def full_kernel(mode, sums):
    """Name of the kernel of the full sweeps: the lockstep sweep's span also
    holds the interleave of the stores before it."""
    if mode != pcx.MODE_BATCHED:
        return "k_persistent"
    return "k_sweep_lockstep" if sums.lockstep else "k_sweep_batched"


def marked_kernel(pir, mode, sums, suffix=""):
    """Name of the kernel of the marked sweeps: batched runs seeded from a
    base store drain each instance's marks in one launch; batched runs whose
    full sweep is lockstep list their marks in queues by default
    (mark_queue)."""
    cfg = pir.get_config()
    if mode == pcx.MODE_BATCHED and sums.v["full_evals"] == 0 and cfg.base_seed != 0:
        return "k_drain_marks" + suffix
    if mode == pcx.MODE_BATCHED and sums.lockstep and cfg.mark_queue != 0:
        return "k_sweep_queue" + suffix
    return "k_sweep_marked" + suffix


def autotune(args, pir, batch, key):
    """Configure the engine from the tuning cache, or run the autotuner on the
    loaded batch (setup, outside the timed region) and cache its choice."""
    path = args.tune_cache or pcx.default_tuning_path()
    cached = None if args.retune else pcx.load_tuning(path, key)
    t0 = time.time()
    if cached:
        best = cached["best"]
        mode = pcx.apply_tuning(pir, best)
        how = "cached"
    else:
        best, points = batch.autotune(reps=2)
        mode = best["mode"]
        how = f"tuned over {len(points)} points in {time.time() - t0:.1f}s"
        try:
            pcx.save_tuning(path, key, best, points)
        except OSError as e:
            log(f"autotune: could not save {path}: {e}")
    log(f"autotune ({how}): {best}")
    return mode, dict(best, source=how)


def cpu_oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib  # test infrastructure: used only as the CPU baseline / checker
    return oracle_lib


GEN_LABEL = ("planted generator: a hidden assignment satisfies every propagator, lb ~ U[0, 5e5], "
             "ub = lb + U[1, 5e5] around it, 10% constants, 25% Boolean variables, 30% of the "
             "propagators in chains of 200; opcodes uniform over add, sub, mul, min, max, eq-reif, "
             "leq-reif")


def useful_lane_fraction(off, n_inst, P, gi):
    """Active (instance, propagator) evaluations / pairs the lockstep full
    sweep loads: every table entry for every slot of every group of gi."""
    if not gi:
        return None
    groups = (n_inst + gi - 1) // gi
    return float(off[-1]) / float(groups * gi * P)


def bench_batched(args):
    V, P, n_inst = args.nvars, args.nprops, 4096
    t0 = time.time()
    p = pcx.gen_params(seed=args.seed, nvars=V, nprops=P)
    op, x, y, z, lb, ub, w = pcx.gen_network(p)
    pir = pcx.PIR(props_per_wave=args.ppw, timing=1, compact_tables=args.compact,
                  stream_tables=args.stream, event_driven=args.events, lockstep=args.lockstep,
                  lds_mode=args.lds)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    rr = pir.fixpoint(root_lb, root_ub)
    slb, sub = pcx.gen_splits(args.seed + 1, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(args.seed + 2, P, n_inst, min(2000, P), P)
    batch = pcx.Batch(pir, n_inst, V, off, idx)
    batch.load(slb, sub)
    log(f"setup {time.time() - t0:.1f}s: root fixpoint sweeps={rr.sweeps} failed={rr.failed}, "
        f"{int(off[-1])} active propagators over {n_inst} instances")

    mode = {"batched": pcx.MODE_BATCHED, "persistent": pcx.MODE_PERSISTENT}[args.mode]
    tuning = None
    if args.autotune:
        mode, tuning = autotune(args, pir, batch, f"cfg2:V{V}:P{P}:n{n_inst}:seed{args.seed}")

    # replay: each step restores the loaded stores from the batch's image
    for _ in range(args.warmup):
        batch.reset()
        batch.run(mode)
    sums = StatSum()
    smax = 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        batch.reset()
        res, st = batch.run(mode)
        sums.add(st)
        smax = max(smax, st.sweeps_max)
    replay_s = time.perf_counter() - t_start
    evals = sums.v["evals"]
    replay_fps = n_inst * args.steps / replay_s
    failed = sum(r.failed for r in res)
    lb_g, ub_g = batch.read()
    log(f"replay: {replay_s / args.steps * 1e3:.2f} ms/step, {replay_fps:.0f} fixpoints/s, "
        f"sweeps_max={smax}, failed={failed}/{n_inst}, kernel {sums.v['kernel_ms'] / args.steps:.2f} ms/step")

    # plain engine on the same batch (one instance at a time)
    plain_fps = None
    lb_p, ub_p = lb_g, ub_g
    if not args.no_plain:
        batch.reset()
        tp = time.perf_counter()
        res_p, st_p = batch.run(pcx.MODE_PLAIN)
        plain_s = time.perf_counter() - tp
        plain_fps = n_inst / plain_s
        lb_p, ub_p = batch.read()
        log(f"plain: {plain_s * 1e3:.0f} ms for {n_inst} instances, {plain_fps:.0f} fixpoints/s")

    out = {
        "config": {"workload": f"batched search-node fixpoints: 4,096 subproblems of one "
                               f"{V:,}-var / {P:,}-propagator network ({GEN_LABEL}), each the "
                               f"network's root fixpoint with 1-20 random splits and "
                               f"{min(2000, P):,}-{P:,} active propagators; a step loads 4,096 new "
                               f"stores from HBM and propagates them",
                   "nvars": V, "nprops": P, "instances": n_inst, "props_per_wave": args.ppw,
                   "mode": {pcx.MODE_BATCHED: "batched", pcx.MODE_PERSISTENT: "persistent"}[mode],
                   "compact_tables": args.compact, "stream_tables": args.stream,
                   "event_driven": args.events, "lockstep": sums.lockstep, "autotuned": tuning},
        "replay": {"fixpoints_per_s": replay_fps, "ms_per_step": replay_s / args.steps * 1e3,
                   "what": "each step restores the same 4,096 stores from the batch's device-resident "
                           "reset image (the image's interleaved copy and emptiness are reused)"},
        "plain_fixpoints_per_s": plain_fps,
        "evals_per_s": evals / replay_s,
        "failed_instances": int(failed),
        "sweeps_max": int(smax),
    }
    out["replay"]["roofline"] = sums.roofline(args.steps, full_kernel(mode, sums),
                                              marked_kernel(pir, mode, sums), "cfg2")

    if args.search_step:
        out["search_step"] = search_step(args, pir, batch, off, idx, res, lb_g, ub_g, V, n_inst)
    if args.formula_ask:
        out["pc_formula_ask"] = formula_ask(pir, batch, op, x, y, z, lb, ub, w, n_inst, V)

    if not args.no_cpu:
        ol = cpu_oracle()
        top, tx, ty, tz = pir.table()
        threads = host_threads()
        sample = []
        t_cpu = 0.0
        k = 0
        idx64 = idx.astype(np.int64)
        # bounded sample: whole instances until about cpu_seconds of CPU work
        while k < n_inst and t_cpu < args.cpu_seconds:
            n = min(threads * 2, n_inst - k)
            rows = slice(k * V, (k + n) * V)
            lo, hi = slb[rows].copy(), sub[rows].copy()
            o = off[k:k + n + 1] - off[k]
            ix = idx64[off[k]:off[k + n]]
            ts = time.perf_counter()
            r = ol.fixpoint_batch(top, tx, ty, tz, lo, hi, V, idx=ix, off=o, threads=threads)
            t_cpu += time.perf_counter() - ts
            sample.append((k, n, lo, hi, r))
            k += n
        mism = 0
        for k0, n, lo, hi, r in sample:
            for j in range(n):
                kk = k0 + j
                g_fail = bool(res[kk].failed)
                if g_fail != bool(r[j].failed):
                    mism += 1
                elif not g_fail:
                    s = slice(kk * V, (kk + 1) * V)
                    if not (np.array_equal(lb_g[s], lo[j * V:(j + 1) * V]) and
                            np.array_equal(ub_g[s], hi[j * V:(j + 1) * V]) and
                            np.array_equal(lb_p[s], lo[j * V:(j + 1) * V]) and
                            np.array_equal(ub_p[s], hi[j * V:(j + 1) * V])):
                        mism += 1
        out["cpu_baseline"] = cpu_baseline(k / t_cpu, threads, "port",
                                           f"first {k} of the {n_inst} instances, one instance per "
                                           f"thread, oracle Gauss-Seidel (-O3)")
        out["parity"] = {"checked_instances": k, "mismatches": mism}
        log(f"cpu: {k} instances in {t_cpu:.1f}s on {threads} threads; parity mismatches {mism}")

    fr = fresh_steps(args, pir, batch, root_lb, root_ub, off, idx, mode, n_inst, V)
    out["value"] = fr.pop("fixpoints_per_s")
    out["unit"] = "fixpoints/s"
    out["ms_per_step"] = fr.pop("ms_per_step")
    out["roofline"] = fr.pop("roofline")
    out["fresh"] = fr
    out["batched_vs_plain"] = out["value"] / plain_fps if plain_fps else None
    batch.close()
    return out


def ingest_kernel(pir, n_inst, V, written):
    """(kernel, algorithmic bytes) of a device load of n_inst stores: with
    base seeding, into stores that hold the base store (a load after a
    drained run puts them back to it), the load reads the stores (8 B per
    value) and writes the `written` values that differ from the base (8 B
    each); without it, it writes the stores and their interleaved copy (8 B
    read + 16 B written per value)."""
    if pir.get_config().base_seed != 0:
        return ("k_ingest_held" if V % 4 == 0 else "k_ingest_rows"), 8 * n_inst * V + 8 * written
    return "k_ingest_interleave4", 24 * n_inst * V


def differing_values(lo, hi, n_inst, V):
    """Values of a store set that differ from its base store (per variable
    the value two of instances 0-2 hold, instance 0's when none agree)."""
    lo2, hi2 = lo.reshape(n_inst, V), hi.reshape(n_inst, V)
    if n_inst < 3:
        return int(((lo2 != lo2[0]) | (hi2 != hi2[0])).sum())
    same = [(lo2[a] == lo2[b]) & (hi2[a] == hi2[b]) for a, b in ((0, 1), (0, 2), (1, 2))]
    first = same[0] | same[1]
    blo = np.where(first, lo2[0], np.where(same[2], lo2[1], lo2[0]))
    bhi = np.where(first, hi2[0], np.where(same[2], hi2[1], hi2[0]))
    return int(((lo2 != blo) | (hi2 != bhi)).sum())


def timed_fresh(args, batch, sets, mode, steps):
    """`steps` steps, each a device load of the next staged store set and a
    run; returns (seconds, StatSum, the last step's results)."""
    sums = StatSum()
    t_start = time.perf_counter()
    for s in range(steps):
        d_lo, d_hi = sets[s % len(sets)]
        batch.load_device(d_lo.ptr, d_hi.ptr)
        res, st = batch.run(mode)
        sums.add(st)
    return time.perf_counter() - t_start, sums, res


def fresh_steps(args, pir, batch, root_lb, root_ub, off, idx, mode, n_inst, V, n_sets=4):
    """The headline steps: each loads a new set of 4,096 stores (splits of the
    root fixpoint drawn with their own seed) from device memory into the batch
    and propagates it.  The sets are staged in HBM before the timed region;
    consecutive steps load different sets.  The last step's results are
    checked on a bounded sample against the oracle.  The same steps with the
    first sweep a full sweep (base seeding off) are timed beside them, and
    their last stores compared with the seeded run's."""
    from pcx.device import DeviceArray
    t0 = time.time()
    sets, written = [], []
    for j in range(n_sets):
        lo, hi = pcx.gen_splits(args.seed + 100 + j, root_lb, root_ub, n_inst, 1, 20)
        sets.append((DeviceArray(lo), DeviceArray(hi)))
        written.append(differing_values(lo, hi, n_inst, V))
        if j == (args.steps - 1) % n_sets:
            last_host = (lo, hi)
        del lo, hi
    log(f"fresh: {n_sets} store sets staged in HBM in {time.time() - t0:.1f}s")
    for s in range(max(1, args.warmup)):
        d_lo, d_hi = sets[s % n_sets]
        batch.load_device(d_lo.ptr, d_hi.ptr)
        batch.run(mode)
    elapsed, sums, res = timed_fresh(args, batch, sets, mode, args.steps)
    fps = n_inst * args.steps / elapsed
    seeded = pir.get_config().base_seed != 0 and sums.v["full_evals"] == 0
    load = ingest_kernel(pir, n_inst, V, sum(written) // n_sets)
    out = {"fixpoints_per_s": fps, "ms_per_step": elapsed / args.steps * 1e3, "store_sets": n_sets,
           "evals_per_s": sums.v["evals"] / elapsed,
           "kernel_ms_per_step": sums.v["kernel_ms"] / args.steps,
           "load_ms_per_step": sums.v["load_ms"] / args.steps,
           "failed_instances": int(sum(r.failed for r in res)),
           "first_sweep": "seeded from the base store" if seeded else "full",
           "roofline": sums.roofline(args.steps, full_kernel(mode, sums), marked_kernel(pir, mode, sums),
                                     "cfg2", load=load),
           "values_differing_from_base": sum(written) // n_sets,
           "what": "each step loads 4,096 new stores from HBM (pcx_batch_load_device: the stores, put "
                   "back to the base store after the last run, get the values that differ from it; the "
                   "emptiness, the base store and the listed differences in the same pass) and "
                   "propagates them; the first sweep evaluates only the propagators adjacent to a "
                   "listed value and those the base store does not satisfy"}
    log(f"fresh: {out['ms_per_step']:.2f} ms/step, {fps:.0f} fixpoints/s, first sweep {out['first_sweep']}")
    lb_s, ub_s = batch.read()
    if not args.no_cpu:
        ol = cpu_oracle()
        top, tx, ty, tz = pir.table()
        k = min(n_inst, args.fresh_check)
        lo, hi = last_host[0][:k * V].copy(), last_host[1][:k * V].copy()
        r = ol.fixpoint_batch(top, tx, ty, tz, lo, hi, V, idx=idx[:off[k]].astype(np.int64),
                              off=off[:k + 1], threads=host_threads())
        mism = 0
        for j in range(k):
            if bool(res[j].failed) != bool(r[j].failed):
                mism += 1
            elif not r[j].failed and not (np.array_equal(lb_s[j * V:(j + 1) * V], lo[j * V:(j + 1) * V]) and
                                          np.array_equal(ub_s[j * V:(j + 1) * V], hi[j * V:(j + 1) * V])):
                mism += 1
        out["parity"] = {"checked_instances": k, "mismatches": mism, "of_step": args.steps}
        log(f"fresh parity: {k} instances of the last step, mismatches {mism}")
    if seeded:
        # the same steps with a full first sweep (lockstep), for comparison
        pir.set_config(base_seed=0)
        for s in range(max(1, args.warmup)):
            d_lo, d_hi = sets[s % n_sets]
            batch.load_device(d_lo.ptr, d_hi.ptr)
            batch.run(mode)
        el_f, sums_f, res_f = timed_fresh(args, batch, sets, mode, args.steps)
        lb_f, ub_f = batch.read()
        ok = ~np.repeat(np.array([bool(x.failed) for x in res]), V)
        agree = (np.array_equal(np.array([bool(x.failed) for x in res_f]), np.array([bool(x.failed) for x in res]))
                 and np.array_equal(lb_f[ok], lb_s[ok]) and np.array_equal(ub_f[ok], ub_s[ok]))
        out["full_first_sweep"] = {
            "fixpoints_per_s": n_inst * args.steps / el_f, "ms_per_step": el_f / args.steps * 1e3,
            "load_ms_per_step": sums_f.v["load_ms"] / args.steps,
            "same_fixpoints_as_seeded": bool(agree),
            "useful_lane_frac": useful_lane_fraction(off, n_inst, len(pir.table()[0]), sums_f.lockstep),
            "roofline": sums_f.roofline(args.steps, full_kernel(mode, sums_f), marked_kernel(pir, mode, sums_f),
                                        "cfg2", load=ingest_kernel(pir, n_inst, V, 0))}
        pir.set_config(base_seed=-1)
        log(f"fresh, full first sweep: {out['full_first_sweep']['ms_per_step']:.2f} ms/step, "
            f"same fixpoints: {agree}")
    for d_lo, d_hi in sets:
        d_lo.close()
        d_hi.close()
    return out


def formula_ask(pir, batch, op, x, y, z, lb, ub, w, n_inst, V):
    """PC's is_extractable over formula trees: the network written as
    FlatZinc, read back by the front-end, which keeps each constraint as PC's
    tree (x = y op z as an equality of a term, a reified comparison as an
    equivalence); one GPU thread per (store, tree) (pcx_batch_ask_formulas).
    On the batch's fixpoints (no store is a solution: each store stops at its
    first tree not asked, as PC's loop does), and on 4,096 copies of the
    planted solution (every tree asked: the full work, and every store must
    come out extractable).  Host parsing is setup; the timed part is the call
    (tree validation, upload, the kernel, the flags back)."""
    t0 = time.time()
    m = pcx.Model(pcx.flatzinc(op, x, y, z, lb, ub))
    nodes, off = m.formulas()
    sol = pcx.Batch(pir, n_inst, V)
    ws = np.tile(w.astype(np.int32), n_inst)
    sol.load(ws, ws)
    setup = time.time() - t0
    out = {"formulas": int(len(off) - 1), "nodes": int(len(nodes)), "instances": n_inst,
           "setup_s": round(setup, 2)}
    for name, b in (("fixpoints", batch), ("solutions", sol)):
        b.ask_formulas(nodes, off)
        ts = []
        for _ in range(3):
            t1 = time.perf_counter()
            ext = b.ask_formulas(nodes, off)
            ts.append(time.perf_counter() - t1)
        t = min(ts)
        out[name] = {"ms": t * 1e3, "extractable": int(ext.sum())}
        log(f"formula ask on the {name}: {out['formulas']} trees x {n_inst} stores in {t * 1e3:.2f} ms, "
            f"{int(ext.sum())} extractable")
    out["solutions"]["trees_per_s"] = (len(off) - 1) * n_inst / (out["solutions"]["ms"] * 1e-3)
    sol.close()
    return out


def search_step(args, pir, parent, off, idx, res, plo, phi, V, n_inst):
    """One step of a search on the device after the timed batch: the 4,096
    propagated subproblems are split on their first unfixed variable into
    8,192 children (pcx_batch_branch), which are propagated starting from
    their parents' fixpoints (marks on the split variable's propagators).
    Reported beside the headline figure; a sample of children is checked
    against the oracle."""
    sizes = np.diff(off)
    coff = np.concatenate([[0], np.cumsum(np.repeat(sizes, 2))]).astype(np.int64)
    cidx = np.concatenate([np.concatenate([idx[off[k]:off[k + 1]]] * 2) for k in range(n_inst)])
    child = pcx.Batch(pir, 2 * n_inst, V, coff, cidx)
    # warm-up: buffers, member bitmaps, graphs
    parent.branch(child, from_fixpoint=True)
    child.run(pcx.MODE_BATCHED)
    t0 = time.perf_counter()
    var = parent.branch(child, from_fixpoint=True)
    t1 = time.perf_counter()
    cres, st = child.run(pcx.MODE_BATCHED)
    t2 = time.perf_counter()
    n_child = 2 * n_inst
    split = int((var >= 0).sum())
    out = {"children": n_child, "parents_split": split, "branch_ms": (t1 - t0) * 1e3,
           "propagate_ms": (t2 - t1) * 1e3, "children_fixpoints_per_s": n_child / (t2 - t0),
           "evals": int(st.evals), "failed_children": int(cres.failed.sum())}
    log(f"search step: {n_child} children of {split} split parents, branch {out['branch_ms']:.2f} ms, "
        f"propagate {out['propagate_ms']:.2f} ms, {int(st.evals)} evaluations")
    if not args.no_cpu:
        ol = cpu_oracle()
        top, tx, ty, tz = pir.table()
        k_max = min(n_inst, 32)
        clo = np.empty(2 * k_max * V, np.int32)
        chi = np.empty(2 * k_max * V, np.int32)
        for k in range(k_max):
            lo, hi = plo[k * V:(k + 1) * V], phi[k * V:(k + 1) * V]
            for j in range(2):
                clo[(2 * k + j) * V:(2 * k + j + 1) * V] = lo
                chi[(2 * k + j) * V:(2 * k + j + 1) * V] = hi
            v = int(var[k])
            if v >= 0:
                mid = int(lo[v]) + (int(hi[v]) - int(lo[v])) // 2
                chi[2 * k * V + v] = mid
                clo[(2 * k + 1) * V + v] = mid + 1
        ro = ol.fixpoint_batch(top, tx, ty, tz, clo, chi, V, idx=cidx[:coff[2 * k_max]].astype(np.int64),
                               off=coff[:2 * k_max + 1], threads=host_threads())
        glo, ghi = child.read()
        mism = 0
        for c in range(2 * k_max):
            if bool(cres[c].failed) != bool(ro[c].failed):
                mism += 1
            elif not ro[c].failed and not (
                    np.array_equal(glo[c * V:(c + 1) * V], clo[c * V:(c + 1) * V]) and
                    np.array_equal(ghi[c * V:(c + 1) * V], chi[c * V:(c + 1) * V])):
                mism += 1
        out["parity"] = {"checked_children": 2 * k_max, "mismatches": mism}
        log(f"search step parity: {2 * k_max} children, mismatches {mism}")
    child.close()
    return out


def bench_single(args):
    V, P = 1000000, 4000000
    p = pcx.gen_params(seed=args.seed, nvars=V, nprops=P)
    op, x, y, z, lb, ub, w = pcx.gen_network(p)
    pir = pcx.PIR(props_per_wave=args.ppw, timing=1, stream_tables=args.stream,
                  event_driven=args.events)
    pir.tell(op, x, y, z, lb, ub)
    batch = pcx.Batch(pir, 1, V)
    batch.load(lb, ub)
    for _ in range(args.warmup):
        batch.reset()
        batch.run(pcx.MODE_PLAIN)
    sums = StatSum()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        batch.reset()
        res, st = batch.run(pcx.MODE_PLAIN)
        sums.add(st)
    elapsed = time.perf_counter() - t_start
    evals = sums.v["evals"]
    lb_g, ub_g = batch.read()
    out = {
        "value": args.steps / elapsed, "unit": "fixpoints/s", "ms_per_step": elapsed / args.steps * 1e3,
        "config": {"workload": f"single 1,000,000-var / 4,000,000-propagator fixpoint from the root "
                               f"store (plain engine; {GEN_LABEL}); each step restores the store "
                               f"from its device-resident image",
                   "nvars": V, "nprops": P, "props_per_wave": args.ppw},
        "evals_per_s": evals / elapsed, "sweeps": int(res[0].sweeps), "failed": bool(res[0].failed),
        "roofline": sums.roofline(args.steps, "k_sweep_plain", "k_sweep_marked_plain", "cfg1"),
    }
    if not args.no_cpu:
        ol = cpu_oracle()
        top, tx, ty, tz = pir.table()
        lo, hi = lb.copy(), ub.copy()
        ol.tell_embed(op, x, lo, hi)
        ts = time.perf_counter()
        r = ol.fixpoint(top, tx, ty, tz, lo, hi)
        t_cpu = time.perf_counter() - ts
        out["cpu_baseline"] = cpu_baseline(1.0 / t_cpu, 1, "port", "the full fixpoint, one thread")
        out["parity"] = {"bit_exact": bool(np.array_equal(lo, lb_g) and np.array_equal(hi, ub_g)),
                         "oracle_sweeps": int(r.sweeps)}
    return out


def bench_bits(args):
    """Bitset store: 200,000 words, 800,000 propagators; single instance and
    a batch of 1,024 bit-removal subproblems (one step = the batch)."""
    V, P, n_inst = 200000, 800000, 1024
    op, x, y, z, w = pcx.gen_bits_network(args.seed, V, P)
    pir = pcx.BitsPIR(props_per_wave=args.ppw, timing=1, stream_tables=args.stream,
                      event_driven=args.events, lockstep=args.lockstep)
    pir.tell(op, x, y, z, w)
    root = w.copy()
    t0 = time.perf_counter()
    r_single = pir.fixpoint(root)
    single_ms = (time.perf_counter() - t0) * 1e3
    inst = pcx.gen_bits_splits(args.seed + 1, root, n_inst, 1, 20)
    batch = pcx.Batch(pir, n_inst, V)
    batch.load_words(inst)
    mode = {"batched": pcx.MODE_BATCHED, "persistent": pcx.MODE_PERSISTENT}[args.mode]
    tuning = None
    if args.autotune:
        mode, tuning = autotune(args, pir, batch, f"cfg3:V{V}:P{P}:n{n_inst}:seed{args.seed}")
    for _ in range(args.warmup):
        batch.reset()
        batch.run(mode)
    sums = StatSum()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        batch.reset()
        res, st = batch.run(mode)
        sums.add(st)
    elapsed = time.perf_counter() - t_start
    evals = sums.v["evals"]
    fps = n_inst * args.steps / elapsed
    words_g = batch.read_words()
    rf = sums.roofline(args.steps, full_kernel(mode, sums) + "_bits", marked_kernel(pir, mode, sums, "_bits"),
                       "cfg3", tight_bytes=8)
    full_first = None
    if mode == pcx.MODE_BATCHED and sums.v["full_evals"] == 0:
        # the same steps with a full first sweep (seeding off): same fixpoints
        pir.set_config(base_seed=0)
        fsums = StatSum()
        batch.reset()
        batch.run(mode)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            batch.reset()
            fres, fst = batch.run(mode)
            fsums.add(fst)
        f_el = time.perf_counter() - t1
        same = (np.array_equal(batch.read_words(), words_g) and
                all(bool(a.failed) == bool(b.failed) for a, b in zip(fres, res)))
        full_first = {"fixpoints_per_s": n_inst * args.steps / f_el, "ms_per_step": f_el / args.steps * 1e3,
                      "same_fixpoints_as_seeded": same,
                      "roofline": fsums.roofline(args.steps, full_kernel(mode, fsums) + "_bits",
                                                 marked_kernel(pir, mode, fsums, "_bits"), "cfg3",
                                                 tight_bytes=8)}
        pir.set_config(base_seed=-1)
    out = {"value": fps, "unit": "fixpoints/s", "ms_per_step": elapsed / args.steps * 1e3,
           "config": {"workload": "bitset store: 1,024 bit-removal subproblems (1-20 removals each) of "
                                  "the root fixpoint of one 200,000-word / 800,000-propagator planted "
                                  "network (add, neq, leq, eq-reif; domain density U[0.3, 1.0] over the "
                                  "values 0-61); each step restores the 1,024 stores from their "
                                  "device-resident image",
                      "nvars": V, "nprops": P, "instances": n_inst,
                      "mode": {pcx.MODE_BATCHED: "batched", pcx.MODE_PERSISTENT: "persistent"}[mode],
                      "lockstep": sums.lockstep, "autotuned": tuning},
           "single_fixpoint_ms": single_ms, "single_sweeps": int(r_single.sweeps),
           "evals_per_s": evals / elapsed, "failed_instances": int(sum(r.failed for r in res)),
           "roofline": rf, "full_first_sweep": full_first}
    if not args.no_cpu:
        ol = cpu_oracle()
        top, tx, ty, tz = pir.table()
        threads = host_threads()
        k = 0
        t_cpu = 0.0
        mism = 0
        while k < n_inst and t_cpu < args.cpu_seconds:
            n = min(threads * 2, n_inst - k)
            wo = inst[k * V:(k + n) * V].copy()
            ts = time.perf_counter()
            ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=threads)
            t_cpu += time.perf_counter() - ts
            for j in range(n):
                g_fail = bool(res[k + j].failed)
                if g_fail != bool(ro[j].failed) or (not g_fail and not np.array_equal(
                        words_g[(k + j) * V:(k + j + 1) * V], wo[j * V:(j + 1) * V])):
                    mism += 1
            k += n
        out["cpu_baseline"] = cpu_baseline(k / t_cpu, threads, "port",
                                           f"first {k} of the {n_inst} instances, one per thread")
        out["parity"] = {"checked_instances": k, "mismatches": mism}
    return out


def bench_serving(args):
    from pcx.serving import make_trace
    t0 = time.time()
    pirs = []

    def embed(op, x, y, z, lb, ub):
        p = pcx.PIR(props_per_wave=args.ppw)
        p.tell(op, x, y, z, lb, ub)
        pirs.append(p)

    tr = make_trace(args.seed, args.requests, rate=args.rate, embed=embed)
    srv = pcx.Server(max_slots=args.slots, sweeps_per_tick=args.sweeps_per_tick,
                     props_per_wave=args.ppw)
    for j, p in enumerate(pirs):
        srv.add_network(p)
        if args.serve_root:               # requests inside the told store start from the root fixpoint
            srv.set_root(j, tr.nets[j][4], tr.nets[j][5])
    n = len(tr)
    sizes = np.array([tr.nprops(i) for i in range(n)])
    log(f"setup {time.time() - t0:.1f}s: {n} requests over {tr.arrival_s[-1]:.2f}s, "
        f"{len(pirs)} networks of {min(len(q[0]) for q in tr.nets)}-{max(len(q[0]) for q in tr.nets)} "
        f"propagators, mean {sizes.mean():.0f} propagators/request")

    def replay(plain, count=n):
        lo, hi = tr.lb_all.copy(), tr.ub_all.copy()
        out, wall = srv.replay(tr.arrival_s[:count], tr.net[:count], lo, hi, tr.store_off[:count + 1],
                               plain=plain)
        return lo, hi, out, wall

    warm = min(n, 200)
    for _ in range(args.warmup):
        replay(False, warm)
        replay(True, warm)
    walls, lat, evals = [], [], 0
    for _ in range(args.steps):
        lo, hi, out, wall = replay(False)
        walls.append(wall)
        lat.extend(o.latency_ms for o in out)
        evals += sum(o.result.evals for o in out)
    st = srv.stats()
    elapsed = sum(walls)
    lat = np.array(lat)
    failed = sum(o.result.failed for o in out)
    log(f"continuous: {n * args.steps / elapsed:.0f} requests/s, p50 {np.percentile(lat, 50):.2f} ms, "
        f"p99 {np.percentile(lat, 99):.2f} ms, max running {st.max_running}, failed {failed}/{n}")
    out_d = {
        "value": n * args.steps / elapsed, "unit": "fixpoints/s", "ms_per_step": elapsed / args.steps * 1e3,
        "config": {"workload": f"serving stream: Poisson arrivals at {args.rate:.0f} requests/s, "
                               f"{n} requests over {len(pirs)} seeded networks of 1k-1M propagators "
                               "(log-uniform, vars = propagators/4), 1-10 splits each of the network's told "
                               "store" + ("; each network's told store registered with the server "
                                          "(pcx_server_set_root), so a request starts from its meet "
                                          "with the root fixpoint" if args.serve_root else ""),
                   "requests": n, "rate": args.rate, "slots": args.slots, "roots": bool(args.serve_root),
                   "sweeps_per_tick": args.sweeps_per_tick, "props_per_wave": args.ppw},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "evals_per_s": evals / elapsed,
        "failed_requests": int(failed),
        "server": st.as_dict(),
    }
    res_c = (lo, hi, out)
    if not args.no_plain:
        lo_p, hi_p, out_p, wall_p = replay(True)
        lat_p = np.array([o.latency_ms for o in out_p])
        out_d["plain"] = {"requests_per_s": n / wall_p, "p50_ms": float(np.percentile(lat_p, 50)),
                          "p99_ms": float(np.percentile(lat_p, 99)), "max_ms": float(lat_p.max())}
        mism = 0
        for i in range(n):
            s = slice(int(tr.store_off[i]), int(tr.store_off[i + 1]))
            a, b = out[i].result, out_p[i].result
            if bool(a.failed) != bool(b.failed) or (not a.failed and not (
                    np.array_equal(lo[s], lo_p[s]) and np.array_equal(hi[s], hi_p[s]))):
                mism += 1
        out_d["plain"]["mismatches_vs_continuous"] = mism
        log(f"plain: {n / wall_p:.0f} requests/s, p50 {np.percentile(lat_p, 50):.2f} ms, "
            f"p99 {np.percentile(lat_p, 99):.2f} ms, mismatches {mism}")
    if not args.no_cpu:
        ol = cpu_oracle()
        threads = host_threads()
        tables = [q.table() for q in pirs]
        lo, hi, out = res_c
        # bounded sample: the requests of networks up to 2^17 propagators (the
        # oracle takes up to a minute for one request of the largest ones), in
        # arrival order, until about cpu_seconds of CPU work
        small = [i for i in range(n) if tr.nprops(i) <= 1 << 17]
        k = 0
        t_cpu = 0.0
        mism = 0
        while k < len(small) and t_cpu < args.cpu_seconds:
            # the next chunk of requests, grouped by network, each group's
            # requests spread over the host threads
            j = min(len(small), k + threads)
            chunk = small[k:j]
            for net in np.unique(tr.net[chunk]):
                group = [i for i in chunk if tr.net[i] == net]
                top, tx, ty, tz = tables[net]
                V = int(tr.store_off[group[0] + 1] - tr.store_off[group[0]])
                clo = np.concatenate([tr.store(i)[0] for i in group])
                chi = np.concatenate([tr.store(i)[1] for i in group])
                ts = time.perf_counter()
                r = ol.fixpoint_batch(top, tx, ty, tz, clo, chi, V, threads=threads)
                t_cpu += time.perf_counter() - ts
                for q, i in enumerate(group):
                    g = out[i].result
                    si = slice(int(tr.store_off[i]), int(tr.store_off[i + 1]))
                    if bool(g.failed) != bool(r[q].failed) or (not g.failed and not (
                            np.array_equal(lo[si], clo[q * V:(q + 1) * V]) and
                            np.array_equal(hi[si], chi[q * V:(q + 1) * V]))):
                        mism += 1
            k = j
        out_d["cpu_baseline"] = cpu_baseline(k / t_cpu, threads, "port",
                                             f"first {k} of the {len(small)} requests on networks of at most "
                                             f"131,072 propagators (of {n}), the requests of each network "
                                             "spread over the threads, oracle Gauss-Seidel (-O3)")
        out_d["parity"] = {"checked_requests": k, "mismatches": mism}
        log(f"cpu: {k} requests in {t_cpu:.1f}s on {threads} threads; parity mismatches {mism}")
    return out_d


def sub_args(args, **kw):
    """Arguments of a workload run beside the headline: bounded steps and CPU
    time, so the whole default line finishes in a few minutes."""
    sa = argparse.Namespace(**vars(args))
    sa.steps = min(args.steps, 10)
    sa.warmup = min(args.warmup, 2)
    sa.cpu_seconds = args.sub_cpu_seconds
    for k, v in kw.items():
        setattr(sa, k, v)
    return sa


def all_workloads(args):
    """The headline (batched search nodes, fresh stores per step) with the
    single-instance, bitset and serving workloads beside it."""
    import gc
    out = bench_batched(args)
    subs = {}
    for name, fn, kw in (("single", bench_single, {}),
                         ("bitset", bench_bits, {"lockstep": -1}),
                         ("serving", bench_serving, {"steps": 1, "warmup": 1})):
        gc.collect()
        t0 = time.time()
        log(f"== {name}")
        subs[name] = fn(sub_args(args, **kw))
        subs[name]["bench_s"] = round(time.time() - t0, 1)
    out["workloads"] = subs
    out["latency_ms"] = dict(subs["serving"]["latency_ms"], workload="serving")
    return out


COPY_CALIBRATION_BYTES = 640 * 50000 * 8


def copy_calibration():
    """Three launches of the library's plain device copy (a batch reset that
    copies the reset image: 640 stores of 50,000 variables, 256,000,000
    bytes read and written) after the workload: the known-size kernel that
    calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    V, n = 50000, 640
    pir = pcx.PIR(base_seed=0, lockstep=0, event_driven=0)     # no kernel the workloads time
    lb = np.zeros(V, np.int32)
    ub = np.full(V, 100, np.int32)
    pir.tell(np.array([pcx.ADD], np.int32), np.array([0], np.int32), np.array([1], np.int32),
             np.array([2], np.int32), lb, ub)
    b = pcx.Batch(pir, n, V)
    b.load(np.tile(lb, n), np.tile(ub, n))
    for _ in range(3):
        b.run(pcx.MODE_BATCHED)
        b.reset()
    b.close()
    log(f"copy calibration: 3 resets of {COPY_CALIBRATION_BYTES:,} bytes")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="0 (default): every workload, the batched search nodes as the headline; "
                         "1-4: one workload's line alone")
    ap.add_argument("--mode", default="batched", choices=["batched", "persistent"],
                    help="engine for the timed batch (the plain engine is always run once beside it)")
    ap.add_argument("--ppw", type=int, default=0, choices=[0, 64, 128, 256, 512, 1024, 2048],
                    help="propagators per wave (0 = engine default)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--compact", type=int, default=-1, choices=[-1, 0, 1],
                    help="per-instance subset tables: -1 auto, 0 shared table via indices, 1 copies")
    ap.add_argument("--stream", type=int, default=-1, choices=[-1, 0, 1],
                    help="table loads: -1 auto, 0 cached, 1 non-temporal")
    ap.add_argument("--autotune", type=int, default=1, choices=[0, 1],
                    help="configs 2 and 3: pick the sweep configuration with the autotuner "
                         "(cached as JSON, see --tune-cache) before the timed steps")
    ap.add_argument("--tune-cache", default=None, help="tuning cache file (default $PCX_TUNING or "
                                                       "~/.cache/pcx/tuning.json)")
    ap.add_argument("--retune", action="store_true", help="ignore a cached tuning")
    ap.add_argument("--events", type=int, default=-1, choices=[-1, 0, 1],
                    help="sweeps after the first: -1 auto, 0 every propagator, 1 marked only")
    ap.add_argument("--nvars", type=int, default=50000,
                    help="config 2: variables of the network (the config's own size by default)")
    ap.add_argument("--nprops", type=int, default=200000,
                    help="config 2: propagators of the network (the config's own size by default)")
    ap.add_argument("--lds", type=int, default=-1, choices=[-1, 0, 1],
                    help="persistent mode: stores staged in LDS (-1 auto, 0 never, 1 when they fit)")
    ap.add_argument("--lockstep", type=int, default=-1, choices=[-1, 0, 16, 32, 64],
                    help="configs 2 and 3, full sweeps: -1 auto, 0 per-instance work items, "
                         "16/32/64 instances per lockstep group")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity sample")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-engine comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--sub-cpu-seconds", type=float, default=8.0,
                    help="CPU-baseline time of each workload beside the headline")
    ap.add_argument("--fresh-check", type=int, default=256,
                    help="headline: instances of the last fresh step checked against the oracle")
    ap.add_argument("--search-step", type=int, default=1, choices=[0, 1],
                    help="config 2: after the timed steps, branch the subproblems on the device and "
                         "propagate the children (reported beside the headline figure)")
    ap.add_argument("--requests", type=int, default=4000, help="serving: requests in the trace")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: Poisson arrival rate (1/s)")
    ap.add_argument("--slots", type=int, default=256, help="serving: requests on the GPU at once")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the workload, three plain device copies of known size (counter calibration)")
    ap.add_argument("--formula-ask", type=int, default=1, choices=[0, 1],
                    help="config 2: PC's is_extractable over formula trees on the batch")
    ap.add_argument("--serve-root", type=int, default=1, choices=[0, 1],
                    help="serving: register each network's told store (requests start from the root fixpoint)")
    ap.add_argument("--sweeps-per-tick", type=int, default=4, help="serving: sweeps between host checks")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; --gpus must be 1 (batched instances would "
              "shard across GPUs with no exchange; a multi-GPU run is not part of this build)",
              file=sys.stderr)
        sys.exit(2)
    if pcx.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    out = {0: all_workloads, 1: bench_single, 2: bench_batched, 3: bench_bits,
           4: bench_serving}[args.config](args)
    if args.copy_calibration:
        copy_calibration()
    line = {"metric": METRIC, "value": out.pop("value"), "unit": out.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": out.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic (seeded generator)", "config": out.pop("config")}
    line.update(out)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
