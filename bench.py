"""Benchmark of the MI355X event-based transport loop.

Workload (default): the GPU-scale lattice problem -- 60 synthetic nuclides
with 20k-100k-point grids, three 20-nuclide materials in a 17x17 pin lattice
(1.26 cm pitch, reflective sides, vacuum at z = +-50 cm), Watt source in the
fuel, 1e7 particles per batch, 100x100x100 track-length flux mesh in FP64.
A step is one batch.  Inputs (nuclear data, geometry, tallies) are resident in
HBM before timing starts; the timed region is W warm-up batches excluded, then
exactly K batches, each ending in a stream synchronization.

The timed batches run the library's default path: event cycles replayed
from hipGraphs, no per-kernel timers.  After them, one more batch runs with a
HIP event recorded on the library's stream around every kernel launch (direct
launches); the per-kernel table and the roofline of the dominant kernel come
from that batch (its measured HBM traffic from the rocprofv3 summary in
profiles/).

Prints one JSON line: particles/s (whole job), XS lookups/s, the roofline,
the learned-index comparison on this workload and on the lookup-dominated
one (300 nuclides x 200k points in one material), and a CPU baseline (the
oracle, history- and event-based, OpenMP over the host's cores, best of three
on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (fixed contract): 48 B per XS lookup (E, sigma_t,
# sigma_a brackets), 16 B per collision (sigma_el bracket), 16 B per scored
# mesh segment (FP64 read + write), 144 B of particle state per event.
B_LOOKUP, B_COLL, B_SEG, B_EVENT = 48, 16, 16, 144


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def xs_lookups(st):
    """Micro lookups of the calculate-xs events (the reference's calculate_xs
    calls, cache skips honoured).  The collision kernel repeats the lookup of
    the sampled nuclide, which the reference reads from its per-particle cache:
    one per collision, not credited."""
    return st["lookups"] - st["collisions"]


def kxs_lookups(st):
    """The calculate-xs lookups done by the k_xs kernel (not the history-based tail)."""
    return xs_lookups(st) - st.get("hist_lookups", 0)


def kernel_bytes(st):
    """Algorithmic bytes moved by each event kernel over the measured region
    (the events of the history-based tail are credited to k_history)."""
    hx, hl, ha, hs = st.get("hist_xs_events", 0), st.get("hist_lookups", 0), st.get("hist_adv_events", 0), \
        st.get("hist_segments", 0)
    hsu, hc = st.get("hist_surf_events", 0), st.get("hist_coll_events", 0)
    return {
        "k_xs": B_LOOKUP * (xs_lookups(st) - hl) + B_EVENT * (st["xs_events"] - hx),
        "k_advance": B_SEG * (st["segments"] - hs) + B_EVENT * (st["adv_events"] - ha),
        "k_surface": B_EVENT * (st["surf_events"] - hsu),
        "k_collision": B_COLL * (st["collisions"] - hc) + B_EVENT * (st["coll_events"] - hc),
        "k_history": B_LOOKUP * hl + B_SEG * hs + B_COLL * hc + B_EVENT * (hx + ha + hsu + hc),
    }


# statistics that are settings, not counts (reported as they stand, not differenced)
CONFIG_STATS = ("blocks", "xs_lanes", "sort", "xs_group", "block", "sort_every", "cycles_per_check", "sort_bits",
                "sort_blocks", "history_below", "history_steps", "fused_advance", "recompute_micro")


def diff(a, b):
    return {k: b[k] - a[k] for k in b}


def run_gpu(d, steps, warmup, learned=False, n_seg=256, epochs=200, options=None, timed_steps=1, particles=None):
    """W untimed warm-up batches, then exactly K timed batches on the default
    launch path (event cycles replayed from hipGraphs, no per-kernel timers);
    then `timed_steps` more batches with a HIP event recorded around every
    kernel launch for the per-kernel times and the roofline (direct launches,
    so these are kept out of the headline)."""
    from openmc_amd import lib
    old = os.getcwd()
    os.chdir(d)
    try:
        lib.init(output=False)
        info = {}
        if particles:
            lib.settings.particles = int(particles)
        for k, v in (options or {}).items():
            lib.set_option(k, v)
        if learned:
            t = time.perf_counter()
            info["learned_max_err"] = lib.learned_build(n_seg, epochs)
            info["learned_build_s"] = time.perf_counter() - t
            info.update({"learned_" + k: v for k, v in lib.learned_stats().items()})
            lib.set_option("learned_index", 1)
        lib.set_option("event_timing", 0)
        lib.set_option("progress_s", 30)
        lib.simulation_init()
        s_first = lib.stats()
        tw = time.perf_counter()
        for k in range(warmup):
            lib.next_batch()
            log("  warm-up batch %d done (%.2f s)" % (k + 1, time.perf_counter() - tw))
        lib.reset_timers()
        s0 = lib.stats()
        t0 = time.perf_counter()
        for k in range(steps):
            lib.next_batch()
            log("  batch %d done (%.2f s)" % (k + 1, time.perf_counter() - t0))
        t1 = time.perf_counter()
        s1 = lib.stats()
        if timed_steps > 0:
            lib.set_option("event_timing", 1)
            lib.reset_timers()
            s1b = lib.stats()
            tk = time.perf_counter()
            for k in range(timed_steps):
                lib.next_batch()
                log("  event-timed batch %d done (%.2f s)" % (k + 1, time.perf_counter() - tk))
            s2 = lib.stats()
        s_last = lib.stats()
        lib.simulation_finalize()
        n = lib.settings.particles
        lib.finalize()
    finally:
        os.chdir(old)
    st = diff(s0, s1)
    info.update({"elapsed_s": t1 - t0, "stats": st, "n_particles": n, "stats_all": diff(s_first, s_last),
                 "config_stats": {k: s1.get(k, 0) for k in CONFIG_STATS}})
    if timed_steps > 0:
        info["timed"] = {"stats": diff(s1b, s2), "steps": timed_steps,
                         "config_stats": {k: s2.get(k, 0) for k in CONFIG_STATS}}
    return info


# HBM traffic per event of each kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes over this workload scaled by the copy-kernel calibration
# (lines "traffic <kernel> <unit> <bytes per unit>").
TRAFFIC_FILES = {1: os.path.join(ROOT, "profiles", "hbm_traffic_lattice_1e7.txt"),
                 2: os.path.join(ROOT, "profiles", "hbm_traffic_cube300_1e7.txt")}
UNIT_COUNTER = {"xs_event": "xs_events", "adv_event": "adv_events", "surf_event": "surf_events",
                "coll_event": "coll_events"}


def traffic_per_launch(kernel, st, launches, config=1):
    """Measured HBM bytes per launch of `kernel` in this run: the profiled
    bytes per event times this run's events per launch (None without one)."""
    path = TRAFFIC_FILES.get(config)
    if not path or not os.path.exists(path) or launches <= 0:
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == "traffic" and parts[1] == kernel:
                c = UNIT_COUNTER[parts[2]]
                return round(float(parts[3]) * (st[c] - st.get("hist_" + c, 0)) / launches, 1)
    return None


def roofline(info, config=1):
    st = info["stats"]
    times = {"k_xs": st["t_xs_ms"], "k_advance": st["t_adv_ms"], "k_surface": st["t_surf_ms"],
             "k_collision": st["t_coll_ms"], "partition_xs": st["t_sort_ms"],
             "partition_split": st["t_split_ms"], "k_init": st["t_init_ms"], "k_history": st.get("t_hist_ms", 0.0)}
    launches = {"k_xs": st["n_xs"], "k_advance": st["n_adv"], "k_surface": st["n_surf"],
                "k_collision": st["n_coll"], "partition_xs": st["n_sort"], "partition_split": st["n_split"],
                "k_init": st["n_init"], "k_history": st.get("n_hist", 0)}
    by = kernel_bytes(st)
    # queue bookkeeping and source sampling carry no algorithmic bytes
    by["partition_xs"] = 0
    by["partition_split"] = 0
    by["k_init"] = 0
    fused = info.get("config_stats", {}).get("fused_advance", 0) > 0
    if fused:
        # one launch does calculate-xs and advance: its bytes are both
        by["k_xs"] += by["k_advance"]
        by["k_advance"] = 0
    dom = max(times, key=times.get)
    # average algorithmic bytes per launch over the average launch duration
    achieved = by[dom] / (times[dom] * 1e-3) / 1e9
    traffic = traffic_per_launch(dom, st, launches[dom], config)
    if fused and dom == "k_xs":
        # the fused kernel's own profile, else the two kernels' profiles added
        traffic = traffic_per_launch("k_xs_adv", st, launches[dom], config)
        if traffic is None:
            xs = traffic_per_launch("k_xs", st, launches[dom], config)
            adv = traffic_per_launch("k_advance", st, launches[dom], config)
            traffic = None if xs is None or adv is None else round(xs + adv, 1)
    per_kernel = {k: {"ms": times[k], "launches": launches[k], "avg_us": 1e3 * times[k] / max(launches[k], 1),
                      "GB_per_s": by[k] / max(times[k], 1e-9) * 1e-6} for k in times}
    total_bytes = sum(by.values())
    whole = total_bytes / (st["transport_ms"] * 1e-3) / 1e9
    name = "k_xs_lane (calculate-xs + advance)" if fused and dom == "k_xs" else dom
    return dom, {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                 "algorithmic_bytes_per_launch": round(by[dom] / max(launches[dom], 1), 1),
                 "kernel": name, "whole_loop_GB_per_s": round(whole, 2)}, per_kernel


def host_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS where the host sets
    it (the host's share of cores), else every core."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(d, target_s=4.0, repeats=3):
    """The oracle on the host cores, history-based and event-based, each on a
    sample of the first particles of batch 1 sized to about target_s, timed
    `repeats` times (best kept)."""
    from oracle.oracle import Oracle
    threads = host_threads()
    path = os.path.join(d, "oracle_problem.txt")

    def timed(event_based, n):
        o = Oracle(path)  # loading the data is not timed
        t = time.perf_counter()
        o.run(event_based=event_based, threads=threads, n_particles=n, n_batches=1, trace_hash=False)
        el = time.perf_counter() - t
        c = o.counters()
        o.close()
        return el, c

    modes = {}
    for name, event_based in (("history", False), ("event", True)):
        # probes of growing size until one takes a tenth of the target, then
        # a sample sized to the target, timed `repeats` times
        n, el = 64, 0.0
        while True:
            el, c = timed(event_based, n)
            log("  cpu %s probe: %d histories in %.2f s" % (name, n, el))
            if el >= 0.1 * target_s or n >= 2_000_000:
                break
            n = int(min(2_000_000, n * min(16.0, max(2.0, 0.1 * target_s / max(el, 1e-3)))))
        n = int(max(n, min(2_000_000, n * target_s / max(el, 1e-3))))
        times = []
        for r in range(repeats):
            el, c = timed(event_based, n)
            times.append(el)
            log("  cpu %s sample %d/%d: %d histories in %.2f s" % (name, r + 1, repeats, n, el))
        best_t = min(times)
        modes[name] = {"particles_per_s": round(n / best_t, 2), "particles": n, "seconds": round(best_t, 3),
                       "seconds_all": [round(t, 3) for t in times],
                       "lookups_per_s": round(c["lookups"] / best_t, 1)}
    best = max(modes, key=lambda k: modes[k]["particles_per_s"])
    return {"value": modes[best]["particles_per_s"], "unit": "particles/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": "the oracle (C++ restatement, -O3, OpenMP %d threads on %s) on the first %d histories of "
                      "batch 1 of the same problem, %s-based (the faster of its two modes); best of %d runs, "
                      "%.2f s" % (threads, cpu_model(), modes[best]["particles"], best, repeats,
                                  modes[best]["seconds"]),
            "modes": modes}


def run_problem_set(dirs, steps, warmup, options=None, tune_path=None, timed_steps=1, learned=False, n_seg=256,
                    epochs=200):
    """The problems of a set in one shared event loop: W warm-up and K timed
    batches on the graph path (every problem advances one batch per step),
    then `timed_steps` event-timed batches for the kernel table; with
    `learned`, on the learned index."""
    from openmc_amd import autotune, lib
    lib.init_problems(dirs, output=False)
    info = {}
    try:
        for k, v in (options or {}).items():
            lib.set_option(k, v)
        if tune_path:
            t = time.perf_counter()
            best, rec = autotune.tune_cached(dirs, tune_path, log=log)
            info["autotune"] = {"options": best, "seconds": round(time.perf_counter() - t, 2),
                                "trials": rec["trials"], "from_cache": bool(rec.get("cached"))}
        if learned:
            t = time.perf_counter()
            info["learned_max_err"] = lib.learned_build(n_seg, epochs)
            info["learned_build_s"] = time.perf_counter() - t
            info.update({"learned_" + k: v for k, v in lib.learned_stats().items()})
            lib.set_option("learned_index", 1)
        lib.set_option("event_timing", 0)
        lib.simulation_init()
        for _ in range(warmup):
            lib.next_batch()
        lib.reset_timers()
        s0 = lib.stats()
        t0 = time.perf_counter()
        for k in range(steps):
            lib.next_batch()
            log("  batch %d done (%.2f s)" % (k + 1, time.perf_counter() - t0))
        t1 = time.perf_counter()
        s1 = lib.stats()
        if timed_steps > 0:
            lib.set_option("event_timing", 1)
            lib.reset_timers()
            s1b = lib.stats()
            for k in range(timed_steps):
                lib.next_batch()
            s2 = lib.stats()
            info["timed"] = {"stats": diff(s1b, s2), "steps": timed_steps,
                             "config_stats": {k: s2.get(k, 0) for k in CONFIG_STATS}}
        info["options"] = {k: s1[k] for k in ("blocks", "block", "sort_every", "cycles_per_check", "sort_bits",
                                              "xs_group", "history_below", "fused_advance", "recompute_micro")}
        info["config_stats"] = {k: s1.get(k, 0) for k in CONFIG_STATS}
        lib.simulation_finalize()
        info["n_particles"] = [lib.problem_info(k)["n_particles"] for k in range(lib.n_problems())]
    finally:
        lib.finalize()
    info.update({"elapsed_s": t1 - t0, "stats": diff(s0, s1)})
    return info


def cpu_baseline_set(problems, dirs, target_s=10.0):
    """The oracle, history-based, on the first particles of batch 1 of the
    problems of a set in turn, until about target_s of CPU work."""
    from oracle.oracle import Oracle
    from oracle.problem import write_oracle_problem
    threads = host_threads()
    done, secs, used = 0, 0.0, 0
    for (model, nucs), d in zip(problems, dirs):
        path = write_oracle_problem(model, nucs, d, data_dir=os.path.join("..", "data"))
        n = min(model.settings.particles, 20000)
        o = Oracle(path)
        t = time.perf_counter()
        o.run(event_based=False, threads=threads, n_particles=n, n_batches=1, trace_hash=False)
        secs += time.perf_counter() - t
        o.close()
        done += n
        used += 1
        if secs >= target_s:
            break
    return {"value": round(done / secs, 2), "unit": "particles/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": "the oracle (C++ restatement, -O3, OpenMP %d threads, history-based) on the first "
                      "min(n, 20000) histories of batch 1 of problems 0-%d of the set, one after another; "
                      "%d histories in %.1f s" % (threads, used - 1, done, secs)}


def bench_problem_set(args):
    """Batched uneven problems: 64 lattice problems from a seed in one event
    loop, against the same problems run one after another."""
    from openmc_amd import configs
    n_batches = args.warmup + args.steps + args.timed_steps + 1
    log("preparing the problem set")
    d, dirs, problems, nucs = configs.problem_set_dir(n_batches=max(n_batches, 2))
    opts = {k: float(v) for k, v in (o.split("=", 1) for o in args.option)}
    tune_path = None if args.no_autotune else os.path.join(d, "autotune.json")
    log("GPU problem set: %d problems, %d warm-up + %d timed batches" % (len(dirs), args.warmup, args.steps))
    rs = run_problem_set(dirs, args.steps, args.warmup, options=opts, tune_path=tune_path,
                         timed_steps=args.timed_steps)
    st = rs["stats"]
    tst = rs["timed"]["stats"]
    total = sum(rs["n_particles"])
    value = total * args.steps / rs["elapsed_s"]
    dom, roof, per_kernel = roofline(rs["timed"], config=3)
    chosen = dict(opts)
    chosen.update(rs.get("autotune", {}).get("options", {}))
    log("GPU problems one at a time (1 warm-up + 1 timed batch each)")
    alone = []
    for k, dd in enumerate(dirs):
        r = run_gpu(dd, 1, 1, options=chosen, timed_steps=0)
        alone.append({"particles": r["n_particles"], "ms_per_batch": round(1e3 * r["elapsed_s"], 3),
                      "particles_per_s": round(r["n_particles"] / r["elapsed_s"], 1)})
    seq_s = sum(a["ms_per_batch"] for a in alone) * 1e-3
    log("GPU problem set on the learned index (the same launch options)")
    rl = run_problem_set(dirs, args.steps, args.warmup, options=chosen, timed_steps=args.timed_steps, learned=True)
    out = {
        "metric": "particles/s and XS lookups/s on 1 MI355X: learned-index vs binary search",
        "value": round(value, 2),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * rs["elapsed_s"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic nuclides from fixed seeds; random source histories",
        "config": {"workload": "batched-%d-lattice-problems-%.3gpps" % (len(dirs), total), "problems": len(dirs),
                   "particles_per_batch": total, "nuclides": len(nucs),
                   "mode": "event-based, one event loop shared by all problems"},
        "xs_lookups_per_s": round(kxs_lookups(tst) / tst["t_xs_ms"] * 1e3, 1),
        "sequential": {"particles_per_s": round(total / seq_s, 2), "seconds_per_batch": round(seq_s, 4)},
        "batching_speedup": round(value / (total / seq_s), 3),
        "per_problem": alone,
        "learned": learned_summary(rs, rl, total, args.steps),
        "launch_options": rs["options"],
        "kernels": per_kernel,
        "events": {k: st.get(k, 0) for k in ("xs_events", "adv_events", "surf_events", "coll_events", "lookups",
                                             "collisions", "segments", "search_probes", "cycles", "hist_launches",
                                             "hist_xs_events", "hist_lookups", "hist_adv_events",
                                             "hist_segments", "hist_surf_events", "hist_coll_events")},
        "roofline": roof,
    }
    if "autotune" in rs:
        out["autotune"] = rs["autotune"]
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, history-based, problem samples)")
        out["cpu_baseline"] = cpu_baseline_set(problems, dirs)
    print(json.dumps(out), flush=True)


# rocprof kernel name -> (name in this script's kernel table, unit of work)
KERNEL_UNITS = {"k_xs_lane": ("k_xs", "xs_event"), "k_xs": ("k_xs", "xs_event"),
                "k_xs_adv": ("k_xs_adv", "xs_event"),
                "k_advance": ("k_advance", "adv_event"), "k_surface": ("k_surface", "surf_event"),
                "k_collision": ("k_collision", "coll_event")}


def pmc_summary(dirs, events=None, copy_bytes=None):
    """Per-kernel sums of the counters in rocprofv3 databases (``--pmc`` or
    ``--kernel-trace`` runs), as text.  With the run's event counts (a JSON
    dict from this script's output) and the copy kernel's size, also the
    calibrated HBM bytes per event of each event kernel ("traffic" lines:
    counter bytes scaled by copy_bytes / what the counter read for the copy)."""
    import collections
    import glob
    import re
    import sqlite3
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    dur = collections.defaultdict(float)
    for d in dirs:
        for f in glob.glob(os.path.join(d, "*_results.db")):
            c = sqlite3.connect(f)
            tables = {r[0] for r in c.execute("select name from sqlite_master")}
            if "counters_collection" in tables:
                for name, cn, val, did in c.execute(
                        "select kernel_name, counter_name, value, dispatch_id from counters_collection"):
                    k = re.sub(r"\(.*", "", name).replace("void ", "").replace("omc::", "")
                    agg[k][cn] += val
                    disp[k].add(did)
            if "kernels" in tables:
                for name, s, e, did in c.execute("select name, start, end, dispatch_id from kernels"):
                    k = re.sub(r"\(.*", "", name).replace("void ", "").replace("omc::", "")
                    dur[k] += (e - s) * 1e-6
                    disp[k].add(did)
    lines = ["%-34s %10s %14s %16s %16s" % ("kernel", "dispatches", "total_ms", "FETCH_KiB/disp", "WRITE_KiB/disp")]
    for k in sorted(disp, key=lambda k: -max(dur.get(k, 0), agg[k].get("FETCH_SIZE", 0))):
        n = max(len(disp[k]), 1)
        lines.append("%-34s %10d %14.3f %16.1f %16.1f" % (k, n, dur.get(k, 0.0), agg[k].get("FETCH_SIZE", 0) / n,
                                                        agg[k].get("WRITE_SIZE", 0) / n))
    # any other counters, per dispatch
    for k in sorted(agg):
        for cn in sorted(agg[k]):
            if cn not in ("FETCH_SIZE", "WRITE_SIZE"):
                lines.append("counter %-30s %-28s %18.1f" % (k, cn, agg[k][cn] / max(len(disp[k]), 1)))
    if events and copy_bytes:
        copy = [k for k in agg if k.startswith("k_copy")]
        if copy:
            ck = copy[0]
            fetch_scale = copy_bytes / (agg[ck]["FETCH_SIZE"] * 1024.0 / len(disp[ck]))
            write_scale = copy_bytes / (agg[ck]["WRITE_SIZE"] * 1024.0 / len(disp[ck]))
            lines.append("copy calibration: %d-byte copy; FETCH scale %.4f, WRITE scale %.4f"
                         % (copy_bytes, fetch_scale, write_scale))
            for k in agg:
                base = re.sub(r"<.*", "", k)
                if base not in KERNEL_UNITS:
                    continue
                name, unit = KERNEL_UNITS[base]
                c = UNIT_COUNTER[unit]
                # the kernel's own events (not those of the history-based tail)
                n_ev = events.get(c, 0) - events.get("hist_" + c, 0)
                if n_ev <= 0:
                    continue
                b = (agg[k]["FETCH_SIZE"] * 1024.0 * fetch_scale + agg[k]["WRITE_SIZE"] * 1024.0 * write_scale)
                lines.append("traffic %s %s %.2f" % (name, unit, b / n_ev))
    return "\n".join(lines)


WORKLOADS = {0: "cube-4nuclides-1e5pps-mesh10^3", 1: "lattice-17x17-60nuclides-1e7pps-mesh100^3",
             2: "cube-1m-300nuclides-200kpoints-1e7pps-mesh100^3",
             4: "slab-5x20cm-20nuclides-per-layer-weight-windows-50x50x100-mesh200^3x10groups"}
EVENT_KEYS = ("xs_events", "adv_events", "surf_events", "coll_events", "lookups", "collisions", "segments",
              "search_probes", "cycles", "hist_launches", "hist_xs_events", "hist_lookups", "hist_adv_events",
              "hist_segments", "hist_surf_events", "hist_coll_events", "scatter_fallthrough")


def learned_summary(plain, lr, n, steps):
    """The learned-index run next to the plain one (same inputs, same launch
    options): whole-batch speedup on the graph path, lookups/s in the
    calculate-xs kernel from the event-timed batches."""
    st, lst = plain["stats"], lr["stats"]
    out = {"particles_per_s": round(n * steps / lr["elapsed_s"], 2),
           "speedup_vs_binary": round(plain["elapsed_s"] / lr["elapsed_s"], 4),
           "probes_per_lookup": round(lst["search_probes"] / max(lst["lookups"], 1), 3),
           "probes_per_lookup_binary": round(st["search_probes"] / max(st["lookups"], 1), 3),
           "max_err": lr.get("learned_max_err"), "mean_err": lr.get("learned_mean_err"),
           "segments": {"base": lr.get("learned_n_seg"), "min": lr.get("learned_min_segments"),
                        "max": lr.get("learned_max_segments"), "total": lr.get("learned_total_segments")},
           "build_s": round(lr.get("learned_build_s", 0), 3)}
    if "timed" in lr and "timed" in plain:
        lt, pt = lr["timed"]["stats"], plain["timed"]["stats"]
        out["xs_lookups_per_s"] = round(kxs_lookups(lt) / lt["t_xs_ms"] * 1e3, 1)
        out["xs_lookups_per_s_binary"] = round(kxs_lookups(pt) / pt["t_xs_ms"] * 1e3, 1)
        out["k_xs_speedup_vs_binary"] = round(pt["t_xs_ms"] / lt["t_xs_ms"] * kxs_lookups(lt) / kxs_lookups(pt), 4)
    return out


def lookup_kernel_rates(d, n=1 << 20, n_seg=256, epochs=200):
    """The macroscopic lookup kernel alone on material 0 of the problem in d:
    n log-uniform energies in random order and sorted, plain and learned
    (the same interleaved lookup code as the calculate-xs kernel), lookups/s
    and the check that both give the same bits."""
    import numpy as np
    from openmc_amd import lib
    rng = np.random.default_rng(7)
    old = os.getcwd()
    os.chdir(d)
    try:
        lib.init(output=False)
        try:
            e_min, e_max, _ = lib.energy_grid()
            E = np.exp(rng.uniform(np.log(e_min), np.log(e_max), n))
            lib.learned_build(n_seg, epochs)
            nn = lib.macro_nuclides(0)
            out = {"energies": n, "nuclides_per_lookup_event": nn}
            for order, Eo in (("random_order", E), ("sorted", np.sort(E))):
                res = {}
                r_plain, ms_p = lib.macro_batch(0, Eo, learned=False, repeat=3)
                r_learn, ms_l = lib.macro_batch(0, Eo, learned=True, repeat=3)
                res["plain_lookups_per_s"] = round(n * nn / (ms_p * 1e-3), 1)
                res["learned_lookups_per_s"] = round(n * nn / (ms_l * 1e-3), 1)
                res["speedup_vs_binary"] = round(ms_p / ms_l, 4)
                res["bitwise_equal"] = bool(np.array_equal(r_plain.view(np.uint64), r_learn.view(np.uint64)))
                out[order] = res
        finally:
            lib.finalize()
    finally:
        os.chdir(old)
    return out


def bench_lookup_workload(args, opts):
    """The lookup-dominated workload the metric's comparison is about: 300
    nuclides with 200k-point grids in one material (1 m cube, 1e7 particles
    per batch), plain and learned, one warm-up and --lookup-steps timed
    batches each on the graph path, one event-timed batch each for the kernel
    split; then the same two with the calculate-xs queue left in particle
    order (no energy sort, the reference's own order) at a tenth of the
    particles; then the lookup kernel alone on random-order and sorted
    energies."""
    from openmc_amd import configs
    steps = args.lookup_steps
    log("preparing the lookup-dominated problem (300 nuclides x 200k points)")
    d, model, nucs = configs.problem_dir(2, n_particles=args.particles, n_batches=1 + steps + 1 + 1 + 1)
    log("GPU plain, lookup-dominated: 1 warm-up + %d timed batches (+1 event-timed)" % steps)
    plain = run_gpu(d, steps, 1, options=opts, timed_steps=1)
    log("GPU learned, lookup-dominated: 1 warm-up + %d timed batches (+1 event-timed)" % steps)
    lr = run_gpu(d, steps, 1, learned=True, options=opts, timed_steps=1)
    n = plain["n_particles"]
    st = plain["stats"]
    tst = plain["timed"]["stats"]
    dom, roof, per_kernel = roofline(plain["timed"], config=2)
    nu = max(1, n // 10)
    log("GPU plain and learned with the queue in particle order (no energy sort), %d particles" % nu)
    pu = run_gpu(d, 1, 1, options=dict(opts, sort=0), timed_steps=1, particles=nu)
    lu = run_gpu(d, 1, 1, learned=True, options=dict(opts, sort=0), timed_steps=1, particles=nu)
    unsorted = dict(learned_summary(pu, lu, nu, 1), particles_per_batch=nu,
                    plain_particles_per_s=round(nu / pu["elapsed_s"], 2))
    log("lookup kernel alone on the same data (random order and sorted)")
    kernel = lookup_kernel_rates(d)
    return {"workload": WORKLOADS[2], "particles_per_batch": n, "nuclides": len(nucs),
            "grid_points": int(sum(x.n_points for x in nucs)), "steps": steps,
            "particles_per_s": round(n * steps / plain["elapsed_s"], 2),
            "xs_lookups_per_s": round(kxs_lookups(tst) / tst["t_xs_ms"] * 1e3, 1),
            "xs_lookups_per_s_transport": round(xs_lookups(st) / plain["elapsed_s"], 1),
            "learned": learned_summary(plain, lr, n, steps),
            "unsorted_queue": unsorted,
            "lookup_kernel": kernel,
            "roofline": roof, "kernels": per_kernel,
            "events": {k: st.get(k, 0) for k in EVENT_KEYS}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 result directories instead of running")
    ap.add_argument("--events", help="with --pmc-summary: JSON line printed by the profiled run")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=10_000_000)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--timed-steps", type=int, default=1,
                    help="batches after the timed ones with per-kernel HIP event timers (kernel table, roofline)")
    ap.add_argument("--no-lookup-workload", action="store_true",
                    help="skip the 300-nuclide lookup-dominated comparison of plain and learned")
    ap.add_argument("--lookup-steps", type=int, default=1,
                    help="timed batches of the lookup-dominated comparison")
    ap.add_argument("--no-autotune", action="store_true",
                    help="with --config 3: keep the default launch options")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="transport option passed to omc_set_option (e.g. sort=0, xs_lanes=8)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="run a plain device copy of this many bytes first (HBM counter calibration)")
    args = ap.parse_args()
    if args.pmc_summary:
        events = None
        if args.events:
            with open(args.events) as f:
                for line in f:
                    if line.startswith("{"):
                        rec = json.loads(line)
                        events = rec.get("events_all_batches") or rec.get("events")
        print(pmc_summary(args.pmc_summary, events, args.copy_calibration or None))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU "
              "design is a per-rank particle split plus one tally reduction per batch)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.config == 3:
        bench_problem_set(args)
        return
    from openmc_amd import configs
    copy_info = None
    if args.copy_calibration > 0:
        from openmc_amd import lib
        ms = lib.copy_kernel(args.copy_calibration, reps=3)
        copy_info = {"bytes": args.copy_calibration, "ms": ms,
                     "GB_per_s": round(2 * args.copy_calibration / (ms * 1e-3) / 1e9, 1)}
        log("copy kernel: %d bytes in %.3f ms (%.0f GB/s read+write)" % (args.copy_calibration, ms,
                                                                          copy_info["GB_per_s"]))
    log("preparing problem (config %d, %d particles/batch)" % (args.config, args.particles))
    kw = {"n_particles": args.particles, "n_batches": args.warmup + args.steps + args.timed_steps + 1}
    d, model, nucs = configs.problem_dir(args.config, **kw)
    from oracle.problem import write_oracle_problem
    if not args.no_cpu_baseline:
        write_oracle_problem(model, nucs, d)

    log("GPU plain binary search: %d warm-up + %d timed batches (+%d event-timed)" % (args.warmup, args.steps,
                                                                                    args.timed_steps))
    opts = {k: float(v) for k, v in (o.split("=", 1) for o in args.option)}
    plain = run_gpu(d, args.steps, args.warmup, options=opts, timed_steps=args.timed_steps)
    if "timed" in plain:
        dom, roof, per_kernel = roofline(plain["timed"], config=args.config)
    else:
        # --timed-steps 0: no event-timed batch, so no per-kernel times
        dom, roof, per_kernel = None, None, None
    n = plain["n_particles"]
    value = n * args.steps / plain["elapsed_s"]
    st = plain["stats"]
    tst = plain["timed"]["stats"] if "timed" in plain else None
    out = {
        "metric": "particles/s and XS lookups/s on 1 MI355X: learned-index vs binary search",
        "value": round(value, 2),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * plain["elapsed_s"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic nuclides from fixed seeds; random source histories",
        "config": {"workload": WORKLOADS.get(args.config, "config%d" % args.config), "particles_per_batch": n,
                   "nuclides": len(nucs), "grid_points": int(sum(x.n_points for x in nucs)),
                   "mode": "event-based, hipGraph-replayed event cycles"},
        "xs_lookups_per_s": round(kxs_lookups(tst) / tst["t_xs_ms"] * 1e3, 1) if tst else None,
        "xs_lookups_per_s_transport": round(xs_lookups(st) / plain["elapsed_s"], 1),
        "events_per_particle": round((st["xs_events"] + st["adv_events"] + st["surf_events"] +
                                      st["coll_events"]) / (n * args.steps), 2),
        "kernel_timing": "per-kernel figures and the roofline from %d further batch(es) with a HIP event "
                         "around every launch (direct launches); the headline from the %d graph-replayed "
                         "batches" % (args.timed_steps, args.steps),
        "kernels": per_kernel,
        "events": {k: st.get(k, 0) for k in EVENT_KEYS},
        # every batch of the plain run (warm-up, timed, event-timed): what a
        # profiler wrapped around a --no-learned run sees
        "events_all_batches": {k: plain["stats_all"].get(k, 0) for k in EVENT_KEYS},
        "roofline": roof,
    }
    if copy_info:
        out["copy_calibration"] = copy_info
    if not args.no_learned:
        log("GPU learned index: %d warm-up + %d timed batches (+%d event-timed)" % (args.warmup, args.steps,
                                                                                  args.timed_steps))
        lr = run_gpu(d, args.steps, args.warmup, learned=True, options=opts, timed_steps=args.timed_steps)
        out["learned"] = learned_summary(plain, lr, n, args.steps)
    if args.config == 1 and not args.no_lookup_workload:
        out["lookup_dominated"] = bench_lookup_workload(args, opts)
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(d)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
