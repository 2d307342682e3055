#!/usr/bin/env python3
"""Benchmark: synapse-updates/s of the per-cycle cortical pipeline (cycle + LTP)
on one MI355X, Config 2 workload: one 256x256-column area in the plain form,
SSC 1x4x32 + pyramidal 2 slices x 2 tufts x 8x64 (142,606,336 synapses), 2%
SDR input with 50% frame-to-frame overlap, learning on.  Input frames are
resident in HBM before the timed span; each timed cycle copies its frame into
the axon space and runs every fused stage.

Prints one JSON line.  --gpus N > 1 exits: this engine runs on one GPU.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X published HBM3E bandwidth
# Bytes of one synapse-update in a fused cycle that materialises every synapse
# buffer the reference keeps (flags rewritten, previous state read, unpacked
# source metadata).  This design does not move them; the figure is reported as
# an equivalent rate beside the roofline, never as its basis.
REF_BYTES_PER_SYNAPSE = 8
REF_BASIS = ("8 B per synapse-update: 3 B source metadata + 1 B strength + 1 B previous "
             "state + 1 B new state + 2 B flags read and written (dendrite/tuft bytes excluded)")
def engine_basis(syn_bs):
    per = " or ".join(f"{b} B" for b in sorted(set(syn_bs)))
    return (f"compulsory bytes of this design's launch: {per} per synapse by tuft (a 2-bit "
            "staged-slice code when the tuft's gather stages several slices, offsets packed into "
            "1 B when |v|,|u| <= 7, one strength-sign bit; states are virtual: none written, "
            "recomputed from a snapshot of the staged slices where read) + 4 B per dendrite + "
            "12 B per pyramidal cell-tuft + the snapshot copy of the staged slices")


def cpu_info():
    """Host CPU model, logical CPUs, and the CPUs this process may run on."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "allowed_cpus": allowed}


def baseline_threads():
    """OpenMP threads of the CPU baseline: every core this job is allotted --
    OMP_NUM_THREADS when the host sets it (16 per GPU on the GPU hosts), else
    every CPU the process may run on."""
    allowed = cpu_info()["allowed_cpus"]
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(int(env), allowed) if env.isdigit() and int(env) > 0 else allowed
PMC_SUMMARY = os.path.join(ROOT, "profiles", "kf_tuft_config2_pmc.md")


def writes_states(stage, kind, t):
    """Whether a check-free launch's tuft t of a layer of this kind writes its
    states: only tuft 0 of a spiny stellate layer (kind 1), and only where its
    fired bits are not stored (chunks below 8 synapses, or a launch staging
    several slices); every other tuft's states are virtual."""
    return kind == 1 and t == 0 and (stage["chunk"] < 8 or stage["staged"] > 1)


def syn_bytes(stage, tuft, kind=2, t=0):
    """Per-synapse HBM bytes of tuft t in an F1 launch.  Check-free gather: a
    2-bit staged-slice code when it stages several slices and nothing for one,
    the offsets (1 B packed, else 2 B) and one strength-sign bit; states are
    virtual (none written), except that tuft 0 of a spiny stellate layer stores
    each synapse's fired bit for its layer's learning (or its 1 B state, see
    writes_states).  Checked gather: the 1 B slice id, 2 B offsets, strength in
    and state out."""
    if stage["fast"]:
        slc = 0 if len(set(tuft.src_slcs)) <= 1 else 0.25
        out = 0
        if kind == 1 and t == 0:
            out = 1 if writes_states(stage, kind, t) else 0.125
        return slc + (1 if stage["packed"] else 2) + 0.125 + out
    return 1 + 2 + 2


def snap_bytes(scheme, tuft):
    """Bytes of a check-free tuft's snapshot (virtual states): its staged source
    slices."""
    sl = scheme.slices[min(tuft.src_slcs)]
    return len(set(tuft.src_slcs)) * sl.v_size * sl.u_size


def stage_bytes(scheme, li, kind, stage):
    """HBM bytes this engine moves in the F1 launches of layer li:
    per synapse syn_bytes(stage, tuft); per dendrite raw/state writes + activity
    read/write (4 B); per cell-tuft 4 B read + 8 B written (pyramidal), per SSC
    cell 1 B energy read.  F2/F3 are reported but not used for the roofline."""
    lay = scheme.layers[li]
    C = scheme.cells(li)
    if kind == 1:
        b = 0
        for ti, t in enumerate(lay.tufts):
            b += int(C * t.syns_per_tft * syn_bytes(stage, t, lay.kind, ti)) + C * t.dens_per_tft * 4
            if lay.kind == 2:
                b += C * 12
            if stage["fast"] and not writes_states(stage, lay.kind, ti):
                b += 2 * snap_bytes(scheme, t)
        if lay.kind == 1:
            b += C
        return b
    return None


def read_traffic():
    """HBM bytes per dominant-kernel launch from the committed rocprof PMC summary."""
    try:
        for line in open(PMC_SUMMARY):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def cpu_baseline(scheme, n_cycles, frames, label="Config 2", warm=2):
    """The oracle (plain-C restatement, OpenMP) timed on this host on a bounded
    sample of the same workload."""
    import numpy as np
    from oracle import pyoracle as P
    from bismit_amd import _lib as L
    threads = baseline_threads()
    P.set_threads(threads)
    o = P.OracleArea(scheme)
    for li, lay in enumerate(scheme.layers):
        cells = scheme.cells(li)
        parts = [L.gen_tuft_layout(0, li, t, cells * ts.syns_per_tft, ts.src_slcs, 6)
                 for t, ts in enumerate(lay.tufts)]
        for k, name in enumerate(["SYN_SRC_SLC_IDS", "SYN_SRC_COL_V_OFFS", "SYN_SRC_COL_U_OFFS"]):
            o.write(name, li, np.concatenate([p[k] for p in parts]))
    rnd = L.xorshift128(0, (n_cycles + warm) * o.n_slots).view(np.int32)
    for c in range(warm):
        o.intake(frames[c % len(frames)])
        o.cycle(rnd[c * o.n_slots:(c + 1) * o.n_slots], 1)
    t0 = time.perf_counter()
    for c in range(warm, n_cycles + warm):
        o.intake(frames[c % len(frames)])
        o.cycle(rnd[c * o.n_slots:(c + 1) * o.n_slots], 1)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": scheme.n_synapses() * n_cycles / dt, "unit": "synapse-updates/s",
            "cores": threads, "kind": "port", "cpu": cpu_info(),
            "ms_per_cycle": 1000.0 * dt / n_cycles,
            "sample": f"{n_cycles} cycles of the same {label} area (after {warm} warm-up cycles), "
                      f"oracle/bsm_oracle.c with OpenMP over {threads} threads, {dt:.1f} s"}


def config1_leg(ctx, n_cycles=1000, cpu=True):
    """Config 1 (64x64 columns, 24x24 glyph input, 2.6M synapses, learning on):
    n_cycles plain cycles on the GPU (frames resident in HBM, HIP events) and,
    as the CPU baseline of that workload, the same cycles on the oracle."""
    from bismit_amd import inputs, recipes
    from bismit_amd.cortex import CorticalArea
    scheme = recipes.config1_area()
    frames = inputs.GlyphStream(side=24, p=0.15, hold=10, seed=0).frames(n_cycles)
    off, ln = scheme.input_range()
    area = CorticalArea(scheme, ctx, rnd_seed=0, layout_seed=0, reach=6)
    dev = ctx.alloc(frames.nbytes)
    ctx.upload(dev, frames)
    area.run(dev, ln, ln, 50)
    area.sync()
    ms = statistics.median(area.time_cycles(dev, ln, ln, n_cycles) for _ in range(3))
    area.close()
    ctx.free(dev)
    out = {"workload": "config1: one 64x64-column area, 24x24 glyph input (p 0.15, new glyph "
                       "every 10 cycles), SSC 1x4x32, pyramidal 2 slices x 2 tufts x 4x32, "
                       "learning on",
           "synapses": scheme.n_synapses(), "cycles": n_cycles,
           "gpu": {"synapse_updates_per_s": scheme.n_synapses() * n_cycles / (ms / 1000.0),
                   "ms_per_cycle": ms / n_cycles, "cycles_per_s": 1000.0 * n_cycles / ms}}
    if cpu:
        out["cpu"] = cpu_baseline(scheme, n_cycles, frames, label="Config 1", warm=0)
    return out


def batched_leg(ctx, n_cycles=200, runs=3, tune=True, serial=False):
    """Config 3: 64 uneven areas in one persistent launch per stage with a
    hipGraph-captured cycle, against the same areas cycled one by one."""
    import numpy as np
    from bismit_amd import inputs, recipes
    from bismit_amd.autotune import candidates
    from bismit_amd.cortex import Batch, CorticalArea
    schemes = recipes.config3_areas(64, seed=0)
    areas = [CorticalArea(s, ctx, rnd_seed=i, layout_seed=i) for i, s in enumerate(schemes)]
    n_syn = sum(s.n_synapses() for s in schemes)
    nf = 64
    devs, lens = [], []
    for i, s in enumerate(schemes):
        f = inputs.SdrStream(s.input_range()[1], 0.02, 0.5, seed=i).frames(nf)
        d = ctx.alloc(f.nbytes)
        ctx.upload(d, f)
        devs.append(d)
        lens.append(f.shape[1])
    # plain per-area form: every area launches its own stages on its own stream
    # (the reference gives each area its own queue); all areas' cycles are
    # enqueued up front so host overhead stays out and streams may overlap.
    # Both stream modes of the plain form are timed and the faster one is the
    # baseline (a second stream per area helps one large area, and hurts when
    # 64 areas' streams share the hardware queues).
    from bismit_amd.cortex import cycle_areas
    for i, a in enumerate(areas):
        a.run(devs[i], lens[i], lens[i], 3)
    ctx.sync()
    n_plain = max(20, n_cycles // 4)
    n_loop = max(10, n_cycles // 10)
    plain_t, loop_t = {}, {}
    for mode in (1, 2):
        for i, a in enumerate(areas):
            a.set_tuning(overlap=mode)
            a.run(devs[i], lens[i], lens[i], 2)
        ctx.sync()
        t0 = time.perf_counter()
        for i, a in enumerate(areas):
            a.run(devs[i], lens[i], lens[i], min(n_plain, nf))
            if n_plain > nf:
                a.run(devs[i], lens[i], lens[i], n_plain - nf)
        for a in areas:
            a.sync()
        ctx.sync()
        plain_t[mode] = (time.perf_counter() - t0) / n_plain
        # the reference's driving pattern (Cortex::cycle): every cycle, each area's
        # intake and cycle enqueued in turn by one host thread (C loop, no Python)
        t0 = time.perf_counter()
        cycle_areas(areas, devs, lens, nf, n_loop)
        for a in areas:
            a.sync()
        ctx.sync()
        loop_t[mode] = (time.perf_counter() - t0) / n_loop
    for a in areas:
        a.set_tuning(overlap=1)
    plain_s = min(plain_t.values())
    plain = n_syn / plain_s
    loop_s = min(loop_t.values())
    loop = n_syn / loop_s
    b = Batch(areas, ctx)
    if serial:
        b.set_tuning(overlap=2)
    b.set_inputs(devs, lens, lens, [s.input_range()[0] for s in schemes], nf)
    best, table = None, []
    if tune:
        for t in candidates():
            b.set_tuning(**t)
            b.cycle(2)
            ms = b.time_cycles(10) / 10
            table.append((ms, t))
        table.sort(key=lambda x: x[0])
        best = table[0][1]
        b.set_tuning(**best)
        if not serial:
            # the schedule at that tiling: independent pyramidal tufts on a side
            # stream (1) or every launch on one stream (2), whichever is faster here
            sched = []
            for ov in (1, 2):
                b.set_tuning(overlap=ov)
                b.cycle(2)
                sched.append((b.time_cycles(20) / 20, ov))
            best = dict(best, overlap=min(sched)[1])
            b.set_tuning(**best)
    b.cycle(5)
    times = [b.time_cycles(n_cycles) for _ in range(runs)]
    ms = statistics.median(times)
    batched = n_syn * n_cycles / (ms / 1000.0)
    # per-stage device time (HIP events around each persistent launch)
    n_prof = 30
    prof = b.profile_stages(n_prof)
    kinds = {1: "kf_tuft", 2: "kf_ssc_post", 3: "kf_pyr_post"}
    stages = []
    plan = b.stage_plan()
    for (k, t), p in zip(prof, plan):
        assert k == p["kind"]
        st = {"kernel": kinds[k], "stream": p["stream"], "avg_us": round(1000.0 * t / n_prof, 2),
              "chunk": p["chunk"], "fast": p["fast"], "packed": p["packed"],
              "tile": [p["tile_v"], p["tile_u"]], "items": p["n_items"]}
        if k == 1:
            st["engine_bytes_per_launch"] = p["alg_bytes"]
            st["hbm_frac"] = round(p["alg_bytes"] / (t / n_prof / 1000.0) / 1e9 /
                                   HBM_PEAK_GBS, 4)
        stages.append(st)
    cyc_alg = sum(st.get("engine_bytes_per_launch", 0) for st in stages)
    b.close()
    for a in areas:
        a.close()
    for d in devs:
        ctx.free(d)
    return {"workload": "config3: 64 areas, sides log-uniform 32..192 (multiples of 8), 1-3 slices "
                        "per layer, 16-64 synapses per dendrite, own 2% SDR input each",
            "synapses": n_syn, "batched_synapse_updates_per_s": batched,
            "plain_per_area_synapse_updates_per_s": plain, "speedup_vs_plain": batched / plain,
            "plain_cortex_loop_synapse_updates_per_s": loop,
            "speedup_vs_cortex_loop": batched / loop,
            "cortex_loop_ms_per_cycle": 1000 * loop_s,
            "batched_ms_per_cycle": ms / n_cycles, "plain_ms_per_cycle": 1000 * plain_s,
            "plain_ms_per_cycle_by_mode": {"side_stream": 1000 * plain_t[1],
                                           "one_stream": 1000 * plain_t[2]},
            "cortex_loop_ms_per_cycle_by_mode": {"side_stream": 1000 * loop_t[1],
                                                 "one_stream": 1000 * loop_t[2]},
            "tuning": best, "stages": stages,
            "hbm_frac": cyc_alg * n_cycles / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
            "hbm_frac_basis": "F1 launches' compulsory bytes (as the roofline) per cycle over the "
                              "whole cycle's time",
            "ref_8B_equivalent_frac": batched * REF_BYTES_PER_SYNAPSE / (HBM_PEAK_GBS * 1e9)}


def serving_leg(ctx, wall_s=20.0, rate=200.0, capacity=1.0e12, **server_kw):
    """Config 5: Poisson session arrivals, admission by HBM footprint and compute
    budget, continuous batching of frames into shared cycles; p50/p99 per-frame
    latency from frame production at the client."""
    from bismit_amd.serve import Server, area_scheme, session_specs
    specs = session_specs(rate=rate, horizon_s=wall_s, seed=0)
    syn = {}
    offered = 0.0   # synapse-updates the arriving sessions ask for over their lifetimes
    for sp in specs:
        for side in sp.sides:
            if side not in syn:
                syn[side] = area_scheme(side, "o").n_synapses()
            offered += syn[side] * sp.n_frames
    # the estimate starts at 0.6 of the static batch's rate (masked cycles and the
    # loop's host work sustained 0.53-0.56 of it in earlier runs) and may rise to it
    srv = Server(ctx, capacity=capacity, initial_capacity=0.6 * capacity, **server_kw)
    st = srv.run(specs, wall_limit_s=wall_s)
    out = st.summary()
    out.update({"arrival_rate_sessions_per_s": rate,
                "offered_synapse_updates_per_s": offered / max(wall_s, 1e-9),
                "offered_basis": "every arriving session's synapses x frames, over the run",
                "capacity_source": "capacity_max: the batched leg's measured synapse-updates/s"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed cycles per run")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--runs", type=int, default=5, help="timed runs (median reported)")
    ap.add_argument("--cpu-cycles", type=int, default=250,
                    help="CPU baseline sample (about 12 s of oracle work at 16 threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--autotune", action="store_true",
                    help="plain form: tile chosen by the autotuner instead of the fixed default")
    ap.add_argument("--tile", type=int, nargs=2, metavar=("V", "U"), default=[2, 64],
                    help="column tile of the plain form (default 2 64: with the one-stream "
                         "schedule and 32-synapse chunks the fastest of 1x64, 2x32, 2x64, 4x16, "
                         "4x32, 4x64, 8x16 and 8x32 on MI355X, 1.5%% ahead of 4x32)")
    ap.add_argument("--chunk", type=int, default=32,
                    help="synapses per lane per chunk of the plain form's synapse stage "
                         "(default 32: about 4%% faster per cycle than 16 at 4x32 tiles)")
    ap.add_argument("--inhib-tile", type=int, default=0,
                    help="side of the inhibition tile (8, 16, 32; 0: the planner's choice)")
    ap.add_argument("--schedule", choices=("one", "side"), default="one",
                    help="plain form: one stream (default), or the independent pyramidal "
                         "tuft on a side stream")
    ap.add_argument("--serial", action="store_true",
                    help="one stream per area/batch (no side stream); for profiler runs")
    ap.add_argument("--batched-only", action="store_true",
                    help="profiling helper: run only the 64-area batched leg and print it")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="profiling helper: one copy-kernel calibration launch, then 3 cycles")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from bismit_amd import inputs, recipes
    from bismit_amd.cortex import Context, CorticalArea

    ctx = Context(0)
    if args.batched_only:
        print(json.dumps({"batched": batched_leg(ctx, n_cycles=args.steps, runs=args.runs,
                                                 serial=args.serial)}),
              flush=True)
        return
    scheme = recipes.config2_area()
    area = CorticalArea(scheme, ctx, rnd_seed=0, layout_seed=0, reach=6)
    in_ofs, in_len = scheme.input_range()
    n_frames = max(args.steps, args.warmup, 8)
    frames = inputs.SdrStream(in_len, 0.02, 0.5, seed=0).frames(n_frames)
    dev = ctx.alloc(frames.nbytes)
    ctx.upload(dev, frames)

    if not args.autotune:
        tuning = dict(tile_v=args.tile[0], tile_u=args.tile[1], chunk=args.chunk, halo=6)
        if args.inhib_tile:
            tuning["inhib_tile"] = args.inhib_tile
        area.set_tuning(**tuning)
    plain_overlap = 2 if (args.serial or args.schedule == "one") else 1
    area.set_tuning(overlap=plain_overlap)
    if args.pmc_probe:
        n = 256 << 20
        a, b = ctx.alloc(n), ctx.alloc(n)
        ms = C_float_copy(ctx, a, b, n)
        print(f"calib copy {n} bytes {ms:.3f} ms", flush=True)
        area.run(dev, in_len, in_len, 3)
        area.sync()
        print("pmc probe done", flush=True)
        return

    if args.autotune:
        # tile geometry by the project autotuner (timed on a scratch copy of the area)
        from bismit_amd.autotune import tune_area
        tiles = [(8, 8), (4, 8), (2, 8), (2, 16), (1, 32), (4, 16), (1, 16)]
        tuning, _ = tune_area(scheme, ctx, cycles=10,
                              cands=[dict(tile_v=tv, tile_u=tu, chunk=32, halo=6)
                                     for tv, tu in tiles])
        area.set_tuning(**tuning)
    area.run(dev, in_len, in_len, args.warmup)
    area.sync()
    times = []
    for r in range(args.runs):
        ctx.sync()
        times.append(area.time_cycles(dev, in_len, in_len, args.steps))
        print(f"# run {r}: {args.steps} cycles in {times[-1]:.2f} ms", file=sys.stderr, flush=True)
    ms = statistics.median(times)
    n_syn = scheme.n_synapses()
    value = n_syn * args.steps / (ms / 1000.0)

    # roofline of the dominant kernel: HIP events around every launch, launched one
    # at a time on the one-stream schedule (each layer's tufts in one launch, as
    # in the committed rocprof summaries), whatever schedule the timed runs used
    n_prof = 50
    schedule = [(p["kind"], p["stream"]) for p in area.stage_plan()]
    area.set_tuning(overlap=2)
    prof = area.profile_stages(n_prof)
    kinds = {1: "kf_tuft", 2: "kf_ssc_post", 3: "kf_pyr_post"}
    plan = area.stage_plan()
    stages = [{"kernel": kinds[k], "layer": scheme.layers[li].name, "stream": p["stream"],
               "avg_us": round(1000.0 * t / n_prof, 2), "chunk": p["chunk"], "fast": p["fast"],
               "packed": p["packed"],
               "tile": [p["tile_v"], p["tile_u"]], "items": p["n_items"],
               "engine_bytes_per_launch": p["alg_bytes"]}
              for (k, li, t), p in zip(prof, plan)]
    groups = {}
    for (k, li, t), p in zip(prof, plan):
        g = groups.setdefault((k, li), [0.0, 0, 0, p])
        g[0] += t
        g[1] += p["alg_bytes"]
        g[2] += 1
    area.set_tuning(overlap=plain_overlap)
    # the host-buffer boundary: each cycle's frame handed over from host memory
    # (bsm_area_intake of a host frame, then bsm_area_cycle), PCIe included
    n_host = min(200, args.steps)
    area.sync()
    t0 = time.perf_counter()
    for c in range(n_host):
        area.intake(frames[c % n_frames])
        area.cycle(1)
    area.sync()
    host_s = (time.perf_counter() - t0) / n_host
    pcie = {"synapse_updates_per_s": n_syn / host_s, "ms_per_cycle": 1000 * host_s,
            "cycles": n_host,
            "how": "host frame per cycle through bsm_area_intake (synchronous copy), then "
                   "bsm_area_cycle with host-drawn rnd; Python host loop"}
    (dk, dl), (dt, alg_tot, n_launch, dstage) = max(groups.items(), key=lambda kv: kv[1][0])
    syn_bs = [syn_bytes(dstage, t, scheme.layers[dl].kind, ti)
              for ti, t in enumerate(scheme.layers[dl].tufts)]
    # the engine's count equals the recipe's
    assert alg_tot == stage_bytes(scheme, dl, dk, dstage), (alg_tot, syn_bs)
    avg_s = dt / n_prof / n_launch / 1000.0
    syn_launch = scheme.cells(dl) * sum(t.syns_per_tft for t in scheme.layers[dl].tufts) / n_launch
    # algorithmic bytes of the launch: what this design must move (syn_bytes per
    # synapse, 4 B per dendrite, 12 B per pyramidal cell-tuft); the PMC traffic of
    # the same launch confirms it (traffic_over_alg)
    alg = alg_tot / n_launch
    achieved = alg / avg_s / 1e9
    frac = achieved / HBM_PEAK_GBS
    if frac > 1.0:
        raise RuntimeError(f"roofline basis moves {achieved:.0f} GB/s, above the {HBM_PEAK_GBS} "
                           "GB/s peak: the byte count is not what the launch moves")
    traffic = read_traffic()
    ref8 = REF_BYTES_PER_SYNAPSE * syn_launch
    # the calibration kernel with the launch's access pattern: read-only streams for
    # the check-free gather (no synapse byte written), else its read streams
    # (syn_bytes - 1) and one write stream
    if dstage["fast"]:
        ceil = stream_ceiling(ctx, n_reads=1, write=False)
    else:
        ceil = stream_ceiling(ctx, n_reads=int(max(syn_bs)) - 1)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(frac, 4),
                "traffic": traffic,
                "traffic_over_alg": round(traffic / alg, 4) if traffic else None,
                "kernel": f"{kinds[dk]} ({scheme.layers[dl].name})",
                "launches_per_cycle": n_launch, "synapse_updates_per_launch": syn_launch,
                "alg_bytes_per_launch": alg, "avg_launch_us": round(avg_s * 1e6, 2),
                "alg_basis": engine_basis(syn_bs),
                "frac_of_stream_ceiling": round(achieved / ceil["gbs"], 4),
                "stream_ceiling": ceil,
                "ref_8B_equivalent": {
                    "bytes_per_launch": ref8,
                    "gbs": round(ref8 / avg_s / 1e9, 1),
                    "basis": REF_BASIS,
                    "note": "the rate a design materialising every reference synapse buffer "
                            "would need for the same synapse-updates/s; it exceeds the HBM peak "
                            "because this design does not move those bytes, so it is no "
                            "roofline"}}

    out = {
        "metric": "synapse-updates/s (cycle+LTP), batched vs plain on 1 GPU; p50/p99 frame latency",
        "value": value,
        "unit": "synapse-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "config2: one 256x256-column area, plain form, 142,606,336 synapses "
                               "(SSC 1x4x32, pyramidal 2 slices x 2 tufts x 8x64), 2% SDR input "
                               "with 50% overlap, learning on",
                   "synapses": n_syn, "runs": args.runs, "tuning": tuning,
                   "device": ctx.info()["name"]},
        "roofline": roofline,
        "stages": stages,
        "timed_schedule": [{"kernel": kinds[k], "stream": st} for k, st in schedule],
        "host_frames_pcie_inclusive": pcie,
    }
    if not args.no_batched:
        out["batched"] = batched_leg(ctx)
        print("# batched leg done", file=sys.stderr, flush=True)
    if not args.no_serve:
        cap = out["batched"]["batched_synapse_updates_per_s"] if "batched" in out else 1.0e12
        out["serving"] = serving_leg(ctx, wall_s=args.serve_seconds, capacity=cap)
        print("# serving leg done", file=sys.stderr, flush=True)
    if not args.no_config1:
        out["config1"] = config1_leg(ctx, cpu=not args.no_cpu)
        print("# config1 leg done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scheme, args.cpu_cycles,
                                           inputs.SdrStream(in_len, 0.02, 0.5, seed=0).frames(8))
    area.close()
    print(json.dumps(out), flush=True)


def stream_ceiling(ctx, n_reads=4, write=True, n=128 << 20, reps=5):
    """Best of `reps` launches of the stream calibration kernel (F1's read
    streams and, for the checked gather, its one write stream; no gather): a
    practical HBM ceiling."""
    import ctypes as C
    from bismit_amd import _lib as L
    src = ctx.alloc(n_reads * n)
    dst = ctx.alloc(n) if write else None
    best = None
    for _ in range(reps):
        ms = C.c_float()
        L.check(ctx.lib.bsm_calib_stream(ctx.h, dst, src, n, n_reads, C.byref(ms)))
        best = ms.value if best is None else min(best, ms.value)
    ctx.free(src)
    if write:
        ctx.free(dst)
    return {"gbs": round((n_reads + (1 if write else 0)) * n / (best / 1000.0) / 1e9, 1),
            "pattern": f"{n_reads} non-temporal 16-B read stream(s)" +
                       (" + 1 16-B write stream" if write else ", nothing written") +
                       f", {n >> 20} MiB each"}


def C_float_copy(ctx, dst, src, n):
    import ctypes as C
    from bismit_amd import _lib as L
    ms = C.c_float()
    L.check(ctx.lib.bsm_calib_copy(ctx.h, dst, src, n, C.byref(ms)))
    return ms.value


if __name__ == "__main__":
    main()
