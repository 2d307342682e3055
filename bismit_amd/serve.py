"""Serving loop: continuous batching of sessions' sensory frames into shared
cycles on one GPU.

Each session owns a cortex of 1-3 areas resident in HBM for its lifetime and
streams 2%-sparse SDR frames at its own rate.  The server
  - admits waiting sessions while their HBM footprint fits the budget (and at
    most `max_sessions` are resident),
  - at every shared cycle steps exactly the areas whose session has a frame
    due, through one dynamic batch (one persistent launch per stage, idle areas
    masked on the device; optionally captured as a hipGraph),
  - reads every stepped area's output axons back through a sampler (one gather
    kernel and one copy per cycle for all sessions), hands them to the session
    (`on_output`), and records per-frame latency = the frame's output readable
    on the host - the frame's arrival time, and aggregate synapse-updates/s,
  - regrows a session's weak synapses after every `regrow_every` of its cycles
    (the reference's CorticalArea::cycle regrows every 201 cycles,
    src/cortex/cortical_area.rs:773-776,792-806), all due sessions in one masked call.
Sessions that arrive while others run join at the next admission epoch; their
frames overlap on the device with everyone else's.  Each area keeps its own rnd
stream, state parity and input, so its result is the same as a plain replay of
the session alone (tests/test_gpu_batch.py checks it).

The reference has no serving mode; its single-threaded flywheel event loop
(src/flywheel.rs:235-377) is the nearest analogue.
"""
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import recipes
from .cortex import Batch, CorticalArea, footprint
from .inputs import SdrStream
from .sampler import SamplerKind, recv_many


@dataclass
class SessionSpec:
    sid: int
    arrival: float          # seconds after start
    sides: List[int]        # one entry per area
    rate_hz: float
    duration_s: float
    seed: int

    @property
    def n_frames(self):
        return max(1, int(self.rate_hz * self.duration_s))


def session_specs(rate=200.0, horizon_s=10.0, seed=0, sides=(64, 256), rates=(30.0, 120.0),
                  durations=(2.0, 30.0), max_areas=3):
    """Poisson arrivals at `rate` sessions/s over `horizon_s`; 1..max_areas areas
    per session with sides uniform over multiples of 8 in `sides`."""
    rng = np.random.default_rng(seed)
    t, out, sid = 0.0, [], 0
    lo, hi = sides[0] // 8, sides[1] // 8
    while True:
        t += rng.exponential(1.0 / rate)
        if t > horizon_s:
            break
        n = int(rng.integers(1, max_areas + 1))
        out.append(SessionSpec(sid, t, [8 * int(rng.integers(lo, hi + 1)) for _ in range(n)],
                               float(rng.uniform(*rates)), float(rng.uniform(*durations)),
                               int(rng.integers(0, 2**31))))
        sid += 1
    return out


def area_scheme(side, name):
    """Session areas use the workload's layer recipe at their own size."""
    return recipes.cortical_area(name, side)


def output_layer(scheme):
    """The layer whose axon slices are a session area's output: the last
    pyramidal layer."""
    return max(i for i, lay in enumerate(scheme.layers) if not lay.spatial)


class Session:
    def __init__(self, spec: SessionSpec, ctx, t_admit, keep_frames=False, sample_outputs=True):
        self.spec = spec
        self.t0 = t_admit      # reset when the session joins the batch
        self.schemes = [area_scheme(s, f"s{spec.sid}a{i}") for i, s in enumerate(spec.sides)]
        self.areas = []
        for i, sc in enumerate(self.schemes):
            self.areas.append(CorticalArea(sc, ctx, rnd_seed=spec.seed + 7919 * i,
                                           layout_seed=spec.seed + 104729 * i, reach=6))
        self.streams = [SdrStream(sc.input_range()[1], 0.02, 0.5, seed=spec.seed + i)
                        for i, sc in enumerate(self.schemes)]
        self.next = 0                      # next frame index
        self.done = 0                      # frames completed
        self.synapses = sum(sc.n_synapses() for sc in self.schemes)
        self.log = [] if keep_frames else None
        # per-cycle output readback (no backpressure: a slow consumer skips frames)
        self.outputs = [a.sampler(SamplerKind.Axons(output_layer(sc)))
                        for a, sc in zip(self.areas, self.schemes)] if sample_outputs else []
        self.last_outputs = [None] * len(self.areas)

    def due(self, k):
        return self.t0 + k / self.spec.rate_hz

    def regrow_seed(self, k):
        """Regrowth seed of area k (its epochs count the session's regrowths)."""
        return self.spec.seed + 15485863 * k + 1

    def finished(self):
        return self.done >= self.spec.n_frames

    def close(self):
        for rx in self.outputs:
            rx.close()
        for a in self.areas:
            a.close()


@dataclass
class ServeStats:
    latencies: list = field(default_factory=list)
    cycles: int = 0
    synapse_updates: float = 0.0
    gpu_busy_s: float = 0.0
    wall_s: float = 0.0
    admitted: int = 0
    completed: int = 0
    arrived: int = 0
    max_resident: int = 0
    rebuilds: int = 0
    mean_active_areas: float = 0.0
    rebuild_s: list = field(default_factory=list)
    setup_s: list = field(default_factory=list)

    def summary(self):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {
            "frames": len(self.latencies),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "synapse_updates_per_s": self.synapse_updates / max(self.wall_s, 1e-9),
            "shared_cycles": self.cycles,
            "sessions_arrived": self.arrived,
            "sessions_admitted": self.admitted,
            "sessions_completed": self.completed,
            "max_resident_sessions": self.max_resident,
            "batch_rebuilds": self.rebuilds,
            "mean_active_areas_per_cycle": self.mean_active_areas,
            "wall_s": self.wall_s,
            "rebuild_ms_max": max(self.rebuild_s, default=0.0) * 1e3,
            "rebuild_ms_mean": float(np.mean(self.rebuild_s) * 1e3) if self.rebuild_s else 0.0,
            "session_setup_ms_mean": float(np.mean(self.setup_s) * 1e3) if self.setup_s else 0.0,
        }


class Server:
    def __init__(self, ctx, hbm_budget_bytes=96 << 30, max_sessions=64, epoch_s=0.05,
                 use_graph=False, keep_frames=False, on_retire=None, setup_workers=2,
                 regrow_every=201, regrow_reach=6, sample_outputs=True, on_output=None):
        self.ctx = ctx
        # Masked cycles are enqueued directly by default: the batch changes with
        # every admission epoch, and capturing plus instantiating its two graphs
        # (about 4 ms, on the frame path) costs more latency than the ~25 us of
        # launches per cycle a graph saves (p99 14.4 -> 11.7 ms, same p50, at
        # 200 sessions/s).
        self.sample_outputs = sample_outputs
        self.on_output = on_output            # on_output(session, [Frame per area])
        self.regrow_every = regrow_every      # 0: regrowth off
        self.regrow_reach = regrow_reach
        self.budget = hbm_budget_bytes
        self.max_sessions = max_sessions
        self.epoch_s = epoch_s
        self.use_graph = use_graph
        self.keep_frames = keep_frames
        self.on_retire = on_retire
        self.resident = []
        self.used = 0
        self.batch = None
        # sessions are set up (areas allocated and laid out on the device) by
        # worker threads while the loop keeps cycling; the ctypes calls release
        # the interpreter lock, and each area works on its own stream
        self.setup_workers = setup_workers
        self._ready = []
        self._ready_lock = threading.Lock()

    def footprint(self, spec):
        return sum(footprint(area_scheme(s, "f")) for s in spec.sides)

    def _rebuild(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        areas = [a for s in self.resident for a in s.areas]
        if areas:
            self.batch = Batch(areas, self.ctx, dynamic=True)

    def _deliver(self, active_sessions):
        """Every stepped area's output frame of this cycle, to its session."""
        rxs = [rx for s in active_sessions for rx in s.outputs]
        frames = recv_many(rxs, wait_for_frame=False)
        k = 0
        for s in active_sessions:
            got = frames[k:k + len(s.outputs)]
            k += len(s.outputs)
            if any(f is None for f in got):
                raise RuntimeError(f"session {s.spec.sid}: output frame missing")
            s.last_outputs = got
            if self.on_output:
                self.on_output(s, got)

    def _regrow_due(self, active_set):
        """Regrowth of the active sessions whose cycle about to complete is a
        multiple of regrow_every (it belongs to that cycle, as upstream)."""
        n = len(self.batch.areas)
        which = np.zeros(n, np.uint8)
        seeds = [0] * n
        epochs = [0] * n
        idx = 0
        for s in self.resident:
            c = s.done + 1
            if id(s) in active_set and c % self.regrow_every == 0:
                for k in range(len(s.areas)):
                    which[idx + k] = 1
                    seeds[idx + k] = s.regrow_seed(k)
                    epochs[idx + k] = c // self.regrow_every - 1
            idx += len(s.areas)
        if which.any():
            self.batch.regrow(seeds, epochs, self.regrow_reach, which)

    def _setup(self, spec, fp, st):
        t0 = time.perf_counter()
        sess = Session(spec, self.ctx, 0.0, self.keep_frames, self.sample_outputs)
        sess.fp = fp
        st.setup_s.append(time.perf_counter() - t0)
        with self._ready_lock:
            self._ready.append(sess)

    def run(self, specs, wall_limit_s=30.0, time_scale=1.0):
        """Serve `specs` (arrival times in seconds, compressed by time_scale) until
        every admitted session finished or wall_limit_s passed."""
        with ThreadPoolExecutor(max_workers=self.setup_workers) as pool:
            return self._run(specs, wall_limit_s, time_scale, pool)

    def _run(self, specs, wall_limit_s, time_scale, pool):
        st = ServeStats()
        pending = sorted(specs, key=lambda s: s.arrival)
        waiting = []
        setting_up = []           # futures of sessions being set up
        t_start = time.perf_counter()
        clock = lambda: time.perf_counter() - t_start
        last_epoch = -1e9
        act_sum = 0
        while True:
            now = clock()
            if now > wall_limit_s:
                break
            while pending and pending[0].arrival * time_scale <= now:
                waiting.append(pending.pop(0))
                st.arrived += 1
            changed = False
            if now - last_epoch >= self.epoch_s:
                last_epoch = now
                retired = [s for s in self.resident if s.finished()]
                if retired:
                    # the batch releases every area before retired ones are freed
                    if self.batch is not None:
                        self.batch.close()
                        self.batch = None
                    for s in retired:
                        if self.on_retire:
                            self.on_retire(s)
                        self.used -= s.fp
                        s.close()
                        st.completed += 1
                    self.resident = [s for s in self.resident if not s.finished()]
                    changed = True
                # admission: HBM is reserved now, the set-up runs in the background
                while waiting and len(self.resident) + len(setting_up) < self.max_sessions:
                    fp = self.footprint(waiting[0])
                    if self.used + fp > self.budget:
                        break
                    spec = waiting.pop(0)
                    self.used += fp
                    setting_up.append(pool.submit(self._setup, spec, fp, st))
                    st.admitted += 1
                # sessions whose set-up finished join the batch
                for f in [f for f in setting_up if f.done()]:
                    f.result()           # re-raise a set-up error here
                    setting_up.remove(f)
                with self._ready_lock:
                    ready, self._ready = self._ready, []
                for sess in ready:
                    sess.t0 = clock()
                    self.resident.append(sess)
                    changed = True
            if changed:
                t_rb = time.perf_counter()
                self._rebuild()
                st.rebuild_s.append(time.perf_counter() - t_rb)
                st.rebuilds += 1
                st.max_resident = max(st.max_resident, len(self.resident))
            if not self.resident:
                if not pending and not waiting and not setting_up:
                    break
                nxt = pending[0].arrival * time_scale if pending else now + 1e-3
                time.sleep(max(0.0, min(nxt - now, 0.01)))
                continue
            # frames due now
            now = clock()
            active_sessions = [s for s in self.resident
                               if s.next < s.spec.n_frames and s.due(s.next) <= now]
            if not active_sessions:
                nxt = min(s.due(s.next) for s in self.resident if s.next < s.spec.n_frames) \
                    if any(s.next < s.spec.n_frames for s in self.resident) else now + 1e-3
                time.sleep(max(0.0, min(nxt - now, 0.005)))
                continue
            n = len(self.batch.areas)
            mask = np.zeros(n, dtype=np.uint8)
            ofs = np.full(n, -1, dtype=np.int64)
            chunks, pos, idx, due = [], 0, 0, []
            syn = 0
            active_set = set(id(s) for s in active_sessions)
            for s in self.resident:
                if id(s) in active_set:
                    for k, a in enumerate(s.areas):
                        fr = s.streams[k].next_frame()
                        if s.log is not None:
                            s.log.append((k, s.next))
                        mask[idx + k] = 1
                        ofs[idx + k] = pos
                        chunks.append(fr)
                        pos += fr.size
                    due.append(s.due(s.next))
                    s.next += 1
                    syn += s.synapses
                idx += len(s.areas)
            frames = np.concatenate(chunks) if chunks else np.zeros(0, np.uint8)
            t0 = clock()
            self.batch.cycle_masked(mask, frames, ofs, graph=self.use_graph)
            if self.regrow_every:
                self._regrow_due(active_set)
            if self.sample_outputs:
                # the frames' event follows the cycle on the batch stream, so
                # receiving them is the wait for the cycle
                self._deliver(active_sessions)
            else:
                self.batch.sync()
            t1 = clock()
            st.gpu_busy_s += t1 - t0
            for s in active_sessions:
                s.done += 1
            st.latencies.extend(t1 - d for d in due)
            st.cycles += 1
            act_sum += int(mask.sum())
            st.synapse_updates += syn
        st.wall_s = clock()
        st.mean_active_areas = act_sum / max(st.cycles, 1)
        # drain: sessions still being set up are closed once ready
        for f in setting_up:
            f.result()
        with self._ready_lock:
            self.resident.extend(self._ready)
            self._ready = []
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for s in self.resident:
            if self.on_retire and s.finished():
                self.on_retire(s)
        for s in self.resident:
            s.close()
        self.resident = []
        return st
