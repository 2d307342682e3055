"""Serving loop: continuous batching of sessions' sensory frames into shared
cycles on one GPU.

Each session owns a cortex of 1-3 areas resident in HBM for its lifetime.  Its
client streams 2%-sparse SDR frames at the session's rate from the moment the
session arrives: frame k exists at arrival + k / rate.  The server
  - queues arriving sessions and admits them in arrival order while their HBM
    footprint fits the budget, their compute demand (synapses x frame rate)
    fits `target_util` of the GPU's synapse-update capacity, and the resident
    sessions keep up with their clients (90th percentile lag below `max_lag_s`);
    a session still waiting after `max_wait_s` is rejected (its client gives up),
  - sets admitted sessions up on worker threads while the loop keeps cycling,
    and adds them to / removes finished ones from one dynamic batch between
    cycles (bsm_batch_add_areas / bsm_batch_remove_areas: no batch rebuild),
  - at every shared cycle steps exactly the sessions with a frame produced and
    not yet stepped (one per session; a session behind its client steps every
    cycle until it has caught up), sending the frames as SDR entries (active
    indices, scattered on the device), one persistent launch per stage with
    idle areas masked on the device,
  - generates the next cycle's frames while the GPU runs the current one, then
    receives every stepped area's output axons through its sampler (one gather
    kernel and one copy per cycle for all sessions) and hands them to the session,
  - regrows a session's weak synapses after every `regrow_every` of its cycles
    (the reference's CorticalArea::cycle regrows every 201 cycles,
    src/cortex/cortical_area.rs:773-776,792-806), all due sessions in one call.
Frame latency is measured from the frame's production at the client to its
output frame readable on the host, so it includes the admission wait, the
set-up and any catch-up.  Each area keeps its own rnd stream, state parity and
input, so its result is the same as a plain replay of the session alone (the
GPU tests check it against the oracle).

The reference has no serving mode; its single-threaded flywheel event loop
(src/flywheel.rs:235-377) is the nearest analogue.
"""
import ctypes
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib as L
from . import recipes
from .cortex import Batch, CorticalArea, footprint
from .inputs import SdrStream, next_packed
from .sampler import Frame, ReceiverSet, SamplerBufferKind, SamplerKind

NO_FRAME = np.uint64(2**64 - 1)


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
    def __init__(self, spec: SessionSpec, ctx, t_arrival, sample_outputs=True):
        self.spec = spec
        self.t_arr = t_arrival      # on the server clock
        self.t_join = None          # joined the batch
        self.schemes = [area_scheme(s, f"s{spec.sid}a{i}") for i, s in enumerate(spec.sides)]
        self.areas = []
        for i, sc in enumerate(self.schemes):
            self.areas.append(CorticalArea(sc, ctx, rnd_seed=spec.seed + 7919 * i,
                                           layout_seed=spec.seed + 104729 * i, reach=6))
        self.streams = [SdrStream(sc.input_range()[1], 0.02, 0.5, seed=spec.seed + i)
                        for i, sc in enumerate(self.schemes)]
        self.entry_bytes = [4 * (st.k + 1) for st in self.streams]
        self.next = 0                      # next frame to step
        self.done = 0                      # frames completed
        self.synapses = sum(sc.n_synapses() for sc in self.schemes)
        # per-cycle output readback (no backpressure: a slow consumer skips frames)
        self.outputs = [a.sampler(SamplerKind.Axons(output_layer(sc)), SamplerBufferKind.DOUBLE)
                        for a, sc in zip(self.areas, self.schemes)] if sample_outputs else []
        self.last_outputs = [None] * len(self.areas)
        self.pos = np.zeros(0, np.int64)   # the areas' positions in the batch
        self.rx = np.zeros(0, np.int64)    # their samplers' positions in the receiver set
        self.fp = 0
        self.demand = 0.0

    def produced(self, k):
        """When the client produces frame k."""
        return self.t_arr + k / self.spec.rate_hz

    def regrow_seed(self, k):
        """Regrowth seed of area k (its epochs count the session's regrowths)."""
        return self.spec.seed + 15485863 * k + 1

    def finished(self):
        return self.done >= self.spec.n_frames

    def close(self):
        for st in self.streams:
            st.close()
        for rx in self.outputs:
            rx.close()
        for a in self.areas:
            a.close()


@dataclass
class ServeStats:
    latencies: list = field(default_factory=list)      # frame production -> output on host
    service: list = field(default_factory=list)        # max(production, join) -> output
    admit_waits: list = field(default_factory=list)
    cycles: int = 0
    synapse_updates: float = 0.0
    gpu_busy_s: float = 0.0
    wall_s: float = 0.0
    arrived: int = 0
    admitted: int = 0
    rejected: int = 0
    completed: int = 0
    max_resident: int = 0
    max_queue: int = 0
    queue_area: float = 0.0       # integral of the queue length over time
    resident_area: float = 0.0
    active_areas: int = 0
    membership_changes: int = 0
    membership_s: list = field(default_factory=list)
    setup_s: list = field(default_factory=list)
    config: dict = field(default_factory=dict)
    host_s: dict = field(default_factory=dict)    # loop phase -> seconds

    def phase(self, name, dt):
        self.host_s[name] = self.host_s.get(name, 0.0) + dt

    def summary(self):
        def pct(x, q):
            return float(np.percentile(np.array(x), q) * 1e3) if x else 0.0
        w = max(self.wall_s, 1e-9)
        out = {
            "frames": len(self.latencies),
            "p50_ms": pct(self.latencies, 50),
            "p99_ms": pct(self.latencies, 99),
            "latency_from": "frame produced at the client (arrival + k / rate) to its output "
                            "frame readable on the host: admission wait, set-up and catch-up "
                            "included",
            "service_p50_ms": pct(self.service, 50),
            "service_p99_ms": pct(self.service, 99),
            "admission_wait_p50_ms": pct(self.admit_waits, 50),
            "admission_wait_p99_ms": pct(self.admit_waits, 99),
            "synapse_updates_per_s": self.synapse_updates / w,
            "gpu_busy_frac": self.gpu_busy_s / w,
            "shared_cycles": self.cycles,
            "mean_active_areas_per_cycle": self.active_areas / max(self.cycles, 1),
            "sessions_arrived": self.arrived,
            "sessions_admitted": self.admitted,
            "sessions_rejected": self.rejected,
            "sessions_completed": self.completed,
            "queue_len_mean": self.queue_area / w,
            "queue_len_max": self.max_queue,
            "resident_sessions_mean": self.resident_area / w,
            "max_resident_sessions": self.max_resident,
            "membership_changes": self.membership_changes,
            "membership_ms_mean": float(np.mean(self.membership_s) * 1e3) if self.membership_s else 0.0,
            "membership_ms_max": max(self.membership_s, default=0.0) * 1e3,
            "session_setup_ms_mean": float(np.mean(self.setup_s) * 1e3) if self.setup_s else 0.0,
            "wall_s": self.wall_s,
        }
        out["host_ms_per_cycle"] = {k: 1e3 * v / max(self.cycles, 1)
                                    for k, v in sorted(self.host_s.items())}
        out.update(self.config)
        return out


class _Cycle:
    """A shared cycle prepared or in flight."""
    __slots__ = ("sessions", "produced", "ofs", "total", "syn", "n_areas")


class Server:
    def __init__(self, ctx, hbm_budget_bytes=None, max_sessions=None, epoch_s=0.05,
                 use_graph=False, on_retire=None, setup_workers=4, regrow_every=201,
                 regrow_reach=6, sample_outputs=True, on_output=None, max_wait_s=0.25,
                 capacity=1.0e12, target_util=0.8, max_lag_s=0.1, initial_capacity=None,
                 cycle_gap_s=0.002, membership_s=0.1):
        self.ctx = ctx
        # Masked cycles are enqueued directly by default: the batch changes with
        # most admission epochs, and capturing plus instantiating its two graphs
        # costs more latency than the launches per cycle a graph saves.
        self.use_graph = use_graph
        self.sample_outputs = sample_outputs
        self.on_output = on_output            # on_output(session, [Frame per area]) after the
                                              # cycle is counted (frame.cycle == session.done);
                                              # the data stays valid until the session's next cycle
        self.on_retire = on_retire
        self.regrow_every = regrow_every      # 0: regrowth off
        self.regrow_reach = regrow_reach
        if hbm_budget_bytes is None:
            hbm_budget_bytes = int(0.75 * ctx.info()["mem_bytes"])
        self.budget = hbm_budget_bytes
        self.max_sessions = max_sessions      # None: HBM and compute decide
        self.epoch_s = epoch_s
        # shared cycles start at least this far apart: frames gather into larger
        # cycles (each costs the host a fixed ~0.5-1 ms to build and launch), at
        # a latency cost of at most the gap
        self.cycle_gap_s = cycle_gap_s
        # joins and retirements are applied together, at most every membership_s
        # (each change re-plans the batch, ~7 ms for 300 resident areas); a
        # finished session waits there unstepped, a set-up one starts late and
        # catches up
        self.membership_s = membership_s
        self.max_wait_s = max_wait_s          # None: sessions wait as long as it takes
        self.capacity = capacity              # synapse-updates/s the GPU sustains at most
        self.initial_capacity = initial_capacity or capacity
        self.capacity_est = self.initial_capacity   # moved by _observe
        self.window_s = 0.5
        self.ramp = 0.1                       # new demand admitted per epoch, of the estimate
        self._win = (0.0, 0.0)
        self._idle_s = 0.0
        self.target_util = target_util
        self.max_lag_s = max_lag_s
        self.setup_workers = setup_workers
        self.resident = []
        self.used = 0
        self.demand = 0.0
        self.batch = None
        self.recvset = ReceiverSet([])
        self._ready = []
        self._ready_lock = threading.Lock()
        self._per_side = {}
        self._buf = np.empty(1 << 20, np.uint8)

    # ------------------------------------------------------------ admission
    def _side(self, side):
        """(footprint bytes, synapses) of one session area of this side."""
        if side not in self._per_side:
            sc = area_scheme(side, "f")
            self._per_side[side] = (footprint(sc), sc.n_synapses())
        return self._per_side[side]

    def footprint(self, spec):
        return sum(self._side(s)[0] for s in spec.sides)

    def session_demand(self, spec):
        return sum(self._side(s)[1] for s in spec.sides) * spec.rate_hz

    def _observe(self, st, now):
        """Capacity estimate, every `window_s`: when resident sessions fall
        behind their clients, what the server delivered in that window is all it
        sustains, and the estimate drops to it; while they keep up with demand
        close to the admission limit, the estimate rises 2% (up to `capacity`)."""
        w0, u0 = self._win
        if now - w0 < self.window_s:
            return
        rate = (st.synapse_updates - u0) / (now - w0)
        if not self._keeping_up(now):
            self.capacity_est = min(self.capacity_est, rate)
        elif self.demand >= 0.95 * self.target_util * self.capacity_est:
            self.capacity_est = min(self.capacity, 1.02 * self.capacity_est)
        self._win = (now, st.synapse_updates)

    def _keeping_up(self, now):
        """Resident sessions (joined over half a second ago) keep up with their
        clients: 90th percentile of how long their next frame has waited."""
        lags = [now - s.produced(s.next) for s in self.resident
                if s.next < s.spec.n_frames and s.t_join is not None and now - s.t_join > 0.25]
        return not lags or float(np.percentile(lags, 90)) <= self.max_lag_s

    def _admit(self, waiting, setting_up, pool, st, now):
        if not self._keeping_up(now):
            return
        budget = self.ramp * self.capacity_est   # demand this epoch may add
        while waiting:
            spec, t_arr = waiting[0]
            if self.max_sessions is not None and \
                    len(self.resident) + len(setting_up) >= self.max_sessions:
                break
            fp = self.footprint(spec)
            d = self.session_demand(spec)
            if self.used + fp > self.budget:
                break
            if self.resident and self.demand + d > self.target_util * self.capacity_est:
                break
            if d > budget and budget < self.ramp * self.capacity_est:
                break
            budget -= d
            waiting.popleft()
            self.used += fp
            self.demand += d
            st.admit_waits.append(now - t_arr)
            setting_up.append(pool.submit(self._setup, spec, t_arr, fp, d, st))
            st.admitted += 1

    def _setup(self, spec, t_arr, fp, d, st):
        t0 = time.perf_counter()
        sess = Session(spec, self.ctx, t_arr, self.sample_outputs)
        sess.fp = fp
        sess.demand = d
        st.setup_s.append(time.perf_counter() - t0)
        with self._ready_lock:
            self._ready.append(sess)

    # ------------------------------------------------------------ membership
    def _membership(self, setting_up, st, now):
        """Between cycles: finished sessions leave the batch, set-up ones join."""
        for f in [f for f in setting_up if f.done()]:
            f.result()           # re-raise a set-up error here
            setting_up.remove(f)
        retired = [s for s in self.resident if s.finished()]
        with self._ready_lock:
            ready, self._ready = self._ready, []
        if not retired and not ready:
            return
        t0 = time.perf_counter()
        if retired:
            for s in retired:
                s.last_outputs = [Frame(self.recvset.outs[i].copy(), s.done) for i in s.rx]
                if self.on_retire:
                    self.on_retire(s)
            self.batch.remove([a for s in retired for a in s.areas])
            gone = set(id(s) for s in retired)
            self.resident = [s for s in self.resident if id(s) not in gone]
            for s in retired:
                self.used -= s.fp
                self.demand -= s.demand
                st.completed += 1
            # freeing device memory waits for the device: off the loop thread
            self._pool.submit(lambda ss=retired: [x.close() for x in ss])
        if ready:
            self.batch.add([a for s in ready for a in s.areas])
            for s in ready:
                s.t_join = now
            self.resident.extend(ready)
        pos = {id(a): i for i, a in enumerate(self.batch.areas)}
        rxs = []
        for s in self.resident:
            s.pos = np.array([pos[id(a)] for a in s.areas], np.int64)
            s.rx = np.arange(len(rxs), len(rxs) + len(s.outputs), dtype=np.int64)
            rxs.extend(s.outputs)
        self.recvset = ReceiverSet(rxs, reuse=self.recvset)
        st.membership_s.append(time.perf_counter() - t0)
        st.membership_changes += 1
        st.max_resident = max(st.max_resident, len(self.resident))

    # ------------------------------------------------------------ cycles
    def _prepare(self, now):
        """The next shared cycle: every resident session with a frame produced
        and not yet stepped; its frames generated into the staging buffer."""
        act = [s for s in self.resident if s.next < s.spec.n_frames and s.produced(s.next) <= now]
        if not act:
            return None
        sizes = [b for s in act for b in s.entry_bytes]
        ofs = np.zeros(len(sizes), np.uint64)
        if len(sizes) > 1:
            ofs[1:] = np.cumsum(sizes[:-1], dtype=np.uint64)
        total = int(ofs[-1]) + sizes[-1]
        if self._buf.size < total:
            self._buf = np.empty(max(total, 2 * self._buf.size), np.uint8)
        next_packed([st for s in act for st in s.streams], self._buf, ofs, sdr=True)
        c = _Cycle()
        c.sessions = act
        c.produced = [s.produced(s.next) for s in act]
        c.ofs = ofs
        c.total = total
        c.syn = sum(s.synapses for s in act)
        c.n_areas = len(sizes)
        for s in act:
            s.next += 1
        return c

    def _enqueue(self, c):
        """Enqueues the cycle, then the regrowth it makes due: a session regrows
        after every `regrow_every` of its cycles, so the regrowth follows this
        cycle on the batch stream, ahead of any later cycle."""
        n = len(self.batch.areas)
        mask = np.zeros(n, np.uint8)
        fo = np.full(n, NO_FRAME, np.uint64)
        pos = np.concatenate([s.pos for s in c.sessions])
        mask[pos] = 1
        fo[pos] = c.ofs | np.uint64(L.FRAME_SDR)
        self.batch.cycle_masked(mask, self._buf[:c.total], fo, graph=self.use_graph)
        if self.regrow_every:
            due = [s for s in c.sessions if s.next % self.regrow_every == 0]
            if due:
                which = np.zeros(n, np.uint8)
                seeds = [0] * n
                epochs = [0] * n
                for s in due:
                    for k, p in enumerate(s.pos):
                        which[p] = 1
                        seeds[p] = s.regrow_seed(k)
                        epochs[p] = s.next // self.regrow_every - 1
                self.batch.regrow(seeds, epochs, self.regrow_reach, which)

    def _receive(self, c, clock):
        """Waits for the cycle and receives its sessions' output frames (the
        oldest each sampler holds; a later cycle may already be running), and
        counts the cycle for its sessions."""
        rs = self.recvset
        t0 = clock()
        frames = None
        if self.sample_outputs:
            idx = np.concatenate([s.rx for s in c.sessions])
            got = rs.recv(wait_for_frame=True, idx=idx)
            t = clock()
            self._stats.phase("receive_frames", t - t0)
            for s in c.sessions:
                if not got[s.rx].all() or (rs.cycles[s.rx] != s.done + 1).any():
                    raise RuntimeError(f"session {s.spec.sid}: output frame missing")
            if self.on_output:
                frames = [[Frame(rs.outs[i], int(rs.cycles[i])) for i in s.rx]
                          for s in c.sessions]
        else:
            self.batch.sync()
            t = clock()
            self._stats.phase("receive_frames", t - t0)
        for s in c.sessions:
            s.done += 1
        return frames, t

    def _membership_due(self, setting_up):
        """A session's set-up ended (it waits to join, or raised), or a resident
        session has received its last frame."""
        with self._ready_lock:
            if self._ready:
                return True
        return any(f.done() for f in setting_up) or any(s.finished() for s in self.resident)

    def _finish(self, c, clock, st):
        """Receives a cycle in flight and accounts it."""
        t0 = clock()
        frames, t = self._receive(c, clock)
        st.phase("receive", clock() - t0)
        t1 = clock()
        self._account(c, frames, t, st)
        self._observe(st, clock())
        st.phase("account", clock() - t1)

    def _account(self, c, frames, t, st):
        """Hands the received frames out and records the cycle's latencies (the
        next cycle is already running)."""
        if frames is not None:
            for s, f in zip(c.sessions, frames):
                self.on_output(s, f)
        prod = np.array(c.produced)
        st.latencies.extend((t - prod).tolist())
        join = np.array([s.t_join for s in c.sessions])
        st.service.extend((t - np.maximum(prod, join)).tolist())
        st.cycles += 1
        st.synapse_updates += c.syn
        st.active_areas += c.n_areas

    # ------------------------------------------------------------ loop
    def run(self, specs, wall_limit_s=30.0, time_scale=1.0):
        """Serve `specs` (arrival times in seconds, scaled by time_scale) until
        every session is done or rejected, or wall_limit_s passed."""
        with ThreadPoolExecutor(max_workers=self.setup_workers) as pool:
            self._pool = pool
            return self._run(specs, wall_limit_s, time_scale, pool)

    def _run(self, specs, wall_limit_s, time_scale, pool):
        st = ServeStats()
        self._stats = st
        self.capacity_est = self.initial_capacity
        self._win = (0.0, 0.0)
        self._idle_s = 0.0
        st.config = {"hbm_budget_gb": self.budget / 2**30, "capacity_max": self.capacity,
                     "capacity_initial": self.initial_capacity,
                     "target_util": self.target_util, "max_wait_s": self.max_wait_s,
                     "max_lag_ms": self.max_lag_s * 1e3, "max_sessions": self.max_sessions}
        self.batch = Batch([], self.ctx, dynamic=True)
        pending = deque(sorted(specs, key=lambda s: s.arrival))
        waiting = deque()
        setting_up = []
        t_start = time.perf_counter()
        clock = lambda: time.perf_counter() - t_start
        last_epoch = -1e9
        t_last = 0.0
        inflight = deque()
        t_enq = -1.0
        t_memb = -1e9
        while True:
            now = clock()
            st.queue_area += len(waiting) * (now - t_last)
            st.resident_area += len(self.resident) * (now - t_last)
            t_last = now
            if now > wall_limit_s:
                break
            while pending and pending[0].arrival * time_scale <= now:
                spec = pending.popleft()
                waiting.append((spec, spec.arrival * time_scale))
                st.arrived += 1
            if self.max_wait_s is not None:
                while waiting and now - waiting[0][1] > self.max_wait_s:
                    waiting.popleft()
                    st.rejected += 1
            st.max_queue = max(st.max_queue, len(waiting))
            if now - last_epoch >= self.epoch_s:
                last_epoch = now
                self._admit(waiting, setting_up, pool, st, now)
            # The next cycle's frames are made while the GPU runs the current one;
            # the next cycle is enqueued as soon as the current one ends, before
            # the host receives the current one's frames (the output samplers hold
            # two frames), so the GPU waits only for the enqueue.  A membership
            # change (a set-up session joins, a finished one leaves) first drains
            # what is in flight.
            t0 = clock()
            if (self._membership_due(setting_up) and
                    (now - t_memb >= self.membership_s or not self.resident)) or \
                    not self.sample_outputs:
                while inflight:
                    self._finish(inflight.popleft(), clock, st)
                self._membership(setting_up, st, clock())
                t_memb = now
            t1 = clock()
            st.phase("membership", t1 - t0)
            if inflight and t1 - t_enq < self.cycle_gap_s:
                time.sleep(self.cycle_gap_s - (t1 - t_enq))
                now = t1 = clock()
            nxt = self._prepare(now)
            t2 = clock()
            st.phase("prepare", t2 - t1)
            if inflight:
                while not self.batch.query():      # the current cycle's end, polled
                    time.sleep(0)
            t3 = clock()
            st.phase("gpu_wait", t3 - t2)
            if nxt is not None:
                t_enq = clock()
                self._enqueue(nxt)
            t4 = clock()
            st.phase("enqueue", t4 - t3)
            if inflight:
                self._finish(inflight.popleft(), clock, st)
            if nxt is not None:
                inflight.append(nxt)
            if inflight:
                continue
            if not self.resident and not pending and not waiting and not setting_up:
                with self._ready_lock:
                    if not self._ready:
                        break
            # idle: sleep until the next frame or arrival is due (bounded)
            wake = [s.produced(s.next) for s in self.resident if s.next < s.spec.n_frames]
            if pending:
                wake.append(pending[0].arrival * time_scale)
            nxt_t = min(wake) if wake else now + 1e-3
            t_sl = clock()
            time.sleep(max(0.0, min(nxt_t - t_sl, 0.002)))
            self._idle_s += clock() - t_sl
        while inflight:
            self._finish(inflight.popleft(), clock, st)
        st.wall_s = clock()
        st.config["capacity_estimate_end"] = self.capacity_est
        st.config["loop_idle_frac"] = self._idle_s / max(st.wall_s, 1e-9)
        busy = ctypes.c_double()
        L.check(self.ctx.lib.bsm_batch_busy_ms(self.batch.h, ctypes.byref(busy)))
        st.gpu_busy_s = busy.value / 1000.0
        # drain: sessions still being set up are closed once ready
        for f in setting_up:
            f.result()
        with self._ready_lock:
            late, self._ready = self._ready, []
        for s in self.resident:
            if s.finished():
                s.last_outputs = [Frame(self.recvset.outs[i].copy(), s.done) for i in s.rx]
                if self.on_retire:
                    self.on_retire(s)
        self.batch.close()
        self.batch = None
        for s in self.resident + late:
            s.close()
        self.resident = []
        self.recvset = ReceiverSet([])
        return st

