"""Flywheel: the request loop that cycles a cortex and serves samplers.

Host mirror of src/flywheel.rs:27-377.  One thread runs `Flywheel.spin()`; other
threads talk to it through queues: a command queue (`Command`: iterate n cycles,
stop, exit) and request/response queue pairs (`Request`: status, current
iteration, area info, a sampler of an area, finish queues).  Between cycles the
loop drains every request queue and answers on the paired response queue, as
`Flywheel::fulfill_requests` does.  A `Sampler` request returns a
`TractReceiver` (bismit_amd.sampler) whose frames the engine reads back after
each cycle on the cycle's own stream.

Sensory input: where the reference's thalamus pulls frames from input
generators inside `Cortex::cycle`, here `inputs` maps an area name to a
callable returning that area's next input frame.

The loop keeps the reference's control semantics, including that a command
other than `None`, `Stop` or `ExitAfterCycling` arriving while cycling ends the
current run (src/flywheel.rs:270-288).
"""
import queue
import time
from dataclasses import dataclass, field, replace
from typing import Any, Callable, Dict, List, Optional, Tuple


# ---------------------------------------------------------------- commands
@dataclass(frozen=True)
class Command:
    """Imperative cycle control (src/flywheel.rs:71-79)."""
    kind: str               # "None", "Iterate", "Stop", "Exit", "ExitAfterCycling"
    n: int = 0              # Iterate: number of cycles (0: until stopped)

    @staticmethod
    def none():
        return Command("None")

    @staticmethod
    def iterate(n):
        return Command("Iterate", int(n))

    @staticmethod
    def stop():
        return Command("Stop")

    @staticmethod
    def exit():
        return Command("Exit")

    @staticmethod
    def exit_after_cycling():
        return Command("ExitAfterCycling")


# ---------------------------------------------------------------- requests
@dataclass(frozen=True)
class Request:
    """Requests for data (src/flywheel.rs:82-90)."""
    kind: str               # "CurrentIter", "Status", "AreaInfo", "Sampler", "FinishQueues"
    area_name: str = ""
    sampler_kind: Any = None
    buffer_kind: Any = None
    backpressure: bool = False

    @staticmethod
    def current_iter():
        return Request("CurrentIter")

    @staticmethod
    def status():
        return Request("Status")

    @staticmethod
    def area_info():
        return Request("AreaInfo")

    @staticmethod
    def sampler(area_name, kind, buffer_kind=None, backpressure=False):
        return Request("Sampler", area_name, kind, buffer_kind, backpressure)

    @staticmethod
    def finish_queues():
        return Request("FinishQueues")


@dataclass
class Response:
    """Responses (src/flywheel.rs:93-106): kind is one of "CycleStarted",
    "CurrentIter", "Status", "Ready", "Motor", "AreaInfo", "SampleProgress",
    "QueuesFinished", "Sampler", "Exiting"; value carries the payload."""
    kind: str
    value: Any = None


@dataclass
class AreaInfo:
    """(src/flywheel.rs:63-68) name, output slice ids and dimensions."""
    name: str
    aff_out_slc_ids: List[int]
    dims: Tuple[int, int]
    n_slices: int
    layers: List[str]


# ---------------------------------------------------------------- status
@dataclass
class Status:
    """Cycle status (src/flywheel.rs:109-170); times in seconds."""
    cycling: bool = False
    cur_cycle: int = 0
    prev_cycles: int = 0
    prev_elapsed: float = 0.0
    cur_start_time: Optional[float] = field(default_factory=time.monotonic)
    cycle_counter: int = 0

    def cur_elapsed(self):
        return time.monotonic() - self.cur_start_time if self.cur_start_time is not None else 0.0

    def cur_cps(self):
        return self._cps(self.cur_cycle, self.cur_elapsed()) if self.cur_start_time else 0.0

    def ttl_cycles(self):
        return (self.cur_cycle + self.prev_cycles) & 0xFFFFFFFF

    def ttl_elapsed(self):
        return self.prev_elapsed + self.cur_elapsed()

    def ttl_cps(self):
        return self._cps(self.ttl_cycles(), self.ttl_elapsed())

    @staticmethod
    def _cps(cycles, elapsed):
        ms = int(elapsed * 1000)
        return cycles / ms * 1000.0 if ms > 0 else 0.0


# ---------------------------------------------------------------- loop
class Flywheel:
    """An event loop for a cortex (src/flywheel.rs:180-377).  `cortex` needs
    `areas` (name -> area with `sampler(...)`, `scheme`), `cycle(inputs)` and
    `finish_queues()` -- bismit_amd.cortex.Cortex has them."""

    def __init__(self, cortex, command_rx: "queue.Queue[Command]", area_name: str,
                 inputs: Optional[Dict[str, Callable[[], Any]]] = None):
        self.command_rx = command_rx
        self.req_res_pairs: List[Tuple[queue.Queue, queue.Queue]] = []
        self.motor_txs: List[queue.Queue] = []
        self.cortex = cortex
        self.cycle_iters_max = 1
        self.area_name = area_name
        self.inputs = dict(inputs or {})
        self.status = Status()
        self.exiting = False

    def add_req_res_pair(self, req_rx: queue.Queue, res_tx: queue.Queue):
        self.req_res_pairs.append((req_rx, res_tx))

    def add_motor_tx(self, motor_tx: queue.Queue):
        self.motor_txs.append(motor_tx)

    def set_input(self, area_name, source: Callable[[], Any]):
        self.inputs[area_name] = source

    def spin(self):
        while not self.exiting:
            self.fulfill_requests()
            cmd = self.command_rx.get()
            if cmd.kind == "Iterate":
                self.cycle_iters_max = cmd.n
            elif cmd.kind == "Exit":
                self.exiting = True
                break
            else:
                continue
            self.status.cur_cycle = 0
            self.status.cur_start_time = time.monotonic()
            self.status.cycling = True
            self.broadcast_status()
            if self.cycle_loop().kind == "Exit":
                self.exiting = True
                break
            self.status.cycling = False
            self.status.prev_cycles += self.status.cur_cycle
            self.status.prev_elapsed += self.status.cur_elapsed()
            self.status.cur_cycle = 0
            self.status.cur_start_time = None
            self.broadcast_status()
        for _, res_tx in self.req_res_pairs:
            res_tx.put(Response("Exiting"))

    def cycle_loop(self) -> Command:
        exit_after_cycling = False
        while True:
            if self.cycle_iters_max != 0 and self.status.cur_cycle >= self.cycle_iters_max:
                break
            frames = {name: src() for name, src in self.inputs.items()}
            self.cortex.cycle(frames)
            self.status.cur_cycle += 1
            self.status.cycle_counter += 1
            try:
                c = self.command_rx.get_nowait()
            except queue.Empty:
                c = None
            if c is not None:
                if c.kind == "Stop":
                    return c
                if c.kind == "ExitAfterCycling":
                    exit_after_cycling = True
                elif c.kind != "None":
                    return c
            self.fulfill_requests()
        return Command.exit() if exit_after_cycling else Command.none()

    def fulfill_requests(self):
        for req_rx, res_tx in self.req_res_pairs:
            while True:
                try:
                    r = req_rx.get_nowait()
                except queue.Empty:
                    break
                if r.kind == "Sampler":
                    area = self.cortex.areas[r.area_name]
                    res_tx.put(Response("Sampler", area.sampler(r.sampler_kind, r.buffer_kind,
                                                                r.backpressure)))
                elif r.kind == "AreaInfo":
                    res_tx.put(Response("AreaInfo", self.area_info()))
                elif r.kind == "Status":
                    res_tx.put(Response("Status", replace(self.status)))
                elif r.kind == "CurrentIter":
                    res_tx.put(Response("CurrentIter", self.status.cur_cycle))
                elif r.kind == "FinishQueues":
                    self.cortex.finish_queues()
                    res_tx.put(Response("QueuesFinished", self.status.cycle_counter))
                else:
                    raise ValueError(f"unknown request {r.kind}")

    def area_info(self) -> AreaInfo:
        area = self.cortex.areas[self.area_name]
        sc = area.scheme
        out = sorted({lay.base_slc + k for lay in sc.layers if not lay.spatial
                      for k in range(lay.depth)})
        return AreaInfo(sc.name, out, (sc.v_size, sc.u_size), len(sc.slices),
                        [lay.name for lay in sc.layers])

    def broadcast_status(self):
        for _, res_tx in self.req_res_pairs:
            res_tx.put(Response("Status", replace(self.status)))
