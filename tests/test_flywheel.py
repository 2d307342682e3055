"""Request loop (flywheel) control logic on the CPU, with a stand-in cortex
that records what the loop asks of it (src/flywheel.rs:235-377 semantics)."""
import queue
import threading
import time

import pytest

from bismit_amd import recipes
from bismit_amd.flywheel import Command, Flywheel, Request
from bismit_amd.sampler import SamplerKind


class FakeArea:
    def __init__(self, scheme):
        self.scheme = scheme
        self.samplers = []

    def sampler(self, kind, buffer_kind, backpressure):
        self.samplers.append((kind, buffer_kind, backpressure))
        return ("rx", kind)


class FakeCortex:
    def __init__(self):
        self.areas = {"v1": FakeArea(recipes.config1_area())}
        self.cycles = []
        self.finished = 0

    def cycle(self, inputs):
        self.cycles.append(dict(inputs))

    def finish_queues(self):
        self.finished += 1


def start(fw):
    t = threading.Thread(target=fw.spin, daemon=True)
    t.start()
    return t


def wait_until(pred, timeout=10.0):
    t_end = time.monotonic() + timeout
    while not pred():
        assert time.monotonic() < t_end, "timed out"
        time.sleep(0.001)


def drain(q, until, timeout=10.0):
    out = []
    while True:
        r = q.get(timeout=timeout)
        out.append(r)
        if r.kind == until:
            return out


def test_iterate_then_exit_with_inputs_and_status():
    cx = FakeCortex()
    cmd, req, res = queue.Queue(), queue.Queue(), queue.Queue()
    n = iter(range(1000))
    fw = Flywheel(cx, cmd, "v1", inputs={"v1": lambda: next(n)})
    fw.add_req_res_pair(req, res)
    t = start(fw)
    cmd.put(Command.iterate(5))
    out = []
    while not out or out[-1].kind != "Status" or out[-1].value.cycling:
        out.append(res.get(timeout=10))      # until the end-of-run status
    cmd.put(Command.exit())
    t.join(10)
    assert not t.is_alive()
    assert [c["v1"] for c in cx.cycles] == [0, 1, 2, 3, 4]
    out += drain(res, "Exiting")
    st = [r.value for r in out if r.kind == "Status"]
    assert st[0].cycling and st[0].cur_cycle == 0
    assert not st[-1].cycling and st[-1].prev_cycles == 5 and st[-1].cycle_counter == 5
    assert st[-1].ttl_cycles() == 5


def test_requests_between_cycles():
    cx = FakeCortex()
    cmd, req, res = queue.Queue(), queue.Queue(), queue.Queue()
    fw = Flywheel(cx, cmd, "v1")
    fw.add_req_res_pair(req, res)
    kind = SamplerKind.Axons(None)
    req.put(Request.status())
    req.put(Request.current_iter())
    req.put(Request.area_info())
    req.put(Request.sampler("v1", kind, None, True))
    t = start(fw)
    cmd.put(Command.iterate(3))
    req.put(Request.finish_queues())
    wait_until(lambda: fw.status.prev_cycles >= 3)
    cmd.put(Command.exit())
    t.join(10)
    assert not t.is_alive()
    out = drain(res, "Exiting")
    kinds = [r.kind for r in out]
    assert kinds[:4] == ["Status", "CurrentIter", "AreaInfo", "Sampler"]
    info = out[2].value
    sc = cx.areas["v1"].scheme
    assert info.name == sc.name and info.dims == (sc.v_size, sc.u_size)
    pyr = [l for l in sc.layers if not l.spatial]
    assert info.aff_out_slc_ids == sorted(l.base_slc + k for l in pyr for k in range(l.depth))
    assert out[3].value == ("rx", kind)
    assert cx.areas["v1"].samplers == [(kind, None, True)]
    assert "QueuesFinished" in kinds and cx.finished == 1
    assert len(cx.cycles) == 3


def test_stop_ends_an_unbounded_run_and_exit_after_cycling():
    cx = FakeCortex()
    cmd = queue.Queue()
    fw = Flywheel(cx, cmd, "v1")
    t = start(fw)
    cmd.put(Command.iterate(0))          # cycle until stopped
    wait_until(lambda: len(cx.cycles) >= 20)
    cmd.put(Command.stop())
    wait_until(lambda: not fw.status.cycling and fw.status.prev_cycles > 0)
    stopped_at = len(cx.cycles)
    assert stopped_at >= 20 and fw.status.prev_cycles == stopped_at
    # a bounded run that is told to exit when it completes
    cmd.put(Command.iterate(7))
    cmd.put(Command.exit_after_cycling())
    t.join(10)
    assert not t.is_alive()
    assert len(cx.cycles) == stopped_at + 7


def test_exit_during_a_run_ends_it():
    """A command other than None/Stop/ExitAfterCycling that arrives while
    cycling ends the run at once (src/flywheel.rs:270-288)."""
    cx = FakeCortex()
    cmd = queue.Queue()
    fw = Flywheel(cx, cmd, "v1")
    cmd.put(Command.iterate(1000000))
    cmd.put(Command.exit())
    t = start(fw)
    t.join(10)
    assert not t.is_alive()
    assert len(cx.cycles) == 1


def test_idle_commands_do_not_cycle():
    cx = FakeCortex()
    cmd = queue.Queue()
    fw = Flywheel(cx, cmd, "v1")
    t = start(fw)
    cmd.put(Command.none())
    cmd.put(Command.stop())
    cmd.put(Command.exit())
    t.join(10)
    assert not t.is_alive()
    assert cx.cycles == []


def test_sampler_kind_variants():
    assert SamplerKind.Axons().buf == "AXN_STATES"
    assert SamplerKind.SynFlagSets(1).buf == "SYN_FLAG_SETS"
    assert SamplerKind.TuftPrevBestDenStatesRaw("iv").layer == "iv"
    with pytest.raises(ValueError):
        SamplerKind.DenStates()
