"""The team kernels read their state descriptor from a 256-slot device ring
filled by stream-ordered copies from pinned host memory; an event recorded
after each launch guards a slot's reuse.  More launches than slots, queued
without synchronizing and alternating between two streams, must each see
their own descriptor: every result equals the same state stepped alone.
The one-lane kernel's steps share the context's hand-over list and status
scratch; on two streams they must still not interleave on the device.
Each launch's work cursors belong to its descriptor slot."""
import numpy as np
import pytest

from pfchem import abi, workloads
from pfchem.lib import Context, DeviceState

pytestmark = pytest.mark.gpu


def test_more_launches_than_descriptor_slots():
    import torch
    wl = workloads.calcite_workload(ncell=96)
    ctx = Context(wl.net, wl.opt)
    assert ctx.try_set_launch(abi.KERNEL_BATCHED_FP64, 4, 128, 4, 1)
    n = wl.state.ncell
    rng = np.random.default_rng(5)
    hosts = []
    for _ in range(300):
        h = wl.state.copy()
        h["total"][:, :n] *= 1.0 + 0.05 * rng.random((1, n))
        hosts.append(h)
    print("case ring: 300 launches on two streams", flush=True)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    devs = [DeviceState(h) for h in hosts]
    torch.cuda.synchronize()
    for i, d in enumerate(devs):
        ctx.step(d, wl.dt, abi.KERNEL_BATCHED_FP64, stream=streams[i % 2])
    torch.cuda.synchronize()
    for i in list(range(0, 300, 13)) + [299]:
        alone = DeviceState(hosts[i])
        ctx.step(alone, wl.dt, abi.KERNEL_BATCHED_FP64)
        torch.cuda.synchronize()
        a, b = alone.to_host(), devs[i].to_host()
        assert np.all(b["status"][0, :n] == 0), i
        for f in ("pri_molal", "total", "mnrl_volfrac"):
            np.testing.assert_array_equal(b[f][:, :n], a[f][:, :n], err_msg="launch %d %s" % (i, f))


def test_one_lane_steps_on_two_streams_share_the_hand_over_list():
    """mixed one-lane steps alternate between two streams, half of them with
    no status array (the context's status scratch): the FP32 pass, the
    hand-over list and the FP64 passes of one step must not interleave with
    another step's on the device; each launch's answer equals the same state
    stepped alone"""
    import torch
    wl = workloads.random_network_workload(ncell=512, spread=0.5, wmax=1.0)
    ctx = Context(wl.net, wl.opt)
    assert ctx.get_launch(abi.KERNEL_BATCHED_MIXED)["lanes_per_cell"] == 1
    n = wl.state.ncell
    rng = np.random.default_rng(11)
    hosts = []
    for _ in range(40):
        h = wl.state.copy()
        h["total"][:, :n] *= 1.0 + 0.05 * rng.random((1, n))
        hosts.append(h)
    print("case one-lane: 40 mixed launches on two streams", flush=True)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    devs = [DeviceState(h) for h in hosts]
    for i in range(1, 40, 2):
        devs[i].t.pop("status", None)  # these use the context's status scratch
    torch.cuda.synchronize()
    for i, d in enumerate(devs):
        ctx.step(d, wl.dt, abi.KERNEL_BATCHED_MIXED, stream=streams[i % 2])
    torch.cuda.synchronize()
    for i in range(40):
        alone = DeviceState(hosts[i])
        ctx.step(alone, wl.dt, abi.KERNEL_BATCHED_MIXED)
        torch.cuda.synchronize()
        a, b = alone.to_host(), devs[i].to_host()
        assert np.all(a["status"][0, :n] == 0), i
        for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
            np.testing.assert_array_equal(b[f][:, :n], a[f][:, :n], err_msg="launch %d %s" % (i, f))
