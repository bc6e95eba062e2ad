"""The launch-shape tuner (pfchem_autotune) on the GPU: it leaves the state
as given, makes its choice the context's shape, records it in the JSON
cache that a new context applies through PFCHEM_TUNE_CACHE, and the tuned
step keeps parity with the FP64 oracle (relative concentration error <=
1e-8).  Covers a fixed network (one-lane kernel and team shapes) and a
network with per-cell active counts (one team per size bucket).  No
reference counterpart: the reference runs one cell at a time on the CPU."""
import json
import os

import numpy as np
import pytest

from pfchem import abi, workloads
from pfchem.lib import Context, DeviceState

import gpu_util as G
import variable_util as V

pytestmark = pytest.mark.gpu

CONC_TOL = 1e-8


def _entries(path):
    with open(path) as f:
        return json.load(f)["entries"]


def _tuned_context_from_env(wl, cache):
    old = os.environ.get("PFCHEM_TUNE_CACHE")
    os.environ["PFCHEM_TUNE_CACHE"] = cache
    try:
        return Context(wl.net, wl.opt)
    finally:
        if old is None:
            del os.environ["PFCHEM_TUNE_CACHE"]
        else:
            os.environ["PFCHEM_TUNE_CACHE"] = old


def test_autotune_fixed_network(tmp_path):
    import torch
    wl = workloads.random_network_workload(ncell=65536)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    cache, table = str(tmp_path / "tune.json"), str(tmp_path / "table.txt")
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(wl.state)
    print("case autotune/random15/mixed", flush=True)
    chosen = ctx.autotune(ds, wl.dt, abi.KERNEL_BATCHED_MIXED, cache_path=cache,
                          table_path=table, quick=True)
    torch.cuda.synchronize()
    print("chosen", chosen, flush=True)
    # the state is as given
    back = ds.to_host()
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        np.testing.assert_array_equal(back[f], wl.state[f])
    assert ctx.get_launch(abi.KERNEL_BATCHED_MIXED) == chosen
    key, bucket = ctx.tune_key()
    es = [e for e in _entries(cache) if e["variant"] == abi.KERNEL_BATCHED_MIXED]
    assert len(es) == 1 and es[0]["key"] == key and es[0]["bucket"] == bucket
    assert es[0]["ms"] <= es[0]["default_ms"] * 1.05
    with open(table) as f:
        rows = [l for l in f if not l.startswith("#")]
    assert len(rows) >= 4 and sum("CHOSEN" in l for l in rows) == 1
    # a second call finds the cached choice
    assert ctx.autotune(ds, wl.dt, abi.KERNEL_BATCHED_MIXED, cache_path=cache) == chosen
    # a new context applies the cache
    ctx2 = _tuned_context_from_env(wl, cache)
    assert ctx2.get_launch(abi.KERNEL_BATCHED_MIXED) == chosen
    got, _ = G.run_gpu(wl, abi.KERNEL_BATCHED_MIXED, ctx=ctx2)
    n = wl.state.ncell
    assert np.all(got["status"][0, :n] == 0)
    errs = G.compare(ref, got, n, net=wl.net)
    print("tuned parity", errs, flush=True)
    for f, e in errs.items():
        assert e <= CONC_TOL, (f, e)


def test_autotune_per_cell_active_counts(tmp_path):
    import torch
    wl = workloads.variable_network_workload(2048)
    ref, _ = V.run_oracle_variable(wl)
    cache = str(tmp_path / "tune.json")
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(wl.state)
    print("case autotune/variable/fp64", flush=True)
    ctx.autotune(ds, wl.dt, abi.KERNEL_BATCHED_FP64, cache_path=cache, quick=True)
    torch.cuda.synchronize()
    es = [e for e in _entries(cache) if e["variant"] == abi.KERNEL_BATCHED_FP64]
    # one entry per size bucket, each that bucket's team size
    assert sorted(e["size_bucket"] for e in es) == [0, 1, 2]
    for e in es:
        assert e["lanes_per_cell"] == (16, 32, 64)[e["size_bucket"]]
    ctx2 = _tuned_context_from_env(wl, cache)
    got, _ = G.run_gpu(wl, abi.KERNEL_BATCHED_FP64, ctx=ctx2)
    n = wl.state.ncell
    assert np.all(got["status"][0, :n] == 0)
    for f in ("pri_molal", "total", "sec_molal"):
        e = float(G.rel_err(got[f][:, :n], ref[f][:, :n]).max())
        print("tuned variable", f, e, flush=True)
        assert e <= CONC_TOL, (f, e)
