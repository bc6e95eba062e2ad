"""The launch-shape tuner's JSON cache (pflotran-elm-interface_amd/csrc/
pfchem_tune.hip), read without a device through pfchem_tune_cache_query:
exact network key first, else the system-size bucket, nearest cell-count
bucket, per variant.  The tuner itself runs on the GPU
(tests/test_gpu_tune.py).  No reference counterpart: the reference runs one
cell at a time on the CPU."""
import json

from pfchem import abi, lib


def _entry(key, bucket, variant, ncell, L, block=256, bpc=4, rows=1, size_bucket=-1):
    return {"key": key, "bucket": bucket, "variant": variant, "size_bucket": size_bucket,
            "ncell": ncell, "lanes_per_cell": L, "rows_per_lane": rows, "block_size": block,
            "blocks_per_cu": bpc, "ms": 1.0, "default_ms": 2.0}


def _write(path, entries):
    # the tuner's own layout: one entry per line
    with open(path, "w") as f:
        f.write('{"pfchem_tune_cache": 1, "entries": [\n')
        f.write(",\n".join(json.dumps(e) for e in entries))
        f.write("\n]}\n")


def test_query_prefers_the_exact_network_then_the_bucket(tmp_path):
    p = str(tmp_path / "tune.json")
    M, F = abi.KERNEL_BATCHED_MIXED, abi.KERNEL_BATCHED_FP64
    _write(p, [_entry("gfx950/256cu/netA", "gfx950/256cu/n16", M, 1 << 20, 1, bpc=0),
               _entry("gfx950/256cu/netA", "gfx950/256cu/n16", M, 1 << 12, 16, block=128),
               _entry("gfx950/256cu/netB", "gfx950/256cu/n16", M, 1 << 20, 8, rows=2),
               _entry("gfx950/256cu/netA", "gfx950/256cu/n16", F, 1 << 20, 32, bpc=2)])
    q = lambda key, bucket, v, n: lib.tune_cache_query(p, key, bucket, v, n)
    # exact key, nearest cell-count bucket
    assert q("gfx950/256cu/netA", "gfx950/256cu/n16", M, 1 << 20)["lanes_per_cell"] == 1
    assert q("gfx950/256cu/netA", "gfx950/256cu/n16", M, 5000)["block_size"] == 128
    assert q("gfx950/256cu/netA", "gfx950/256cu/n16", F, 100)["blocks_per_cu"] == 2
    # another network of the same size bucket takes the bucket's nearest entry
    got = q("gfx950/256cu/netC", "gfx950/256cu/n16", M, 1 << 19)
    assert got is not None and got["lanes_per_cell"] in (1, 8)
    # nothing for another bucket or device
    assert q("gfx950/256cu/netC", "gfx950/256cu/n32", M, 1 << 20) is None
    assert q("gfx942/304cu/netA", "gfx942/304cu/n16", M, 1 << 20) is None


def test_query_skips_per_bucket_entries_and_bad_lines(tmp_path):
    p = str(tmp_path / "tune.json")
    M = abi.KERNEL_BATCHED_MIXED
    with open(p, "w") as f:
        f.write('{"pfchem_tune_cache": 1, "entries": [\n')
        f.write(json.dumps(_entry("k", "b", M, 1 << 18, 64, size_bucket=2)) + ",\n")
        f.write('{"key": "k", "bucket": "b", "variant": 2}\n')   # incomplete: ignored
        f.write("]}\n")
    assert lib.tune_cache_query(p, "k", "b", M, 1 << 18) is None
    assert lib.tune_cache_query(str(tmp_path / "missing.json"), "k", "b", M, 1) is None
