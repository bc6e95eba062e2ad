"""Host-side checks that need no GPU: the C ABI library loads and exports
every function include/pfchem.h declares, the ctypes mirror has the C
layout (compared with what gcc computes from the header), argument errors
come back as return codes, and the host helpers (workload generators,
state containers) behave."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from pfchem import abi, flops, workloads
from pfchem import lib as PL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfchem.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfchem_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(PL.LIB_PATH):
        pytest.skip("libpfchem.so not built")
    return PL.load()


def test_exports_every_declared_function(lib):
    names = header_functions()
    assert "pfchem_create" in names and "pfchem_step" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_codes_without_a_device(lib):
    for code in (abi.PFCHEM_OK, abi.PFCHEM_ERR_INVALID, abi.PFCHEM_ERR_HIP,
                 abi.PFCHEM_ERR_UNSUPPORTED, abi.PFCHEM_ERR_NOMEM, abi.PFCHEM_ERR_CELLS):
        msg = lib.pfchem_strerror(code)
        assert msg and msg != b"unknown pfchem error"
    assert lib.pfchem_strerror(99) == b"unknown pfchem error"
    assert lib.pfchem_last_error(None) == b"null context"
    st = abi.State()
    assert lib.pfchem_step(None, 1, C.byref(st), 1.0, abi.KERNEL_BATCHED_MIXED, None) == \
        abi.PFCHEM_ERR_INVALID
    assert lib.pfchem_step_host(None, 1, C.byref(st), 1.0, abi.KERNEL_BATCHED_MIXED) == \
        abi.PFCHEM_ERR_INVALID
    cfg = abi.Launch(16, 256, 4, 1)
    assert lib.pfchem_set_launch(None, abi.KERNEL_BATCHED_MIXED, C.byref(cfg)) == \
        abi.PFCHEM_ERR_INVALID
    assert lib.pfchem_create(None, None, 0, C.byref(C.c_void_p())) == abi.PFCHEM_ERR_INVALID
    k = abi.Constraint()
    assert lib.pfchem_equilibrate(None, 1, C.byref(st), C.byref(k), None, None) == \
        abi.PFCHEM_ERR_INVALID


STRUCTS = {"pfchem_network": abi.Network, "pfchem_options": abi.Options,
           "pfchem_state": abi.State, "pfchem_launch": abi.Launch, "pfchem_somdec": abi.Somdec,
           "pfchem_constraint": abi.Constraint}


def test_ctypes_layout_matches_the_header():
    """sizeof / offsetof of every ABI struct as gcc lays them out"""
    src = ['#include <stddef.h>', '#include <stdio.h>', '#include "pfchem.h"', 'int main(void) {']
    for cname, cls in STRUCTS.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    src.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(cls, fname).offset, \
                (cname, fname)


def test_product_library_does_not_link_the_oracle():
    so = PL.LIB_PATH
    if not os.path.exists(so):
        pytest.skip("libpfchem.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True,
                         text=True).stdout
    assert "orc_" not in out
    needed = subprocess.run(["readelf", "-d", so], capture_output=True, text=True).stdout
    assert "liboracle" not in needed


def test_workloads_are_seeded_and_deterministic():
    a = workloads.calcite_workload(64)
    b = workloads.calcite_workload(64)
    for name in a.state.arrays:
        np.testing.assert_array_equal(a.state.arrays[name], b.state.arrays[name])
    c = workloads.calcite_workload(64, seed=1)
    assert not np.array_equal(a.state.arrays["total"], c.state.arrays["total"])
    r = workloads.random_network_workload(ncell=256)
    net = r.net
    assert (net.naqcomp, net.neqcplx, net.nkinmnrl) == (15, 60, 5)
    assert all(2 <= len(s) <= 4 for s in net.eqcplx_species)
    assert all(abs(v) in (1.0, 2.0, 3.0) for st in net.eqcplx_stoich for v in st)
    # per-cell active mineral count in 0..5
    active = (r.state.arrays["mnrl_area"][:, :256] > 0).sum(axis=0)
    assert active.min() == 0 and active.max() == 5


def test_state_subset_and_padding():
    wl = workloads.calcite_workload(32)
    idx = np.array([5, 3, 31, 0])
    sub = wl.state.subset(idx)
    assert sub.ncell == 4
    np.testing.assert_array_equal(sub.arrays["total"][:, :4], wl.state.arrays["total"][:, idx])
    padded = workloads.calcite_workload(32, ld=40)
    assert padded.state.ld == 40
    np.testing.assert_array_equal(padded.state.arrays["total"][:, :32],
                                  wl.state.arrays["total"][:, :32])


def test_flop_model_counts():
    """F_it of the calcite network by hand: complexes OH-, CO3--, CO2(aq),
    CaHCO3+, CaCO3(aq) on 1, 2, 2, 2, 3 components (water is not one),
    n = 3, one mineral of 3 components active"""
    wl = workloads.calcite_workload(4)
    n = 3
    f_tot = sum(4 * m + 4 + 2 * m * m for m in (1, 2, 2, 2, 3)) + 2 * n + 2 * n * n
    f_acc = n + n * n
    f_min = 4 * 3 + 20 + 2 * 9
    f_lu = 4 * n * n + (2.0 / 3.0) * n ** 3 + 2 * n * n
    f_upd = 12 * n
    assert flops.per_iteration_flops(wl.net, wl.opt, 1) == pytest.approx(
        f_tot + f_acc + f_min + f_lu + f_upd)
