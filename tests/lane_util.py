"""Host build of the network-specialized one-lane kernel (test infrastructure).

The library generates the kernel's device source from a network
(pfchem_prepare_kernels); these helpers compile that same source for the CPU
with tests/lane_host/lane_host_shim.h, where one thread plays a wavefront of
one lane, and run it over a CellState.  The arithmetic, the RStep / RReact
state machine, the escalation and the write-out are the generated code's own;
only the cross-lane primitives reduce to their one-lane meaning.
"""
import os
import subprocess
import tempfile

import numpy as np

from pfchem import lib

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_DIR = os.path.join(HERE, "lane_host")
FIELDS = ("pri_molal", "total", "immobile", "sec_molal", "pri_act_coef", "sec_act_coef",
          "ln_act_h2o", "mnrl_volfrac", "mnrl_area", "mnrl_rate", "porosity", "saturation",
          "den_kg", "volume", "temperature", "num_iterations", "num_sub_steps", "status")


def build_host(net, opt, workdir):
    """generate the device source of (net, opt) and compile it for the host;
    returns the executable's path"""
    src = os.path.join(workdir, "lane_dev.hip")
    rc = lib.prepare_kernels(net, opt, src)
    assert rc == 0, "one-lane kernel does not take this network"
    with open(src) as f:
        text = f.read().replace("#include <hip/hip_runtime.h>", "")
    cpp = os.path.join(workdir, "lane_dev.cpp")
    with open(cpp, "w") as f:
        f.write(text)
        f.write('\n#include "lane_host_main.cpp"\n')
    exe = os.path.join(workdir, "lane_host")
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-w", "-I", HOST_DIR,
                    "-include", os.path.join(HOST_DIR, "lane_host_shim.h"), cpp, "-o", exe],
                   check=True)
    return exe


def run_host(exe, state, dt, f32=False, escalate=True):
    """run the host build over a copy of `state`; returns the output state"""
    out = state.copy()
    with tempfile.TemporaryDirectory() as d:
        for name in FIELDS:
            a = out.arrays.get(name)
            if a is not None and a.shape[0] > 0:
                np.ascontiguousarray(a).tofile(os.path.join(d, name + ".bin"))
        subprocess.run([exe, d, str(state.ncell), str(state.ld), repr(float(dt)),
                        "1" if f32 else "0", "1" if escalate else "0"], check=True)
        for name in FIELDS:
            p = os.path.join(d, name + ".out")
            if os.path.exists(p):
                a = out.arrays[name]
                a[...] = np.fromfile(p, dtype=a.dtype).reshape(a.shape)
    return out


ADJ_FIELDS = ("pri_molal", "total", "pri_act_coef", "sec_act_coef", "ln_act_h2o", "mnrl_volfrac",
              "mnrl_area", "porosity", "saturation", "den_kg", "volume", "temperature")


def build_host_adjoint(net, opt, workdir):
    """the generated adjoint kernel (PFLANE_ADJOINT) compiled for the host"""
    src = os.path.join(workdir, "lane_dev.hip")
    rc = lib.prepare_kernels(net, opt, src)
    assert rc == 0, "one-lane kernel does not take this network"
    with open(src) as f:
        text = f.read()
    cpp = os.path.join(workdir, "lane_adj.cpp")
    with open(cpp, "w") as f:
        f.write("#define PFLANE_ADJOINT 1\n")
        f.write(text)
        f.write('\n#include "lane_host_adjoint.cpp"\n')
    exe = os.path.join(workdir, "lane_host_adjoint")
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-w", "-I", HOST_DIR,
                    "-include", os.path.join(HOST_DIR, "lane_host_shim.h"), cpp, "-o", exe],
                   check=True)
    return exe


def run_host_adjoint(exe, net, state, dt, wv, ws, exact=True):
    """(per-cell gradient [npar, ncell], f per cell, status per cell)"""
    npar = net.nkinmnrl + net.neqcplx
    nf = net.naqcomp + net.nkinmnrl + 1
    n = state.ncell
    with tempfile.TemporaryDirectory() as d:
        for name in ADJ_FIELDS:
            a = state.arrays.get(name)
            if a is not None and a.shape[0] > 0:
                np.ascontiguousarray(a).tofile(os.path.join(d, name + ".bin"))
        np.ascontiguousarray(wv, dtype=np.float64).tofile(os.path.join(d, "wv.bin"))
        np.ascontiguousarray(ws, dtype=np.float64).tofile(os.path.join(d, "ws.bin"))
        subprocess.run([exe, d, str(n), str(state.ld), repr(float(dt)), "1" if exact else "0",
                        str(npar), str(nf)], check=True)
        g = np.fromfile(os.path.join(d, "grad.out"), dtype=np.float64).reshape(max(npar, 1), n)
        f = np.fromfile(os.path.join(d, "fcell.out"), dtype=np.float64)
        st = np.fromfile(os.path.join(d, "status.out"), dtype=np.int32)
    return g[:npar], f, st
