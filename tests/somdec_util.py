"""Carbon and nitrogen inventories of a SomDec (ELM-CN cascade) state, for
the conservation checks of the oracle and of the HIP kernels."""
import numpy as np


def carbon_nitrogen(net, arrays, n):
    """C and N (mol per cell) in the SOM and litter pools and in CO2, NH4+,
    NO3- and N2O, for the first n cells of a state laid out as
    workloads.somdec_workload writes it (SOM1-4, Lit1C-3C, Lit1N-3N)"""
    a = arrays
    sd = net.somdec
    psv = a["porosity"][0, :n] * a["saturation"][0, :n] * 1000.0 * a["volume"][0, :n]
    imm = a["immobile"][:, :n] * a["volume"][0, :n]
    tot = a["total"][:, :n] * psv
    nc_som = sd.arrays["upstream_nc"][3:7]     # SOM1-4 fixed N:C (reactions 4-7)
    C = imm[0:4].sum(0) + imm[4:7].sum(0) + tot[0]
    N = (imm[0:4] * nc_som[:, None]).sum(0) + imm[7:10].sum(0) + tot[1] + tot[2] + 2 * tot[3]
    return C, N
