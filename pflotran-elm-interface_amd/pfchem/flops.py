"""Algorithmic work model of one RReact Newton iteration (used for the
roofline figures of the reaction step).

Counting rule: FMA = 2 flops; add / mul / div / sqrt / exp / log = 1 flop.
Per Newton iteration of a cell with n components:
    F_it = F_act + F_tot + F_acc + F_min + F_LU + F_upd
    F_act = 8 (n + nc)            only if activity coefficients are updated
                                  every Newton iteration
    F_tot = sum_k (4 m_k + 4 + 2 m_k^2) + 2 n + 2 n^2   (m_k species of complex k)
    F_acc = n + n^2
    F_min = sum_{active minerals} (4 nu + 20 + 2 nu^2)  (nu species of mineral)
    F_LU  = 4 n^2 + (2/3) n^3 + 2 n^2
    F_upd = 12 n
Work of a launch = sum over cells of (FP64 reference iterations x F_it), so
extra iterations a reduced-precision solve may take are not credited.
"""
from __future__ import annotations

import numpy as np

from . import abi


def per_iteration_flops(net, opt, active_minerals: np.ndarray) -> np.ndarray:
    """F_it per cell; active_minerals is a bool array (nkinmnrl, ncell)."""
    n = net.ncomp
    nc = net.neqcplx
    f_act = 8.0 * (net.naqcomp + nc) if opt.act_coef_update_frequency == abi.ACT_COEF_FREQUENCY_NEWTON_ITER else 0.0
    f_tot = sum(4.0 * len(sp) + 4.0 + 2.0 * len(sp) ** 2 for sp in net.eqcplx_species)
    f_tot += 2.0 * net.naqcomp + 2.0 * net.naqcomp ** 2
    f_acc = n + n * n
    f_lu = 4.0 * n * n + (2.0 / 3.0) * n ** 3 + 2.0 * n * n
    f_upd = 12.0 * n
    per_min = np.array([4.0 * len(sp) + 20.0 + 2.0 * len(sp) ** 2 for sp in net.kinmnrl_species])
    f_min = (per_min[:, None] * active_minerals).sum(axis=0) if len(per_min) else 0.0
    return f_act + f_tot + f_acc + f_lu + f_upd + f_min


def state_bytes_per_cell(net) -> float:
    """HBM bytes the step must read and write per cell (SoA FP64 state read
    once and written once; network excluded)."""
    naq, nc, nm, nim = net.naqcomp, net.neqcplx, net.nkinmnrl, net.nimmobile
    read = 8 * (3 * naq + 2 * nc + 2 * nm + nim + 5)
    write = 8 * (2 * naq + nc + 2 * nm + nim) + 4 * 3
    return float(read + write)
