"""FP64 restatement of the OSRT transport step (test infrastructure: only
tests/ and bench.py's checks use it).

The implicit system of one transport step, per aqueous component
(pmc_subsurface_osrt.F90:258-333):
  rhs = fixed_accum / dt + boundary inflow      (fixed_accum :259-272,
                                                 RTCalculateRHS_t1
                                                 reactive_transport.F90:1239-1465)
  M   = psv / dt + connection coefficients       (RTCalculateTransportMatrix
                                                 :1469-1703)
with psv = porosity sat 1000 volume, the coefficients of TFluxCoef
(transport.F90:495-565) and the boundary ones of TFluxCoefBC (:569-609),
solved directly (scipy's sparse LU; the reference's deck uses SOLVER
DIRECT or its KSP to a tolerance).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

RT_MIN_SATURATION = 1e-40  # reactive_transport_aux.F90:19


def flux_coef(q, cod, area):
    """TFluxCoef: (T_up, T_dn) with the 1000 L/m^3 factor"""
    q = np.asarray(q, dtype=np.float64)
    t_up = np.where(q > 0.0, cod + q, cod) * area * 1000.0
    t_dn = np.where(q > 0.0, -cod, -cod + q) * area * 1000.0
    return t_up, t_dn


def transport_system(ncell, id_up, id_dn, area, vel, cod, bc_cell, bc_area, bc_vel, bc_cod,
                     por, sat, vol, dt):
    """(M as CSR, accumulation coefficient per cell, T_up of each boundary
    connection)"""
    acc = por * sat * 1000.0 * vol / dt
    rows, cols, vals = [np.arange(ncell)], [np.arange(ncell)], [acc]
    t_up, t_dn = flux_coef(vel, cod, area)
    # interior: both sides wet (check_upwind_saturation)
    wet = (sat[id_dn] >= RT_MIN_SATURATION) & (sat[id_up] >= RT_MIN_SATURATION)
    t_up = np.where(wet, t_up, 0.0)
    t_dn = np.where(wet, t_dn, 0.0)
    rows += [id_up, id_up, id_dn, id_dn]
    cols += [id_up, id_dn, id_dn, id_up]
    vals += [t_up, t_dn, -t_dn, -t_up]
    b_up, b_dn = flux_coef(bc_vel, bc_cod, bc_area)
    bwet = sat[bc_cell] >= RT_MIN_SATURATION
    b_up = np.where(bwet, b_up, 0.0)
    b_dn = np.where(bwet, b_dn, 0.0)
    rows.append(bc_cell)
    cols.append(bc_cell)
    vals.append(-b_dn)
    # a dry cell (no accumulation, no fluxes) has an empty row in the
    # reference's matrix; here it keeps its totals
    dry = sat < RT_MIN_SATURATION
    rows.append(np.flatnonzero(dry))
    cols.append(np.flatnonzero(dry))
    vals.append(np.ones(int(dry.sum())))
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ncell, ncell))
    return M, acc, b_up


def transport_step(ncell, id_up, id_dn, area, vel, cod, bc_cell, bc_area, bc_vel, bc_cod,
                   bc_total, por, sat, vol, dt, total):
    """totals of the new time level, [naq, ncell]"""
    M, acc, b_up = transport_system(ncell, id_up, id_dn, area, vel, cod, bc_cell, bc_area,
                                    bc_vel, bc_cod, por, sat, vol, dt)
    lu = spla.splu(M.tocsc())
    out = np.empty_like(total)
    dry = sat < RT_MIN_SATURATION
    for k in range(total.shape[0]):
        rhs = acc * total[k]
        np.add.at(rhs, bc_cell, b_up * bc_total[k])
        rhs[dry] = total[k][dry]
        out[k] = lu.solve(rhs)
    return out
