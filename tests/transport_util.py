"""OSRT transport cases for the tests and the bench (host arrays): a
structured box with a uniform Darcy flux along x plus a random part on
every face, diffusion coefficients over distance per connection, a west
face with inflow of fixed totals, an east face with outflow, and optionally
a few dry cells."""
import numpy as np

from pfchem.grid import StructuredGrid


def box_case(nx, ny, nz, naq, seed=3, q=1e-6, dry_frac=0.0, dt=3600.0):
    rng = np.random.default_rng(seed)
    g = StructuredGrid(nx, ny, nz, 1.0, 1.0, 1.0)
    n = g.ncell
    up, dn, area, dist, axis = g.connections()
    vel = np.where(axis == 0, q, 0.0) + q * 0.3 * rng.standard_normal(up.size)
    cod = 0.25 * 1e-9 / dist * rng.uniform(0.5, 2.0, up.size) + 1e-8 * rng.uniform(0, 1, up.size)
    wc, wa, wd = g.face("west")
    ec, ea, ed = g.face("east")
    bc_cell = np.concatenate([wc, ec])
    bc_area = np.concatenate([wa, ea])
    bc_vel = np.concatenate([np.full(wc.size, q), np.full(ec.size, -q)])
    bc_cod = np.concatenate([0.25 * 1e-9 / wd, np.zeros(ec.size)])
    bc_total = np.vstack([np.concatenate([10.0 ** rng.uniform(-6, -2, wc.size), np.zeros(ec.size)])
                          for _ in range(naq)])
    por = rng.uniform(0.2, 0.4, n)
    sat = rng.uniform(0.5, 1.0, n)
    if dry_frac > 0:
        sat[rng.choice(n, size=max(1, int(dry_frac * n)), replace=False)] = 0.0
    vol = np.ones(n)
    total = 10.0 ** rng.uniform(-7, -3, (naq, n))
    return dict(grid=g, ncell=n, id_up=up, id_dn=dn, area=area, vel=vel, cod=cod,
                bc_cell=bc_cell, bc_area=bc_area, bc_vel=bc_vel, bc_cod=bc_cod,
                bc_total=bc_total, por=por, sat=sat, vol=vol, total=total, dt=dt)


def oracle_step(c):
    import oracle_transport as OT
    return OT.transport_step(c["ncell"], c["id_up"], c["id_dn"], c["area"], c["vel"], c["cod"],
                             c["bc_cell"], c["bc_area"], c["bc_vel"], c["bc_cod"], c["bc_total"],
                             c["por"], c["sat"], c["vol"], c["dt"], c["total"])


def device_coefs(c):
    import torch
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    return dict(area=t(c["area"]), velocity=t(c["vel"]), coef_over_dist=t(c["cod"]),
                bc_area=t(c["bc_area"]), bc_velocity=t(c["bc_vel"]),
                bc_coef_over_dist=t(c["bc_cod"]), bc_total=t(c["bc_total"]))
