"""Structured-grid connections for the OSRT transport step (host side).

The reference builds its connection sets from the discretization
(grid_structured.F90: internal connections between face neighbours, id_up
the lower index, and boundary connections on the domain faces).  This
module does the same for an nx x ny x nz box of uniform cells, natural
(x fastest) cell order, and returns the arrays pfchem_transport_setup and
pfchem_transport_step take.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class StructuredGrid:
    nx: int
    ny: int
    nz: int
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def cell(self, i, j, k):
        return i + self.nx * (j + self.ny * k)

    def connections(self):
        """internal connections (id_up, id_dn, area, dist, axis): face
        neighbours along x, y and z, id_up the lower cell"""
        ups, dns, area, dist, axis = [], [], [], [], []
        idx = np.arange(self.ncell).reshape(self.nz, self.ny, self.nx)
        for ax, (a, d) in enumerate(((self.dy * self.dz, self.dx), (self.dx * self.dz, self.dy),
                                     (self.dx * self.dy, self.dz))):
            if ax == 0:
                u, v = idx[:, :, :-1], idx[:, :, 1:]
            elif ax == 1:
                u, v = idx[:, :-1, :], idx[:, 1:, :]
            else:
                u, v = idx[:-1, :, :], idx[1:, :, :]
            ups.append(u.ravel())
            dns.append(v.ravel())
            area.append(np.full(u.size, a))
            dist.append(np.full(u.size, d))
            axis.append(np.full(u.size, ax))
        return (np.concatenate(ups).astype(np.int64), np.concatenate(dns).astype(np.int64),
                np.concatenate(area), np.concatenate(dist), np.concatenate(axis))

    def face(self, name):
        """(cells, area, half-distance) of a domain face: west/east (x),
        south/north (y), bottom/top (z)"""
        idx = np.arange(self.ncell).reshape(self.nz, self.ny, self.nx)
        sel = {"west": (idx[:, :, 0], self.dy * self.dz, 0.5 * self.dx),
               "east": (idx[:, :, -1], self.dy * self.dz, 0.5 * self.dx),
               "south": (idx[:, 0, :], self.dx * self.dz, 0.5 * self.dy),
               "north": (idx[:, -1, :], self.dx * self.dz, 0.5 * self.dy),
               "bottom": (idx[0, :, :], self.dx * self.dy, 0.5 * self.dz),
               "top": (idx[-1, :, :], self.dx * self.dy, 0.5 * self.dz)}[name]
        cells = sel[0].ravel().astype(np.int64)
        return cells, np.full(cells.size, sel[1]), np.full(cells.size, sel[2])
