"""Staggered uniform-raster grid (reference solver/core/grid.py:1-86).

Layout and sign convention are the reference's, because every array crossing the
C-ABI uses it (grid.py:9-32):

* cells ``(ny, nx)``: ``h``, ``z``, ``eta = h + z``
* x-faces ``(ny, nx+1)``: ``qx[i, j]`` between cells ``(i, j-1)`` and ``(i, j)``,
  positive toward increasing ``j``
* y-faces ``(ny+1, nx)``: ``qy[i, j]`` between cells ``(i-1, j)`` and ``(i, j)``,
  positive toward increasing ``i``
"""

from __future__ import annotations

from dataclasses import dataclass

GRAVITY = 9.81  # grid.py:44
H_DRY = 1.0e-3  # grid.py:49


@dataclass(frozen=True)
class Grid:
    ny: int
    nx: int
    dx: float

    @property
    def shape(self) -> tuple[int, int]:
        return (self.ny, self.nx)

    @property
    def qx_shape(self) -> tuple[int, int]:
        return (self.ny, self.nx + 1)

    @property
    def qy_shape(self) -> tuple[int, int]:
        return (self.ny + 1, self.nx)

    @property
    def cell_area(self) -> float:
        return self.dx * self.dx

    @property
    def n_cells(self) -> int:
        return self.ny * self.nx
