"""Grids and boundary conditions (oracle restatement of src/physics/grid.py)."""

import torch


def _along(axis, s, nd=2):
    idx = [slice(None)] * nd
    idx[axis] = s
    return tuple(idx)


class Grid1D:
    """grid.py:23-43."""

    def __init__(self, xmin, xmax, ncells, ng=2):
        self.xmin, self.xmax, self.ncells, self.ng = xmin, xmax, ncells, ng
        self.dx = (xmax - xmin) / ncells
        self.ntotal = ncells + 2 * ng
        i = torch.arange(self.ntotal, dtype=torch.float64)
        self.x = xmin + (i - ng + 0.5) * self.dx

    def outflow(self, U):
        """grid.py:49-62 (in place)."""
        ng, n = self.ng, self.ncells
        for v in U.values():
            v[:ng] = v[ng].unsqueeze(0).expand(ng, *v.shape[1:])
            v[ng + n:] = v[ng + n - 1].unsqueeze(0).expand(ng, *v.shape[1:])
        return U

    def periodic(self, U):
        """grid.py:64-71."""
        ng, n = self.ng, self.ncells
        for v in U.values():
            v[:ng] = v[n:n + ng]
            v[n + ng:] = v[ng:2 * ng]
        return U

    def constant(self, U, U_old):
        """grid.py:73-87."""
        ng, n = self.ng, self.ncells
        for k, v in U.items():
            v[:ng] = U_old[k][:ng]
            v[ng + n:] = U_old[k][ng + n:]
        return U


def bc_cells_axis(v, axis, n, ng, kind, ref=None):
    """Cell-centred ghosts along one axis, in place   grid.py:93-113."""
    if kind == "outflow":
        first = v[_along(axis, slice(ng, ng + 1))]
        last = v[_along(axis, slice(ng + n - 1, ng + n))]
        v[_along(axis, slice(None, ng))] = first
        v[_along(axis, slice(ng + n, None))] = last
    elif kind == "periodic":
        v[_along(axis, slice(None, ng))] = v[_along(axis, slice(n, n + ng))]
        v[_along(axis, slice(n + ng, None))] = v[_along(axis, slice(ng, 2 * ng))]
    elif kind == "constant":
        if ref is None:
            raise ValueError("constant BC needs a reference array")
        v[_along(axis, slice(None, ng))] = ref[_along(axis, slice(None, ng))]
        v[_along(axis, slice(ng + n, None))] = ref[_along(axis, slice(ng + n, None))]
    else:
        raise ValueError(f"unknown boundary condition {kind!r}")


def bc_faces_normal(v, axis, n, ng, kind):
    """Face-staggered ghosts along the face's own normal, in place; the
    periodic seam face ng+n is collapsed onto face ng   grid.py:116-143."""
    if kind == "outflow":
        first = v[_along(axis, slice(ng, ng + 1))]
        last = v[_along(axis, slice(ng + n, ng + n + 1))]
        v[_along(axis, slice(None, ng))] = first
        v[_along(axis, slice(ng + n + 1, None))] = last
    elif kind == "periodic":
        v[_along(axis, slice(ng + n, ng + n + 1))] = v[_along(axis, slice(ng, ng + 1))]
        v[_along(axis, slice(None, ng))] = v[_along(axis, slice(n, n + ng))]
        v[_along(axis, slice(ng + n + 1, None))] = v[_along(axis, slice(ng + 1, ng + 1 + ng))]
    elif kind == "constant":
        pass
    else:
        raise ValueError(f"unknown boundary condition {kind!r}")


class Grid2D:
    """grid.py:146-266: cells (nxt, nyt), x-faces (nxt+1, nyt), y-faces
    (nxt, nyt+1), corners (nxt+1, nyt+1); y is the contiguous axis."""

    def __init__(self, xmin, xmax, nx, ymin, ymax, ny, ng=2):
        self.xmin, self.xmax, self.nx = xmin, xmax, nx
        self.ymin, self.ymax, self.ny = ymin, ymax, ny
        self.ng = ng
        self.dx = (xmax - xmin) / nx
        self.dy = (ymax - ymin) / ny
        self.nxt, self.nyt = nx + 2 * ng, ny + 2 * ng
        i = torch.arange(self.nxt, dtype=torch.float64)
        j = torch.arange(self.nyt, dtype=torch.float64)
        self.x = xmin + (i - ng + 0.5) * self.dx
        self.y = ymin + (j - ng + 0.5) * self.dy
        self.xf = xmin + (torch.arange(self.nxt + 1, dtype=torch.float64) - ng) * self.dx
        self.yf = ymin + (torch.arange(self.nyt + 1, dtype=torch.float64) - ng) * self.dy
        self.X = self.x.unsqueeze(1).expand(self.nxt, self.nyt)
        self.Y = self.y.unsqueeze(0).expand(self.nxt, self.nyt)
        self.phys = (slice(ng, ng + nx), slice(ng, ng + ny))

    @property
    def shape(self):
        return (self.nxt, self.nyt)

    def bc_cells(self, U, bc_x="outflow", bc_y="outflow", U_ref=None):
        """x first, then y   grid.py:229-240."""
        for k, v in U.items():
            ref = None if U_ref is None else U_ref[k]
            bc_cells_axis(v, 0, self.nx, self.ng, bc_x, ref)
            bc_cells_axis(v, 1, self.ny, self.ng, bc_y, ref)
        return U

    def bc_faces(self, Bxf, Byf, bc_x="outflow", bc_y="outflow"):
        """grid.py:242-255."""
        bc_faces_normal(Bxf, 0, self.nx, self.ng, bc_x)
        bc_cells_axis(Bxf, 1, self.ny, self.ng, bc_y)
        bc_cells_axis(Byf, 0, self.nx, self.ng, bc_x)
        bc_faces_normal(Byf, 1, self.ny, self.ng, bc_y)
        return Bxf, Byf
