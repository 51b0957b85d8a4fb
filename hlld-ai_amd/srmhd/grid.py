"""Grid2D mirror (src/physics/grid.py:146-266).

Coordinates are float64 CPU tensors computed exactly as the reference does
(initial data are evaluated on them on the CPU, then moved); the state
lives on ``device``.  ``apply_bc_cc``/``apply_bc_face`` are element copies
on whatever device the arrays live on (grid.py:93-143).
"""

import torch


def _sl(axis, s, nd=2):
    out = [slice(None)] * nd
    out[axis] = s
    return tuple(out)


def _bc_cc_axis(v, axis, n, ng, kind):
    if kind == "outflow":
        v[_sl(axis, slice(None, ng))] = v[_sl(axis, slice(ng, ng + 1))]
        v[_sl(axis, slice(ng + n, None))] = v[_sl(axis, slice(ng + n - 1, ng + n))]
    elif kind == "periodic":
        v[_sl(axis, slice(None, ng))] = v[_sl(axis, slice(n, n + ng))].clone()
        v[_sl(axis, slice(n + ng, None))] = v[_sl(axis, slice(ng, 2 * ng))].clone()
    else:
        raise ValueError(f"unknown boundary condition {kind!r}")


def _bc_face_axis(v, axis, n, ng, kind):
    if kind == "outflow":
        v[_sl(axis, slice(None, ng))] = v[_sl(axis, slice(ng, ng + 1))]
        v[_sl(axis, slice(ng + n + 1, None))] = v[_sl(axis, slice(ng + n, ng + n + 1))]
    elif kind == "periodic":
        v[_sl(axis, slice(ng + n, ng + n + 1))] = v[_sl(axis, slice(ng, ng + 1))]
        v[_sl(axis, slice(None, ng))] = v[_sl(axis, slice(n, n + ng))].clone()
        v[_sl(axis, slice(ng + n + 1, None))] = v[_sl(axis, slice(ng + 1, ng + 1 + ng))].clone()
    else:
        raise ValueError(f"unknown boundary condition {kind!r}")


class Grid2D:
    def __init__(self, xmin, xmax, nx, ymin, ymax, ny, ng=2, device="cuda"):
        self.xmin, self.xmax, self.nx = xmin, xmax, nx
        self.ymin, self.ymax, self.ny = ymin, ymax, ny
        self.ng = ng
        self.device = torch.device(device)
        self.dx = (xmax - xmin) / nx
        self.dy = (ymax - ymin) / ny
        self.nxt = nx + 2 * ng
        self.nyt = ny + 2 * ng
        i = torch.arange(self.nxt, dtype=torch.float64)
        j = torch.arange(self.nyt, dtype=torch.float64)
        self.x = xmin + (i - ng + 0.5) * self.dx
        self.y = ymin + (j - ng + 0.5) * self.dy
        self.xf = xmin + (torch.arange(self.nxt + 1, dtype=torch.float64) - ng) * self.dx
        self.yf = ymin + (torch.arange(self.nyt + 1, dtype=torch.float64) - ng) * self.dy
        self.X = self.x.unsqueeze(1).expand(self.nxt, self.nyt)
        self.Y = self.y.unsqueeze(0).expand(self.nxt, self.nyt)
        self.phys = (slice(ng, ng + nx), slice(ng, ng + ny))
        self.phys_xf = (slice(ng, ng + nx + 1), slice(ng, ng + ny))
        self.phys_yf = (slice(ng, ng + nx), slice(ng, ng + ny + 1))

    @property
    def shape(self):
        return (self.nxt, self.nyt)

    def apply_bc_cc(self, U, bc_x="outflow", bc_y="outflow"):
        for v in U.values():
            _bc_cc_axis(v, 0, self.nx, self.ng, bc_x)
            _bc_cc_axis(v, 1, self.ny, self.ng, bc_y)
        return U

    def apply_bc_face(self, Bxf, Byf, bc_x="outflow", bc_y="outflow"):
        _bc_face_axis(Bxf, 0, self.nx, self.ng, bc_x)
        _bc_cc_axis(Bxf, 1, self.ny, self.ng, bc_y)
        _bc_cc_axis(Byf, 0, self.nx, self.ng, bc_x)
        _bc_face_axis(Byf, 1, self.ny, self.ng, bc_y)
        return Bxf, Byf

    def div_b(self, Bxf, Byf):
        return (Bxf[1:, :] - Bxf[:-1, :]) / self.dx + (Byf[:, 1:] - Byf[:, :-1]) / self.dy
