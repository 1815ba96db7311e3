"""Drop-in for xpsi/cellmesh/mesh_tools.pyx: allocate_cells (:892-1099)."""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def allocate_cells(num_cells, min_sqrt_num_cells, max_sqrt_num_cells, mass, r_s, R_eq, Omega, zeta,
                   epsilon, star_shape_ind, superRadius, cedeRadius, superColatitude, cedeColatitude,
                   superAzimuth, holeColatitude, holeRadius, holeAzimuth):
    """-> (super_sqrt_num_cells, super_num_cells, cede_sqrt_num_cells, cede_num_cells)"""
    D.require_gpu()
    L = lib()
    ws = D.empty((L.xh_allocate_cells_workspace_bytes() + 7) // 8)
    out = np.zeros(4)
    rc = L.xh_allocate_cells(int(num_cells), int(min_sqrt_num_cells), int(max_sqrt_num_cells), float(mass),
                             float(r_s), float(R_eq), float(Omega), float(zeta), float(epsilon),
                             int(star_shape_ind), float(superRadius), float(cedeRadius),
                             float(superColatitude), float(cedeColatitude), float(superAzimuth),
                             float(holeColatitude), float(holeRadius), float(holeAzimuth),
                             out.ctypes.data, D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    check(rc, "xh_allocate_cells")
    return tuple(float(v) for v in out)
