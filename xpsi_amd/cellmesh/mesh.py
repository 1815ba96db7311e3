"""Drop-ins for xpsi/cellmesh/mesh.pyx construct_spot_cellMesh (:23-417) and
xpsi/cellmesh/polar_mesh.pyx construct_polar_cellMesh (:23-430): same
signatures, returning (theta, phi, r, cellAreas, maxEmissionAngle,
cos_gamma, effGrav)."""
import ctypes as C

from .._lib import lib, check, XHError
from .. import _device as D


def _mesh(polar, numCell, sqrt_numCell, mass, r_s, R_eq, zeta, epsilon, star_shape_ind, cedeRadius,
          cedeColatitude, superRadius, superColatitude, superAzimuth):
    S = int(sqrt_numCell)
    theta, phi, area = D.empty((S, S)), D.empty((S, S)), D.empty((S, S))
    r, maxA, cg, eg = D.empty(S), D.empty(S), D.empty(S), D.empty(S)
    L = lib()
    ws = D.empty((L.xh_cell_mesh_workspace_bytes(S) + 7) // 8)
    st = C.c_int(0)
    rc = L.xh_construct_cell_mesh(int(polar), float(numCell), S, float(mass), float(r_s), float(R_eq),
                                  float(zeta), float(epsilon), int(star_shape_ind), float(cedeRadius),
                                  float(cedeColatitude), float(superRadius), float(superColatitude),
                                  float(superAzimuth), D.ptr(theta), D.ptr(phi), D.ptr(r), D.ptr(area),
                                  D.ptr(maxA), D.ptr(cg), D.ptr(eg), D.ptr(ws), ws.numel() * 8,
                                  C.byref(st), D.stream_ptr())
    check(rc, "xh_construct_cell_mesh")
    if st.value != 0:
        raise XHError("xh_construct_cell_mesh: mesh status %d" % st.value)
    return tuple(a.cpu().numpy() for a in (theta, phi, r, area, maxA, cg, eg))


def construct_spot_cellMesh(numThreads, numCell, sqrt_numCell, mass, r_s, R_eq, zeta, epsilon,
                            star_shape_ind, cedeRadius, cedeColatitude, superRadius, superColatitude,
                            superAzimuth):
    return _mesh(0, numCell, sqrt_numCell, mass, r_s, R_eq, zeta, epsilon, star_shape_ind, cedeRadius,
                 cedeColatitude, superRadius, superColatitude, superAzimuth)


def construct_polar_cellMesh(numThreads, numCell, sqrt_numCell, mass, r_s, R_eq, zeta, epsilon,
                             star_shape_ind, cedeRadius, cedeColatitude, superRadius, superColatitude,
                             superAzimuth):
    return _mesh(1, numCell, sqrt_numCell, mass, r_s, R_eq, zeta, epsilon, star_shape_ind, cedeRadius,
                 cedeColatitude, superRadius, superColatitude, superAzimuth)
