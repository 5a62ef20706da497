"""Small tetrahedral geometry and direction sets for the moment tests.

Cells are independent random tetrahedra (the moments only read per-cell face
data: outward unit normals, face areas and the cell volume, in the layouts of
the reference's normals.bin / squares.bin / volume.bin, cuda/cuda_struct.cu:
137-147). Face f is the face opposite vertex f.
"""
import math

import numpy as np

FACE_VERTS = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))


def random_tetra_cells(n, rng, spread=10.0):
    """Returns (normals [n][4][3], areas [n][4], volume [n], centroids [n][4][3])."""
    origin = rng.uniform(-spread, spread, size=(n, 1, 3))
    v = origin + rng.normal(size=(n, 4, 3))
    normals = np.empty((n, 4, 3))
    areas = np.empty((n, 4))
    cents = np.empty((n, 4, 3))
    for f, (a, b, c) in enumerate(FACE_VERTS):
        cr = np.cross(v[:, b] - v[:, a], v[:, c] - v[:, a])
        ln = np.linalg.norm(cr, axis=1)
        nrm = cr / ln[:, None]
        cen = (v[:, a] + v[:, b] + v[:, c]) / 3.0
        inward = np.einsum("ij,ij->i", nrm, v[:, f] - cen) > 0  # points towards the opposite vertex
        nrm[inward] *= -1.0
        normals[:, f], areas[:, f], cents[:, f] = nrm, 0.5 * ln, cen
    e1, e2, e3 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]
    volume = np.abs(np.einsum("ij,ij->i", e1, np.cross(e2, e3))) / 6.0
    return normals, areas, volume, cents


def six_axis_directions():
    """+x, -x, +y, -y, +z, -z with equal weights 4 pi / 6 and full_area 4 pi.
    The rule integrates 1, x_k and x_h x_k over the sphere exactly."""
    xyz = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    w = np.full(6, 4.0 * math.pi / 6.0)
    return xyz, w, 4.0 * math.pi


def linear_flux_illum(cents, base=2.0, slope=3.0):
    """Face intensities on the six-axis set whose face flux is (slope x_f, 0, 0)
    and whose face pressure tensor is base/3 Id + slope x_f e_x e_x^T, x_f the
    face centroid's x: the discrete divergences are then exactly (slope) and
    (slope, 0, 0) by the divergence theorem (the reference's TestDivStream
    check, solve_module/illum/Illum_async.cpp:1351-1387)."""
    n = cents.shape[0]
    illum = np.full((6, n, 4), base)
    illum[0] += 6.0 * slope * cents[:, :, 0]
    return illum


def moment_scales(illum, w, F, areas, volume):
    """Per-cell magnitudes that bound each moment: e = max_f sum_j w_j |I_jf| / F
    bounds energy, stream and impuls; e * sum_f area_f / volume bounds the divergences."""
    e = np.max(np.einsum("j,jcf->cf", w, np.abs(illum)) / F, axis=1)
    d = None if areas is None else e * areas.sum(axis=1) / volume
    return e, d
