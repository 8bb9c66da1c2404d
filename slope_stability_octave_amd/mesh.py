"""Structured P2 slope meshes (setup path, numpy).

* ``mesh_P2_2D`` follows mesh_P2_2D.m:2-263: the same node numbering, element
  order and local node order, and the Dirichlet mask of :258-261 (fixed base,
  rollers on the two vertical sides).
* ``mesh_P2_3D`` follows mesh_P2_3D.m:2-420: lower block over the full
  length, upper block over the slope part, x mirrored, six tetrahedra per hex
  (:203-219) and the mask of :415-418.  Segment counts are integers per axis
  (the reference takes one count per metre, N_h), so meshes of any DOF count
  can be built.  ``mesh_P2_box`` is the same generator without the upper block.

Node ids are 0-based; ``elem`` is ``n_p x n_e``; ``Q`` is ``dim x n_n`` bool.
"""
from __future__ import annotations

import math

import numpy as np


def _oround(x: float) -> int:
    """Octave round (half away from zero)."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def _linspace(a, b, n):
    """Octave linspace: a + i*(b-a)/(n-1), last element exactly b."""
    if n == 1:
        return np.array([b], dtype=np.float64)
    d = (b - a) / (n - 1)
    out = a + np.arange(n, dtype=np.float64) * d
    out[-1] = b
    return out


def mesh_P2_2D(h, x1, x2, x3, y1, y2):
    """2D slope: foundation of height y1 over x1+x2+x3, slope of height y2 and
    horizontal extent x2, crest length x1 (left), toe berm x3 (right)."""
    Nx12 = _oround((x1 + x2) / h)
    Nx3 = _oround(x3 / h)
    Nx = Nx12 + Nx3
    Ny1 = _oround(y1 / h)
    Ny2 = _oround(y2 / h)
    Ny = Ny1 + Ny2
    coord_x = np.concatenate([_linspace(0, x1 + x2, 2 * Nx12 + 1), _linspace(x1 + x2, x1 + x2 + x3, 2 * Nx3 + 1)[1:]])
    coord_y = np.concatenate([_linspace(0, y1, 2 * Ny1 + 1), _linspace(y1, y1 + y2, 2 * Ny2 + 1)[1:]])
    n_n = (2 * Ny1 + 1) * (2 * Nx + 1) + 2 * Ny2 * (2 * Nx12 + 1)
    coord = np.zeros((2, n_n))
    C = -np.ones((2 * Nx + 1, 2 * Ny + 1), dtype=np.int64)
    k = 0
    for j in range(2 * Ny1 + 1):
        C[:, j] = np.arange(k, k + 2 * Nx + 1)
        coord[0, k:k + 2 * Nx + 1] = coord_x
        coord[1, k:k + 2 * Nx + 1] = coord_y[j]
        k += 2 * Nx + 1
    for j in range(2 * Ny1 + 1, 2 * Ny + 1):
        x_max = x1 + x2 * (y1 + y2 - coord_y[j]) / y2
        cx = _linspace(0, x_max, 2 * Nx12 + 1)
        C[: 2 * Nx12 + 1, j] = np.arange(k, k + 2 * Nx12 + 1)
        coord[0, k:k + 2 * Nx12 + 1] = cx
        coord[1, k:k + 2 * Nx12 + 1] = coord_y[j]
        k += 2 * Nx12 + 1
    assert k == n_n

    def cells(nx, jr):
        out = []
        for j in jr:
            for i in range(1, nx + 1):
                # 1-based (i, j) of the reference loops
                out.append([C[2 * i - 2, 2 * j - 2], C[2 * i, 2 * j - 2], C[2 * i - 2, 2 * j],
                            C[2 * i - 1, 2 * j - 1], C[2 * i - 2, 2 * j - 1], C[2 * i - 1, 2 * j - 2]])
                out.append([C[2 * i, 2 * j], C[2 * i - 2, 2 * j], C[2 * i, 2 * j - 2],
                            C[2 * i - 1, 2 * j - 1], C[2 * i, 2 * j - 1], C[2 * i - 1, 2 * j]])
        return out

    elem = np.array(cells(Nx, range(1, Ny1 + 1)) + cells(Nx12, range(Ny1 + 1, Ny + 1)), dtype=np.int64).T
    x_max = coord[0].max() - 1e-9
    Q = np.zeros((2, n_n), dtype=bool)
    Q[0] = (coord[0] > 0) & (coord[1] > 0) & (coord[0] < x_max)
    Q[1] = coord[1] > 0
    return coord, elem, Q


# six tetrahedra per hexahedron, hex vertices V1..V8 =
# (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1); mesh_P2_3D.m:203-219
_TETS = [(1, 2, 4, 6), (1, 4, 5, 6), (4, 5, 6, 8), (2, 3, 4, 6), (3, 6, 7, 4), (4, 6, 7, 8)]
_HEX = {1: (0, 0, 0), 2: (1, 0, 0), 3: (1, 1, 0), 4: (0, 1, 0), 5: (0, 0, 1), 6: (1, 0, 1), 7: (1, 1, 1), 8: (0, 1, 1)}
# P2 local order: 4 vertices then edges (0,1) (1,2) (0,2) (1,3) (2,3) (0,3)
_EDGES = [(0, 1), (1, 2), (0, 2), (1, 3), (2, 3), (0, 3)]


def _tet_offsets():
    """(6, 10, 3) doubled-grid offsets of the P2 nodes of each tetrahedron."""
    out = np.zeros((6, 10, 3), dtype=np.int64)
    for t, verts in enumerate(_TETS):
        v = [np.array(_HEX[k]) * 2 for k in verts]
        for a in range(4):
            out[t, a] = v[a]
        for m, (a, b) in enumerate(_EDGES):
            out[t, 4 + m] = (v[a] + v[b]) // 2
    return out


def _structured_3d(Nx, Ny, Nz, N1x, N1y, Lx, y1, y2, z, x1, x2, x3, slope):
    """Lower block (Nx x N1y cells) plus, when `slope`, the upper block over
    the last Nx-N1x x-cells (N_y-N1y rows).  Returns coord, elem, grid index
    of every node."""
    Ny_tot = Ny if slope else N1y
    gx, gy, gz = 2 * Nx + 1, 2 * Ny_tot + 1, 2 * Nz + 1
    node = -np.ones((gx, gy, gz), dtype=np.int64)
    # lower block, x fastest then y then z (the reference's C1 numbering)
    iy_lo = 2 * N1y + 1
    ix, iy, iz = np.meshgrid(np.arange(gx), np.arange(iy_lo), np.arange(gz), indexing="ij")
    order = np.lexsort((ix.ravel(order="F"), iy.ravel(order="F"), iz.ravel(order="F")))
    lo = np.stack([ix.ravel(order="F"), iy.ravel(order="F"), iz.ravel(order="F")], 1)[order]
    node[lo[:, 0], lo[:, 1], lo[:, 2]] = np.arange(lo.shape[0])
    nodes = [lo]
    if slope:
        ix, iy, iz = np.meshgrid(np.arange(2 * N1x, gx), np.arange(iy_lo, gy), np.arange(gz), indexing="ij")
        up = np.stack([ix.ravel(order="F"), iy.ravel(order="F"), iz.ravel(order="F")], 1)
        up = up[np.lexsort((up[:, 0], up[:, 1], up[:, 2]))]
        node[up[:, 0], up[:, 1], up[:, 2]] = lo.shape[0] + np.arange(up.shape[0])
        nodes.append(up)
    g = np.vstack(nodes)
    n_n = g.shape[0]
    cx = _linspace(0, Lx, gx)
    cy = _linspace(0, y1 + (y2 if slope else 0.0), gy)
    cz = _linspace(0, z, gz)
    coord = np.zeros((3, n_n))
    coord[0] = Lx - cx[g[:, 0]]
    coord[1] = cy[g[:, 1]]
    coord[2] = cz[g[:, 2]]
    if slope:
        up = g[:, 1] > 2 * N1y
        # x of upper-block nodes, mesh_P2_3D.m:36-40 (relative height above y1)
        s = _linspace(-x2 - x3, 0, 2 * (Nx - N1x) + 1)[g[up, 0] - 2 * N1x]
        rel = cy[g[up, 1] - 2 * N1y]
        coord[0, up] = Lx - (Lx + s * (1 - x2 / ((x2 + x3) * y2) * rel))
    # hex cells in x-fastest order (lower and upper blocks interleaved by y as
    # the reference's logical-mask extraction yields)
    cells = []
    for kz in range(Nz):
        for ky in range(Ny_tot):
            xs = range(Nx) if ky < N1y else range(N1x, Nx)
            for kx in xs:
                cells.append((kx, ky, kz))
    cells = np.array(cells, dtype=np.int64)
    off = _tet_offsets()  # (6,10,3)
    base = 2 * cells[:, None, None, :] + off[None]  # (n_c, 6, 10, 3)
    elem = node[base[..., 0], base[..., 1], base[..., 2]].reshape(-1, 10).T
    assert (elem >= 0).all()
    return coord, np.ascontiguousarray(elem), n_n


def mesh_P2_3D(nx1, nx2, nx3, ny1, ny2, nz, x1, x2, x3, y1, y2, z):
    """3D slope with integer cell counts per part: nx1 cells over the toe
    berm x1, nx2 over the slope x2, nx3 over the crest x3, ny1 over the
    foundation height y1, ny2 over the slope height y2, nz over the width z.
    Uniform spacing requires x1/nx1 == x2/nx2 == x3/nx3 and y1/ny1 == y2/ny2
    (as the reference generator, which uses one count per metre)."""
    Nx = nx1 + nx2 + nx3
    Ny = ny1 + ny2
    coord, elem, n_n = _structured_3d(Nx, Ny, nz, nx1, ny1, x1 + x2 + x3, y1, y2, z, x1, x2, x3, True)
    Q = np.zeros((3, n_n), dtype=bool)
    Q[0] = (coord[0] > 0) & (coord[1] > 0) & (coord[0] < x1 + x2 + x3)
    Q[1] = coord[1] > 0
    Q[2] = (coord[2] > 0) & (coord[1] > 0) & (coord[2] < z)
    return coord, elem, Q


def mesh_P2_box(nx, ny, nz, Lx, Ly, Lz):
    """Structured box [0,Lx]x[0,Ly]x[0,Lz], nx*ny*nz hexes x 6 P2 tetrahedra,
    with the slope generator's boundary rule (fixed base y=0, normal rollers
    on the x and z faces)."""
    coord, elem, n_n = _structured_3d(nx, ny, nz, 0, ny, Lx, Ly, 0.0, Lz, 0.0, 0.0, 0.0, False)
    Q = np.zeros((3, n_n), dtype=bool)
    Q[0] = (coord[0] > 0) & (coord[1] > 0) & (coord[0] < Lx)
    Q[1] = coord[1] > 0
    Q[2] = (coord[2] > 0) & (coord[1] > 0) & (coord[2] < Lz)
    return coord, elem, Q


def load_mesh_P2(data, boundary_type=0):
    """Mesh in the HDF5 layout of the reference's meshes (datasets node, elem,
    face, boundary, material; 0-based), as read by load_mesh_P2.m:25-57:
    Dirichlet mask from the boundary labels (1,2 -> x; 3,4 -> z; 5 -> y, or
    all components when boundary_type), then the axis permutation [1 3 2]
    that makes the second coordinate vertical.  ``data`` is a mapping of the
    five arrays as stored in the file (row-major: node n_n x 3, elem n_e x 10,
    face n_f x 6).  Returns coord (3 x n_n), elem (10 x n_e), surf, Q (3 x n_n),
    material (n_e)."""
    node = np.asarray(data["node"], dtype=np.float64).T
    elem = np.asarray(data["elem"]).astype(np.int64).T
    face = np.asarray(data["face"]).astype(np.int64).T
    boundary = np.asarray(data["boundary"]).ravel()
    material = np.asarray(data["material"]).ravel().astype(np.int64)
    Q = np.ones(node.shape, dtype=bool)
    for lab, comp in ((1, 0), (2, 0), (3, 1), (4, 1)):
        Q[comp, np.unique(face[:, boundary == lab])] = False
    f5 = np.unique(face[:, boundary == 5])
    if boundary_type:
        Q[:, f5] = False
    else:
        Q[2, f5] = False
    coord = node[[0, 2, 1], :]
    Q = Q[[0, 2, 1], :]
    return coord, elem, face, Q, material
