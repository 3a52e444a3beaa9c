"""Brute-force global assembly of the P1 operators on the refined macro mesh (test helper).

The refined mesh is built from the 24 micro-tetrahedra around a micro-vertex
(src/hyteg/p1functionspace/P1Elements.hpp:86-136); tests check separately that
these micro-tetrahedra tile every macro-cell conformingly, so the global matrix
below is an independent statement of "the P1 stiffness / mass matrix of the
uniformly refined mesh" against which the matrix-free operators are compared.
"""
import numpy as np
import scipy.sparse as sp

DOFF = np.array([[0, 0, 0], [-1, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [-1, 1, 0], [1, -1, 0],
                 [0, 0, 1], [-1, 0, 1], [0, -1, 1], [1, -1, 1], [0, 0, -1], [0, 1, -1], [1, 0, -1],
                 [-1, 1, -1]])
C, W, E, N, S, NW, SE, TC, TW, TS, TSE, BC, BN, BE, BNW = range(15)
TETS24 = [
    (C, BC, BE, BN), (C, S, SE, TS), (C, W, NW, TW), (C, N, E, TC),
    (C, W, BC, S), (C, E, SE, BE), (C, N, NW, BN), (C, TS, TC, TW),
    (C, BC, BN, BNW), (C, W, S, TS), (C, E, SE, TSE), (C, NW, N, TC),
    (C, BC, S, SE), (C, W, NW, BNW), (C, E, BN, N), (C, TC, TS, TSE),
    (C, W, BC, BNW), (C, E, BE, BN), (C, TC, TW, NW), (C, SE, TS, TSE),
    (C, BC, BE, SE), (C, BN, BNW, NW), (C, E, TSE, TC), (C, W, TS, TW)]


def micro_tets(level):
    """All micro-tetrahedra of one macro-cell as (n,4,3) integer index coordinates."""
    w = 2 ** level + 1
    tets = set()
    for z in range(w):
        for y in range(w - z):
            for x in range(w - y - z):
                p = np.array([x, y, z])
                for t in TETS24:
                    q = p + DOFF[list(t)]
                    if (q >= 0).all() and (q.sum(axis=1) <= w - 1).all():
                        tets.add(tuple(sorted(map(tuple, q))))
    return np.array(sorted(tets), dtype=np.int64)


def tet_volume(P):
    return abs(np.linalg.det(np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]]))) / 6.0


def local_matrix(P, form):
    J = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], axis=1)
    vol = abs(np.linalg.det(J)) / 6.0
    if form == 1:
        return vol / 20.0 * (np.ones((4, 4)) + np.eye(4))
    Ginv = np.linalg.inv(J)
    g = np.vstack([-Ginv.sum(axis=0), Ginv])
    return vol * g @ g.T


def assemble(mesh, level, form=0):
    """Global matrix over the oracle's owned-DoF numbering (matched by coordinates)."""
    coords = mesh.owned_coords(level)
    key = {tuple(np.round(c * 2 ** 20).astype(np.int64)): i for i, c in enumerate(coords)}
    vx = mesh.owned_coords(0)
    nV, nE, nF, nC = mesh.counts()
    w = 2 ** level + 1
    tets = micro_tets(level)
    rows, cols, vals = [], [], []
    for c in range(nC):
        V = vx[mesh.cell_vertices(c)]
        for t in tets:
            P = V[0] + (t[:, 0:1] * (V[1] - V[0]) + t[:, 1:2] * (V[2] - V[0]) + t[:, 2:3] * (V[3] - V[0])) / (w - 1)
            ids = [key[tuple(np.round(p * 2 ** 20).astype(np.int64))] for p in P]
            K = local_matrix(P, form)
            for a in range(4):
                for b in range(4):
                    rows.append(ids[a]); cols.append(ids[b]); vals.append(K[a, b])
    n = len(coords)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
