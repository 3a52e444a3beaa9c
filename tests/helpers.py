"""Shared setup for the parity tests: the same mesh built in the oracle and in hyteg_amd."""
import os
import sys

import numpy as np

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def read_gmsh(path):
    """Nodes and tetrahedra of a Gmsh 2.2 ASCII file (test-side reader for the oracle)."""
    lines = open(path).read().split("\n")
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    ids, X = [], []
    for k in range(n):
        p = lines[i + 2 + k].split()
        ids.append(int(p[0]))
        X.append([float(v) for v in p[1:4]])
    idmap = {v: k for k, v in enumerate(ids)}
    j = lines.index("$Elements")
    m = int(lines[j + 1])
    cells = []
    for k in range(m):
        p = [int(v) for v in lines[j + 2 + k].split()]
        if p[1] == 4:
            nt = p[2]
            cells.append([idmap[v] for v in p[3 + nt:3 + nt + 4]])
    return np.array(X), np.array(cells, dtype=np.int32)


def mesh_pair(kind, host_ctx=None):
    """(oracle Mesh, hyteg_amd PrimitiveStorage) for kind = ('cube', n) or ('gmsh', filename)."""
    import hyteg_amd as H
    import oracle as O
    if kind[0] == "cube":
        n = kind[1]
        om = O.Mesh.cuboid(n)
        S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), n, n, n), host_ctx)
    else:
        path = os.path.join(MESHES, kind[1])
        X, C = read_gmsh(path)
        om = O.Mesh.from_arrays(X, C)
        S = H.PrimitiveStorage(H.MeshInfo.fromGmshFile(path), host_ctx)
    return om, S


def sine_rhs(X):
    return 3 * np.pi ** 2 * np.prod(np.sin(np.pi * X), axis=1)


def sine_exact(X):
    return np.prod(np.sin(np.pi * X), axis=1)


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale
