"""Mixed-element box meshes for the tests: an nx x ny x nz grid of hexahedral
blocks in which each (i, j) column is one of
  'hex'   - one hexahedron per block,
  'wedge' - two wedges per block, split through the block's (x, y) diagonal
            (the same diagonal all the way up, so triangles meet triangles),
  'pyr'   - six pyramids per block with their apex at the block centre,
  'tet'   - six tetrahedra per block: one block face's pyramid with its base
            split into two triangles is not conforming, so a tetrahedron column
            is only placed where all four side neighbours are 'tet' columns or
            the boundary (used by the all-tetrahedron consistency case).
Every element is written in VTK node order with positive volume in the
reference's formulas; interior nodes are jittered.  Boundary faces (quads or
triangles) go to six patches x-, x+, y-, y+, z-, z+."""
import numpy as np

import netflux_amd as nf

HEX_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (3, 2, 6, 7), (3, 0, 4, 7), (1, 2, 6, 5)]


def _normal(P, f):
    p = P[list(f)]
    if len(f) == 3:
        return np.cross(p[1] - p[0], p[2] - p[0])
    return np.cross(p[2] - p[0], p[3] - p[1])


def mixed_box(n=(4, 3, 3), kinds=("hex", "wedge", "pyr"), jitter=0.1, seed=7, lengths=(1.0, 1.0, 1.0),
              patches=("SlipWall",) * 6, column_kind=None):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    h = np.array(lengths, float) / np.array([nx, ny, nz], float)
    grid = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij"), -1)
    X = (grid * h).reshape(-1, 3).astype(float)
    idx = np.arange(len(X)).reshape(nx + 1, ny + 1, nz + 1)
    inner = np.ones((nx + 1, ny + 1, nz + 1), bool)
    inner[[0, -1], :, :] = inner[:, [0, -1], :] = inner[:, :, [0, -1]] = False
    X[inner.ravel()] += (rng.random((int(inner.sum()), 3)) - 0.5) * 2 * jitter * h
    nodes = [x for x in X]

    def kind(i, j):
        if column_kind is not None:
            return column_kind(i, j)
        return kinds[(i + 2 * j) % len(kinds)]

    cells, bfaces = [], {s: [] for s in range(6)}

    def add_boundary(face_nodes, i, j, k, c):
        """face_nodes on the block boundary of block (i,j,k)? -> patch list"""
        pts = np.array([nodes[v] for v in face_nodes])
        for s, (axis, end) in enumerate([(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)]):
            b = (i, j, k)[axis]
            lim = 0 if end == 0 else (nx, ny, nz)[axis] - 1
            if b != lim:
                continue
            target = 0.0 if end == 0 else lengths[axis]
            if np.allclose(pts[:, axis], target):
                bfaces[s].append(list(face_nodes))

    for i in range(nx):
        for j in range(ny):
            kd = kind(i, j)
            for k in range(nz):
                v = [idx[i, j, k], idx[i + 1, j, k], idx[i + 1, j + 1, k], idx[i, j + 1, k],
                     idx[i, j, k + 1], idx[i + 1, j, k + 1], idx[i + 1, j + 1, k + 1], idx[i, j + 1, k + 1]]
                P = np.array([nodes[q] for q in v])
                ctr = P.mean(axis=0)
                if kd == "hex":
                    cells.append(v)
                    for f in HEX_FACES:
                        add_boundary([v[q] for q in f], i, j, k, ctr)
                elif kd == "pyr":
                    nodes.append(ctr.copy())
                    apex = len(nodes) - 1
                    for f in HEX_FACES:
                        base = [v[q] for q in f]
                        # VTK pyramid: base normal toward the apex
                        pts = np.array([nodes[q] for q in base])
                        nrm = np.cross(pts[2] - pts[0], pts[3] - pts[1])
                        if np.dot(nrm, ctr - pts.mean(axis=0)) < 0:
                            base = base[::-1]
                        cells.append(base + [apex])
                        add_boundary([v[q] for q in f], i, j, k, ctr)
                elif kd == "wedge":
                    # triangles (0,1,2) / (0,2,3) of the bottom square, lifted
                    for tri in ((0, 1, 2), (0, 2, 3)):
                        bot = [v[q] for q in tri]
                        top = [v[q + 4] for q in tri]
                        pb = np.array([nodes[q] for q in bot])
                        # VTK wedge: (0,1,2) normal points away from (3,4,5)
                        if np.dot(np.cross(pb[1] - pb[0], pb[2] - pb[0]), np.array([0, 0, 1.0])) > 0:
                            bot, top = [bot[0], bot[2], bot[1]], [top[0], top[2], top[1]]
                        cells.append(bot + top)
                        add_boundary(bot, i, j, k, ctr)
                        add_boundary(top, i, j, k, ctr)
                        for a, b in ((0, 1), (1, 2), (2, 0)):
                            add_boundary([bot[a], bot[b], top[b], top[a]], i, j, k, ctr)
                elif kd == "tet":
                    # Kuhn split along the 0-6 diagonal (the six tetrahedra
                    # of the cube's coordinate orderings)
                    for path in ((1, 2), (1, 5), (3, 2), (3, 7), (4, 5), (4, 7)):
                        t = [v[0], v[path[0]], v[path[1]], v[6]]
                        pt = np.array([nodes[q] for q in t])
                        if np.dot(np.cross(pt[1] - pt[0], pt[2] - pt[0]), pt[3] - pt[0]) < 0:
                            t = [t[0], t[2], t[1], t[3]]
                        cells.append(t)
                        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
                            add_boundary([t[q] for q in f], i, j, k, ctr)
    patch_faces = [bfaces[s] for s in range(6)]
    return nf.Mesh.from_elements(np.array(nodes), cells, patch_faces, list(patches))
