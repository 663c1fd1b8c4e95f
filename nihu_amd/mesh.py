"""Meshes for the BEM engine: seeded icospheres, ellipsoids, cubed spheres
(quadrilaterals), OFF files.

A mesh is a pair of arrays, ``nodes`` (n_nodes x 3, float64) and ``tris``
(n_tri x 3, int32), elements numbered by row (NiHu numbers elements by
insertion order, nihu/core/mesh.hpp:137-193).  Triangles are oriented so that
(X1 - X0) x (X2 - X0) points outwards.

Icosphere: the 20 faces of an icosahedron are subdivided ``level`` times
depth first (each parent's four children are consecutive, so consecutive
element indices form compact patches), with edge midpoints projected to the
unit sphere.  Levels 3/4/5 give 1280/5120/20480 triangles and 642/2562/10242
vertices.  Radial jitter: every vertex is scaled by ``1 + U(-jitter, jitter)``
drawn from ``numpy.random.default_rng(seed)`` (PCG64) in vertex order.
"""

from __future__ import annotations

import numpy as np

NIHU_TRIA_1_ID = 32303  # element type id of tria_1_elem, nihu/core/element.hpp:112-118
NIHU_QUAD_1_ID = 32404  # quad_1_elem (space dimension 3, quad domain, 4 nodes)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    faces = []
    for a, b, c in f:
        n = np.cross(v[b] - v[a], v[c] - v[a])
        if np.dot(n, v[a] + v[b] + v[c]) < 0:
            b, c = c, b
        faces.append((a, b, c))
    return v, faces


def icosphere(level: int, radius: float = 1.0, jitter: float = 0.0, seed: int = 0):
    """Return (nodes, tris) of a subdivided icosahedron on a sphere of ``radius``."""
    if level < 0:
        raise ValueError("level must be >= 0")
    v0, faces = _icosahedron()
    verts = [tuple(p) for p in v0]
    mid_cache: dict[tuple[int, int], int] = {}

    def midpoint(a, b):
        key = (a, b) if a < b else (b, a)
        idx = mid_cache.get(key)
        if idx is None:
            p = np.add(verts[a], verts[b]) * 0.5
            p = p / np.linalg.norm(p)
            idx = len(verts)
            verts.append(tuple(p))
            mid_cache[key] = idx
        return idx

    out = []

    def subdivide(a, b, c, lev):
        if lev == 0:
            out.append((a, b, c))
            return
        ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
        subdivide(a, ab, ca, lev - 1)
        subdivide(ab, b, bc, lev - 1)
        subdivide(ca, bc, c, lev - 1)
        subdivide(ab, bc, ca, lev - 1)

    for a, b, c in faces:
        subdivide(a, b, c, level)
    nodes = np.asarray(verts, dtype=np.float64) * radius
    if jitter:
        rng = np.random.default_rng(seed)
        u = rng.uniform(-jitter, jitter, size=nodes.shape[0])
        nodes = nodes * (1.0 + u)[:, None]
    return np.ascontiguousarray(nodes), np.ascontiguousarray(np.asarray(out, dtype=np.int32))


def jittered_icosphere(level: int, seed: int = 0):
    """The configuration meshes: unit icosphere, +-1% seeded radial jitter."""
    return icosphere(level, 1.0, 0.01, seed)


def ellipsoid(level: int = 5, seed: int = 0, jitter: float = 0.01):
    """Icosphere scaled by seeded semi-axes U[0.5, 1.5] (drawn first from the
    same generator), then radially jittered by 1 + U(-jitter, jitter)."""
    nodes, tris = icosphere(level, 1.0, 0.0)
    rng = np.random.default_rng(seed)
    axes = rng.uniform(0.5, 1.5, size=3)
    u = rng.uniform(-jitter, jitter, size=nodes.shape[0])
    nodes = nodes * axes[None, :] * (1.0 + u)[:, None]
    return np.ascontiguousarray(nodes), tris


def cubed_sphere(n: int, radius: float = 1.0, jitter: float = 0.0, seed: int = 0):
    """Quadrilateral mesh of a sphere: each face of the cube [-1, 1]^3 cut
    into n x n squares, nodes projected radially to the sphere (so the
    quadrilaterals are warped bilinear elements), 6 n^2 elements and
    6 n^2 + 2 nodes.  Corners ordered so that (X1 - X0) x (X3 - X0) points
    outwards; elements numbered face by face, row by row.  Radial jitter as
    for the icosphere: 1 + U(-jitter, jitter) per node in node order."""
    index = {}
    nodes = []

    def node(p):
        key = tuple(int(round(v)) for v in p)
        if key not in index:
            index[key] = len(nodes)
            nodes.append(key)
        return index[key]

    quads = []
    # (normal axis, sign): the two in-plane axes (u, v) with u x v along the outward normal
    for axis, sign in ((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)):
        u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
        if sign < 0:
            u_ax, v_ax = v_ax, u_ax
        for j in range(n):
            for i in range(n):
                corner = []
                for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    p = [0, 0, 0]
                    p[axis] = sign * n
                    p[u_ax] = 2 * (i + di) - n
                    p[v_ax] = 2 * (j + dj) - n
                    corner.append(node(p))
                quads.append(corner)
    X = np.array(nodes, dtype=np.float64)
    X /= np.linalg.norm(X, axis=1)[:, None]
    X *= radius
    if jitter > 0.0:
        u = np.random.default_rng(seed).uniform(-jitter, jitter, size=len(X))
        X = X * (1.0 + u)[:, None]
    return np.ascontiguousarray(X), np.ascontiguousarray(np.array(quads, dtype=np.int32))


def jittered_cubed_sphere(n: int, seed: int = 0):
    """Unit cubed sphere with +-1% seeded radial jitter."""
    return cubed_sphere(n, 1.0, 0.01, seed)


def mixed_cubed_sphere(n: int, seed: int = 0, every: int = 3, jitter: float = 0.01):
    """A closed surface with both element types: the cubed sphere of n x n
    quadrilaterals per face (radial jitter +-jitter from seed; 0 gives NiHu's
    blown-up cube sphere) with every `every`-th quadrilateral split along its
    diagonal (0, 2) into two triangles of the same orientation.  Returns
    (nodes, table) with NiHu's element table rows [type id, n0, n1, n2, n3]
    (n3 = 0 for triangles), the two types interleaved."""
    nodes, quads = cubed_sphere(n, 1.0, jitter, seed)
    rows = []
    for e, q in enumerate(quads):
        if e % every == 0:
            rows.append([NIHU_TRIA_1_ID, q[0], q[1], q[2], 0])
            rows.append([NIHU_TRIA_1_ID, q[0], q[2], q[3], 0])
        else:
            rows.append([NIHU_QUAD_1_ID, q[0], q[1], q[2], q[3]])
    return nodes, np.array(rows, dtype=np.int32)


def table_faces(table):
    """Node indices of each row of NiHu's element table (3 or 4 per row)."""
    return [list(r[1:4]) if r[0] == NIHU_TRIA_1_ID else list(r[1:5]) for r in np.asarray(table)]


def read_off(path):
    """Read an OFF file of triangles or of quadrilaterals
    (nihu/interface/read_off_mesh.hpp:75-105); returns (nodes, elements) with
    3 or 4 node indices per element.

    Raises ValueError for a bad header, truncated data, faces that are
    neither triangles nor quadrilaterals (NiHu raises std::runtime_error in the
    same places) and for files that mix the two (read_off_elements reads
    those)."""
    nodes, elems = _read_off_faces(path)
    sizes = {len(e) for e in elems}
    if len(sizes) > 1:
        raise ValueError("triangles and quadrilaterals in one file: use read_off_elements")
    npe = sizes.pop() if sizes else 3
    return nodes, np.array(elems, dtype=np.int32).reshape(len(elems), npe)


def read_off_elements(path):
    """Read an OFF file of triangles, quadrilaterals or both into NiHu's
    element table, as read_off_data does with the tags tria_1 and quad_1
    (nihu/interface/read_off_mesh.hpp:75-105): rows [type id, n0, n1, n2, n3],
    32303 for a triangle (n3 = 0) and 32404 for a quadrilateral, in file
    order.  Returns (nodes, table); nbem.Mesh.from_elements takes the table."""
    nodes, elems = _read_off_faces(path)
    table = np.zeros((len(elems), 5), dtype=np.int32)
    for e, f in enumerate(elems):
        table[e, 0] = NIHU_TRIA_1_ID if len(f) == 3 else NIHU_QUAD_1_ID
        table[e, 1:1 + len(f)] = f
    return nodes, table


def _read_off_faces(path):
    with open(path, "r") as fh:
        tok = fh.read().split()
    if not tok or tok[0] != "OFF":
        raise ValueError("Possibly invalid off file")
    try:
        n_nodes, n_elem = int(tok[1]), int(tok[2])
    except (IndexError, ValueError):
        raise ValueError("Error reading number of mesh entries")
    pos = 4
    try:
        nodes = np.array(tok[pos:pos + 3 * n_nodes], dtype=np.float64).reshape(n_nodes, 3)
    except ValueError:
        raise ValueError("Error reading mesh nodes")
    pos += 3 * n_nodes
    elems = []
    for e in range(n_elem):
        if pos >= len(tok):
            raise ValueError("Error reading mesh elements")
        nv = int(tok[pos])
        if nv not in (3, 4):
            raise ValueError("Invalid number of vertices for current element tag vector")
        if pos + 1 + nv > len(tok):
            raise ValueError("Error reading mesh elements")
        elems.append([int(t) for t in tok[pos + 1:pos + 1 + nv]])
        pos += 1 + nv
    return nodes, elems


def write_off(path, nodes, elems):
    """elems: rows of 3 or 4 node indices (a list may mix the two)."""
    with open(path, "w") as fh:
        fh.write("OFF\n%d %d 0\n" % (len(nodes), len(elems)))
        for p in nodes:
            fh.write("%.17g %.17g %.17g\n" % tuple(p))
        for t in elems:
            fh.write(("%d" + " %d" * len(t)) % ((len(t),) + tuple(t)) + "\n")


def to_nihu_elements(elems):
    """NiHu's element matrix rows [type_id, n0, n1, n2(, n3)]
    (nihu/core/mesh.hpp:137-193): 32303 for triangles, 32404 for
    quadrilaterals."""
    elems = np.asarray(elems)
    tid = NIHU_TRIA_1_ID if elems.shape[1] == 3 else NIHU_QUAD_1_ID
    return np.hstack([np.full((len(elems), 1), tid, dtype=np.int64), elems.astype(np.int64)])


def from_nihu_elements(elements):
    """Node indices from NiHu's element matrix: all tria_1 (32303) or all
    quad_1 (32404) rows."""
    elements = np.asarray(elements)
    if elements.ndim != 2 or elements.shape[1] < 4 or len(elements) == 0:
        raise ValueError("only tria_1 (32303) or quad_1 (32404) elements are supported")
    if np.all(elements[:, 0] == NIHU_TRIA_1_ID):
        return np.ascontiguousarray(elements[:, 1:4].astype(np.int32))
    if np.all(elements[:, 0] == NIHU_QUAD_1_ID) and elements.shape[1] >= 5:
        return np.ascontiguousarray(elements[:, 1:5].astype(np.int32))
    raise ValueError("a table with both tria_1 (32303) and quad_1 (32404) rows: use nbem.Mesh.from_elements")


def sphere_boundary(r0: float, n0: int):
    """NiHu's blown-up cube sphere (nihu/matlab/meshing/create_sphere_boundary.m):
    the boundary of the cube [-1, 1]^3 with 2 n0 quadrilaterals along each
    edge, nodes moved radially onto the sphere of radius r0 (node * max|x_i| /
    |x|, i.e. node / |node| on the cube's surface), normals outwards;
    6 (2 n0)^2 elements.  The node set and the elements are those of the
    MATLAB generator (numbering and corner start may differ; NiHu's
    quad_1 integrals do not depend on them beyond rounding)."""
    return cubed_sphere(2 * n0, radius=r0)


def create_slab(corners, n1: int, n2: int):
    """A bilinear patch of n1 x n2 quadrilaterals
    (nihu/matlab/meshing/create_slab.m with four corner rows): the grid of
    [-1, 1]^2, n1 segments along the first reference axis, n2 along the
    second, mapped by the bilinear shape functions of corners (4 x 3, ordered
    (-1,-1), (1,-1), (1,1), (-1,1)); nodes row by row, element (i, j) =
    (n(i, j), n(i+1, j), n(i+1, j+1), n(i, j+1))."""
    R = np.asarray(corners, dtype=np.float64).reshape(4, 3)
    xi = np.linspace(-1.0, 1.0, n1 + 1)
    eta = np.linspace(-1.0, 1.0, n2 + 1)
    E, X = np.meshgrid(eta, xi, indexing="ij")        # node (i, j) at row j * (n1 + 1) + i
    X, E = X.ravel(), E.ravel()
    phi = np.stack([(1 - X) * (1 - E), (1 + X) * (1 - E), (1 + X) * (1 + E), (1 - X) * (1 + E)], axis=1) / 4.0
    nodes = phi @ R
    idx = lambda i, j: j * (n1 + 1) + i
    quads = [[idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)] for j in range(n2) for i in range(n1)]
    return np.ascontiguousarray(nodes), np.ascontiguousarray(np.array(quads, dtype=np.int32))


def select_elements(nodes, elems, node_predicate):
    """Elements all of whose nodes satisfy node_predicate(nodes) (a boolean
    per node): mesh_select(..., 'all') of nihu/matlab/meshing/mesh_select.m.
    Returns (nodes, kept elements) with the node table unchanged."""
    ok = np.asarray(node_predicate(np.asarray(nodes)), dtype=bool)
    keep = ok[np.asarray(elems)].all(axis=1)
    return nodes, np.ascontiguousarray(np.asarray(elems)[keep])


def edge_size(nodes, elems):
    """Longest edge of each element (nihu/matlab/meshing/mesh_edge_size.m)."""
    X = np.asarray(nodes)[np.asarray(elems)]
    m = X.shape[1]
    return np.max(np.stack([np.linalg.norm(X[:, (a + 1) % m] - X[:, a], axis=1) for a in range(m)], axis=1), axis=1)


def mesh_kmax(nodes, elems, ratio: float = 8.0):
    """Largest wavenumber per element at `ratio` elements per wavelength:
    2 pi / (edge size * ratio) (nihu/matlab/meshing/mesh_kmax.m)."""
    return 2.0 * np.pi / (edge_size(nodes, elems) * ratio)
