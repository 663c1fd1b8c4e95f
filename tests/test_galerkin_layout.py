"""CPU checks of the engine's host-side Galerkin set-up (no GPU calls): the
canonical singular rules and the node-overlap scan agree with the oracle, the
singular pair list is every node-sharing pair, and the wave tiles and the
column-slot schedule of the nodal regular kernel are consistent (every
element in exactly one lane, every corner's local row and slot holds its node,
every evicted node is the slot's previous occupant) -- for triangles and for
quadrilaterals."""

import ctypes as C

import numpy as np
import pytest

from tests import bemref

I32 = C.POINTER(C.c_int32)
D = C.POINTER(C.c_double)
TILE, CS, KMAX = 64, 8, 8


@pytest.fixture(scope="module")
def raw():
    from nihu_amd import nbem
    l = nbem.lib()
    l.nbg_singular_rule.argtypes = [C.c_int, D, C.c_int]
    l.nbg_overlap.argtypes = [I32, I32, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    l.nbg_singular_pairs.argtypes = [C.c_int, C.c_int, I32, I32, I32, I32]
    l.nbg_singular_pairs.restype = C.c_int64
    l.nbg_galerkin_layout.argtypes = [C.c_int, I32, D, C.c_int, I32, I32, I32, I32, I32, I32]
    l.nbg_quad_singular_rule.argtypes = [C.c_int, D, C.c_int]
    l.nbg_overlap_n.argtypes = [I32, I32, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    l.nbg_singular_pairs_n.argtypes = [C.c_int, C.c_int, I32, C.c_int, I32, I32, I32]
    l.nbg_singular_pairs_n.restype = C.c_int64
    l.nbg_galerkin_layout_n.argtypes = [C.c_int, I32, C.c_int, C.c_int, D, C.c_int, I32, I32, I32, I32, I32, I32]
    return l


def _i(a):
    return a.ctypes.data_as(I32)


@pytest.mark.parametrize("match", [0, 1, 2])
def test_singular_rule_matches_oracle(orc, raw, match):
    buf = np.zeros((1536, 5))
    n = raw.nbg_singular_rule(match, buf.ctypes.data_as(D), 1536)
    ref = orc.singular_rule(match)
    assert n == len(ref)
    assert np.abs(buf[:n] - ref).max() <= 1e-16


def test_overlap_scan_matches_oracle(orc, raw):
    rng = np.random.default_rng(5)
    i1, i2 = C.c_int(), C.c_int()
    for _ in range(2000):
        a = rng.choice(6, 3, replace=False).astype(np.int32)
        b = rng.choice(6, 3, replace=False).astype(np.int32)
        n = raw.nbg_overlap(_i(a), _i(b), C.byref(i1), C.byref(i2))
        assert (n, i1.value, i2.value) == orc.overlap(a, b), (a, b)


def _mesh(level=2):
    from nihu_amd.mesh import jittered_icosphere
    nodes, tris = jittered_icosphere(level, seed=1)
    return nodes, np.ascontiguousarray(tris, dtype=np.int32)


def test_singular_pairs_are_node_sharing_pairs(orc, raw):
    nodes, tris = _mesh(2)
    n = len(tris)
    cnt = raw.nbg_singular_pairs(len(nodes), n, _i(tris), None, None, None)
    pe, pf, info = (np.zeros(cnt, np.int32) for _ in range(3))
    assert raw.nbg_singular_pairs(len(nodes), n, _i(tris), _i(pe), _i(pf), _i(info)) == cnt
    assert np.all(pe[:n] == np.arange(n)) and np.all(pf[:n] == np.arange(n)) and np.all(info[:n] == 2)
    ref = {(e, f) for e in range(n) for f in range(n) if e != f and set(tris[e]) & set(tris[f])}
    got = set(zip(pe[n:].tolist(), pf[n:].tolist()))
    assert got == ref and len(got) == cnt - n
    for e, f, inf in zip(pe[n:], pf[n:], info[n:]):
        num, i1, i2 = orc.overlap(tris[e], tris[f])
        assert inf == (1 if num > 1 else 0) | (i1 << 4) | (i2 << 8)


def _layout(raw, nodes, tris, chunk):
    n = len(tris)
    c = np.ascontiguousarray(nodes[tris].mean(1).T)   # [3][n]
    tiles = raw.nbg_galerkin_layout(n, _i(tris), c.ctypes.data_as(D), chunk, None, None, None, None, None, None)
    order = np.zeros(n, np.int32)
    elem, rows = (np.zeros(tiles * TILE, np.int32) for _ in range(2))
    contrib = np.zeros(tiles * TILE * 2, np.uint32)
    col = np.zeros(4 * n, np.int32)
    fin = np.zeros(((n + chunk - 1) // chunk) * CS, np.int32)
    assert raw.nbg_galerkin_layout(n, _i(tris), c.ctypes.data_as(D), chunk, _i(order), _i(elem),
                                   contrib.ctypes.data_as(I32), _i(rows), _i(col), _i(fin)) == tiles
    return tiles, order, elem, contrib.reshape(-1, 2), rows, col.reshape(n, 4), fin.reshape(-1, CS)


@pytest.mark.parametrize("shuffle", [False, True])
def test_wave_tiles_cover_each_element_once(raw, shuffle):
    nodes, tris = _mesh(3)
    if shuffle:   # a mesh without locality: the Morton order takes over
        tris = np.ascontiguousarray(tris[np.random.default_rng(0).permutation(len(tris))])
    tiles, order, elem, contrib, rows, _, _ = _layout(raw, nodes, tris, 128)
    assert sorted(order.tolist()) == list(range(len(tris)))
    used = elem[elem >= 0]
    assert sorted(used.tolist()) == list(range(len(tris)))
    # tiles take elements in sequence order
    assert np.all(used == order)
    for t in range(tiles):
        r = rows[t * TILE:(t + 1) * TILE]
        nr = int((r >= 0).sum())
        assert np.all(r[:nr] >= 0) and np.all(r[nr:] == -1) and len(set(r[:nr].tolist())) == nr
        # each (lane, corner) of the tile contributes to exactly the row of its node
        seen = set()
        for l in range(TILE):
            bytes8 = [(int(contrib[t * TILE + l][q >> 2]) >> (8 * (q & 3))) & 255 for q in range(8)]
            cs = [b for b in bytes8 if b != 255]
            assert bytes8[len(cs):] == [255] * (8 - len(cs)) and len(cs) <= KMAX
            if l >= nr:
                assert not cs
            for b in cs:
                lane, a = b >> 2, b & 3
                e = elem[t * TILE + lane]
                assert a < 3 and e >= 0 and tris[e][a] == r[l]
                seen.add((lane, a))
        assert seen == {(l, a) for l in range(TILE) if elem[t * TILE + l] >= 0 for a in range(3)}
    # a mesh with locality fills its tiles
    if not shuffle:
        assert tiles == (len(tris) + TILE - 1) // TILE


@pytest.mark.parametrize("chunk", [64, 128, 512])
def test_column_slot_schedule_replays(raw, chunk):
    nodes, tris = _mesh(3)
    n = len(tris)
    _, order, _, _, _, col, fin = _layout(raw, nodes, tris, chunk)
    flushed = {}
    for c0 in range(0, n, chunk):
        slot = [-1] * CS
        for i in range(c0, min(n, c0 + chunk)):
            e = order[i]
            sl = [col[i, 3] & 31, (col[i, 3] >> 5) & 31, (col[i, 3] >> 10) & 31]
            assert len(set(sl)) == 3 and max(sl) < CS
            for b in range(3):
                v = col[i, b]
                if v >= 0:
                    assert slot[sl[b]] == v       # evict the current occupant
                    flushed[v] = flushed.get(v, 0) + 1
                    slot[sl[b]] = -1
                if slot[sl[b]] == -1:
                    slot[sl[b]] = tris[e][b]
                assert slot[sl[b]] == tris[e][b]
        assert slot == fin[c0 // chunk].tolist()
        for v in slot:
            if v >= 0:
                flushed[v] = flushed.get(v, 0) + 1
    # every node of the mesh is flushed at least once, and on average few times
    assert set(flushed) == set(range(len(nodes)))
    assert sum(flushed.values()) < 2.5 * len(nodes) * max(1, 512 // chunk)


# ------------------------------------------------------------ quadrilaterals
QCS = 6


@pytest.mark.parametrize("match", [0, 1, 2])
def test_quad_singular_rule_matches_oracle(orc, raw, match):
    buf = np.zeros((2048, 5))
    n = raw.nbg_quad_singular_rule(match, buf.ctypes.data_as(D), 2048)
    ref = orc.quad_singular_rule(match)
    assert n == len(ref)
    assert np.abs(buf[:n] - ref).max() <= 4e-16
    assert raw.nbg_quad_singular_rule(match, buf.ctypes.data_as(D), n - 1) == -1


def test_quad_overlap_scan_matches_oracle(orc, raw):
    rng = np.random.default_rng(6)
    i1, i2 = C.c_int(), C.c_int()
    for _ in range(3000):
        a = rng.choice(7, 4, replace=False).astype(np.int32)
        b = rng.choice(7, 4, replace=False).astype(np.int32)
        n = raw.nbg_overlap_n(_i(a), _i(b), 4, C.byref(i1), C.byref(i2))
        assert (n, i1.value, i2.value) == orc.overlap_quads(a, b), (a, b)


def _qmesh(n=4):
    from nihu_amd.mesh import jittered_cubed_sphere
    nodes, quads = jittered_cubed_sphere(n, seed=1)
    return nodes, np.ascontiguousarray(quads, dtype=np.int32)


def test_quad_singular_pairs_are_node_sharing_pairs(orc, raw):
    nodes, quads = _qmesh(4)
    n = len(quads)
    cnt = raw.nbg_singular_pairs_n(len(nodes), n, _i(quads), 4, None, None, None)
    pe, pf, info = (np.zeros(cnt, np.int32) for _ in range(3))
    assert raw.nbg_singular_pairs_n(len(nodes), n, _i(quads), 4, _i(pe), _i(pf), _i(info)) == cnt
    assert np.all(pe[:n] == np.arange(n)) and np.all(info[:n] == 2)
    ref = {(e, f) for e in range(n) for f in range(n) if e != f and set(quads[e]) & set(quads[f])}
    got = set(zip(pe[n:].tolist(), pf[n:].tolist()))
    assert got == ref and len(got) == cnt - n
    for e, f, inf in zip(pe[n:], pf[n:], info[n:]):
        num, i1, i2 = orc.overlap_quads(quads[e], quads[f])
        assert inf == (1 if num > 1 else 0) | (i1 << 4) | (i2 << 8)


def _qlayout(raw, nodes, quads, chunk):
    n = len(quads)
    c = np.ascontiguousarray(nodes[quads].mean(1).T)
    tiles = raw.nbg_galerkin_layout_n(n, _i(quads), 4, QCS, c.ctypes.data_as(D), chunk, None, None, None, None,
                                      None, None)
    order = np.zeros(n, np.int32)
    elem = np.zeros(tiles * TILE, np.int32)
    rows = np.zeros(tiles * TILE * 2, np.int32)          # two local rows per lane
    contrib = np.zeros(tiles * TILE * 2 * 2, np.uint32)
    col = np.zeros(5 * n, np.int32)
    fin = np.zeros(((n + chunk - 1) // chunk) * QCS, np.int32)
    assert raw.nbg_galerkin_layout_n(n, _i(quads), 4, QCS, c.ctypes.data_as(D), chunk, _i(order), _i(elem),
                                     contrib.ctypes.data_as(I32), _i(rows), _i(col), _i(fin)) == tiles
    return tiles, order, elem, contrib.reshape(-1, 2), rows, col.reshape(n, 5), fin.reshape(-1, QCS)


def test_quad_layout_rejects_bad_arguments(raw):
    nodes, quads = _qmesh(2)
    c = np.ascontiguousarray(nodes[quads].mean(1).T)
    for npe, cs in ((4, 3), (4, 9), (5, 6), (2, 6)):
        assert raw.nbg_galerkin_layout_n(len(quads), _i(quads), npe, cs, c.ctypes.data_as(D), 64, None, None,
                                         None, None, None, None) == -1


def test_quad_wave_tiles_and_slot_schedule(raw):
    nodes, quads = _qmesh(6)
    n = len(quads)
    for chunk in (64, 256):
        tiles, order, elem, contrib, rows, col, fin = _qlayout(raw, nodes, quads, chunk)
        assert sorted(order.tolist()) == list(range(n))
        used = elem[elem >= 0]
        assert np.all(used == order)
        # 64 quadrilaterals touch about 81 nodes: the two rows per lane let
        # the tiles fill their lanes
        assert tiles <= (n + TILE - 1) // TILE + 1
        for t in range(tiles):
            r = rows[2 * t * TILE:2 * (t + 1) * TILE]
            nr = int((r >= 0).sum())
            assert len(set(r[:nr].tolist())) == nr and np.all(r[nr:] == -1)
            seen = set()
            for l in range(2 * TILE):
                bytes8 = [(int(contrib[2 * t * TILE + l][q >> 2]) >> (8 * (q & 3))) & 255 for q in range(8)]
                cs = [b for b in bytes8 if b != 255]
                for b in cs:
                    lane, a = b >> 2, b & 3
                    e = elem[t * TILE + lane]
                    assert e >= 0 and quads[e][a] == r[l]
                    seen.add((lane, a))
            assert seen == {(l, a) for l in range(TILE) if elem[t * TILE + l] >= 0 for a in range(4)}
        flushed = {}
        for c0 in range(0, n, chunk):
            slot = [-1] * QCS
            for i in range(c0, min(n, c0 + chunk)):
                e = order[i]
                sl = [(col[i, 4] >> (5 * b)) & 31 for b in range(4)]
                assert len(set(sl)) == 4 and max(sl) < QCS
                for b in range(4):
                    v = col[i, b]
                    if v >= 0:
                        assert slot[sl[b]] == v
                        flushed[v] = flushed.get(v, 0) + 1
                        slot[sl[b]] = -1
                    if slot[sl[b]] == -1:
                        slot[sl[b]] = quads[e][b]
                    assert slot[sl[b]] == quads[e][b]
            assert slot == fin[c0 // chunk].tolist()
            for v in slot:
                if v >= 0:
                    flushed[v] = flushed.get(v, 0) + 1
        assert set(flushed) == set(range(len(nodes)))


@pytest.mark.parametrize("npe", [3, 4])
def test_element_dof_map_layout(raw, npe):
    """The dof map of a constant field in a pair with a nodal one names the
    element at every corner: each local row is one element with its npe
    corners as contributors, and every corner of a trial element shares one
    column slot (one eviction per element)."""
    if npe == 3:
        nodes, el = _mesh(3)
    else:
        nodes, el = _qmesh(6)
    n = len(el)
    dm = np.ascontiguousarray(np.repeat(np.arange(n, dtype=np.int32)[:, None], npe, axis=1))
    c = np.ascontiguousarray(nodes[el].mean(1).T)
    cs, rpl, chunk = (8, 1, 128) if npe == 3 else (QCS, 2, 64)
    tiles = raw.nbg_galerkin_layout_n(n, _i(dm), npe, cs, c.ctypes.data_as(D), chunk, None, None, None, None,
                                      None, None)
    order = np.zeros(n, np.int32)
    elem = np.zeros(tiles * TILE, np.int32)
    rows = np.zeros(tiles * TILE * rpl, np.int32)
    contrib = np.zeros(tiles * TILE * rpl * 2, np.uint32)
    col = np.zeros((npe + 1) * n, np.int32)
    fin = np.zeros(((n + chunk - 1) // chunk) * cs, np.int32)
    assert raw.nbg_galerkin_layout_n(n, _i(dm), npe, cs, c.ctypes.data_as(D), chunk, _i(order), _i(elem),
                                     contrib.ctypes.data_as(I32), _i(rows), _i(col), _i(fin)) == tiles
    contrib = contrib.reshape(-1, 2)
    col = col.reshape(n, npe + 1)
    assert tiles == (n + (TILE - (npe == 4)) - 1) // (TILE - (npe == 4))
    for t in range(tiles):
        for l in range(TILE):
            e = elem[t * TILE + l]
            if e < 0:
                continue
            # the lane's element is one local row with its npe corners
            r = [j for j in range(rpl * TILE) if rows[t * TILE * rpl + j] == e]
            assert len(r) == 1
            bytes8 = [(int(contrib[t * TILE * rpl + r[0]][q >> 2]) >> (8 * (q & 3))) & 255 for q in range(8)]
            assert sorted(b for b in bytes8 if b != 255) == [l * 4 + a for a in range(npe)]
    for i in range(n):
        sl = [(col[i, npe] >> (5 * b)) & 31 for b in range(npe)]
        assert len(set(sl)) == 1
        assert (col[i, 1:npe] == -1).all()
