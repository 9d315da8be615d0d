"""The element clusters of the brick-local stage assembly
(sodhip_element_clusters), checked on the host.

The fused p=3 FP64 stage sums element residuals per cluster in LDS, finishes
the nodes all of whose elements lie in one cluster there, and gathers partial
records for the rest.  These tests check the tables that drive it: every
element in exactly one cluster, 2x2x2 bricks on box meshes, local indices
that point back at the element's own nodes, interior / surface kinds that
follow from the incidence, and a surface incidence that lists each partial
record once.  An assembly simulated with the tables reproduces the plain
node-ordered sum: bit for bit on interior nodes (the same terms in the same
ascending element order), to rounding on surface nodes."""
import ctypes as C

import numpy as np
import pytest

from helpers import make_case


def _clusters(mesh):
    from sod2d_amd._lib import check, lib
    connec = np.asfortranarray(mesh.connec, dtype=np.int32)
    lw = np.ascontiguousarray(mesh.lpoin_w, dtype=np.int32)
    ne, npoin = connec.shape[0], mesh.npoin
    sizes = np.zeros(4, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    args = (3, ne, npoin, p(connec), lw.size, p(lw), p(sizes))
    check(lib().sodhip_element_clusters(*args, *([None] * 8)), "element_clusters")
    ncl, nloc, ns, nsrc = (int(x) for x in sizes)
    t = dict(cel=np.zeros(ncl * 8, np.int32), cloc=np.zeros(ncl * 8 * 64, np.int32),
             cnb=np.zeros(ncl + 1, np.int32), cnode=np.zeros(nloc, np.int32),
             ckind=np.zeros(nloc, np.int32), snode=np.zeros(ns, np.int32),
             srow=np.zeros(ns + 1, np.int32), ssrc=np.zeros(max(nsrc, 1), np.int32))
    check(lib().sodhip_element_clusters(*args, *[p(t[k]) for k in ("cel", "cloc", "cnb", "cnode",
                                                                  "ckind", "snode", "srow",
                                                                  "ssrc")]), "element_clusters")
    t["ssrc"] = t["ssrc"][:nsrc]
    t["ncl"] = ncl
    return t


def _lexperm(oracle):
    T = oracle.Tables(3)
    return (T.gI - 1) + 4 * ((T.gJ - 1) + 4 * (T.gK - 1))  # lexicographic position of node a


@pytest.mark.parametrize("n,periodic", [(4, True), (6, True), (3, True), (5, False), (2, True)])
def test_cluster_tables(oracle, n, periodic):
    g = make_case(3, n, periodic=periodic)
    m = g.mesh
    t = _clusters(m)
    ncl, cel = t["ncl"], t["cel"].reshape(-1, 8)
    members = cel[cel >= 0]
    assert np.array_equal(np.sort(members), np.arange(m.nelem))  # each element once
    for row in cel:  # ascending within a cluster, padding at the end
        k = row[row >= 0]
        assert np.all(np.diff(k) > 0) and np.all(row[len(k):] == -1)
    if periodic and n % 2 == 0:
        # 2x2x2 bricks: 8 elements and 7^3 nodes each (6^3 when one brick
        # wraps the whole periodic box)
        assert ncl == m.nelem // 8 and np.all(cel >= 0)
        assert np.all(np.diff(t["cnb"]) == (343 if n > 2 else 216))
    # local indices point back at the element's own nodes
    lex = _lexperm(oracle)
    connec0 = m.connec - 1
    cnb, cnode, ckind = t["cnb"], t["cnode"], t["ckind"]
    cloc = t["cloc"].reshape(ncl, 8, 64)
    for c in range(ncl):
        loc_nodes = cnode[cnb[c]:cnb[c + 1]]
        assert np.all(np.diff(loc_nodes) > 0)
        for s in range(8):
            e = cel[c, s]
            if e < 0:
                continue
            assert np.array_equal(loc_nodes[cloc[c, s, lex]], connec0[e])
    # kinds: interior <=> every element of the node is in the cluster
    inc = np.bincount(connec0.ravel(), minlength=m.npoin)
    working = np.zeros(m.npoin, bool)
    working[m.lpoin_w - 1] = True
    for c in range(ncl):
        es = cel[c][cel[c] >= 0]
        cnt = np.bincount(connec0[es].ravel(), minlength=m.npoin)
        for l in range(cnb[c], cnb[c + 1]):
            nd = cnode[l]
            want = 2 if not working[nd] else (0 if cnt[nd] == inc[nd] else 1)
            assert ckind[l] == want, (c, l)
    # surface incidence: the working nodes not interior anywhere, in list
    # order, each with its partial records once, in cluster order
    interior = np.zeros(m.npoin, bool)
    interior[cnode[ckind == 0]] = True
    expect = [x for x in m.lpoin_w if not interior[x - 1]]
    assert np.array_equal(t["snode"], expect)
    recs = np.sort(np.nonzero(ckind == 1)[0])
    assert np.array_equal(np.sort(t["ssrc"]), recs)
    for i, nd in enumerate(t["snode"]):
        r = t["ssrc"][t["srow"][i]:t["srow"][i + 1]]
        assert np.all(cnode[r] == nd - 1) and np.all(np.diff(r) > 0)


@pytest.mark.parametrize("n", [4, 3])
def test_cluster_assembly_matches_node_ordered_sum(oracle, n):
    """Random element-node contributions assembled (a) per node in ascending
    element order, as the node-ordered gather does, and (b) with the cluster
    tables: per cluster in its element order, interior nodes finished there,
    surface nodes as the sum of their partial records."""
    g = make_case(3, n, periodic=True)
    m = g.mesh
    t = _clusters(m)
    rng = np.random.default_rng(3)
    contrib = rng.normal(size=(m.nelem, 64))  # by element-local node a
    connec0 = m.connec - 1
    plain = np.zeros(m.npoin)
    for e in range(m.nelem):  # ascending element: the reference's serial order
        for a in range(64):
            plain[connec0[e, a]] += contrib[e, a]
    lex = _lexperm(oracle)
    ncl, cel, cnb, cnode, ckind = t["ncl"], t["cel"].reshape(-1, 8), t["cnb"], t["cnode"], \
        t["ckind"]
    cloc = t["cloc"].reshape(ncl, 8, 64)
    out = np.full(m.npoin, np.nan)
    rec = np.zeros(cnode.size)
    for c in range(ncl):
        acc = np.zeros(cnb[c + 1] - cnb[c])
        for s in range(8):
            e = cel[c, s]
            if e < 0:
                continue
            for a in range(64):
                acc[cloc[c, s, lex[a]]] += contrib[e, a]
        for l in range(acc.size):
            if ckind[cnb[c] + l] == 0:
                out[cnode[cnb[c] + l]] = acc[l]
            elif ckind[cnb[c] + l] == 1:
                rec[cnb[c] + l] = acc[l]
    for i, nd in enumerate(t["snode"]):
        s = 0.0
        for r in t["ssrc"][t["srow"][i]:t["srow"][i + 1]]:
            s += rec[r]
        out[nd - 1] = s
    w = m.lpoin_w - 1
    interior = np.zeros(m.npoin, bool)
    interior[cnode[ckind == 0]] = True
    assert np.array_equal(out[interior], plain[interior])
    assert np.all(np.isfinite(out[w]))
    assert np.abs(out[w] - plain[w]).max() <= 1e-14 * np.abs(plain[w]).max()


def test_cluster_tables_reject_repeated_nodes():
    """A periodic mesh one element wide lists nodes twice in its element: no
    clusters (the stage then keeps the node-ordered gather)."""
    from sod2d_amd._lib import lib
    g = make_case(3, 1, periodic=True)
    m = g.mesh
    connec = np.asfortranarray(m.connec, dtype=np.int32)
    lw = np.ascontiguousarray(m.lpoin_w, dtype=np.int32)
    sizes = np.zeros(4, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = lib().sodhip_element_clusters(3, connec.shape[0], m.npoin, p(connec), lw.size, p(lw),
                                      p(sizes), *([None] * 8))
    assert rc != 0
