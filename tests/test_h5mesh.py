"""SOD2D HDF5 mesh and restart files (sod2d_amd/h5mesh.py over libsod2d_h5.so).

The reader is pinned against files built by an independent writer, HDF5's own
h5import tool, from raw arrays in the layout restated below from the
reference's mesh partitioner and loader (mod_hdf5.f90); the writer is pinned
by h5dump's view of what it wrote (names, element types, C-order shapes)
against the same table; the restart conventions (the swapped mu_e / mu_t
names, the element-order scatter of element-node fields, the periodic copy)
are checked on brute-force restatements; the derived state after a load is
checked through identities of the ideal-gas state, not by retyping it.
The order tables at p = 2 are the 27-node hexahedra of VTK
(VTK_TRIQUADRATIC_HEXAHEDRON) and gmsh (its 27-node hexahedron), as documented
by those projects.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from sod2d_amd import case
from sod2d_amd import h5mesh as H
from sod2d_amd._lib import SodhipError

H5IMPORT = shutil.which("h5import") or "/opt/conda/bin/h5import"
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
need_tools = pytest.mark.skipif(not (os.path.exists(H5IMPORT) and os.path.exists(H5DUMP)),
                                reason="HDF5 command-line tools (h5import, h5dump) not installed")

# h5import classes of the reference's element types (mod_hdf5.f90:27-33):
# int4 = H5T_NATIVE_INTEGER, int8 = H5T_STD_I64LE, uint1 = H5T_STD_U8LE, real8, real4
_CLS = {"i4": ("IN", 32, "STD", "<i4"), "i8": ("IN", 64, "STD", "<i8"),
        "u1": ("UIN", 8, "STD", "u1"), "f8": ("FP", 64, "IEEE", "<f8"),
        "f4": ("FP", 32, "IEEE", "<f4")}


def h5import(path, datasets, work):
    """Build an HDF5 file with h5import from (name, array, type) triples; a 2-D
    array is stored with its C-order shape."""
    args = []
    for k, (name, a, t) in enumerate(datasets):
        cls, bits, arch, dt = _CLS[t]
        a = np.ascontiguousarray(a, dtype=dt)
        b, c = os.path.join(work, f"d{k}.bin"), os.path.join(work, f"d{k}.cfg")
        a.tofile(b)
        lines = [f"PATH {name}", f"INPUT-CLASS {cls}", f"INPUT-SIZE {bits}"]
        if bits > 8:
            lines.append("INPUT-BYTE-ORDER LE")
        lines += [f"RANK {a.ndim}", "DIMENSION-SIZES " + " ".join(str(s) for s in a.shape),
                  f"OUTPUT-CLASS {cls}", f"OUTPUT-SIZE {bits}", f"OUTPUT-ARCHITECTURE {arch}",
                  "OUTPUT-BYTE-ORDER LE"]
        with open(c, "w") as f:
            f.write("\n".join(lines) + "\n")
        args.append([b, "-c", c])
    # h5import takes at most 30 inputs a call; later calls add to the file
    for k in range(0, len(args), 25):
        cmd = [H5IMPORT] + sum(args[k:k + 25], []) + ["-o", str(path)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr


def dump_type(path, name):
    """(kind, bytes, C-order dims) of a dataset as h5dump reports it."""
    r = subprocess.run([H5DUMP, "-H", "-d", name, str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    t = re.search(r"DATATYPE\s+H5T_(STD|IEEE)_([IUF])(\d+)LE", r.stdout)
    d = re.search(r"DATASPACE\s+SIMPLE\s*\{\s*\(([^)]*)\)", r.stdout)
    assert t and d, r.stdout
    return (t.group(2).lower(), int(t.group(3)) // 8,
            tuple(int(x) for x in d.group(1).split(",")))


def vtk_connec(m, tabs):
    """An element's nodes in VTK point order, part-local and 0-based (the
    partitioner's connecVTK - 1, mod_hdf5.f90:6030-6037): VTK point
    vtk2ijk(c) is local node a2ijk(c)."""
    a_of_v = np.empty(len(tabs["a2ijk"]), np.int64)
    a_of_v[np.asarray(tabs["vtk2ijk"]) - 1] = np.asarray(tabs["a2ijk"]) - 1
    return (m.connec_orig[:, a_of_v] - 1).ravel(order="C")


def mesh_datasets(parts, porder, tabs):
    """The mesh file of the partitioner for the given parts (BoxMeshes), as
    (name, array, type): create_hdf5_groups_datasets_in_meshFile_from_tool
    (mod_hdf5.f90:219-346), the connectivity and periodic groups (:349-400,
    :523-555), the parallel data read by load_parallel_data_hdf5 (:2841-2950).
    Index values are local to each part; lists are concatenated in part order."""
    nn = (porder + 1) ** 3
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0)  # noqa: E731
    N = sum(m.npoin for m in parts)
    E = sum(m.nelem for m in parts)
    pairs = [H.periodic_pairs(m) if m.periodic else np.zeros((0, 2), np.int32) for m in parts]
    ds = [
        ("/globalIds/globalIdSrl", np.arange(1, N + 1), "i8"),
        ("/globalIds/globalIdPar", np.arange(1, N + 1) + 1000, "i8"),
        ("/globalIds/elemGid", np.arange(1, E + 1) * 3, "i4"),
        # Fortran dims (ndime, nodes): C order (nodes, 3), a node's x, y, z together
        ("/Coords/Points", cat([m.coord for m in parts]).reshape(N, 3), "f8"),
        ("/Connectivity/connecParOrig", cat([m.connec_orig.ravel(order="C") for m in parts]), "i4"),
        ("/Connectivity/connecParWork", cat([m.connec.ravel(order="C") for m in parts]), "i4"),
        ("/Connectivity/numWorkingNodesRankPar", [m.npoin_w for m in parts], "i4"),
        ("/Connectivity/numOwnedNodesRankPar", [m.npoin for m in parts], "i4"),
        ("/Connectivity/workingNodesPar", cat([m.lpoin_w for m in parts]), "i4"),
        ("/Connectivity/ownedNodesPar", cat([np.arange(1, m.npoin + 1) for m in parts]), "i4"),
        ("/order/porder", [porder], "i4"),
    ]
    for k in ("a2ijk", "gmsh2ijk", "vtk2ijk", "a2ij", "gmsh2ij", "vtk2ij"):
        ds.append((f"/order/{k}", tabs[k], "i4"))
    ds += [
        ("/meshOutputInfo/isLinealOutput", [0], "u1"),
        ("/meshOutputInfo/nnodeVTK", [nn], "i4"),
        ("/meshOutputInfo/numVTKElemsPerMshElem", [1], "i4"),
        ("/meshOutputInfo/numElemsVTKMshRank", [m.nelem for m in parts], "i4"),
        ("/meshOutputInfo/sizeConnecVTKMshRank", [m.nelem * nn for m in parts], "i4"),
        ("/VTKHDF/Connectivity", cat([vtk_connec(m, tabs) for m in parts]), "i8"),
    ]
    if any(m.periodic for m in parts):
        ds += [("/Periodic_data/nPerRankPar", [p.shape[0] for p in pairs], "i4"),
               ("/Periodic_data/masSlaRankPar1", cat([p[:, 0] for p in pairs]), "i4"),
               ("/Periodic_data/masSlaRankPar2", cat([p[:, 1] for p in pairs]), "i4")]
    if len(parts) >= 2:
        ns = np.cumsum([0] + [m.npoin for m in parts])
        es = np.cumsum([0] + [m.nelem for m in parts])
        k = len(parts)
        ds += [("/Parallel_data/rankNodeStart", ns[:-1] + 1, "i8"),
               ("/Parallel_data/rankNodeEnd", ns[1:], "i8"),
               ("/Parallel_data/rankElemStart", es[:-1] + 1, "i4"),
               ("/Parallel_data/rankElemEnd", es[1:], "i4"),
               # one neighbour per part, two nodes each: part r sends nodes 2r+1, 2r+2
               ("/Parallel_data/numRanksWithComms", [1] * k, "i4"),
               ("/Parallel_data/numNodesToComm", [2] * k, "i4"),
               ("/Parallel_data/ranksToComm", [(r + 1) % k for r in range(k)], "i4"),
               ("/Parallel_data/commsMemPosInLoc", [1] * k, "i4"),
               ("/Parallel_data/commsMemSize", [2] * k, "i4"),
               ("/Parallel_data/commsMemPosInNgb", [1] * k, "i4"),
               ("/Parallel_data/nodesToComm", np.arange(1, 2 * k + 1), "i4")]
    return ds


def same_mesh(a, b):
    for k in ("nelem", "npoin", "npoin_w"):
        assert getattr(a, k) == getattr(b, k), k
    for k in ("connec_orig", "connec", "lpoin_w", "coord"):
        x, y = getattr(a, k), getattr(b, k)
        assert x.shape == y.shape and np.array_equal(x, y), k


# ------------------------------------------------------------------ library
def test_h5_library_exports_every_declared_symbol():
    L = C.CDLL(H.H5LIB_PATH)
    names = H.declared_symbols()
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n


# ------------------------------------------------------------- order tables
def _positions(p):
    """Geometric (i, j, k) of the table index c: k, i, j visited in sod2d's
    1-D order 0, p, 1, ..., p-1 (set_allocate_hexahedronHO_ijk_indices)."""
    s = [0, p] + list(range(1, p))
    return [(s[i], s[j], s[k]) for k in range(p + 1) for i in range(p + 1) for j in range(p + 1)]


def test_order_tables_p2_are_the_27_node_hexahedra_of_vtk_and_gmsh():
    t = H.order_tables(2)
    pos = _positions(2)
    vtk = [None] * 27
    gmsh = [None] * 27
    for c, x in enumerate(pos):
        vtk[t["vtk2ijk"][c] - 1] = x
        gmsh[t["gmsh2ijk"][c] - 1] = x
    V = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2)]
    mid = lambda a, b: tuple((V[a][d] + V[b][d]) // 2 for d in range(3))  # noqa: E731
    # VTK_TRIQUADRATIC_HEXAHEDRON: vertices, edges (0-1,1-2,2-3,3-0, 4-5,5-6,6-7,7-4,
    # 0-4,1-5,2-6,3-7), faces x=0, x=2, y=0, y=2, z=0, z=2, centre
    vedges = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
              (0, 4), (1, 5), (2, 6), (3, 7)]
    vfaces = [(0, 1, 1), (2, 1, 1), (1, 0, 1), (1, 2, 1), (1, 1, 0), (1, 1, 2)]
    assert vtk == V + [mid(a, b) for a, b in vedges] + vfaces + [(1, 1, 1)]
    # gmsh 27-node hexahedron: vertices, edges (0-1,0-3,0-4,1-2,1-5,2-3,2-6,3-7,4-5,4-7,
    # 5-6,6-7), faces z=0, y=0, x=0, x=2, y=2, z=2, centre
    gedges = [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (4, 7),
              (5, 6), (6, 7)]
    gfaces = [(1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 1, 1), (1, 2, 1), (1, 1, 2)]
    assert gmsh == V + [mid(a, b) for a, b in gedges] + gfaces + [(1, 1, 1)]
    # quadrilaterals (i, j visited in sod2d order): VTK_BIQUADRATIC_QUAD and the
    # gmsh 9-node quadrangle have the same numbering
    s = [0, 2, 1]
    q = [(s[i], s[j]) for i in range(3) for j in range(3)]
    want = [(0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1), (1, 1)]
    for k in ("vtk2ij", "gmsh2ij", "a2ij"):
        got = [None] * 9
        for c, x in enumerate(q):
            got[t[k][c] - 1] = x
        assert got == want, k


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_order_tables_are_permutations_with_the_lagrange_cell_structure(p):
    t = H.order_tables(p)
    nn, nb, m = (p + 1) ** 3, (p + 1) ** 2, p - 1
    for k in ("a2ijk", "gmsh2ijk", "vtk2ijk"):
        assert sorted(t[k]) == list(range(1, nn + 1)), k
    for k in ("a2ij", "gmsh2ij", "vtk2ij"):
        assert sorted(t[k]) == list(range(1, nb + 1)), k
    T = case.basis_tables(p)
    assert np.array_equal(t["a2ijk"], T.a2ijk) and np.array_equal(t["gmsh2ijk"], T.a2ijk)
    assert np.array_equal(t["a2ij"], t["gmsh2ij"])
    # VTK numbering: 8 vertices, 12 edges of m nodes (each run by increasing
    # parameter), 6 faces of m*m (lexicographic), m^3 body nodes (i fastest)
    pos = _positions(p)
    at = {t["vtk2ijk"][c] - 1: x for c, x in enumerate(pos)}
    nbdy = lambda x: sum(v in (0, p) for v in x)  # noqa: E731
    assert all(nbdy(at[v]) == 3 for v in range(8))
    assert all(nbdy(at[v]) == 2 for v in range(8, 8 + 12 * m))
    assert all(nbdy(at[v]) == 1 for v in range(8 + 12 * m, 8 + 12 * m + 6 * m * m))
    body = [at[v] for v in range(8 + 12 * m + 6 * m * m, nn)]
    assert body == [(i, j, k) for k in range(1, p) for j in range(1, p) for i in range(1, p)]
    for e in range(12):
        run = [at[8 + e * m + r] for r in range(m)]
        axis = [d for d in range(3) if run[0][d] not in (0, p)]
        assert len(axis) == 1 and [x[axis[0]] for x in run] == list(range(1, p))


# ----------------------------------------------------------- mesh: reader
@need_tools
@pytest.mark.parametrize("p,n,periodic", [(3, 2, True), (2, (2, 3, 1), False), (4, 1, True)])
def test_reader_loads_a_file_built_by_h5import(tmp_path, p, n, periodic):
    m = case.box_mesh(p, n, periodic=periodic)
    f = tmp_path / "mesh.hdf"
    h5import(f, mesh_datasets([m], p, H.order_tables(p)), str(tmp_path))
    r = H.read_mesh(f)
    same_mesh(r, m)
    info = r.extra["info"]
    assert (info["nparts"], info["periodic"], info["lineal_output"]) == (1, int(periodic), 0)
    assert info["total_nodes"] == m.npoin and info["total_elems"] == m.nelem
    assert np.array_equal(r.extra["gid_srl"], np.arange(1, m.npoin + 1))
    assert np.array_equal(r.extra["gid_par"], np.arange(1, m.npoin + 1) + 1000)
    assert np.array_equal(r.extra["elem_gid"], np.arange(1, m.nelem + 1) * 3)
    assert np.array_equal(r.extra["owned_nodes"], np.arange(1, m.npoin + 1))
    assert np.array_equal(r.extra["connec_vtk"], vtk_connec(m, H.order_tables(p)))
    if periodic:
        ms = r.extra["mas_sla"]
        assert ms.shape[0] == m.npoin - m.npoin_w > 0
        # every slave of connecParOrig appears with its connecParWork master
        assert np.array_equal(ms, H.periodic_pairs(m))
        co, cw = m.connec_orig.ravel(), m.connec.ravel()
        d = dict(zip(co, cw))
        assert all(d[s] == mm for mm, s in ms)
    else:
        assert r.extra["mas_sla"].shape == (0, 2)


@need_tools
def test_reader_slices_each_part_of_a_two_part_file(tmp_path):
    a = case.box_mesh(3, 2, periodic=True)
    b = case.box_mesh(3, (1, 2, 3), periodic=False, lengths=(1.0, 2.0, 3.0))
    f = tmp_path / "two.hdf"
    h5import(f, mesh_datasets([a, b], 3, H.order_tables(3)), str(tmp_path))
    for part, m in enumerate((a, b)):
        r = H.read_mesh(f, part)
        same_mesh(r, m)
        info = r.extra["info"]
        assert info["nparts"] == 2
        assert info["node_start"] == (1 if part == 0 else a.npoin + 1)
        assert info["elem_start"] == (1 if part == 0 else a.nelem + 1)
        off = 0 if part == 0 else a.npoin
        assert np.array_equal(r.extra["gid_srl"], np.arange(1, m.npoin + 1) + off)
        assert np.array_equal(r.extra["comm"]["nodes_to_comm"], [2 * part + 1, 2 * part + 2])
        assert list(r.extra["comm"]["ranks_to_comm"]) == [1 - part]
        assert r.extra["mas_sla"].shape[0] == (a.npoin - a.npoin_w if part == 0 else 0)
    with pytest.raises(SodhipError, match="part 2 of a mesh of 2 parts"):
        H.read_mesh(f, 2)


@need_tools
def test_reader_rejects_bad_files(tmp_path):
    m = case.box_mesh(3, 2, periodic=True)
    tabs = H.order_tables(3)
    with pytest.raises(SodhipError, match="cannot open"):
        H.read_mesh(tmp_path / "missing.hdf")
    # a connectivity entry outside 1..npoin
    ds = mesh_datasets([m], 3, tabs)
    bad = [(n, (np.where(np.arange(len(a)) == 5, m.npoin + 1, a) if n.endswith("Orig") else a), t)
           for n, a, t in ds]
    h5import(tmp_path / "bad1.hdf", bad, str(tmp_path))
    with pytest.raises(SodhipError, match="connecParOrig.*outside"):
        H.read_mesh(tmp_path / "bad1.hdf")
    # a foreign element node order
    t2 = dict(tabs)
    t2["a2ijk"] = tabs["a2ijk"][::-1].copy()
    h5import(tmp_path / "bad2.hdf", mesh_datasets([m], 3, t2), str(tmp_path))
    with pytest.raises(SodhipError, match="a2ijk"):
        H.read_mesh(tmp_path / "bad2.hdf")
    # a missing dataset
    h5import(tmp_path / "bad3.hdf", [d for d in ds if d[0] != "/order/porder"], str(tmp_path))
    with pytest.raises(SodhipError, match="/order/porder not found"):
        H.read_mesh(tmp_path / "bad3.hdf")


# ----------------------------------------------------------- mesh: writer
@need_tools
def test_writer_layout_matches_the_partitioner(tmp_path):
    a = case.box_mesh(3, 2, periodic=True)
    b = case.box_mesh(3, (1, 2, 1), periodic=False)
    f = tmp_path / "w.hdf"
    H.write_mesh(f, [a, b])
    # the same file built by h5import from the restated layout, compared
    # dataset by dataset through h5dump (type, shape, values)
    g = tmp_path / "ref.hdf"
    ref = mesh_datasets([a, b], 3, H.order_tables(3))
    h5import(g, ref, str(tmp_path))
    comm_free = {"/Parallel_data/numRanksWithComms", "/Parallel_data/numNodesToComm",
                 "/Parallel_data/ranksToComm", "/Parallel_data/commsMemPosInLoc",
                 "/Parallel_data/commsMemSize", "/Parallel_data/commsMemPosInNgb",
                 "/Parallel_data/nodesToComm", "/globalIds/globalIdPar", "/globalIds/elemGid"}
    for name, arr, t in ref:
        got = dump_type(f, name)
        kind, size = t[0], int(t[1])
        if name in comm_free:
            assert got[:2] == (kind, size), name  # values differ: no comms, default ids
            continue
        assert got == (kind, size, np.asarray(arr).shape), name
    for name, arr, t in ref:
        if name in comm_free:
            continue
        va = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", str(f)], capture_output=True,
                            text=True).stdout
        vb = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", str(g)], capture_output=True,
                            text=True).stdout
        body = lambda s: s[s.index("DATA {"):]  # noqa: E731
        assert body(va) == body(vb), name
    # and it reads back to the same meshes
    for part, m in enumerate((a, b)):
        same_mesh(H.read_mesh(f, part), m)


def test_writer_round_trip_and_case_set_up(tmp_path):
    g = case.build_case(3, 3, periodic=True, lengths=(2 * np.pi,) * 3)
    f = tmp_path / "tgv.hdf"
    H.write_mesh(f, g.mesh)
    h = H.load_case(f)
    same_mesh(h.mesh, g.mesh)
    for k in ("He", "gpvol", "Ml", "invMl", "helem", "helem_l"):
        assert np.array_equal(getattr(h, k), getattr(g, k)), k
    assert h.vol_tot == g.vol_tot


# ----------------------------------------------------------------- restart
def _fields(m, seed=3):
    r = np.random.default_rng(seed)
    n, ne, nn = m.npoin, m.nelem, m.connec_orig.shape[1]
    return dict(rho=r.uniform(0.9, 1.1, n), u=np.asfortranarray(r.normal(0, 0.3, (n, 3))),
                pr=r.uniform(0.9, 1.1, n), E=r.uniform(2.0, 3.0, n),
                mu_e=np.asfortranarray(r.uniform(0, 1e-3, (ne, nn))),
                mu_t=np.asfortranarray(r.uniform(0, 1e-4, (ne, nn))),
                walave_u=np.asfortranarray(r.normal(0, 0.1, (n, 3))))


def _scatter_gather(m, v):
    """Brute force of the restart's element-node round trip: node value = the
    value of the last (element, node) pair in element order that touches it."""
    co = m.connec_orig
    node = {}
    for e in range(m.nelem):
        for a in range(co.shape[1]):
            node[co[e, a]] = v[e, a]
    return np.asfortranarray(np.vectorize(node.__getitem__)(co).astype(np.float64))


@pytest.mark.parametrize("rp_vtk", [8, 4])
def test_restart_round_trip(tmp_path, rp_vtk):
    m = case.box_mesh(2, 2, periodic=False)
    f = _fields(m)
    path = tmp_path / H.restart_file_name(H.restart_base_name(str(tmp_path) + "/", "box", 1), 1)
    assert path.name == "restart_box-1_1.h5"
    H.write_restart(path, m, f, time=1.25, istep=17, rp_vtk=rp_vtk)
    r = H.read_restart(path, m, walave=True)
    cast = (lambda x: x) if rp_vtk == 8 else (lambda x: x.astype(np.float32).astype(np.float64))
    assert (r["time"], r["istep"], r["rp_vtk"]) == (float(cast(np.float64(1.25))), 17, rp_vtk)
    for k in ("rho", "u", "pr", "E", "walave_u"):
        assert np.array_equal(r[k], cast(f[k])), k
    for k in ("mu_e", "mu_t"):
        assert np.array_equal(r[k], cast(_scatter_gather(m, f[k]))), k


@need_tools
def test_restart_names_types_and_reader_against_h5import(tmp_path):
    m = case.box_mesh(3, 2, periodic=True)
    f = {k: v for k, v in _fields(m, seed=5).items() if k != "walave_u"}
    w = tmp_path / "ours.h5"
    H.write_restart(w, m, f, time=0.5, istep=3, rp_vtk=4)
    n = m.npoin
    # names (initializeNameFields, mod_saveFields.f90:265-275): mu_e is stored
    # under "mut" and mu_t under "mue"; fields as real(rp_vtk), istep int4
    for name in ("rho", "u_x", "u_y", "u_z", "pr", "E", "mut", "mue"):
        assert dump_type(w, name) == ("f", 4, (n,)), name
    assert dump_type(w, "time") == ("f", 4, (1,))
    assert dump_type(w, "istep") == ("i", 4, (1,))
    assert not H.exists(w, "walave_u_x")
    # a file written by h5import in real(8): node mu_e values of our choosing,
    # slave nodes holding junk that the load must overwrite with the master's
    r = np.random.default_rng(9)
    node = {k: r.uniform(1, 2, n) for k in ("rho", "u_x", "u_y", "u_z", "pr", "E", "mut", "mue")}
    ms = H.periodic_pairs(m)
    g = tmp_path / "ref.h5"
    h5import(g, [(k, v, "f8") for k, v in node.items()] +
             [("time", [2.5], "f8"), ("istep", [41], "i4")], str(tmp_path))
    got = H.read_restart(g, m, walave=True)
    assert (got["time"], got["istep"], got["rp_vtk"]) == (2.5, 41, 8)
    fix = {k: v.copy() for k, v in node.items()}
    for k in ("rho", "u_x", "u_y", "u_z", "pr", "E"):
        fix[k][ms[:, 1] - 1] = fix[k][ms[:, 0] - 1]
    assert np.array_equal(got["rho"], fix["rho"])
    assert np.array_equal(got["u"], np.stack([fix["u_x"], fix["u_y"], fix["u_z"]], 1))
    assert np.array_equal(got["walave_u"], got["u"])  # no walave_u in the file: u
    assert np.array_equal(got["mu_e"], node["mut"][m.connec_orig - 1])
    assert np.array_equal(got["mu_t"], node["mue"][m.connec_orig - 1])


def test_restart_state_identities():
    r = np.random.default_rng(2)
    n, gamma, Rgas = 257, 1.4, 287.0
    rho = r.uniform(0.5, 2.0, n)
    u = np.asfortranarray(r.normal(0, 50.0, (n, 3)))
    K = r.uniform(5e4, 2e5, n)
    pr = K * rho ** gamma  # on the isentrope p = K rho^gamma
    E = pr / (gamma - 1.0) + 0.5 * rho * (u * u).sum(1)
    s = H.restart_state(dict(rho=rho, u=u, pr=pr, E=E), gamma, Rgas)
    assert np.allclose(s["q"], rho[:, None] * u, rtol=1e-15, atol=0)
    assert np.allclose(s["e_int"], pr / (rho * (gamma - 1.0)), rtol=1e-12)
    assert np.allclose(s["Tem"] * rho * Rgas, pr, rtol=1e-15)
    assert np.allclose(s["csound"] ** 2, gamma * Rgas * s["Tem"], rtol=1e-14)
    assert np.allclose(s["eta"], rho / (gamma - 1.0) * np.log(K), rtol=1e-12)
    assert np.allclose(s["machno"] * s["csound"], np.linalg.norm(u, axis=1), rtol=1e-14)
    assert np.array_equal(s["rho"], rho) and np.array_equal(s["E"], E)
