"""SOD2D HDF5 mesh and restart files (libsod2d_h5.so, the C-ABI in include/sod2d_h5.h).

read_mesh() loads one partition ("part") of a mesh written by the reference's
partitioner into a BoxMesh (connecParOrig, connecParWork, workingNodesPar,
coordPar; the periodic master/slave pairs, owned nodes, global ids, order
tables and comm lists in ``mesh.extra``), as load_hdf5_meshFile does for one
rank (mod_hdf5.f90:1472-1587).  load_case() adds the element tables, geometry,
lumped mass and element lengths (the set-up of build_case), so a TGVRun can
start from a mesh file.  write_mesh() writes BoxMeshes in the same layout.

save_restart() / load_restart() write and read restart files with the
reference's names and conventions (save_hdf5_restartFile /
load_hdf5_restartFile, mod_hdf5.f90:4088-4338) and rebuild the derived state
(CFDSolverBase_eval_vars_after_load_hdf5_resultsFile).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import case as case_mod
from ._lib import SodhipError, check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
H5LIB_PATH = os.path.join(HERE, "libsod2d_h5.so")

P, I, D, LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong


class MeshInfo(C.Structure):
    """sodh5_mesh_info."""
    _fields_ = [("porder", I), ("nnode", I), ("npbou", I), ("nparts", I), ("total_nodes", LL),
                ("total_elems", I), ("node_start", LL), ("elem_start", I), ("nelem", I),
                ("npoin", I), ("npoin_w", I), ("npoin_owned", I), ("periodic", I), ("nper", I),
                ("lineal_output", I), ("nnode_vtk", I), ("nvtk_per_elem", I), ("nelem_vtk", I),
                ("size_connec_vtk", I), ("nranks_comm", I), ("nnodes_comm", I),
                ("boundaries", I), ("mapped_faces", I)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Part(C.Structure):
    """sodh5_part."""
    _fields_ = [("nelem", I), ("npoin", I), ("npoin_w", I), ("npoin_owned", I), ("nper", I),
                ("connec_orig", P), ("connec_work", P), ("working_nodes", P),
                ("owned_nodes", P), ("coord", P), ("mas_sla", P), ("gid_srl", P),
                ("gid_par", P), ("elem_gid", P), ("nranks_comm", I), ("nnodes_comm", I),
                ("ranks_to_comm", P), ("comms_pos_loc", P), ("comms_pos_ngb", P),
                ("comms_size", P), ("nodes_to_comm", P)]


_SIGS = {
    "sodh5_last_error": (I, [C.c_char_p, I]),
    "sodh5_exists": (I, [C.c_char_p, C.c_char_p]),
    "sodh5_dataset_info": (I, [C.c_char_p, C.c_char_p, P, P, P, P, P]),
    "sodh5_mesh_info_read": (I, [C.c_char_p, I, C.POINTER(MeshInfo)]),
    "sodh5_mesh_read": (I, [C.c_char_p, I, C.POINTER(MeshInfo)] + [P] * 21),
    "sodh5_mesh_write": (I, [C.c_char_p, I, I, C.POINTER(Part), I] + [P] * 6),
    "sodh5_restart_write": (I, [C.c_char_p, I, I, I, I, P, P, P, P, P, P, P, P, D, I]),
    "sodh5_restart_read": (I, [C.c_char_p, LL, I, I, I, P, I, P] + [P] * 10),
    "sodh5_restart_state": (I, [I, D, D] + [P] * 10),
}

_h5 = None


def h5lib():
    """Load libsod2d_h5.so (raises if it has not been built)."""
    global _h5
    if _h5 is None:
        if not os.path.exists(H5LIB_PATH):
            raise SodhipError(f"{H5LIB_PATH} is missing: build it with `python -m sod2d_amd.build`")
        L = C.CDLL(H5LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _h5 = L
    return _h5


def declared_symbols():
    """Names declared in include/sod2d_h5.h (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "sod2d_h5.h")
    return sorted(set(re.findall(r"\b(sodh5_[a-z0-9_]+)\s*\(", open(hdr).read())))


def _check(rc: int, what: str):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        h5lib().sodh5_last_error(buf, 1024)
        raise SodhipError(f"{what} failed (code {rc}): {buf.value.decode(errors='replace')}")


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _b(path) -> bytes:
    return os.fsencode(path)


# ---------------------------------------------------------------------- inspection
def exists(path, name: str) -> bool:
    r = h5lib().sodh5_exists(_b(path), name.encode())
    if r < 0:
        _check(1, f"exists({path})")
    return r == 1


def dataset_info(path, name: str) -> dict:
    """rank, dims (C order), kind ('i'/'u'/'f'/'other') and item size of one dataset."""
    rank, cls, size, sg = I(), I(), I(), I()
    dims = (LL * 4)()
    _check(h5lib().sodh5_dataset_info(_b(path), name.encode(), C.byref(rank), dims, C.byref(cls),
                                      C.byref(size), C.byref(sg)), f"dataset_info({name})")
    kind = {0: "i" if sg.value else "u", 1: "f"}.get(cls.value, "other")
    return dict(rank=rank.value, dims=tuple(dims[i] for i in range(rank.value)), kind=kind,
                size=size.value)


# ------------------------------------------------------------------------- meshes
def mesh_info(path, part: int = 0) -> MeshInfo:
    info = MeshInfo()
    _check(h5lib().sodh5_mesh_info_read(_b(path), part, C.byref(info)), f"mesh_info({path})")
    return info


def read_mesh(path, part: int = 0) -> case_mod.BoxMesh:
    """Part ``part`` of a SOD2D mesh file as a BoxMesh.  The file's element-local
    node order (/order/a2ijk) must be the one of this library's element tables
    for its order (the partitioner's a2ijk = gmsh2ijk): the kernels address an
    element's nodes through those tables."""
    info = mesh_info(path, part)
    nn, nb, ne, n = info.nnode, info.npbou, info.nelem, info.npoin
    z = lambda *s: np.zeros(s, dtype=np.int32, order="F")  # noqa: E731
    tabs = {k: z(nn) for k in ("a2ijk", "gmsh2ijk", "vtk2ijk")}
    tabs.update({k: z(nb) for k in ("a2ij", "gmsh2ij", "vtk2ij")})
    co, cw = z(ne, nn), z(ne, nn)
    lw, own = z(info.npoin_w), z(info.npoin_owned)
    coord = np.zeros((n, 3), order="F")
    ms = z(info.nper, 2)
    gs, gp = np.zeros(n, np.int64), np.zeros(n, np.int64)
    eg = z(ne)
    cvtk = np.zeros(info.size_connec_vtk, np.int64)
    comm = {k: z(info.nranks_comm) for k in ("ranks_to_comm", "comms_pos_loc", "comms_pos_ngb",
                                            "comms_size")}
    comm["nodes_to_comm"] = z(info.nnodes_comm)
    _check(h5lib().sodh5_mesh_read(
        _b(path), part, C.byref(info), *[_p(tabs[k]) for k in ("a2ijk", "gmsh2ijk", "vtk2ijk",
                                                              "a2ij", "gmsh2ij", "vtk2ij")],
        _p(co), _p(cw), _p(lw), _p(own), _p(coord), _p(ms), _p(gs), _p(gp), _p(eg), _p(cvtk),
        _p(comm["ranks_to_comm"]), _p(comm["comms_pos_loc"]), _p(comm["comms_pos_ngb"]),
        _p(comm["comms_size"]), _p(comm["nodes_to_comm"])), f"mesh_read({path}, part {part})")
    ours = order_tables(info.porder)
    if not np.array_equal(tabs["a2ijk"], ours["a2ijk"]):
        raise SodhipError(f"{path}: /order/a2ijk is not this library's element node order for "
                          f"porder {info.porder}")
    extra = dict(info=info.as_dict(), part=part, mas_sla=ms, owned_nodes=own, gid_srl=gs,
                 gid_par=gp, elem_gid=eg, connec_vtk=cvtk, order=tabs, comm=comm, source=str(path))
    return case_mod.BoxMesh(info.porder, ("hdf5", str(path), part), bool(info.periodic), ne, n,
                            info.npoin_w, co, cw, lw, coord, extra)


def order_tables(porder: int) -> dict:
    """The six /order tables of porder (sodhip_order_tables of libsod2d_hip.so)."""
    n1 = porder + 1
    t = {k: np.zeros(n1 ** 3, np.int32) for k in ("a2ijk", "gmsh2ijk", "vtk2ijk")}
    t.update({k: np.zeros(n1 ** 2, np.int32) for k in ("a2ij", "gmsh2ij", "vtk2ij")})
    check(lib().sodhip_order_tables(porder, *[_p(t[k]) for k in ("a2ijk", "gmsh2ijk", "vtk2ijk",
                                                                 "a2ij", "gmsh2ij", "vtk2ij")]),
          "sodhip_order_tables")
    return t


def periodic_pairs(mesh: case_mod.BoxMesh) -> np.ndarray:
    """masSlaRankPar (nper,2) of a mesh: every node that connecParWork replaces
    (slave, in connecParOrig) with its master, one row per slave, slaves ascending."""
    if "mas_sla" in mesh.extra:
        return mesh.extra["mas_sla"]
    co, cw = mesh.connec_orig.ravel(order="F"), mesh.connec.ravel(order="F")
    d = co != cw
    sla, idx = np.unique(co[d], return_index=True)
    mas = cw[d][idx]
    return np.asfortranarray(np.stack([mas, sla], axis=1).astype(np.int32))


def _part_of(mesh: case_mod.BoxMesh, keep: list) -> Part:
    def arr(a, dtype):
        a = np.asfortranarray(np.asarray(a, dtype=dtype))
        keep.append(a)
        return _p(a)
    ms = periodic_pairs(mesh)
    own = mesh.extra.get("owned_nodes")
    if own is None:
        own = np.arange(1, mesh.npoin + 1, dtype=np.int32)
    p = Part()
    p.nelem, p.npoin, p.npoin_w = mesh.nelem, mesh.npoin, mesh.npoin_w
    p.npoin_owned, p.nper = len(own), ms.shape[0]
    p.connec_orig = arr(mesh.connec_orig, np.int32)
    p.connec_work = arr(mesh.connec, np.int32)
    p.working_nodes = arr(mesh.lpoin_w, np.int32)
    p.owned_nodes = arr(own, np.int32)
    p.coord = arr(mesh.coord, np.float64)
    p.mas_sla = arr(ms, np.int32)
    for k in ("gid_srl", "gid_par"):
        if k in mesh.extra:
            setattr(p, k, arr(mesh.extra[k], np.int64))
    if "elem_gid" in mesh.extra:
        p.elem_gid = arr(mesh.extra["elem_gid"], np.int32)
    comm = mesh.extra.get("comm")
    if comm is not None and len(comm["ranks_to_comm"]):
        p.nranks_comm = len(comm["ranks_to_comm"])
        for k in ("ranks_to_comm", "comms_pos_loc", "comms_pos_ngb", "comms_size"):
            setattr(p, k, arr(comm[k], np.int32))
    if comm is not None and len(comm["nodes_to_comm"]):
        p.nnodes_comm = len(comm["nodes_to_comm"])
        p.nodes_to_comm = arr(comm["nodes_to_comm"], np.int32)
    return p


def write_mesh(path, meshes, porder: int | None = None) -> None:
    """Writes one BoxMesh, or a list of them as the parts of one file, in the
    SOD2D layout.  Periodic pairs come from connecParOrig vs connecParWork
    (or mesh.extra['mas_sla']); owned nodes default to all of the part's nodes."""
    if isinstance(meshes, case_mod.BoxMesh):
        meshes = [meshes]
    porder = porder or meshes[0].porder
    if any(m.porder != porder for m in meshes):
        raise SodhipError("all parts must have the same order")
    keep: list = []
    parts = (Part * len(meshes))(*[_part_of(m, keep) for m in meshes])
    t = order_tables(porder)
    _check(h5lib().sodh5_mesh_write(
        _b(path), porder, len(meshes), parts, int(any(m.periodic for m in meshes)),
        *[_p(t[k]) for k in ("a2ijk", "gmsh2ijk", "vtk2ijk", "a2ij", "gmsh2ij", "vtk2ij")]),
        f"mesh_write({path})")


def load_case(path, part: int = 0, spectral_lengths: bool = True) -> case_mod.Geometry:
    """A mesh file's part with its tables, geometry, lumped mass and element
    lengths: the set-up build_case does for a generated box."""
    m = read_mesh(path, part)
    T = case_mod.basis_tables(m.porder)
    He, gpvol = case_mod.elem_jacobian(m.porder, m.connec_orig, m.coord)
    Ml, invMl = case_mod.lumped_mass(m.porder, m.connec, gpvol, m.npoin, m.lpoin_w)
    helem, helem_l = case_mod.char_length(m.porder, m.connec_orig, m.coord, spectral_lengths)
    return case_mod.Geometry(T, m, He, gpvol, Ml, invMl, helem, helem_l,
                             float(gpvol.sum(dtype=np.float64)))


# ------------------------------------------------------------------------ restart
def restart_base_name(res_path: str, mesh_name: str, nranks: int) -> str:
    """set_hdf5_restartFile_baseName (mod_hdf5.f90:128-139)."""
    return f"{res_path}restart_{mesh_name}-{nranks}_"


def restart_file_name(base: str, cnt: int) -> str:
    """set_hdf5_restartFile_name (mod_hdf5.f90:3609-3618): base + cnt + '.h5'."""
    return f"{base}{cnt}.h5"


def write_restart(path, mesh: case_mod.BoxMesh, fields: dict, time: float, istep: int,
                  rp_vtk: int = 4) -> None:
    """Host fields rho, u (npoin,3), pr, E, mu_e (nelem,nnode) and optionally
    mu_t, walave_u into a restart file stored as real(rp_vtk)."""
    f = {k: np.asfortranarray(np.asarray(v, dtype=np.float64)) for k, v in fields.items()}
    co = np.asfortranarray(mesh.connec_orig, dtype=np.int32)
    _check(h5lib().sodh5_restart_write(
        _b(path), rp_vtk, mesh.npoin, mesh.nelem, co.shape[1], _p(co), _p(f["rho"]), _p(f["u"]),
        _p(f["pr"]), _p(f["E"]), _p(f["mu_e"]), _p(f.get("mu_t")), _p(f.get("walave_u")),
        float(time), int(istep)), f"restart_write({path})")


def read_restart(path, mesh: case_mod.BoxMesh, walave: bool = False) -> dict:
    """The node range of ``mesh`` (its part) of a restart file: rho, u, pr, E,
    mu_e, mu_t (and walave_u), time, istep and the stored precision rp_vtk."""
    n, ne = mesh.npoin, mesh.nelem
    co = np.asfortranarray(mesh.connec_orig, dtype=np.int32)
    nn = co.shape[1]
    ms = periodic_pairs(mesh) if mesh.periodic else np.zeros((0, 2), np.int32, order="F")
    start = int(mesh.extra.get("info", {}).get("node_start", 1))
    out = dict(rho=np.zeros(n), u=np.zeros((n, 3), order="F"), pr=np.zeros(n), E=np.zeros(n),
               mu_e=np.zeros((ne, nn), order="F"), mu_t=np.zeros((ne, nn), order="F"))
    if walave:
        out["walave_u"] = np.zeros((n, 3), order="F")
    t, it, rp = D(), I(), I()
    _check(h5lib().sodh5_restart_read(
        _b(path), start, n, ne, nn, _p(co), ms.shape[0], _p(np.asfortranarray(ms)),
        _p(out["rho"]), _p(out["u"]), _p(out["pr"]), _p(out["E"]), _p(out["mu_e"]),
        _p(out["mu_t"]), _p(out.get("walave_u")), C.byref(t), C.byref(it), C.byref(rp)),
        f"restart_read({path})")
    out.update(time=t.value, istep=it.value, rp_vtk=rp.value)
    return out


def restart_state(fields: dict, gamma: float, Rgas: float) -> dict:
    """The TGVRun initial-state dict after a restart load: rho, u, pr, E as
    loaded and q, e_int, Tem, csound, eta, machno derived as
    CFDSolverBase_eval_vars_after_load_hdf5_resultsFile (CFDSolverBase.f90:2552-2599)."""
    n = fields["rho"].shape[0]
    rho, pr, E = (np.ascontiguousarray(fields[k], dtype=np.float64) for k in ("rho", "pr", "E"))
    u = np.asfortranarray(fields["u"], dtype=np.float64)
    q = np.zeros((n, 3), order="F")
    e_int, Tem, cs, eta, mach = (np.zeros(n) for _ in range(5))
    _check(h5lib().sodh5_restart_state(n, gamma, Rgas, _p(rho), _p(u), _p(pr), _p(E), _p(q),
                                       _p(e_int), _p(Tem), _p(cs), _p(eta), _p(mach)),
           "restart_state")
    return dict(rho=rho, u=u, pr=pr, E=E, q=q, e_int=e_int, Tem=Tem, csound=cs, eta=eta,
                machno=mach)


def save_restart(path, run, rp_vtk: int = 4) -> None:
    """The run's current level (rho, u, pr, E, mu_e; mu_t = 0), time and step
    count into a restart file (CFDSolverBase_saveRestartFile)."""
    n = run.npoin
    fields = dict(rho=run.rho[1].cpu().numpy(),
                  u=run.u[1].cpu().numpy().reshape((n, 3), order="F"),
                  pr=run.pr[1].cpu().numpy(), E=run.E[1].cpu().numpy(),
                  mu_e=run.mu_e.cpu().numpy().reshape((run.nelem, run.nnode), order="F"))
    write_restart(path, run.geo.mesh, fields, run.time, run.istep, rp_vtk)


def load_restart(path, geometry: case_mod.Geometry, params) -> tuple[dict, float, int]:
    """(initial state for TGVRun(initial=...), time, istep) from a restart file."""
    f = read_restart(path, geometry.mesh)
    return restart_state(f, params.gamma_gas, params.Rgas), f["time"], f["istep"]
