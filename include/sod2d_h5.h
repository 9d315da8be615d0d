/*
 * sod2d_h5.h -- C-ABI of the SOD2D HDF5 mesh and restart file I/O
 * (libsod2d_h5.so, host code over the HDF5 C library; no device code).
 *
 * The layout is the one the reference's mesh partitioner writes and its
 * solver reads (src/lib_sod2d/sources/mod_hdf5.f90): one file holds every
 * mesh partition ("part"), each part's elements and nodes in one contiguous
 * range of the global datasets.  A reader asks for one part, as one rank of
 * the reference does (load_hdf5_meshFile, mod_hdf5.f90:1472-1587); a file
 * written for one rank has a single part 0.
 *
 * Arrays are in the reference's layout: column-major, the reference's
 * dimension order, 1-based index values (connecParOrig(nelem,nnode),
 * coordPar(npoin,3), masSlaRankPar(nper,2)).  Any output pointer may be
 * NULL: that dataset is not read.  Every function returns 0 or a SODH5_ERR_*
 * code; sodh5_last_error gives the message of the calling thread's last
 * error.  The reference aborts (MPI_Abort) where these return an error.
 */
#ifndef SOD2D_H5_H
#define SOD2D_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define SODH5_OK 0
#define SODH5_ERR_ARG 1
#define SODH5_ERR_IO 2      /* HDF5 call failed: missing file, dataset or wrong shape */
#define SODH5_ERR_FORMAT 3  /* the file's contents break the layout's invariants */

int sodh5_last_error(char *buf, int len);

/* Sizes of part `part` of a mesh file, from
 *   /globalIds/globalIdSrl, /globalIds/elemGid lengths (totals, mod_hdf5.f90:1487-1509)
 *   /meshOutputInfo/*                                    (load_meshOutputInfo_hdf5, :3487-3529)
 *   /order/porder                                        (load_porder_hdf5, :3531-3591)
 *   /Parallel_data/rank{Node,Elem}{Start,End}, numRanksWithComms, numNodesToComm
 *                                        for files of >= 2 parts (load_parallel_data_hdf5, :2841-2950)
 *   /Connectivity/num{Working,Owned}NodesRankPar         (load_connectivity_hdf5, :2695-2811)
 *   /Periodic_data/nPerRankPar, when the group exists     (load_periodic_data_hdf5, :3045-3112)
 * nparts is the length of /Connectivity/numWorkingNodesRankPar. */
typedef struct sodh5_mesh_info {
    int porder, nnode, npbou;
    int nparts;
    long long total_nodes;   /* totalNumNodesPar */
    int total_elems;         /* totalNumElements */
    long long node_start;    /* rankNodeStart (1-based) of the part */
    int elem_start;          /* rankElemStart (1-based) */
    int nelem, npoin;        /* numElemsRankPar, numNodesRankPar */
    int npoin_w, npoin_owned;/* numWorkingNodesRankPar, numOwnedNodesRankPar */
    int periodic, nper;      /* isMeshPeriodic, nPerRankPar */
    int lineal_output, nnode_vtk, nvtk_per_elem, nelem_vtk, size_connec_vtk;
    int nranks_comm, nnodes_comm;  /* numRanksWithComms, numNodesToComm */
    int boundaries, mapped_faces;  /* /Boundary_data, /MappedFaces_data present (not read here) */
} sodh5_mesh_info;

int sodh5_mesh_info_read(const char *path, int part, sodh5_mesh_info *info);

/* The arrays of part `part` (info from sodh5_mesh_info_read on the same file):
 *   a2ijk, gmsh2ijk, vtk2ijk (nnode); a2ij, gmsh2ij, vtk2ij (npbou)    /order/*
 *   connec_orig, connec_work (nelem,nnode)   /Connectivity/connecParOrig|Work, element-major
 *                                            in the file (mod_hdf5.f90:2716-2741)
 *   working_nodes (npoin_w), owned_nodes (npoin_owned)   /Connectivity/workingNodesPar|ownedNodesPar
 *   coord (npoin,3)        /Coords/Points, (/VTKHDF/Points for a lineal-output mesh), a
 *                          (ndime x nodes) real(8) array in Fortran order (:3380-3406)
 *   mas_sla (nper,2)       /Periodic_data/masSlaRankPar1|2 (master, slave)
 *   gid_srl, gid_par (npoin), elem_gid (nelem)   /globalIds/* (:3442-3485)
 *   connec_vtk (size_connec_vtk)                 /VTKHDF/Connectivity (int8 in the file, :2813-2839)
 *   ranks_to_comm, comms_pos_loc, comms_pos_ngb, comms_size (nranks_comm),
 *   nodes_to_comm (nnodes_comm)                  /Parallel_data/* of the part
 */
int sodh5_mesh_read(const char *path, int part, const sodh5_mesh_info *info,
                    int *a2ijk, int *gmsh2ijk, int *vtk2ijk, int *a2ij, int *gmsh2ij, int *vtk2ij,
                    int *connec_orig, int *connec_work, int *working_nodes, int *owned_nodes,
                    double *coord, int *mas_sla, long long *gid_srl, long long *gid_par,
                    int *elem_gid, long long *connec_vtk, int *ranks_to_comm, int *comms_pos_loc,
                    int *comms_pos_ngb, int *comms_size, int *nodes_to_comm);

/* One part of a mesh to write.  Index values are local to the part (1-based),
 * as the partitioner stores them; mas_sla, elem_gid, the global ids and the
 * comm lists may be NULL when their counts are 0 (the ids then default to
 * the node's / element's global position). */
typedef struct sodh5_part {
    int nelem, npoin, npoin_w, npoin_owned, nper;
    const int *connec_orig, *connec_work;   /* (nelem,nnode) */
    const int *working_nodes, *owned_nodes;
    const double *coord;                    /* (npoin,3) */
    const int *mas_sla;                     /* (nper,2) */
    const long long *gid_srl, *gid_par;     /* (npoin) or NULL */
    const int *elem_gid;                    /* (nelem) or NULL */
    int nranks_comm, nnodes_comm;
    const int *ranks_to_comm, *comms_pos_loc, *comms_pos_ngb, *comms_size, *nodes_to_comm;
} sodh5_part;

/* Writes a mesh file of nparts parts in the layout above (the datasets of
 * create_hdf5_groups_datasets_in_meshFile_from_tool, mod_hdf5.f90:219-346,
 * with the reference's element types: int4, int8 for the global ids and the
 * VTK connectivity, uint8 for isLinealOutput, real(8) coordinates).  The
 * element output is high-order (isLinealOutput = 0, nnodeVTK = nnode, one VTK
 * cell per element, /VTKHDF/Connectivity = the element's nodes in vtk2ijk
 * order, 0-based, global).  /Periodic_data is written when any part has
 * nper > 0 or `periodic` is set; /Parallel_data when nparts >= 2.  The order
 * tables are the caller's (sodhip_order_tables of libsod2d_hip.so).
 * Boundary and mapped-face data are not written. */
int sodh5_mesh_write(const char *path, int porder, int nparts, const sodh5_part *parts,
                     int periodic, const int *a2ijk, const int *gmsh2ijk, const int *vtk2ijk,
                     const int *a2ij, const int *gmsh2ij, const int *vtk2ij);

/* Restart files (save_hdf5_restartFile / load_hdf5_restartFile,
 * mod_hdf5.f90:4088-4338; names from initializeNameFields,
 * mod_saveFields.f90:265-275).  Node fields over all total_nodes nodes:
 *   rho, u_x, u_y, u_z, pr, E; mu_e node-scattered under the name "mut" and
 *   mu_t under "mue" (the reference's name table swaps the two names, and
 *   its save and load both use it, so this keeps the names as written);
 *   walave_u_{x,y,z} when given; time (one value) and istep (int4).
 * Fields are stored as real(rp_vtk): rp_vtk 4 (the reference's default
 * USE_RP_VTK) or 8.  The element-node fields are scattered to nodes through
 * connecParOrig in element order, the last element writing a shared node
 * winning (mod_hdf5.f90:4116-4122).
 * The writer writes a whole file (one part, node_start 1): a run over several
 * parts gathers its node ranges first. */
int sodh5_restart_write(const char *path, int rp_vtk, int npoin, int nelem, int nnode,
                        const int *connec_orig, const double *rho, const double *u,
                        const double *pr, const double *E, const double *mu_e,
                        const double *mu_t, const double *walave_u, double time, int istep);

/* Reads the node range [node_start, node_start + npoin) of a restart file,
 * converting from the stored precision; gathers mu_e(nelem,nnode) and mu_t
 * back through connecParOrig; copies master values onto the nper periodic
 * slaves (mas_sla, may be NULL when nper = 0) for rho, u, pr, E and walave_u
 * (mod_hdf5.f90:4297-4321).  walave_u: when the file has no walave_u_x it
 * receives u (:4273-4292).  rp_vtk_out (may be NULL): 4 or 8, the stored
 * precision. */
int sodh5_restart_read(const char *path, long long node_start, int npoin, int nelem, int nnode,
                       const int *connec_orig, int nper, const int *mas_sla, double *rho,
                       double *u, double *pr, double *E, double *mu_e, double *mu_t,
                       double *walave_u, double *time, int *istep, int *rp_vtk_out);

/* The derived state after a restart load (CFDSolverBase_eval_vars_after_
 * load_hdf5_resultsFile, CFDSolverBase.f90:2552-2599), on all npoin nodes:
 *   q = rho u, e_int = E/rho - |u|^2/2, Tem = pr/(rho Rgas),
 *   csound = sqrt(gamma pr/rho), eta = rho/(gamma-1) log(pr/rho^gamma),
 *   machno = |u|/csound.  u, q are (npoin,3). */
int sodh5_restart_state(int npoin, double gamma, double Rgas, const double *rho, const double *u,
                        const double *pr, const double *E, double *q, double *e_int, double *Tem,
                        double *csound, double *eta, double *machno);

/* Shape and element type of one dataset (C order dims, rank <= 4):
 * cls 0 integer, 1 float, 2 other; size in bytes; is_signed for integers.
 * Used to check a file's layout against the reference's. */
int sodh5_dataset_info(const char *path, const char *name, int *rank, long long *dims, int *cls,
                       int *size, int *is_signed);

/* 1 if the link (group or dataset) exists, 0 if not, < 0 on error. */
int sodh5_exists(const char *path, const char *name);

#ifdef __cplusplus
}
#endif
#endif
