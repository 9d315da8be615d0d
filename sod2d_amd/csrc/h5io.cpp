// SOD2D HDF5 mesh and restart files (host code over the HDF5 C library).
//
// The datasets, their element types and their Fortran-order shapes are the
// reference's (src/lib_sod2d/sources/mod_hdf5.f90): a Fortran dataset of dims
// (d1, d2) appears to the C API as (d2, d1), so /Coords/Points (ndime, nodes)
// is read here as (nodes, 3) rows.  Every partition ("part") owns one
// contiguous range of elements and nodes in the global datasets; the per-part
// lists (working and owned nodes, periodic pairs, comm lists) are
// concatenated in part order, and a part's slice starts at the sum of the
// counts of the parts before it.
#include "../../include/sod2d_h5.h"

#include <hdf5.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define H5_TRY(expr)                                                                        \
    do {                                                                                    \
        int rc_ = (expr);                                                                   \
        if (rc_) return rc_;                                                                \
    } while (0)

// HDF5's own error stack printing is turned off: errors come back as codes.
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid() = default;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    ~Hid()
    {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
};

bool link_exists(hid_t f, const char *name)
{
    // each path component must exist before the next one is asked for
    std::string p(name), pre;
    size_t i = 0;
    while (i < p.size()) {
        size_t j = p.find('/', i + 1);
        if (j == std::string::npos) j = p.size();
        pre = p.substr(0, j);
        if (pre != "/" && H5Lexists(f, pre.c_str(), H5P_DEFAULT) <= 0) return false;
        i = j;
    }
    return true;
}

int open_ro(const char *path, Hid &f)
{
    if (!path) return fail(SODH5_ERR_ARG, "NULL file name");
    quiet();
    f.id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    f.close = H5Fclose;
    if (!f.ok()) return fail(SODH5_ERR_IO, "cannot open HDF5 file %s", path);
    return SODH5_OK;
}

// dims of a dataset (C order)
int dset_dims(hid_t f, const char *name, int want_rank, hsize_t *dims)
{
    if (!link_exists(f, name)) return fail(SODH5_ERR_IO, "dataset %s not found", name);
    Hid d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(SODH5_ERR_IO, "cannot open dataset %s", name);
    Hid s(H5Dget_space(d.id), H5Sclose);
    int r = H5Sget_simple_extent_ndims(s.id);
    if (r != want_rank)
        return fail(SODH5_ERR_FORMAT, "dataset %s has rank %d, expected %d", name, r, want_rank);
    H5Sget_simple_extent_dims(s.id, dims, nullptr);
    return SODH5_OK;
}

int dset_len(hid_t f, const char *name, long long *n)
{
    hsize_t d[1];
    H5_TRY(dset_dims(f, name, 1, d));
    *n = (long long)d[0];
    return SODH5_OK;
}

// rows [off, off+n) of a rank-1 (ncol = 0) or rank-2 (ncol columns) dataset,
// converted to memory type mt
int read_rows(hid_t f, const char *name, hid_t mt, long long off, long long n, int ncol, void *buf)
{
    const int rank = ncol ? 2 : 1;
    hsize_t dims[2];
    H5_TRY(dset_dims(f, name, rank, dims));
    if (ncol && dims[1] != (hsize_t)ncol)
        return fail(SODH5_ERR_FORMAT, "dataset %s has %llu columns, expected %d", name,
                    (unsigned long long)dims[1], ncol);
    if (off < 0 || n < 0 || (hsize_t)(off + n) > dims[0])
        return fail(SODH5_ERR_FORMAT, "dataset %s has %llu rows; rows %lld..%lld requested", name,
                    (unsigned long long)dims[0], off, off + n - 1);
    if (n == 0) return SODH5_OK;
    Hid d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
    Hid s(H5Dget_space(d.id), H5Sclose);
    hsize_t start[2] = {(hsize_t)off, 0}, count[2] = {(hsize_t)n, (hsize_t)ncol};
    if (H5Sselect_hyperslab(s.id, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0)
        return fail(SODH5_ERR_IO, "hyperslab of %s", name);
    Hid m(H5Screate_simple(rank, count, nullptr), H5Sclose);
    if (H5Dread(d.id, mt, m.id, s.id, H5P_DEFAULT, buf) < 0)
        return fail(SODH5_ERR_IO, "reading %s", name);
    return SODH5_OK;
}

int read_i32(hid_t f, const char *name, long long off, long long n, int *buf)
{
    return read_rows(f, name, H5T_NATIVE_INT, off, n, 0, buf);
}

int read_one_i32(hid_t f, const char *name, long long off, int *v)
{
    return read_i32(f, name, off, 1, v);
}

// sum of a per-part count dataset over the parts before `part`, and its entry at part
int part_slice(hid_t f, const char *name, int part, int nparts, long long *off, int *cnt)
{
    std::vector<int> c(nparts);
    H5_TRY(read_i32(f, name, 0, nparts, c.data()));
    long long o = 0;
    for (int i = 0; i < part; ++i) o += c[i];
    if (c[part] < 0) return fail(SODH5_ERR_FORMAT, "%s[%d] = %d < 0", name, part, c[part]);
    *off = o;
    *cnt = c[part];
    return SODH5_OK;
}

int check_range(const char *what, const int *v, long long n, long long lo, long long hi)
{
    for (long long i = 0; i < n; ++i)
        if (v[i] < lo || v[i] > hi)
            return fail(SODH5_ERR_FORMAT, "%s[%lld] = %d outside %lld..%lld", what, i, v[i], lo, hi);
    return SODH5_OK;
}

// ------------------------------------------------------------------ writing
int write_ds(hid_t f, const char *name, hid_t ft, hid_t mt, int rank, const hsize_t *dims,
             const void *buf)
{
    Hid s(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Hid d(H5Dcreate2(f, name, ft, s.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(SODH5_ERR_IO, "cannot create dataset %s", name);
    size_t total = 1;
    for (int i = 0; i < rank; ++i) total *= dims[i];
    if (total && H5Dwrite(d.id, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
        return fail(SODH5_ERR_IO, "writing %s", name);
    return SODH5_OK;
}

int write_i32(hid_t f, const char *name, const std::vector<int> &v)
{
    hsize_t n = v.size();
    return write_ds(f, name, H5T_STD_I32LE, H5T_NATIVE_INT, 1, &n, v.data());
}

int write_i64(hid_t f, const char *name, const std::vector<long long> &v)
{
    hsize_t n = v.size();
    return write_ds(f, name, H5T_STD_I64LE, H5T_NATIVE_LLONG, 1, &n, v.data());
}

int make_group(hid_t f, const char *name)
{
    Hid g(H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(SODH5_ERR_IO, "cannot create group %s", name);
    return SODH5_OK;
}

}  // namespace

extern "C" int sodh5_last_error(char *buf, int len)
{
    if (!buf || len <= 0) return SODH5_ERR_ARG;
    std::snprintf(buf, (size_t)len, "%s", g_err);
    return SODH5_OK;
}

extern "C" int sodh5_exists(const char *path, const char *name)
{
    Hid f;
    if (open_ro(path, f)) return -1;
    if (!name) return -1;
    return link_exists(f.id, name) ? 1 : 0;
}

extern "C" int sodh5_dataset_info(const char *path, const char *name, int *rank, long long *dims,
                                  int *cls, int *size, int *is_signed)
{
    Hid f;
    H5_TRY(open_ro(path, f));
    if (!name || !link_exists(f.id, name)) return fail(SODH5_ERR_IO, "dataset %s not found", name);
    Hid d(H5Dopen2(f.id, name, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return fail(SODH5_ERR_IO, "%s is not a dataset", name);
    Hid s(H5Dget_space(d.id), H5Sclose);
    int r = H5Sget_simple_extent_ndims(s.id);
    if (r < 0 || r > 4) return fail(SODH5_ERR_FORMAT, "dataset %s has rank %d", name, r);
    hsize_t dd[4] = {0, 0, 0, 0};
    H5Sget_simple_extent_dims(s.id, dd, nullptr);
    if (rank) *rank = r;
    if (dims)
        for (int i = 0; i < r; ++i) dims[i] = (long long)dd[i];
    Hid t(H5Dget_type(d.id), H5Tclose);
    H5T_class_t c = H5Tget_class(t.id);
    if (cls) *cls = c == H5T_INTEGER ? 0 : (c == H5T_FLOAT ? 1 : 2);
    if (size) *size = (int)H5Tget_size(t.id);
    if (is_signed) *is_signed = c == H5T_INTEGER ? (H5Tget_sign(t.id) == H5T_SGN_2) : 0;
    return SODH5_OK;
}

// -------------------------------------------------------------------- mesh
extern "C" int sodh5_mesh_info_read(const char *path, int part, sodh5_mesh_info *info)
{
    if (!info) return fail(SODH5_ERR_ARG, "NULL info");
    Hid f;
    H5_TRY(open_ro(path, f));
    const hid_t F = f.id;
    sodh5_mesh_info I;
    std::memset(&I, 0, sizeof(I));
    // totals from the global-id datasets (load_hdf5_meshFile, mod_hdf5.f90:1487-1509)
    long long n;
    H5_TRY(dset_len(F, "/globalIds/globalIdSrl", &I.total_nodes));
    H5_TRY(dset_len(F, "/globalIds/elemGid", &n));
    if (n > INT_MAX) return fail(SODH5_ERR_FORMAT, "%lld elements", n);
    I.total_elems = (int)n;
    // porder (load_porder_hdf5, :3531-3591; get_porder_values)
    H5_TRY(read_one_i32(F, "/order/porder", 0, &I.porder));
    if (I.porder < 1 || I.porder > 16)
        return fail(SODH5_ERR_FORMAT, "porder %d in %s", I.porder, path);
    I.nnode = (I.porder + 1) * (I.porder + 1) * (I.porder + 1);
    I.npbou = (I.porder + 1) * (I.porder + 1);
    // partitions
    H5_TRY(dset_len(F, "/Connectivity/numWorkingNodesRankPar", &n));
    I.nparts = (int)n;
    if (part < 0 || part >= I.nparts)
        return fail(SODH5_ERR_ARG, "part %d of a mesh of %d parts", part, I.nparts);
    // meshOutputInfo (:3487-3529)
    unsigned char lin = 0;
    H5_TRY(read_rows(F, "/meshOutputInfo/isLinealOutput", H5T_NATIVE_UCHAR, 0, 1, 0, &lin));
    I.lineal_output = lin == 1;
    H5_TRY(read_one_i32(F, "/meshOutputInfo/nnodeVTK", 0, &I.nnode_vtk));
    H5_TRY(read_one_i32(F, "/meshOutputInfo/numVTKElemsPerMshElem", 0, &I.nvtk_per_elem));
    H5_TRY(read_one_i32(F, "/meshOutputInfo/numElemsVTKMshRank", part, &I.nelem_vtk));
    H5_TRY(read_one_i32(F, "/meshOutputInfo/sizeConnecVTKMshRank", part, &I.size_connec_vtk));
    // the part's element and node ranges (load_parallel_data_hdf5, :2841-2950)
    if (I.nparts >= 2) {
        long long ns, ne;
        H5_TRY(read_rows(F, "/Parallel_data/rankNodeStart", H5T_NATIVE_LLONG, part, 1, 0, &ns));
        H5_TRY(read_rows(F, "/Parallel_data/rankNodeEnd", H5T_NATIVE_LLONG, part, 1, 0, &ne));
        int es, ee;
        H5_TRY(read_one_i32(F, "/Parallel_data/rankElemStart", part, &es));
        H5_TRY(read_one_i32(F, "/Parallel_data/rankElemEnd", part, &ee));
        H5_TRY(read_one_i32(F, "/Parallel_data/numRanksWithComms", part, &I.nranks_comm));
        H5_TRY(read_one_i32(F, "/Parallel_data/numNodesToComm", part, &I.nnodes_comm));
        if (ns < 1 || ne < ns - 1 || ne > I.total_nodes || ne - ns + 1 > INT_MAX)
            return fail(SODH5_ERR_FORMAT, "part %d nodes %lld..%lld of %lld", part, ns, ne,
                        I.total_nodes);
        if (es < 1 || ee < es - 1 || ee > I.total_elems)
            return fail(SODH5_ERR_FORMAT, "part %d elements %d..%d of %d", part, es, ee,
                        I.total_elems);
        I.node_start = ns;
        I.npoin = (int)(ne - ns + 1);
        I.elem_start = es;
        I.nelem = ee - es + 1;
    } else {
        if (I.total_nodes > INT_MAX) return fail(SODH5_ERR_FORMAT, "%lld nodes", I.total_nodes);
        I.node_start = 1;
        I.npoin = (int)I.total_nodes;
        I.elem_start = 1;
        I.nelem = I.total_elems;
    }
    if ((long long)I.nelem * I.nnode > INT_MAX)
        return fail(SODH5_ERR_FORMAT, "%d elements of %d nodes exceed 32-bit indexing", I.nelem,
                    I.nnode);
    H5_TRY(read_one_i32(F, "/Connectivity/numWorkingNodesRankPar", part, &I.npoin_w));
    H5_TRY(read_one_i32(F, "/Connectivity/numOwnedNodesRankPar", part, &I.npoin_owned));
    // periodic data (load_periodic_data_hdf5, :3045-3112): the group's presence decides
    I.periodic = link_exists(F, "/Periodic_data");
    if (I.periodic) H5_TRY(read_one_i32(F, "/Periodic_data/nPerRankPar", part, &I.nper));
    I.boundaries = link_exists(F, "/Boundary_data");
    I.mapped_faces = link_exists(F, "/MappedFaces_data");
    if (I.npoin_w < 0 || I.npoin_w > I.npoin || I.npoin_owned < 0 || I.npoin_owned > I.npoin ||
        I.nper < 0 || I.nper > I.npoin)
        return fail(SODH5_ERR_FORMAT, "part %d: %d working, %d owned, %d periodic of %d nodes",
                    part, I.npoin_w, I.npoin_owned, I.nper, I.npoin);
    *info = I;
    return SODH5_OK;
}

extern "C" int sodh5_mesh_read(const char *path, int part, const sodh5_mesh_info *info,
                               int *a2ijk, int *gmsh2ijk, int *vtk2ijk, int *a2ij, int *gmsh2ij,
                               int *vtk2ij, int *connec_orig, int *connec_work, int *working_nodes,
                               int *owned_nodes, double *coord, int *mas_sla, long long *gid_srl,
                               long long *gid_par, int *elem_gid, long long *connec_vtk,
                               int *ranks_to_comm, int *comms_pos_loc, int *comms_pos_ngb,
                               int *comms_size, int *nodes_to_comm)
{
    if (!info) return fail(SODH5_ERR_ARG, "NULL info");
    // the sizes are read again so that a stale or foreign info cannot size a read
    sodh5_mesh_info I;
    H5_TRY(sodh5_mesh_info_read(path, part, &I));
    if (I.nnode != info->nnode || I.nelem != info->nelem || I.npoin != info->npoin ||
        I.npoin_w != info->npoin_w || I.npoin_owned != info->npoin_owned || I.nper != info->nper ||
        I.size_connec_vtk != info->size_connec_vtk || I.nranks_comm != info->nranks_comm ||
        I.nnodes_comm != info->nnodes_comm)
        return fail(SODH5_ERR_ARG, "info does not describe part %d of %s", part, path);
    Hid f;
    H5_TRY(open_ro(path, f));
    const hid_t F = f.id;
    const int nn = I.nnode, nb = I.npbou, ne = I.nelem, np = I.npoin;
    // /order (load_porder_hdf5)
    if (a2ijk) H5_TRY(read_i32(F, "/order/a2ijk", 0, nn, a2ijk));
    if (gmsh2ijk) H5_TRY(read_i32(F, "/order/gmsh2ijk", 0, nn, gmsh2ijk));
    if (vtk2ijk) H5_TRY(read_i32(F, "/order/vtk2ijk", 0, nn, vtk2ijk));
    if (a2ij) H5_TRY(read_i32(F, "/order/a2ij", 0, nb, a2ij));
    if (gmsh2ij) H5_TRY(read_i32(F, "/order/gmsh2ij", 0, nb, gmsh2ij));
    if (vtk2ij) H5_TRY(read_i32(F, "/order/vtk2ij", 0, nb, vtk2ij));
    // connectivities: element-major in the file, (nelem,nnode) column-major here
    // (load_connectivity_hdf5, :2707-2741)
    const char *cname[2] = {"/Connectivity/connecParOrig", "/Connectivity/connecParWork"};
    int *cout[2] = {connec_orig, connec_work};
    std::vector<int> tmp;
    for (int c = 0; c < 2; ++c) {
        if (!cout[c]) continue;
        tmp.resize((size_t)ne * nn);
        H5_TRY(read_i32(F, cname[c], (long long)(I.elem_start - 1) * nn, (long long)ne * nn,
                        tmp.data()));
        H5_TRY(check_range(cname[c], tmp.data(), (long long)ne * nn, 1, np));
        for (int e = 0; e < ne; ++e)
            for (int a = 0; a < nn; ++a) cout[c][e + (size_t)ne * a] = tmp[(size_t)e * nn + a];
    }
    long long off;
    int cnt;
    if (working_nodes) {
        H5_TRY(part_slice(F, "/Connectivity/numWorkingNodesRankPar", part, I.nparts, &off, &cnt));
        H5_TRY(read_i32(F, "/Connectivity/workingNodesPar", off, cnt, working_nodes));
        H5_TRY(check_range("workingNodesPar", working_nodes, cnt, 1, np));
    }
    if (owned_nodes) {
        H5_TRY(part_slice(F, "/Connectivity/numOwnedNodesRankPar", part, I.nparts, &off, &cnt));
        H5_TRY(read_i32(F, "/Connectivity/ownedNodesPar", off, cnt, owned_nodes));
        H5_TRY(check_range("ownedNodesPar", owned_nodes, cnt, 1, np));
    }
    // coordinates: (nodes, 3) rows in C order (load_coordinates_hdf5, :3380-3406)
    if (coord) {
        std::vector<double> xyz((size_t)np * 3);
        const char *pts = I.lineal_output ? "/VTKHDF/Points" : "/Coords/Points";
        H5_TRY(read_rows(F, pts, H5T_NATIVE_DOUBLE, I.node_start - 1, np, 3, xyz.data()));
        for (int i = 0; i < np; ++i)
            for (int d = 0; d < 3; ++d) coord[i + (size_t)np * d] = xyz[(size_t)i * 3 + d];
    }
    // periodic master/slave pairs (load_periodic_data_hdf5)
    if (mas_sla && I.periodic) {
        H5_TRY(part_slice(F, "/Periodic_data/nPerRankPar", part, I.nparts, &off, &cnt));
        H5_TRY(read_i32(F, "/Periodic_data/masSlaRankPar1", off, cnt, mas_sla));
        H5_TRY(read_i32(F, "/Periodic_data/masSlaRankPar2", off, cnt, mas_sla + cnt));
        H5_TRY(check_range("masSlaRankPar", mas_sla, 2LL * cnt, 1, np));
    }
    // global ids (load_globalIds_hdf5, :3442-3485)
    if (gid_srl)
        H5_TRY(read_rows(F, "/globalIds/globalIdSrl", H5T_NATIVE_LLONG, I.node_start - 1, np, 0,
                         gid_srl));
    if (gid_par)
        H5_TRY(read_rows(F, "/globalIds/globalIdPar", H5T_NATIVE_LLONG, I.node_start - 1, np, 0,
                         gid_par));
    if (elem_gid) H5_TRY(read_i32(F, "/globalIds/elemGid", I.elem_start - 1, ne, elem_gid));
    // VTK connectivity (load_vtk_connectivity_hdf5, :2813-2839)
    if (connec_vtk)
        H5_TRY(read_rows(F, "/VTKHDF/Connectivity", H5T_NATIVE_LLONG,
                         (long long)(I.elem_start - 1) * I.nvtk_per_elem * I.nnode_vtk,
                         I.size_connec_vtk, 0, connec_vtk));
    // the part's comm lists (load_parallel_data_hdf5)
    if (I.nparts >= 2) {
        int *outs[4] = {ranks_to_comm, comms_pos_loc, comms_size, comms_pos_ngb};
        const char *names[4] = {"/Parallel_data/ranksToComm", "/Parallel_data/commsMemPosInLoc",
                                "/Parallel_data/commsMemSize", "/Parallel_data/commsMemPosInNgb"};
        bool any = false;
        for (auto *o : outs) any = any || o;
        if (any) {
            H5_TRY(part_slice(F, "/Parallel_data/numRanksWithComms", part, I.nparts, &off, &cnt));
            for (int k = 0; k < 4; ++k)
                if (outs[k]) H5_TRY(read_i32(F, names[k], off, cnt, outs[k]));
        }
        if (nodes_to_comm) {
            H5_TRY(part_slice(F, "/Parallel_data/numNodesToComm", part, I.nparts, &off, &cnt));
            H5_TRY(read_i32(F, "/Parallel_data/nodesToComm", off, cnt, nodes_to_comm));
            H5_TRY(check_range("nodesToComm", nodes_to_comm, cnt, 1, np));
        }
    }
    return SODH5_OK;
}

extern "C" int sodh5_mesh_write(const char *path, int porder, int nparts, const sodh5_part *parts,
                                int periodic, const int *a2ijk, const int *gmsh2ijk,
                                const int *vtk2ijk, const int *a2ij, const int *gmsh2ij,
                                const int *vtk2ij)
{
    if (!path || !parts || nparts < 1) return fail(SODH5_ERR_ARG, "no parts to write");
    if (porder < 1 || porder > 16) return fail(SODH5_ERR_ARG, "porder %d", porder);
    if (!a2ijk || !gmsh2ijk || !vtk2ijk || !a2ij || !gmsh2ij || !vtk2ij)
        return fail(SODH5_ERR_ARG, "all six order tables are required");
    const int nn = (porder + 1) * (porder + 1) * (porder + 1), nb = (porder + 1) * (porder + 1);
    long long tn = 0, te = 0;
    for (int p = 0; p < nparts; ++p) {
        const sodh5_part &P = parts[p];
        if (P.nelem < 0 || P.npoin < 0 || P.npoin_w < 0 || P.npoin_owned < 0 || P.nper < 0 ||
            P.nranks_comm < 0 || P.nnodes_comm < 0)
            return fail(SODH5_ERR_ARG, "part %d: negative count", p);
        if ((P.nelem && (!P.connec_orig || !P.connec_work)) || (P.npoin && !P.coord) ||
            (P.npoin_w && !P.working_nodes) || (P.npoin_owned && !P.owned_nodes) ||
            (P.nper && !P.mas_sla) ||
            (P.nranks_comm && (!P.ranks_to_comm || !P.comms_pos_loc || !P.comms_pos_ngb ||
                               !P.comms_size)) ||
            (P.nnodes_comm && !P.nodes_to_comm))
            return fail(SODH5_ERR_ARG, "part %d: missing array", p);
        if (P.nper) periodic = 1;
        if (P.nelem) {
            H5_TRY(check_range("connec_orig", P.connec_orig, (long long)P.nelem * nn, 1, P.npoin));
            H5_TRY(check_range("connec_work", P.connec_work, (long long)P.nelem * nn, 1, P.npoin));
        }
        tn += P.npoin;
        te += P.nelem;
    }
    if (te > INT_MAX || te * nn > (long long)1 << 62) return fail(SODH5_ERR_ARG, "too large");
    quiet();
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(SODH5_ERR_IO, "cannot create %s", path);
    const hid_t F = f.id;
    for (const char *g : {"/globalIds", "/Coords", "/Connectivity", "/order", "/meshOutputInfo",
                          "/VTKHDF"})
        H5_TRY(make_group(F, g));
    // global ids and coordinates over all parts, in part order
    std::vector<long long> srl, par;
    std::vector<int> egid;
    std::vector<double> xyz;
    srl.reserve(tn);
    par.reserve(tn);
    egid.reserve(te);
    xyz.reserve(3 * tn);
    long long noff = 0, eoff = 0;
    for (int p = 0; p < nparts; ++p) {
        const sodh5_part &P = parts[p];
        for (int i = 0; i < P.npoin; ++i) {
            srl.push_back(P.gid_srl ? P.gid_srl[i] : noff + i + 1);
            par.push_back(P.gid_par ? P.gid_par[i] : noff + i + 1);
            for (int d = 0; d < 3; ++d) xyz.push_back(P.coord[i + (size_t)P.npoin * d]);
        }
        for (int e = 0; e < P.nelem; ++e) egid.push_back(P.elem_gid ? P.elem_gid[e] : (int)(eoff + e + 1));
        noff += P.npoin;
        eoff += P.nelem;
    }
    H5_TRY(write_i64(F, "/globalIds/globalIdSrl", srl));
    H5_TRY(write_i64(F, "/globalIds/globalIdPar", par));
    H5_TRY(write_i32(F, "/globalIds/elemGid", egid));
    {
        hsize_t dims[2] = {(hsize_t)tn, 3};
        H5_TRY(write_ds(F, "/Coords/Points", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, 2, dims, xyz.data()));
    }
    // connectivity, element-major; per-part node lists concatenated
    std::vector<int> co, cw, wn, on, nwn, non;
    for (int p = 0; p < nparts; ++p) {
        const sodh5_part &P = parts[p];
        for (int e = 0; e < P.nelem; ++e)
            for (int a = 0; a < nn; ++a) {
                co.push_back(P.connec_orig[e + (size_t)P.nelem * a]);
                cw.push_back(P.connec_work[e + (size_t)P.nelem * a]);
            }
        wn.insert(wn.end(), P.working_nodes, P.working_nodes + P.npoin_w);
        on.insert(on.end(), P.owned_nodes, P.owned_nodes + P.npoin_owned);
        nwn.push_back(P.npoin_w);
        non.push_back(P.npoin_owned);
    }
    H5_TRY(write_i32(F, "/Connectivity/connecParOrig", co));
    H5_TRY(write_i32(F, "/Connectivity/connecParWork", cw));
    H5_TRY(write_i32(F, "/Connectivity/numWorkingNodesRankPar", nwn));
    H5_TRY(write_i32(F, "/Connectivity/numOwnedNodesRankPar", non));
    H5_TRY(write_i32(F, "/Connectivity/workingNodesPar", wn));
    H5_TRY(write_i32(F, "/Connectivity/ownedNodesPar", on));
    // order tables
    H5_TRY(write_i32(F, "/order/porder", std::vector<int>{porder}));
    H5_TRY(write_i32(F, "/order/a2ijk", std::vector<int>(a2ijk, a2ijk + nn)));
    H5_TRY(write_i32(F, "/order/gmsh2ijk", std::vector<int>(gmsh2ijk, gmsh2ijk + nn)));
    H5_TRY(write_i32(F, "/order/vtk2ijk", std::vector<int>(vtk2ijk, vtk2ijk + nn)));
    H5_TRY(write_i32(F, "/order/a2ij", std::vector<int>(a2ij, a2ij + nb)));
    H5_TRY(write_i32(F, "/order/gmsh2ij", std::vector<int>(gmsh2ij, gmsh2ij + nb)));
    H5_TRY(write_i32(F, "/order/vtk2ij", std::vector<int>(vtk2ij, vtk2ij + nb)));
    // high-order element output: one VTK cell of nnode nodes per element
    {
        hsize_t one = 1;
        unsigned char lin = 0;
        H5_TRY(write_ds(F, "/meshOutputInfo/isLinealOutput", H5T_STD_U8LE, H5T_NATIVE_UCHAR, 1,
                        &one, &lin));
    }
    H5_TRY(write_i32(F, "/meshOutputInfo/nnodeVTK", std::vector<int>{nn}));
    H5_TRY(write_i32(F, "/meshOutputInfo/numVTKElemsPerMshElem", std::vector<int>{1}));
    std::vector<int> nev, scv;
    for (int p = 0; p < nparts; ++p) {
        nev.push_back(parts[p].nelem);
        scv.push_back(parts[p].nelem * nn);
    }
    H5_TRY(write_i32(F, "/meshOutputInfo/numElemsVTKMshRank", nev));
    H5_TRY(write_i32(F, "/meshOutputInfo/sizeConnecVTKMshRank", scv));
    // VTK connectivity: the element's nodes in VTK order, part-local, 0-based.
    // vtk2ijk and a2ijk are indexed by the same ijk position, so VTK point v is
    // local node a2ijk(c) where vtk2ijk(c) = v.
    {
        std::vector<int> a_of_v(nn, -1);
        for (int c = 0; c < nn; ++c) {
            int v = vtk2ijk[c] - 1, a = a2ijk[c] - 1;
            if (v < 0 || v >= nn || a < 0 || a >= nn || a_of_v[v] >= 0)
                return fail(SODH5_ERR_ARG, "vtk2ijk/a2ijk are not permutations of 1..%d", nn);
            a_of_v[v] = a;
        }
        std::vector<long long> cv;
        cv.reserve((size_t)te * nn);
        for (int p = 0; p < nparts; ++p) {
            const sodh5_part &P = parts[p];
            for (int e = 0; e < P.nelem; ++e)
                for (int v = 0; v < nn; ++v)
                    cv.push_back((long long)P.connec_orig[e + (size_t)P.nelem * a_of_v[v]] - 1);
        }
        H5_TRY(write_i64(F, "/VTKHDF/Connectivity", cv));
    }
    if (periodic) {
        H5_TRY(make_group(F, "/Periodic_data"));
        std::vector<int> np_, m1, m2;
        for (int p = 0; p < nparts; ++p) {
            const sodh5_part &P = parts[p];
            np_.push_back(P.nper);
            m1.insert(m1.end(), P.mas_sla, P.mas_sla + P.nper);
            m2.insert(m2.end(), P.mas_sla + P.nper, P.mas_sla + 2 * (size_t)P.nper);
        }
        H5_TRY(write_i32(F, "/Periodic_data/nPerRankPar", np_));
        H5_TRY(write_i32(F, "/Periodic_data/masSlaRankPar1", m1));
        H5_TRY(write_i32(F, "/Periodic_data/masSlaRankPar2", m2));
    }
    if (nparts >= 2) {
        H5_TRY(make_group(F, "/Parallel_data"));
        std::vector<long long> ns, ne_;
        std::vector<int> es, ee, nrc, nnc, rtc, pl, pn, cs, ntc;
        long long n0 = 1;
        int e0 = 1;
        for (int p = 0; p < nparts; ++p) {
            const sodh5_part &P = parts[p];
            ns.push_back(n0);
            ne_.push_back(n0 + P.npoin - 1);
            es.push_back(e0);
            ee.push_back(e0 + P.nelem - 1);
            n0 += P.npoin;
            e0 += P.nelem;
            nrc.push_back(P.nranks_comm);
            nnc.push_back(P.nnodes_comm);
            for (int i = 0; i < P.nranks_comm; ++i) {
                rtc.push_back(P.ranks_to_comm[i]);
                pl.push_back(P.comms_pos_loc[i]);
                pn.push_back(P.comms_pos_ngb[i]);
                cs.push_back(P.comms_size[i]);
            }
            ntc.insert(ntc.end(), P.nodes_to_comm, P.nodes_to_comm + P.nnodes_comm);
        }
        H5_TRY(write_i64(F, "/Parallel_data/rankNodeStart", ns));
        H5_TRY(write_i64(F, "/Parallel_data/rankNodeEnd", ne_));
        H5_TRY(write_i32(F, "/Parallel_data/rankElemStart", es));
        H5_TRY(write_i32(F, "/Parallel_data/rankElemEnd", ee));
        H5_TRY(write_i32(F, "/Parallel_data/numRanksWithComms", nrc));
        H5_TRY(write_i32(F, "/Parallel_data/numNodesToComm", nnc));
        H5_TRY(write_i32(F, "/Parallel_data/ranksToComm", rtc));
        H5_TRY(write_i32(F, "/Parallel_data/commsMemPosInLoc", pl));
        H5_TRY(write_i32(F, "/Parallel_data/commsMemPosInNgb", pn));
        H5_TRY(write_i32(F, "/Parallel_data/commsMemSize", cs));
        H5_TRY(write_i32(F, "/Parallel_data/nodesToComm", ntc));
    }
    if (H5Fflush(F, H5F_SCOPE_GLOBAL) < 0) return fail(SODH5_ERR_IO, "flushing %s", path);
    return SODH5_OK;
}

// ----------------------------------------------------------------- restart
namespace {
// restart dataset names (initializeNameFields, mod_saveFields.f90:265-275):
// the element-node fields: mu_e is stored as "mut", mu_t as "mue"
const char *const kRho = "rho", *const kUx = "u_x", *const kUy = "u_y", *const kUz = "u_z",
                  *const kPr = "pr", *const kE = "E", *const kMuE = "mut", *const kMuT = "mue",
                  *const kWx = "walave_u_x", *const kWy = "walave_u_y", *const kWz = "walave_u_z";
}  // namespace

extern "C" int sodh5_restart_write(const char *path, int rp_vtk, int npoin, int nelem, int nnode,
                                   const int *connec_orig, const double *rho, const double *u,
                                   const double *pr, const double *E, const double *mu_e,
                                   const double *mu_t, const double *walave_u, double time,
                                   int istep)
{
    if (!path || npoin < 0 || nelem < 0 || nnode < 1)
        return fail(SODH5_ERR_ARG, "restart_write: bad sizes");
    if (rp_vtk != 4 && rp_vtk != 8) return fail(SODH5_ERR_ARG, "rp_vtk %d (4 or 8)", rp_vtk);
    if (!rho || !u || !pr || !E || !mu_e || (nelem && !connec_orig))
        return fail(SODH5_ERR_ARG, "restart_write: missing field");
    if (nelem) H5_TRY(check_range("connec_orig", connec_orig, (long long)nelem * nnode, 1, npoin));
    quiet();
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(SODH5_ERR_IO, "cannot create %s", path);
    const hid_t F = f.id;
    const hid_t ft = rp_vtk == 4 ? H5T_IEEE_F32LE : H5T_IEEE_F64LE;
    const hsize_t n = (hsize_t)npoin;
    auto put = [&](const char *name, const double *v) {
        return write_ds(F, name, ft, H5T_NATIVE_DOUBLE, 1, &n, v);
    };
    H5_TRY(put(kRho, rho));
    H5_TRY(put(kUx, u));
    H5_TRY(put(kUy, u + (size_t)npoin));
    H5_TRY(put(kUz, u + 2 * (size_t)npoin));
    H5_TRY(put(kPr, pr));
    H5_TRY(put(kE, E));
    // element-node fields to nodes, element order, the last writer of a node wins
    // (save_hdf5_restartFile, mod_hdf5.f90:4116-4122)
    std::vector<double> aux(npoin, 0.0);
    auto scatter = [&](const double *v) {
        std::fill(aux.begin(), aux.end(), 0.0);
        if (v)
            for (int e = 0; e < nelem; ++e)
                for (int a = 0; a < nnode; ++a)
                    aux[connec_orig[e + (size_t)nelem * a] - 1] = v[e + (size_t)nelem * a];
    };
    scatter(mu_e);
    H5_TRY(put(kMuE, aux.data()));
    scatter(mu_t);
    H5_TRY(put(kMuT, aux.data()));
    if (walave_u) {
        H5_TRY(put(kWx, walave_u));
        H5_TRY(put(kWy, walave_u + (size_t)npoin));
        H5_TRY(put(kWz, walave_u + 2 * (size_t)npoin));
    }
    hsize_t one = 1;
    H5_TRY(write_ds(F, "time", ft, H5T_NATIVE_DOUBLE, 1, &one, &time));
    H5_TRY(write_ds(F, "istep", H5T_STD_I32LE, H5T_NATIVE_INT, 1, &one, &istep));
    if (H5Fflush(F, H5F_SCOPE_GLOBAL) < 0) return fail(SODH5_ERR_IO, "flushing %s", path);
    return SODH5_OK;
}

extern "C" int sodh5_restart_read(const char *path, long long node_start, int npoin, int nelem,
                                  int nnode, const int *connec_orig, int nper, const int *mas_sla,
                                  double *rho, double *u, double *pr, double *E, double *mu_e,
                                  double *mu_t, double *walave_u, double *time, int *istep,
                                  int *rp_vtk_out)
{
    if (npoin < 0 || nelem < 0 || nnode < 1 || nper < 0 || node_start < 1)
        return fail(SODH5_ERR_ARG, "restart_read: bad sizes");
    if ((mu_e || mu_t) && nelem && !connec_orig)
        return fail(SODH5_ERR_ARG, "restart_read: mu_e/mu_t need connec_orig");
    if (nper && !mas_sla) return fail(SODH5_ERR_ARG, "restart_read: nper > 0 needs mas_sla");
    if (nelem && connec_orig)
        H5_TRY(check_range("connec_orig", connec_orig, (long long)nelem * nnode, 1, npoin));
    if (nper) H5_TRY(check_range("mas_sla", mas_sla, 2LL * nper, 1, npoin));
    Hid f;
    H5_TRY(open_ro(path, f));
    const hid_t F = f.id;
    const long long off = node_start - 1;
    auto get = [&](const char *name, double *v) {
        return read_rows(F, name, H5T_NATIVE_DOUBLE, off, npoin, 0, v);
    };
    if (rp_vtk_out) {
        int cls, size, sg;
        int r = 0;
        long long d[4];
        H5_TRY(sodh5_dataset_info(path, kRho, &r, d, &cls, &size, &sg));
        *rp_vtk_out = size;
    }
    if (time) H5_TRY(read_rows(F, "time", H5T_NATIVE_DOUBLE, 0, 1, 0, time));
    if (istep) H5_TRY(read_one_i32(F, "istep", 0, istep));
    if (rho) H5_TRY(get(kRho, rho));
    if (u) {
        H5_TRY(get(kUx, u));
        H5_TRY(get(kUy, u + (size_t)npoin));
        H5_TRY(get(kUz, u + 2 * (size_t)npoin));
    }
    if (pr) H5_TRY(get(kPr, pr));
    if (E) H5_TRY(get(kE, E));
    // node values back to element nodes through connecParOrig (:4288-4295)
    std::vector<double> aux(npoin);
    const char *en[2] = {kMuE, kMuT};
    double *eo[2] = {mu_e, mu_t};
    for (int k = 0; k < 2; ++k) {
        if (!eo[k]) continue;
        H5_TRY(get(en[k], aux.data()));
        for (int e = 0; e < nelem; ++e)
            for (int a = 0; a < nnode; ++a)
                eo[k][e + (size_t)nelem * a] = aux[connec_orig[e + (size_t)nelem * a] - 1];
    }
    if (walave_u) {
        if (link_exists(F, kWx)) {
            H5_TRY(get(kWx, walave_u));
            H5_TRY(get(kWy, walave_u + (size_t)npoin));
            H5_TRY(get(kWz, walave_u + 2 * (size_t)npoin));
        } else {
            // no walave_u in the file: it starts from u (:4289-4291)
            std::vector<double> tmp(npoin);
            for (int d = 0; d < 3; ++d) {
                H5_TRY(get(d == 0 ? kUx : (d == 1 ? kUy : kUz), tmp.data()));
                std::copy(tmp.begin(), tmp.end(), walave_u + (size_t)npoin * d);
            }
        }
    }
    // periodic slaves take their master's values (:4297-4321)
    for (int i = 0; i < nper; ++i) {
        const size_t m = mas_sla[i] - 1, s = mas_sla[i + nper] - 1;
        if (rho) rho[s] = rho[m];
        if (pr) pr[s] = pr[m];
        if (E) E[s] = E[m];
        for (int d = 0; d < 3; ++d) {
            if (u) u[s + (size_t)npoin * d] = u[m + (size_t)npoin * d];
            if (walave_u) walave_u[s + (size_t)npoin * d] = walave_u[m + (size_t)npoin * d];
        }
    }
    return SODH5_OK;
}

extern "C" int sodh5_restart_state(int npoin, double gamma, double Rgas, const double *rho,
                                   const double *u, const double *pr, const double *E, double *q,
                                   double *e_int, double *Tem, double *csound, double *eta,
                                   double *machno)
{
    if (npoin < 0 || !rho || !u || !pr || !E) return fail(SODH5_ERR_ARG, "restart_state: args");
    const size_t n = (size_t)npoin;
    for (size_t i = 0; i < n; ++i) {
        const double r = rho[i], p = pr[i];
        const double ux = u[i], uy = u[i + n], uz = u[i + 2 * n];
        const double uu = ux * ux + uy * uy + uz * uz;
        if (q) {
            q[i] = r * ux;
            q[i + n] = r * uy;
            q[i + 2 * n] = r * uz;
        }
        if (e_int) e_int[i] = E[i] / r - 0.5 * uu;
        if (Tem) Tem[i] = p / (r * Rgas);
        const double c = std::sqrt(gamma * p / r);
        if (csound) csound[i] = c;
        if (eta) eta[i] = (r / (gamma - 1.0)) * std::log(p / std::pow(r, gamma));
        if (machno) machno[i] = std::sqrt(uu) / c;
    }
    return SODH5_OK;
}
