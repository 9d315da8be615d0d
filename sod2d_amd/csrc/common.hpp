// common.hpp -- shared definitions of the sod2d_hip library (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sod2d_hip.h"

namespace sodhip {

// ---------------------------------------------------------------- errors
int set_error(int code, const char *fmt, ...);

#define SODHIP_TRY_HIP(expr)                                                               \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::sodhip::set_error(SODHIP_ERR_HIP, "%s failed: %s (%s:%d)", #expr,     \
                                       hipGetErrorString(_e), __FILE__, __LINE__);         \
    } while (0)

#define SODHIP_REQUIRE(cond, ...)                                                          \
    do {                                                                                   \
        if (!(cond)) return ::sodhip::set_error(SODHIP_ERR_ARG, __VA_ARGS__);              \
    } while (0)

// ---------------------------------------------------------------- element sizes
constexpr int kMaxP = 4; // device kernels are instantiated for porder 2..4

// values per node projection record of the fused LPS stage: the upper triangle
// t11 t12 t13 t22 t23 t33 of the projected tensor 2S - 2/3 div u I, then
// ProjMass (grad rho) and ProjEner (grad T); 96 bytes
constexpr int kProjRec = 12;

// size of the allocation-free stage's record scratch, per GPU
#ifndef SODHIP_NOALLOC_SCRATCH_MIB
#define SODHIP_NOALLOC_SCRATCH_MIB 64
#endif

inline int nnode_of(int p) { return (p + 1) * (p + 1) * (p + 1); }

// ---------------------------------------------------------------- device buffers
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    // set while a captured graph holds the address: the buffer may then not be
    // reallocated, and a request beyond its size is an error
    bool pinned = false;
};

// Node-ordered incidence of a connectivity: node n (0-based) owns the slots
// row[n]..row[n+1]-1, one per (e,a) with connec(e,a) == n+1, in ascending (e,a)
// order -- the order in which the reference's serial element loop adds element
// contributions.  Element kernels write the residual of (e,a) to the scratch
// position e*nnode + perm[a] (whole elements, so the writes are coalesced;
// perm groups an element's nodes by their (k, j, i) line indices so that the
// four cache lines of an element hold one k-plane each); src[slot] is that
// position, read by the per-node gathers in slot order.
// Element clusters of a p=3 mesh for the brick-local assembly of the fused
// stage: a block evaluates the (up to) kClusterG elements of a cluster in a
// row and sums their residuals in LDS per local node.  Nodes all of whose
// elements lie in the cluster ("interior") are finished by the element kernel
// itself (mass inverse and stage combine); the others ("surface") get one
// partial record per cluster that contains them, which the surface gather
// sums in cluster order.
constexpr int kClusterG = 8;      // elements per cluster (a 2x2x2 brick on a box mesh)
constexpr int kClusterLMax = 344; // local nodes per cluster (343 for a 2x2x2 brick at p=3)
constexpr int kClusterBlockMax = 16; // clusters one block of the element kernel walks
struct ClusterTables {
    int ncl = 0;          // clusters
    int nloc = 0;         // local nodes over all clusters (= partial records)
    int *cel = nullptr;   // ncl*kClusterG element ids in evaluation order, -1 = none
    int *cloc = nullptr;  // (ncl*kClusterG, 64): local index of the node at each
                          // lexicographic position of the element
    int *cnb = nullptr;   // ncl+1: first local node of each cluster
    int *cnode = nullptr; // nloc: global node (0-based), ascending within a cluster
    int *ckind = nullptr; // nloc: 0 interior working node, 1 surface working node, 2 other
    int ns = 0;           // working nodes finished by the surface gather
    int *snode = nullptr; // ns: node ids, 1-based (working-list order)
    int *srow = nullptr;  // ns+1
    int *ssrc = nullptr;  // record indices (cnb[c] + l), ascending cluster per node
};

struct MeshCache {
    const int *connec_key = nullptr; // device pointer of the caller's connec
    int porder = 0, nelem = 0, npoin = 0;
    int *connecT = nullptr; // (nelem, nnode) element-major, 0-based node ids
    int *row = nullptr;     // npoin+1
    int *src = nullptr;     // nelem*nnode, scratch position of each slot
    int *perm = nullptr;    // nnode, element-local scratch position of node a
    // host copies (for derived tables)
    std::vector<int> h_row, h_src, h_cT;
    std::vector<unsigned char> h_wflag; // working flags of the current list
    // clusters of the current working list (p=3), built on first use
    std::unique_ptr<ClusterTables> clusters;
    // the incidence restricted to a working-node list, in list order:
    // working node i owns wsrc[wrow[i]..wrow[i+1]-1]
    const int *lpw_key = nullptr; // device pointer of the caller's lpoin_w
    int npoin_w = 0;
    int *wrow = nullptr, *wsrc = nullptr;
    unsigned char *wflag = nullptr; // npoin: 1 where the node is in the list
};

struct PresentEntry {
    void *dev = nullptr;
    size_t bytes = 0;
};

struct ProfSlot {
    double total_ms = 0.0;
    long long count = 0;
};

constexpr int kProfKinds = 5;

} // namespace sodhip

struct sodhip_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<uintptr_t, sodhip::PresentEntry> present; // host start address -> mirror
    std::vector<std::unique_ptr<sodhip::MeshCache>> meshes;
    sodhip::DevBuf scratch;  // per element-node residual partials
    sodhip::DevBuf scratch2; // reduction partials of the step (the sensor's normalisation)
    sodhip::DevBuf diag;     // reduction partials read back by the host (dt, diagnostics)
    sodhip::DevBuf hostbuf;  // pinned staging (small reductions)
    bool prof = false;
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    sodhip::ProfSlot slots[sodhip::kProfKinds];
    int num_cu = 256;
    // brick-local assembly for the fused p=3 FP64 stage, opt-in
    // (SODHIP_STAGE_CLUSTERS=1 in the environment at sodhip_create): its LDS
    // accumulator costs the element kernel a quarter of its resident blocks,
    // more than the smaller gather saves (measured 0.79 -> 1.50 ms element,
    // 0.37 -> 0.29 ms gather per stage at 64^3)
    bool cluster_stage = false;
    // launch shape of the fused stage (sodhip_set_stage_config): element-kernel
    // blocks per CU, 16-byte geometry rows when aligned, gather block size
    int stage_bpc = 16;
    int stage_geo16 = 1;
    int gather_threads = 256;
};

namespace sodhip {

// grows ctx->scratch to at least `bytes`
int ensure_buf(DevBuf &b, size_t bytes);
// mesh cache lookup/build for (connec, porder, nelem, npoin)
int get_mesh(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec,
             MeshCache **out);
// working-node incidence of mesh m for the list lpoin_w (1-based, npoin_w entries)
int get_working(sodhip_ctx ctx, MeshCache *m, int npoin_w, const int *lpoin_w);
// lookup only (allocates nothing): the cached tables of (connec, porder, nelem,
// npoin) whose working incidence was built for lpoin_w; an error if a call that
// may allocate has not built them yet
int find_mesh(sodhip_ctx ctx, int porder, int nelem, int npoin, const int *connec, int npoin_w,
              const int *lpoin_w, const MeshCache **out);
// the element clusters of mesh m for its current working list (p=3), built
// on first use (allocates); find_clusters only looks them up
int get_clusters(sodhip_ctx ctx, MeshCache *m, const ClusterTables **out);
void free_clusters(MeshCache *m);
// maps the allocation-free stage's scratch array on `device` (the current
// device), once per device (ops_stage.hip)
int noalloc_init(int device);
// updateF's EOS on the working nodes of m (after get_working) and Sutherland's
// law with a uniform factor on all its nodes, in one pass (ops_node.hip)
int eos_stage_sutherland(sodhip_ctx ctx, const MeshCache *m, double gamma_gas, double Rgas,
                         double mu_factor, const double *rho, const double *q, const double *E,
                         double *u, double *e_int, double *pr, double *aux_h, double *Tem,
                         double *mu_fluid);

// launch timing helpers: prof_begin returns a token (or -1); prof_end closes it
int prof_begin(sodhip_ctx ctx, int kind);
void prof_end(sodhip_ctx ctx, int token);

// XCD-aware grid: number of blocks for a grid-stride element loop
int elem_grid(sodhip_ctx ctx, int nelem, int blocks_per_cu);
// grid of the lane-per-node element loops (xcd_elems): a multiple of 8, at
// most blocks_per_cu blocks per CU
int diag_grid(sodhip_ctx ctx, int nelem, int blocks_per_cu = 32);


// first stage of the next step stored by a stage's gather/update (ops_stage.hip):
// y_next = nl1 * y + 0 * 0 + ngdt * R
struct NextStage {
    double nl1, ngdt;
    double *rho, *q, *E;
};
// gather/update of one stage over the working nodes list[0..nw) (1-based)
// with their records src[row[i]..row[i+1]) (ops_stage.hip)
template <typename TS>
int gather_update_on(sodhip_ctx ctx, int npoin, int nw, const int *list, const int *row,
                     const int *src, const TS *scr, int rw, const double *invMl, double fact,
                     double dt, double lam1, double lams, double gam, const double *rho1,
                     const double *q1, const double *E1, double *rho, double *q, double *E,
                     double *Rmass, double *Rmom, double *Rener, const NextStage *next);
// convection and the LPS-stabilised viscous operator (full_diff_stab_ijk) of
// the fused LPS stage at p=3, FP64, from the conserved state, with the node
// projection records prj (kProjRec per node) and comp_tau's tau; records of
// five per element-node into out (ops_element.hip)
int rhs_lps_elements(sodhip_ctx ctx, const MeshCache *m, const double *He, const double *gpvol,
                     const double *dl, const int *inv, const int *gI, const int *gJ, const int *gK,
                     const double *rho, const double *q, const double *E, const double *mu_e,
                     const double *prj, const double *tau, double gamma, double Rgas, double Cp,
                     double Pr, double mu_factor, double c_rho, double c_ener, double c_lps,
                     double *out);
// the fused p=3 FP64 stage's element kernel in cluster mode: interior nodes
// updated in place, surface partial records into rec (ops_element.hip)
int rhs_cluster_stage(sodhip_ctx ctx, const MeshCache *m, const ClusterTables *cl,
                      const double *He, const double *gpvol, const double *dl, const int *inv,
                      const int *gI, const int *gJ, const int *gK, const double *mu_e,
                      double gamma, double Rgas, double Cp, double Pr, double mu_factor,
                      const double *mu_factor_e, double c_rho, double c_ener, double p_ref,
                      const double *invMl, double fact, double dt, double lam1, double lams,
                      double gam, const double *rho1, const double *q1, const double *E1,
                      double *rho, double *q, double *E, double *Rmass, double *Rmom,
                      double *Rener, const NextStage *next, double *rec);
// gather/update of one stage from element records of rw (5 or 6) values
// (ops_stage.hip)
template <typename TS>
int gather_update(sodhip_ctx ctx, const MeshCache *m, int npoin, int npoin_w,
                  const int *lpoin_w, const TS *scr, int rw, size_t EN, const double *invMl,
                  double fact, double dt, double lam1, double lams, double gam,
                  const double *rho1, const double *q1, const double *E1, double *rho, double *q,
                  double *E, double *Rmass, double *Rmom, double *Rener, const NextStage *next);
// full_convec_ijk element records (ops_element.hip)
template <typename T>
int conv_elements(sodhip_ctx ctx, int porder, const MeshCache *m, const T *He, const T *gpvol,
                  const T *dl, const int *inv, const int *gI, const int *gJ, const int *gK,
                  const T *u, const T *q, const T *rho, const T *pr, const T *E, T *out);
} // namespace sodhip
