// setup.cpp -- host-side case set-up: element tables, structured box meshes in
// SOD2D's partition format, element geometry, lumped mass and element lengths.
//
// These produce the inputs of the hot-path operators the way the reference's
// set-up does (the reference computes them once in CFDSolverBase_run); they run
// on the CPU and never touch the device.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "common.hpp"

namespace {

using sodhip::set_error;

// Run f(begin,end) over [0,n) on the host's cores.
void parallel_for(long n, const std::function<void(long, long)> &f)
{
    unsigned hw = std::thread::hardware_concurrency();
    long nt = std::max(1u, std::min(hw ? hw : 1u, 32u));
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    long chunk = (n + nt - 1) / nt;
    for (long t = 0; t < nt; ++t) {
        long b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(f, b, e);
    }
    for (auto &t : th) t.join();
}

// Legendre polynomial P_n(x) by the three-term recurrence.
double legendre(int n, double x)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) return p0;
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// GLL nodes (ascending) and weights: interior nodes are the roots of P_p',
// found by Newton on (1-x^2) P_p'(x) = p (P_{p-1}(x) - x P_p(x)), whose
// derivative is -p (p+1) P_p(x), from Chebyshev-Gauss-Lobatto guesses.
void gll_nodes(int p, double *x, double *w)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i <= p; ++i) x[i] = -std::cos(pi * i / p);
    for (int i = 1; i < p; ++i) {
        double xi = x[i];
        for (int it = 0; it < 100; ++it) {
            double pp = legendre(p, xi), pm = legendre(p - 1, xi);
            double f = p * (pm - xi * pp);
            double df = -p * (p + 1.0) * pp;
            double dx = f / df;
            xi -= dx;
            if (std::fabs(dx) < 1e-17) break;
        }
        x[i] = xi;
    }
    x[0] = -1.0;
    x[p] = 1.0;
    for (int i = 0; i < (p + 1) / 2; ++i) { // exact symmetry
        double s = 0.5 * (x[p - i] - x[i]);
        x[i] = -s;
        x[p - i] = s;
    }
    if (p % 2 == 0) x[p / 2] = 0.0;
    for (int i = 0; i <= p; ++i) {
        double pp = legendre(p, x[i]);
        w[i] = 2.0 / (p * (p + 1.0) * pp * pp);
    }
}

// Derivative matrix of the GLL Lagrange basis in geometric order:
// D[i][j] = l_j'(x_i) (closed form for Legendre-Gauss-Lobatto points).
void gll_dmat(int p, const double *x, double *D)
{
    int n1 = p + 1;
    for (int i = 0; i <= p; ++i)
        for (int j = 0; j <= p; ++j) {
            double v;
            if (i != j)
                v = legendre(p, x[i]) / (legendre(p, x[j]) * (x[i] - x[j]));
            else if (i == 0)
                v = -p * (p + 1.0) / 4.0;
            else if (i == p)
                v = p * (p + 1.0) / 4.0;
            else
                v = 0.0;
            D[i * n1 + j] = v;
        }
}

// sod2d 1-D index s (0-based) -> geometric position along the axis:
// s = 0 -> 0 (xi=-1), s = 1 -> p (xi=+1), s >= 2 -> s-1 (interior, ascending).
inline int s2g(int p, int s) { return s == 0 ? 0 : (s == 1 ? p : s - 1); }

// gmsh high-order hexahedron node order: vertices, edge-interior nodes (gmsh
// edge order), face-interior nodes (gmsh face order; each face's interior is a
// gmsh quadrilateral of order p-2 placed from the face's first vertex along its
// first and last edges), then the interior as a hexahedron of order p-2.
void gmsh_quad_order(int p, std::vector<std::array<int, 2>> &out)
{
    out.clear();
    if (p < 0) return;
    out.push_back({0, 0});
    if (p == 0) return;
    const std::array<int, 2> v[4] = {{0, 0}, {p, 0}, {p, p}, {0, p}};
    out.push_back(v[1]);
    out.push_back(v[2]);
    out.push_back(v[3]);
    const int ed[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    for (auto &e : ed)
        for (int i = 1; i < p; ++i) {
            int du = (v[e[1]][0] - v[e[0]][0]) / p, dv = (v[e[1]][1] - v[e[0]][1]) / p;
            out.push_back({v[e[0]][0] + i * du, v[e[0]][1] + i * dv});
        }
    if (p >= 2) {
        std::vector<std::array<int, 2>> inner;
        gmsh_quad_order(p - 2, inner);
        for (auto &q : inner) out.push_back({q[0] + 1, q[1] + 1});
    }
}

} // namespace

namespace sodhip {

void gmsh_hex_order(int p, std::vector<std::array<int, 3>> &out)
{
    out.clear();
    out.push_back({0, 0, 0});
    if (p == 0) return;
    const std::array<int, 3> v[8] = {{0, 0, 0}, {p, 0, 0}, {p, p, 0}, {0, p, 0},
                                      {0, 0, p}, {p, 0, p}, {p, p, p}, {0, p, p}};
    for (int c = 1; c < 8; ++c) out.push_back(v[c]);
    if (p == 1) return;
    const int ed[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                           {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
    for (auto &e : ed)
        for (int i = 1; i < p; ++i) {
            std::array<int, 3> n;
            for (int d = 0; d < 3; ++d) n[d] = v[e[0]][d] + i * (v[e[1]][d] - v[e[0]][d]) / p;
            out.push_back(n);
        }
    const int fc[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {0, 4, 7, 3},
                          {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
    std::vector<std::array<int, 2>> face;
    gmsh_quad_order(p - 2, face);
    for (auto &f : fc) {
        for (auto &q : face) {
            std::array<int, 3> n;
            for (int d = 0; d < 3; ++d) {
                int u = (v[f[1]][d] - v[f[0]][d]) / p, w = (v[f[3]][d] - v[f[0]][d]) / p;
                n[d] = v[f[0]][d] + u * (q[0] + 1) + w * (q[1] + 1);
            }
            out.push_back(n);
        }
    }
    std::vector<std::array<int, 3>> inner;
    gmsh_hex_order(p - 2, inner);
    for (auto &q : inner) out.push_back({q[0] + 1, q[1] + 1, q[2] + 1});
}

// Element tables of order p in the reference's conventions.
struct Tables {
    int p, n1, nn;
    std::vector<int> a2ijk, inv, gI, gJ, gK;  // 1-based values
    std::vector<std::array<int, 3>> gcoord;   // gmsh (i,j,k) of local node a (0-based a)
    std::vector<int> local_of;                // (i + n1*(j + n1*k)) -> local node a (0-based)
    std::vector<double> xi, w, D;             // geometric order
};

static int make_tables(int p, Tables &t)
{
    if (p < 1 || p > 8) return set_error(SODHIP_ERR_ARG, "porder %d outside 1..8", p);
    t.p = p;
    t.n1 = p + 1;
    t.nn = t.n1 * t.n1 * t.n1;
    gmsh_hex_order(p, t.gcoord);
    t.local_of.assign(t.nn, -1);
    for (int a = 0; a < t.nn; ++a) {
        auto &c = t.gcoord[a];
        t.local_of[c[0] + t.n1 * (c[1] + t.n1 * c[2])] = a;
    }
    t.a2ijk.resize(t.nn);
    t.inv.resize(t.nn);
    t.gI.resize(t.nn);
    t.gJ.resize(t.nn);
    t.gK.resize(t.nn);
    int c = 0;
    for (int k = 0; k < t.n1; ++k)
        for (int i = 0; i < t.n1; ++i)
            for (int j = 0; j < t.n1; ++j) {
                int a = t.local_of[s2g(p, i) + t.n1 * (s2g(p, j) + t.n1 * s2g(p, k))];
                t.a2ijk[c++] = a + 1;
                t.inv[i + t.n1 * (j + t.n1 * k)] = a + 1;
                t.gI[a] = i + 1;
                t.gJ[a] = j + 1;
                t.gK[a] = k + 1;
            }
    t.xi.resize(t.n1);
    t.w.resize(t.n1);
    t.D.resize(t.n1 * t.n1);
    gll_nodes(p, t.xi.data(), t.w.data());
    gll_dmat(p, t.xi.data(), t.D.data());
    return SODHIP_OK;
}

} // namespace sodhip

using namespace sodhip;

extern "C" int sodhip_basis_tables(int porder, int *a2ijk, int *invAtoIJK, int *gmshAtoI,
                                   int *gmshAtoJ, int *gmshAtoK, double *xgp, double *wgp,
                                   double *dNgp, double *dlxigp_ip)
{
    Tables t;
    int rc = make_tables(porder, t);
    if (rc) return rc;
    const int nn = t.nn, n1 = t.n1, p = porder;
    if (a2ijk) std::copy(t.a2ijk.begin(), t.a2ijk.end(), a2ijk);
    if (invAtoIJK) std::copy(t.inv.begin(), t.inv.end(), invAtoIJK);
    if (gmshAtoI) std::copy(t.gI.begin(), t.gI.end(), gmshAtoI);
    if (gmshAtoJ) std::copy(t.gJ.begin(), t.gJ.end(), gmshAtoJ);
    if (gmshAtoK) std::copy(t.gK.begin(), t.gK.end(), gmshAtoK);
    for (int a = 0; a < nn; ++a) {
        int g[3] = {t.gcoord[a][0], t.gcoord[a][1], t.gcoord[a][2]};
        if (xgp)
            for (int d = 0; d < 3; ++d) xgp[a + nn * d] = t.xi[g[d]];
        if (wgp) wgp[a] = t.w[g[0]] * t.w[g[1]] * t.w[g[2]];
        if (dlxigp_ip)
            for (int d = 0; d < 3; ++d)
                for (int ii = 0; ii < n1; ++ii)
                    dlxigp_ip[a + nn * (d + 3 * ii)] = t.D[g[d] * n1 + s2g(p, ii)];
    }
    if (dNgp) {
        // dNgp(d, b, a): derivative along d of the shape function of node b at
        // node a -- nonzero only when b lies on a's d-line.
        for (size_t i = 0; i < (size_t)3 * nn * nn; ++i) dNgp[i] = 0.0;
        for (int a = 0; a < nn; ++a)
            for (int b = 0; b < nn; ++b)
                for (int d = 0; d < 3; ++d) {
                    bool on_line = true;
                    for (int o = 0; o < 3; ++o)
                        if (o != d && t.gcoord[a][o] != t.gcoord[b][o]) on_line = false;
                    if (on_line)
                        dNgp[d + 3 * (b + (size_t)nn * a)] = t.D[t.gcoord[a][d] * n1 + t.gcoord[b][d]];
                }
    }
    return SODHIP_OK;
}

namespace {

// VTK Lagrange point number (1-based) of the node at position (i,j,k) of a
// hexahedron of order p, in the convention of the reference's
// vtkHigherOrderHexahedron_pointIndexFromIJK (mod_ijk_indices.f90:360-494):
// the 8 vertices, the edges along i, then along j (4 each, edge nodes by
// increasing parameter), the 4 edges along k (in vertex order), the faces
// normal to i, j, k (lower face first, in-face lexicographic) and the body.
int corner_no(int i, int j) { return i ? (j ? 2 : 1) : (j ? 3 : 0); }

int vtk_hex_no(int p, int i, int j, int k)
{
    const int ib = i == 0 || i == p, jb = j == 0 || j == p, kb = k == 0 || k == p;
    const int m = p - 1;
    if (ib + jb + kb == 3) return 1 + corner_no(i, j) + (k ? 4 : 0);
    int n = 9;
    if (ib + jb + kb == 2) {
        if (!ib) return n + (i - 1) + (j ? 2 * m : 0) + (k ? 4 * m : 0);
        if (!jb) return n + (j - 1) + (i ? m : 3 * m) + (k ? 4 * m : 0);
        return n + 8 * m + (k - 1) + m * corner_no(i, j);
    }
    n += 12 * m;
    if (ib + jb + kb == 1) {
        if (ib) return n + (j - 1) + m * (k - 1) + (i ? m * m : 0);
        n += 2 * m * m;
        if (jb) return n + (i - 1) + m * (k - 1) + (j ? m * m : 0);
        n += 2 * m * m;
        return n + (i - 1) + m * (j - 1) + (k ? m * m : 0);
    }
    n += 6 * m * m;
    return n + (i - 1) + m * ((j - 1) + m * (k - 1));
}

// the same for a quadrilateral (vtkHigherOrderQuadrilateral_pointIndexFromIJ,
// mod_ijk_indices.f90:496-570)
int vtk_quad_no(int p, int i, int j)
{
    const int ib = i == 0 || i == p, jb = j == 0 || j == p, m = p - 1;
    if (ib + jb == 2) return 1 + corner_no(i, j);
    if (ib + jb == 1) return 5 + (!ib ? (i - 1) + (j ? 2 * m : 0) : (j - 1) + (i ? m : 3 * m));
    return 5 + 4 * m + (i - 1) + m * (j - 1);
}

} // namespace

extern "C" int sodhip_order_tables(int porder, int *a2ijk, int *gmsh2ijk, int *vtk2ijk, int *a2ij,
                                   int *gmsh2ij, int *vtk2ij)
{
    Tables t;
    int rc = make_tables(porder, t);
    if (rc) return rc;
    const int p = porder, n1 = p + 1;
    // hexahedra: positions visited k, i, j in sod2d's 1-D order
    // (set_allocate_hexahedronHO_ijk_indices, mod_ijk_indices.f90:61-114)
    int c = 0;
    for (int k = 0; k < n1; ++k)
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n1; ++j, ++c) {
                if (a2ijk) a2ijk[c] = t.a2ijk[c];
                if (gmsh2ijk) gmsh2ijk[c] = t.a2ijk[c];
                if (vtk2ijk) vtk2ijk[c] = vtk_hex_no(p, s2g(p, i), s2g(p, j), s2g(p, k));
            }
    // quadrilaterals: positions visited i, j (set_allocate_quadrilateralHO_ij_indices,
    // :116-154); a2ij = gmsh2ij as the partitioner sets it (mod_meshConversorTool.f90:161-162)
    std::vector<std::array<int, 2>> q;
    gmsh_quad_order(p, q);
    std::vector<int> qlocal(n1 * n1, -1);
    for (size_t a = 0; a < q.size(); ++a) qlocal[q[a][0] + n1 * q[a][1]] = (int)a;
    c = 0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j, ++c) {
            const int g = qlocal[s2g(p, i) + n1 * s2g(p, j)] + 1;
            if (a2ij) a2ij[c] = g;
            if (gmsh2ij) gmsh2ij[c] = g;
            if (vtk2ij) vtk2ij[c] = vtk_quad_no(p, s2g(p, i), s2g(p, j));
        }
    return SODHIP_OK;
}

extern "C" int sodhip_box_mesh_sizes(int porder, int nx, int ny, int nz, int periodic, int *nelem,
                                     int *npoin, int *npoin_w)
{
    SODHIP_REQUIRE(porder >= 1 && porder <= 8, "porder %d outside 1..8", porder);
    SODHIP_REQUIRE(nx > 0 && ny > 0 && nz > 0, "box dimensions must be positive");
    long long ne = (long long)nx * ny * nz;
    long long NX = (long long)nx * porder + 1, NY = (long long)ny * porder + 1,
              NZ = (long long)nz * porder + 1;
    long long np = NX * NY * NZ;
    long long nw = periodic ? (NX - 1) * (NY - 1) * (NZ - 1) : np;
    SODHIP_REQUIRE(ne * nnode_of(porder) < (1LL << 31) && np < (1LL << 31),
                   "mesh too large for 32-bit indices");
    if (nelem) *nelem = (int)ne;
    if (npoin) *npoin = (int)np;
    if (npoin_w) *npoin_w = (int)nw;
    return SODHIP_OK;
}

namespace {
inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
inline double uniform_pm1(uint64_t seed, uint64_t node, int axis)
{
    uint64_t h = splitmix64(seed ^ splitmix64(node * 3 + axis + 1));
    return ((h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}
} // namespace

extern "C" int sodhip_box_mesh(int porder, int nx, int ny, int nz, int periodic,
                               const double *lengths, double jitter, unsigned long long seed,
                               int *connec_orig, int *connec_work, int *lpoin_w, double *coord)
{
    int nelem, npoin, npoin_w;
    int rc = sodhip_box_mesh_sizes(porder, nx, ny, nz, periodic, &nelem, &npoin, &npoin_w);
    if (rc) return rc;
    SODHIP_REQUIRE(!(periodic && jitter > 0.0), "jitter is only supported on non-periodic boxes");
    Tables t;
    rc = make_tables(porder, t);
    if (rc) return rc;
    const int p = porder, nn = t.nn;
    const long NX = (long)nx * p + 1, NY = (long)ny * p + 1, NZ = (long)nz * p + 1;
    const int ne[3] = {nx, ny, nz};
    const long NN3[3] = {NX, NY, NZ};
    const double L[3] = {lengths ? lengths[0] : 1.0, lengths ? lengths[1] : 1.0,
                         lengths ? lengths[2] : 1.0};
    // 1-D node positions per axis
    std::vector<double> pos[3];
    for (int d = 0; d < 3; ++d) {
        pos[d].resize(NN3[d]);
        double h = L[d] / ne[d];
        for (long I = 0; I < NN3[d]; ++I) {
            long ex = I / p, li = I % p;
            if (I == NN3[d] - 1) {
                ex = ne[d];
                li = 0;
            }
            pos[d][I] = ex * h + 0.5 * (t.xi[li] + 1.0) * h;
        }
        pos[d][NN3[d] - 1] = L[d];
    }
    auto nid = [&](long I, long J, long K) { return I + NX * (J + NY * K); };
    if (coord) {
        parallel_for(NZ, [&](long kb, long ke) {
            for (long K = kb; K < ke; ++K)
                for (long J = 0; J < NY; ++J)
                    for (long I = 0; I < NX; ++I) {
                        long n = nid(I, J, K);
                        long idx[3] = {I, J, K};
                        for (int d = 0; d < 3; ++d) {
                            double x = pos[d][idx[d]];
                            long ii = idx[d];
                            if (jitter > 0.0 && ii > 0 && ii < NN3[d] - 1) {
                                double hmin = std::min(pos[d][ii] - pos[d][ii - 1],
                                                       pos[d][ii + 1] - pos[d][ii]);
                                x += jitter * hmin * uniform_pm1(seed, (uint64_t)n, d);
                            }
                            coord[n + (size_t)npoin * d] = x;
                        }
                    }
        });
    }
    if (connec_orig || connec_work) {
        parallel_for(nelem, [&](long eb, long ee) {
            for (long e = eb; e < ee; ++e) {
                long ex = e % nx, ey = (e / nx) % ny, ez = e / ((long)nx * ny);
                for (int a = 0; a < nn; ++a) {
                    long I = ex * p + t.gcoord[a][0], J = ey * p + t.gcoord[a][1],
                         K = ez * p + t.gcoord[a][2];
                    size_t off = (size_t)a * nelem + e;
                    if (connec_orig) connec_orig[off] = (int)(nid(I, J, K) + 1);
                    if (connec_work) {
                        long Iw = I, Jw = J, Kw = K;
                        if (periodic) {
                            if (Iw == NX - 1) Iw = 0;
                            if (Jw == NY - 1) Jw = 0;
                            if (Kw == NZ - 1) Kw = 0;
                        }
                        connec_work[off] = (int)(nid(Iw, Jw, Kw) + 1);
                    }
                }
            }
        });
    }
    if (lpoin_w) {
        long c = 0;
        for (long K = 0; K < NZ; ++K)
            for (long J = 0; J < NY; ++J)
                for (long I = 0; I < NX; ++I) {
                    bool slave = periodic && (I == NX - 1 || J == NY - 1 || K == NZ - 1);
                    if (!slave) lpoin_w[c++] = (int)(nid(I, J, K) + 1);
                }
    }
    return SODHIP_OK;
}

extern "C" int sodhip_elem_jacobian(int porder, int nelem, int npoin, const int *connec,
                                    const double *coord, double *He, double *gpvol)
{
    Tables t;
    int rc = make_tables(porder, t);
    if (rc) return rc;
    SODHIP_REQUIRE(connec && coord && He && gpvol, "null argument");
    const int n1 = t.n1, nn = t.nn;
    std::atomic<long> bad(-1);
    parallel_for(nelem, [&](long eb, long ee) {
        std::vector<double> x(3 * nn);
        for (long e = eb; e < ee; ++e) {
            for (int a = 0; a < nn; ++a) {
                long n = connec[(size_t)a * nelem + e] - 1;
                for (int d = 0; d < 3; ++d) x[a + nn * d] = coord[n + (size_t)npoin * d];
            }
            for (int g = 0; g < nn; ++g) {
                const auto &gc = t.gcoord[g];
                double J[3][3]; // J[i][j] = d x_j / d xi_i
                for (int i = 0; i < 3; ++i) {
                    for (int j = 0; j < 3; ++j) J[i][j] = 0.0;
                    for (int m = 0; m < n1; ++m) {
                        int c[3] = {gc[0], gc[1], gc[2]};
                        c[i] = m;
                        int b = t.local_of[c[0] + n1 * (c[1] + n1 * c[2])];
                        double dm = t.D[gc[i] * n1 + m];
                        for (int j = 0; j < 3; ++j) J[i][j] += dm * x[b + nn * j];
                    }
                }
                double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
                double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
                double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
                double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
                if (!(det > 0.0)) bad = e;
                double r = 1.0 / det;
                double inv[3][3];
                inv[0][0] = c00 * r;
                inv[1][0] = c01 * r;
                inv[2][0] = c02 * r;
                inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
                inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
                inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
                inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
                inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
                inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
                // He(i,j) = (J^-1)(i,j) with J(i,j) = dx_j/dxi_i, i.e. He(i,j) = dxi_j/dx_i
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        He[i + 3 * (j + 3 * ((size_t)g + (size_t)nn * e))] = inv[i][j];
                double wg = t.w[gc[0]] * t.w[gc[1]] * t.w[gc[2]];
                gpvol[g + (size_t)nn * e] = wg * det;
            }
        }
    });
    if (bad >= 0)
        return set_error(SODHIP_ERR_ARG, "non-positive Jacobian determinant in element %ld",
                         (long)bad + 1);
    return SODHIP_OK;
}

extern "C" int sodhip_lumped_mass(int porder, int nelem, int npoin, const int *connec_work,
                                  const double *gpvol, int npoin_w, const int *lpoin_w,
                                  double *Ml, double *invMl)
{
    SODHIP_REQUIRE(porder >= 1 && porder <= 8, "porder %d outside 1..8", porder);
    const int nn = nnode_of(porder);
    if (Ml) {
        for (int n = 0; n < npoin; ++n) Ml[n] = 0.0;
        for (long e = 0; e < nelem; ++e)
            for (int a = 0; a < nn; ++a)
                Ml[connec_work[(size_t)a * nelem + e] - 1] += gpvol[a + (size_t)nn * e];
    }
    if (invMl) {
        SODHIP_REQUIRE(Ml && lpoin_w, "invMl needs Ml and lpoin_w");
        for (int n = 0; n < npoin; ++n) invMl[n] = 0.0;
        for (int i = 0; i < npoin_w; ++i) invMl[lpoin_w[i] - 1] = 1.0 / Ml[lpoin_w[i] - 1];
    }
    return SODHIP_OK;
}

extern "C" int sodhip_char_length(int porder, int nelem, int npoin, const int *connec_orig,
                                  const double *coord, double *helem, double *helem_l)
{
    SODHIP_REQUIRE(porder >= 1 && porder <= 8, "porder %d outside 1..8", porder);
    const int nn = nnode_of(porder);
    // corner-edge pairs of the gmsh hexahedron (its 8 vertices are local nodes 0..7)
    static const int ed[12][2] = {{1, 0}, {2, 1}, {3, 2}, {0, 3}, {5, 4}, {6, 5},
                                  {7, 6}, {4, 7}, {4, 0}, {5, 1}, {6, 2}, {7, 3}};
    parallel_for(nelem, [&](long eb, long ee) {
        std::vector<double> x(3 * nn);
        for (long e = eb; e < ee; ++e) {
            for (int a = 0; a < nn; ++a) {
                long n = connec_orig[(size_t)a * nelem + e] - 1;
                for (int d = 0; d < 3; ++d) x[3 * a + d] = coord[n + (size_t)npoin * d];
            }
            if (helem) {
                double h = 1.0e12;
                for (auto &p : ed) {
                    double dx = x[3 * p[0]] - x[3 * p[1]], dy = x[3 * p[0] + 1] - x[3 * p[1] + 1],
                           dz = x[3 * p[0] + 2] - x[3 * p[1] + 2];
                    h = std::min(h, std::sqrt(dx * dx + dy * dy + dz * dz));
                }
                helem[e] = h;
            }
            if (helem_l) {
                for (int a = 0; a < nn; ++a) {
                    double h = 1.0e12;
                    for (int b = 0; b < nn; ++b) {
                        if (a == b) continue;
                        double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1],
                               dz = x[3 * a + 2] - x[3 * b + 2];
                        h = std::min(h, std::sqrt(dx * dx + dy * dy + dz * dz));
                    }
                    helem_l[e + (size_t)nelem * a] = h;
                }
            }
        }
    });
    return SODHIP_OK;
}

extern "C" int sodhip_elem_per_node(int nelem, int nnode, int npoin, const int *connec,
                                    int *lelpn)
{
    SODHIP_REQUIRE(nelem > 0 && nnode > 0 && npoin > 0, "invalid sizes");
    SODHIP_REQUIRE(connec && lelpn, "null array argument");
    for (int n = 0; n < npoin; ++n) lelpn[n] = 0;
    for (long a = 0; a < nnode; ++a)
        for (long e = 0; e < nelem; ++e) {
            const int v = connec[a * nelem + e];
            SODHIP_REQUIRE(v >= 1 && v <= npoin, "connec entry %d outside 1..npoin=%d", v, npoin);
            lelpn[v - 1] += 1;
        }
    return SODHIP_OK;
}
