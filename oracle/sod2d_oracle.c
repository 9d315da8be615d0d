/*
 * sod2d_oracle.c -- CPU FP64 restatement of SOD2D's compressible SEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see sod2d_oracle.h).  Every function follows the
 * reference loop nest it cites, in the same summation and evaluation order, so
 * that its results equal a serial FP64 build of the reference up to the
 * compiler's freedom (built here with -ffp-contract=off, no -ffast-math).
 * References are relative to the LIUJUNLE97/sod2d source tree.
 */
#include "sod2d_oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NDIME 3

/* column-major accessors on the reference's dummy-argument layouts */
#define CONNEC(e, a) (connec[(size_t)(a) * nelem + (e)] - 1)
#define HE(i, j, g, e) He[(i) + 3 * ((j) + 3 * ((size_t)(g) + (size_t)ngaus * (e)))]
#define GPVOL(g, e) gpvol[(g) + (size_t)ngaus * (e)]
#define DL(g, d, ii) dlxigp_ip[(g) + ngaus * ((d) + 3 * (ii))]
#define INV(i, j, k) (invAtoIJK[(i) + n1 * ((j) + n1 * (k))] - 1)
#define NODAL(v, n, d) v[(n) + (size_t)npoin * (d)]
#define ELGP(v, e, g) v[(e) + (size_t)nelem * (g)]

/* ------------------------------------------------------------------------ */
/* Basis: mod_maths.f90                                                      */
/* ------------------------------------------------------------------------ */

/* VALEPO: Legendre polynomial of degree n and its first two derivatives
 * (mod_maths.f90:112-147). */
void or_valepo(int n, double x, double *y_, double *dy_, double *d2y_)
{
    double y = 1.0, dy = 0.0, d2y = 0.0;
    if (n == 0) { *y_ = y; *dy_ = dy; *d2y_ = d2y; return; }
    y = x; dy = 1.0; d2y = 0.0;
    if (n == 1) { *y_ = y; *dy_ = dy; *d2y_ = d2y; return; }
    double yp = 1.0, dyp = 0.0, d2yp = 0.0;
    for (int i = 2; i <= n; ++i) {
        double c1 = (double)i;
        double c2 = 2.0 * c1 - 1.0;
        double c4 = c1 - 1.0;
        double ym = y;
        y = (c2 * x * y - c4 * yp) / c1;
        yp = ym;
        double dym = dy;
        dy = (c2 * x * dy - c4 * dyp + c2 * yp) / c1;
        dyp = dym;
        double d2ym = d2y;
        d2y = (c2 * x * d2y - c4 * d2yp + 2.0 * c2 * dyp) / c1;
        d2yp = d2ym;
    }
    *y_ = y; *dy_ = dy; *d2y_ = d2y;
}

/* ZELEGL: Gauss-Lobatto-Legendre nodes by Newton iteration on P_n'
 * (mod_maths.f90:41-83).  et[0..n], vn[0..n]. */
void or_zelegl(int n, double *et, double *vn)
{
    if (n == 0) return;
    int n2 = (n - 1) / 2;
    double sn = (double)(2 * n - 4 * n2 - 3);
    et[0] = -1.0;
    et[n] = 1.0;
    vn[0] = sn;
    vn[n] = 1.0;
    if (n == 1) return;
    et[n2 + 1] = 0.0;
    double y, dy, d2y;
    or_valepo(n, 0.0, &y, &dy, &d2y);
    vn[n2 + 1] = y;
    if (n == 2) return;
    const double v_pi = 2.0 * asin(1.0); /* mod_constants.f90: v_pi = 2*asin(1) */
    double c = v_pi / (double)n;
    for (int i = 1; i <= n2; ++i) {
        double etx = cos(c * (double)i);
        for (int it = 1; it <= 8; ++it) {
            or_valepo(n, etx, &y, &dy, &d2y);
            etx = etx - dy / d2y;
        }
        et[i] = -etx;
        et[n - i] = etx;
        vn[i] = y * sn;
        vn[n - i] = y;
    }
}

/* WELEGL: GLL weights from P_n at the nodes (mod_maths.f90:85-110). */
void or_welegl(int n, const double *et, const double *vn, double *wt)
{
    (void)et;
    if (n == 0) return;
    int n2 = (n - 1) / 2;
    double dn = (double)n;
    double c = 2.0 / (dn * (dn + 1.0));
    for (int i = 0; i <= n2; ++i) {
        double y = vn[i];
        double wtx = c / (y * y);
        wt[i] = wtx;
        wt[n - i] = wtx;
    }
    if (n - 1 == 2 * n2) return;
    double y = vn[n2 + 1];
    wt[n2 + 1] = c / (y * y);
}

/* getGaussLobattoLegendre_weights_and_roots (mod_maths.f90:28-37) */
void or_gll(int p, double *xi, double *w)
{
    double vn[32];
    or_zelegl(p, xi, vn);
    or_welegl(p, xi, vn, w);
}

/* the sod2d 1-D node order: lorder = [1, p+1, 2, ..., p] (1-based) */
static void lorder0(int p, int *lo)
{
    lo[0] = 0;
    lo[1] = p;
    for (int i = 2; i <= p; ++i) lo[i] = i - 1;
}

/* eval_lagrangePoly (mod_maths.f90:243-281) */
void or_lagrange(int p, const double *xi, double xp, double *l)
{
    int lo[32];
    lorder0(p, lo);
    for (int i = 0; i <= p; ++i) {
        double v = 1.0;
        for (int j = 0; j <= p; ++j)
            if (j != lo[i]) v = v * ((xp - xi[j]) / (xi[lo[i]] - xi[j]));
        l[i] = v;
    }
}

/* eval_lagrangePolyDeriv (mod_maths.f90:283-323) */
void or_lagrange_deriv(int p, const double *xi, double xp, double *dl)
{
    int lo[32];
    lorder0(p, lo);
    for (int i = 0; i <= p; ++i) {
        double s = 0.0;
        for (int j = 0; j <= p; ++j) {
            double aux = 1.0;
            if (j != lo[i]) {
                for (int m = 0; m <= p; ++m)
                    if (m != j && m != lo[i]) aux = aux * (xp - xi[m]) / (xi[lo[i]] - xi[m]);
                s = s + (1.0 / (xi[lo[i]] - xi[j])) * aux;
            }
        }
        dl[i] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Index maps: mod_ijk_indices.f90, elem_hex.f90, elem_qua.f90, mod_geom.f90 */
/* ------------------------------------------------------------------------ */

/* elem_hex.f90:17-18 and elem_qua.f90:14 (1-based corner ids) */
static const int hex_edges[12][2] = {{1, 2}, {1, 4}, {1, 5}, {2, 3}, {2, 6}, {3, 4},
                                     {3, 7}, {4, 8}, {5, 6}, {5, 8}, {6, 7}, {7, 8}};
static const int hex_faces[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {1, 5, 8, 4},
                                    {2, 3, 7, 6}, {3, 4, 8, 7}, {5, 6, 7, 8}};
static const int quad_edges[4][2] = {{1, 2}, {2, 3}, {3, 4}, {4, 1}};

/* genHighOrderQuadGmsh (mod_ijk_indices.f90:629-653); t is (p+1)^2 x 2, column-major */
static void gmsh_quad(int p, int *t)
{
    int nr = (p + 1) * (p + 1);
#define TQ(r, c) t[(r) + nr * (c)]
    TQ(0, 0) = 0; TQ(0, 1) = 0;
    if (p > 0) {
        TQ(1, 0) = p; TQ(1, 1) = 0;
        TQ(2, 0) = p; TQ(2, 1) = p;
        TQ(3, 0) = 0; TQ(3, 1) = p;
        if (p > 1) {
            int inode = 4;
            for (int ie = 0; ie < 4; ++ie) {
                int i0 = quad_edges[ie][0] - 1, i1 = quad_edges[ie][1] - 1;
                int u0 = (TQ(i1, 0) - TQ(i0, 0)) / p, u1 = (TQ(i1, 1) - TQ(i0, 1)) / p;
                for (int i = 1; i <= p - 1; ++i) {
                    TQ(inode, 0) = TQ(i0, 0) + i * u0;
                    TQ(inode, 1) = TQ(i0, 1) + i * u1;
                    ++inode;
                }
            }
            int ns = (p - 1) * (p - 1);
            int *sub = (int *)malloc(sizeof(int) * 2 * ns);
            gmsh_quad(p - 2, sub);
            for (int i = 0; i < ns; ++i) {
                TQ(inode, 0) = sub[i] + 1;
                TQ(inode, 1) = sub[i + ns] + 1;
                ++inode;
            }
            free(sub);
        }
    }
#undef TQ
}

/* genHighOrderHexGmsh (mod_ijk_indices.f90:574-627); t is (p+1)^3 x 3, column-major */
void or_gmsh_hex_table(int p, int *t)
{
    int nr = (p + 1) * (p + 1) * (p + 1);
#define TH(r, c) t[(r) + nr * (c)]
    TH(0, 0) = 0; TH(0, 1) = 0; TH(0, 2) = 0;
    if (p > 0) {
        static const int corners[7][3] = {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1},
                                          {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (int c = 0; c < 7; ++c)
            for (int d = 0; d < 3; ++d) TH(c + 1, d) = corners[c][d] * p;
        if (p > 1) {
            int inode = 8;
            for (int ie = 0; ie < 12; ++ie) {
                int i0 = hex_edges[ie][0] - 1, i1 = hex_edges[ie][1] - 1;
                int u[3];
                for (int d = 0; d < 3; ++d) u[d] = (TH(i1, d) - TH(i0, d)) / p;
                for (int i = 1; i <= p - 1; ++i) {
                    for (int d = 0; d < 3; ++d) TH(inode, d) = TH(i0, d) + i * u[d];
                    ++inode;
                }
            }
            int nf = (p - 1) * (p - 1);
            int *face = (int *)malloc(sizeof(int) * 2 * nf);
            gmsh_quad(p - 2, face);
            for (int iface = 0; iface < 6; ++iface) {
                int i0 = hex_faces[iface][0] - 1, i1 = hex_faces[iface][1] - 1,
                    i3 = hex_faces[iface][3] - 1;
                int u[3], v[3];
                for (int d = 0; d < 3; ++d) {
                    u[d] = (TH(i1, d) - TH(i0, d)) / p;
                    v[d] = (TH(i3, d) - TH(i0, d)) / p;
                }
                for (int i = 0; i < nf; ++i) {
                    for (int d = 0; d < 3; ++d)
                        TH(inode, d) = TH(i0, d) + u[d] * (face[i] + 1) + v[d] * (face[i + nf] + 1);
                    ++inode;
                }
            }
            free(face);
            int nv = (p - 1) * (p - 1) * (p - 1);
            int *vol = (int *)malloc(sizeof(int) * 3 * nv);
            or_gmsh_hex_table(p - 2, vol);
            for (int i = 0; i < nv; ++i) {
                for (int d = 0; d < 3; ++d) TH(inode, d) = vol[i + nv * d] + 1;
                ++inode;
            }
            free(vol);
        }
    }
#undef TH
}

/* set_allocate_hexahedronHO_ijk_indices -> gmsh2ijk (mod_ijk_indices.f90:27-51,61-114);
 * the converter sets a2ijk = gmsh2ijk (mod_meshConversorTool.f90:160-161). */
void or_a2ijk(int p, int *a2ijk)
{
    int n1 = p + 1, nn = n1 * n1 * n1;
    int *t = (int *)malloc(sizeof(int) * 3 * nn);
    or_gmsh_hex_table(p, t);
    int *lut = (int *)malloc(sizeof(int) * nn); /* gmshHexahedraHO_ijkTable(i,j,k) */
    for (int r = 0; r < nn; ++r) lut[t[r] + n1 * (t[r + nn] + n1 * t[r + 2 * nn])] = r + 1;
    int s2g[32];
    s2g[0] = 0;
    s2g[1] = p;
    for (int j = 1; j <= p - 1; ++j) s2g[j + 1] = j;
    int pidx = 0;
    for (int kp = 0; kp < n1; ++kp)
        for (int ip = 0; ip < n1; ++ip)
            for (int jp = 0; jp < n1; ++jp) {
                int i = s2g[ip], j = s2g[jp], k = s2g[kp];
                a2ijk[pidx++] = lut[i + n1 * (j + n1 * k)];
            }
    free(lut);
    free(t);
}

/* atoIJKInverse (mod_geom.f90:127-148) */
void or_atoijk_inverse(int p, const int *a2ijk, int *inv, int *gI, int *gJ, int *gK)
{
    int n1 = p + 1, c = 0;
    for (int k = 1; k <= n1; ++k)
        for (int i = 1; i <= n1; ++i)
            for (int j = 1; j <= n1; ++j) {
                int a = a2ijk[c++];
                inv[(i - 1) + n1 * ((j - 1) + n1 * (k - 1))] = a;
                gI[a - 1] = i;
                gJ[a - 1] = j;
                gK[a - 1] = k;
            }
}

/* GaussLobattoLegendre_hex (quadrature_rules.f90:51-78) followed by hex_highorder /
 * TripleTensorProduct on the GLL grid (elem_hex.f90:22-43; mod_maths.f90:325-363),
 * as driven by CFDSolverBase_evalShapeFunctions (CFDSolverBase.f90:1524-1572). */
void or_shape(int p, const int *a2ijk, double *xgp, double *wgp, double *Ngp, double *dNgp,
              double *dlxigp_ip)
{
    int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    double xi[32], w1d[32];
    int lo[32];
    or_gll(p, xi, w1d);
    lorder0(p, lo);
    int inode = 0;
    for (int k = 0; k < n1; ++k)
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n1; ++j) {
                int a = a2ijk[inode++] - 1;
                xgp[a] = xi[lo[i]];
                xgp[a + ngaus] = xi[lo[j]];
                xgp[a + 2 * ngaus] = xi[lo[k]];
                wgp[a] = w1d[lo[i]] * w1d[lo[j]] * w1d[lo[k]];
            }
    for (int g = 0; g < ngaus; ++g) {
        double s = xgp[g], t = xgp[g + ngaus], z = xgp[g + 2 * ngaus];
        double lx[32], ly[32], lz[32], dlx[32], dly[32], dlz[32];
        or_lagrange(p, xi, s, lx);
        or_lagrange(p, xi, t, ly);
        or_lagrange(p, xi, z, lz);
        or_lagrange_deriv(p, xi, s, dlx);
        or_lagrange_deriv(p, xi, t, dly);
        or_lagrange_deriv(p, xi, z, dlz);
        for (int ii = 0; ii < n1; ++ii) {
            DL(g, 0, ii) = dlx[ii];
            DL(g, 1, ii) = dly[ii];
            DL(g, 2, ii) = dlz[ii];
        }
        int c = 0;
        for (int k = 0; k < n1; ++k)
            for (int i = 0; i < n1; ++i)
                for (int j = 0; j < n1; ++j) {
                    int a = a2ijk[c++] - 1;
                    Ngp[g + (size_t)ngaus * a] = lx[i] * ly[j] * lz[k];
                    dNgp[0 + 3 * (a + (size_t)nnode * g)] = dlx[i] * ly[j] * lz[k];
                    dNgp[1 + 3 * (a + (size_t)nnode * g)] = lx[i] * dly[j] * lz[k];
                    dNgp[2 + 3 * (a + (size_t)nnode * g)] = lx[i] * ly[j] * dlz[k];
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* elem_jacobian (jacobian_oper.f90:7-140), 3-D branch */
void or_elem_jacobian(int p, int nelem, int npoin, const int *connec, const double *coord,
                      const double *dNgp, const double *wgp, double *gpvol, double *He)
{
    int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    for (int e = 0; e < nelem; ++e) {
        for (int g = 0; g < ngaus; ++g) {
            double Je[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double dot = 0.0;
                    for (int a = 0; a < nnode; ++a)
                        dot = dot + (dNgp[i + 3 * (a + (size_t)nnode * g)] *
                                     NODAL(coord, CONNEC(e, a), j));
                    Je[i][j] = 0.0 + dot;
                }
            double det = (Je[0][0] * Je[1][1] * Je[2][2] + Je[0][1] * Je[1][2] * Je[2][0] +
                          Je[0][2] * Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1] * Je[0][2] -
                          Je[2][1] * Je[1][2] * Je[0][0] - Je[2][2] * Je[1][0] * Je[0][1]);
            double a[9], b[9];
            a[0] = Je[1][1] * Je[2][2] - Je[2][1] * Je[1][2];
            a[1] = Je[1][0] * Je[2][2] - Je[2][0] * Je[1][2];
            a[2] = Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1];
            a[3] = Je[0][1] * Je[2][2] - Je[2][1] * Je[0][2];
            a[4] = Je[0][0] * Je[2][2] - Je[2][0] * Je[0][2];
            a[5] = Je[0][0] * Je[2][1] - Je[2][0] * Je[0][1];
            a[6] = Je[0][1] * Je[1][2] - Je[1][1] * Je[0][2];
            a[7] = Je[0][0] * Je[1][2] - Je[1][0] * Je[0][2];
            a[8] = Je[0][0] * Je[1][1] - Je[1][0] * Je[0][1];
            a[1] = -a[1];
            a[3] = -a[3];
            a[5] = -a[5];
            a[7] = -a[7];
            for (int i = 0; i < 9; ++i) b[i] = a[i];
            b[1] = a[3];
            b[2] = a[6];
            b[3] = a[1];
            b[5] = a[7];
            b[6] = a[2];
            b[7] = a[5];
            for (int i = 0; i < 9; ++i) b[i] = (1.0 / det) * b[i];
            HE(0, 0, g, e) = b[0];
            HE(0, 1, g, e) = b[1];
            HE(0, 2, g, e) = b[2];
            HE(1, 0, g, e) = b[3];
            HE(1, 1, g, e) = b[4];
            HE(1, 2, g, e) = b[5];
            HE(2, 0, g, e) = b[6];
            HE(2, 1, g, e) = b[7];
            HE(2, 2, g, e) = b[8];
            GPVOL(g, e) = wgp[g] * det;
        }
    }
}

/* lumped_mass_spectral (mass_matrix.f90:11-42), serial element order */
void or_lumped_mass(int p, int nelem, int npoin, const int *connec, const double *gpvol,
                    double *Ml)
{
    int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    for (int n = 0; n < npoin; ++n) Ml[n] = 0.0;
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nnode; ++a) Ml[CONNEC(e, a)] += GPVOL(a, e);
}

/* invMl over working nodes, 0 elsewhere (CFDSolverBase.f90:1771-1782) */
void or_inv_ml(int npoin, int npoin_w, const int *lpoin_w, const double *Ml, double *invMl)
{
    for (int n = 0; n < npoin; ++n) invMl[n] = 0.0;
    for (int i = 0; i < npoin_w; ++i) invMl[lpoin_w[i] - 1] = 1.0 / Ml[lpoin_w[i] - 1];
}

/* char_length + get_hexa_edge_distances (mod_geom.f90:13-37; elem_hex.f90:73-92) */
void or_char_length(int p, int nelem, int npoin, const int *connec, const double *coord,
                    double *helem)
{
    (void)p;
    static const int ed[12][2] = {{2, 1}, {3, 2}, {4, 3}, {1, 4}, {6, 5}, {7, 6},
                                  {8, 7}, {5, 8}, {5, 1}, {6, 2}, {7, 3}, {8, 4}};
    for (int e = 0; e < nelem; ++e) {
        double dist2 = 1000000000000.0;
        for (int ie = 0; ie < 12; ++ie) {
            int na = CONNEC(e, ed[ie][0] - 1), nb = CONNEC(e, ed[ie][1] - 1);
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = NODAL(coord, na, k) - NODAL(coord, nb, k);
            double aux = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            dist2 = fmin(dist2, aux);
        }
        helem[e] = dist2;
    }
}

/* char_length_spectral (mod_geom.f90:39-66) -> he_l(nelem,nnode) */
void or_char_length_spectral(int p, int nelem, int npoin, const int *connec, const double *coord,
                             double *helem_l)
{
    int n1 = p + 1, nnode = n1 * n1 * n1;
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nnode; ++a) {
            double aux = 1000000000000.0;
            int na = CONNEC(e, a);
            for (int b = 0; b < nnode; ++b) {
                if (a == b) continue;
                int nb = CONNEC(e, b);
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = NODAL(coord, na, k) - NODAL(coord, nb, k);
                double dist2 = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                aux = fmin(dist2, aux);
            }
            helem_l[e + (size_t)nelem * a] = aux;
        }
}

/* ------------------------------------------------------------------------ */
/* Operators                                                                 */
/* ------------------------------------------------------------------------ */

static inline void atomic_add(double *p, double v, int par)
{
    if (par) {
#pragma omp atomic
        *p += v;
    } else {
        *p += v;
    }
}

/* full_convec_ijk (elem_convec.f90:227-470) */
void or_full_convec_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                        const int *gI, const int *gJ, const int *gK, const double *u,
                        const double *q, const double *rho, const double *pr, const double *E,
                        double *Rmass, double *Rmom, double *Rener, int initialize, double fact,
                        int nthreads)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    if (initialize) {
        for (int n = 0; n < npoin; ++n) {
            Rmass[n] = 0.0;
            Rener[n] = 0.0;
            for (int d = 0; d < 3; ++d) NODAL(Rmom, n, d) = 0.0;
        }
    }
    const double aux_fact = fact;
    const int par = nthreads > 1;
    if (nnode > 128) return;
#pragma omp parallel for schedule(static) num_threads(par ? nthreads : 1) if (par)
    for (int e = 0; e < nelem; ++e) {
        int ip[128];
        double ul[128][3], ql[128][3], rhol[128], prl[128], El[128], REl[128], kl[128], Rkl[128];
        double Re_mass[128], Re_ener[128], Re_mom[128][3];
        for (int a = 0; a < nnode; ++a) {
            ip[a] = CONNEC(e, a);
            kl[a] = 0.0;
            for (int d = 0; d < 3; ++d)
                kl[a] = kl[a] + NODAL(u, ip[a], d) * NODAL(u, ip[a], d) * 0.5;
            rhol[a] = rho[ip[a]];
            El[a] = E[ip[a]];
            REl[a] = rho[ip[a]] * E[ip[a]];
            Rkl[a] = rho[ip[a]] * kl[a];
            prl[a] = pr[ip[a]];
            for (int d = 0; d < 3; ++d) {
                ul[a][d] = NODAL(u, ip[a], d);
                ql[a][d] = NODAL(q, ip[a], d);
            }
        }
        for (int g = 0; g < ngaus; ++g) {
            double dlxi[16], dleta[16], dlzeta[16];
            for (int ii = 0; ii < n1; ++ii) {
                dlxi[ii] = DL(g, 0, ii);
                dleta[ii] = DL(g, 1, ii);
                dlzeta[ii] = DL(g, 2, ii);
            }
            int I = gI[g] - 1, J = gJ[g] - 1, K = gK[g] - 1;
            double gRho[3] = {0}, gP[3] = {0}, gE[3] = {0}, gRE[3] = {0}, gk[3] = {0}, gRk[3] = {0};
            double gU[3][3] = {{0}}, gQ[3][3] = {{0}}, gFe[3][3] = {{0}}, gFue[3][3] = {{0}};
            double gFk[3][3] = {{0}}, gFuk[3][3] = {{0}};
            double gF[3][3][3] = {{{0}}}, gFuu[3][3][3] = {{{0}}};
            for (int ii = 0; ii < n1; ++ii) {
                int a1 = INV(ii, J, K), a2 = INV(I, ii, K), a3 = INV(I, J, ii);
                gRho[0] = gRho[0] + dlxi[ii] * rhol[a1];
                gRho[1] = gRho[1] + dleta[ii] * rhol[a2];
                gRho[2] = gRho[2] + dlzeta[ii] * rhol[a3];
                gP[0] = gP[0] + dlxi[ii] * prl[a1];
                gP[1] = gP[1] + dleta[ii] * prl[a2];
                gP[2] = gP[2] + dlzeta[ii] * prl[a3];
                gE[0] = gE[0] + dlxi[ii] * (El[a1]);
                gE[1] = gE[1] + dleta[ii] * (El[a2]);
                gE[2] = gE[2] + dlzeta[ii] * (El[a3]);
                gRE[0] = gRE[0] + dlxi[ii] * (REl[a1]);
                gRE[1] = gRE[1] + dleta[ii] * (REl[a2]);
                gRE[2] = gRE[2] + dlzeta[ii] * (REl[a3]);
                gk[0] = gk[0] + dlxi[ii] * (kl[a1]);
                gk[1] = gk[1] + dleta[ii] * (kl[a2]);
                gk[2] = gk[2] + dlzeta[ii] * (kl[a3]);
                gRk[0] = gRk[0] + dlxi[ii] * (Rkl[a1]);
                gRk[1] = gRk[1] + dleta[ii] * (Rkl[a2]);
                gRk[2] = gRk[2] + dlzeta[ii] * (Rkl[a3]);
                for (int i = 0; i < 3; ++i) {
                    gU[i][0] = gU[i][0] + dlxi[ii] * ul[a1][i];
                    gU[i][1] = gU[i][1] + dleta[ii] * ul[a2][i];
                    gU[i][2] = gU[i][2] + dlzeta[ii] * ul[a3][i];
                    gQ[i][0] = gQ[i][0] + dlxi[ii] * ql[a1][i];
                    gQ[i][1] = gQ[i][1] + dleta[ii] * ql[a2][i];
                    gQ[i][2] = gQ[i][2] + dlzeta[ii] * ql[a3][i];
                    gFe[i][0] = gFe[i][0] + dlxi[ii] * rhol[a1] * El[a1] * ul[a1][i];
                    gFe[i][1] = gFe[i][1] + dleta[ii] * rhol[a2] * El[a2] * ul[a2][i];
                    gFe[i][2] = gFe[i][2] + dlzeta[ii] * rhol[a3] * El[a3] * ul[a3][i];
                    gFue[i][0] = gFue[i][0] + dlxi[ii] * El[a1] * ul[a1][i];
                    gFue[i][1] = gFue[i][1] + dleta[ii] * El[a2] * ul[a2][i];
                    gFue[i][2] = gFue[i][2] + dlzeta[ii] * El[a3] * ul[a3][i];
                    gFk[i][0] = gFk[i][0] + dlxi[ii] * rhol[a1] * kl[a1] * ul[a1][i];
                    gFk[i][1] = gFk[i][1] + dleta[ii] * rhol[a2] * kl[a2] * ul[a2][i];
                    gFk[i][2] = gFk[i][2] + dlzeta[ii] * rhol[a3] * kl[a3] * ul[a3][i];
                    gFuk[i][0] = gFuk[i][0] + dlxi[ii] * kl[a1] * ul[a1][i];
                    gFuk[i][1] = gFuk[i][1] + dleta[ii] * kl[a2] * ul[a2][i];
                    gFuk[i][2] = gFuk[i][2] + dlzeta[ii] * kl[a3] * ul[a3][i];
                    for (int j = 0; j < 3; ++j) {
                        gF[i][j][0] = gF[i][j][0] + dlxi[ii] * ql[a1][i] * ul[a1][j];
                        gF[i][j][1] = gF[i][j][1] + dleta[ii] * ql[a2][i] * ul[a2][j];
                        gF[i][j][2] = gF[i][j][2] + dlzeta[ii] * ql[a3][i] * ul[a3][j];
                        gFuu[i][j][0] = gFuu[i][j][0] + dlxi[ii] * ul[a1][i] * ul[a1][j];
                        gFuu[i][j][1] = gFuu[i][j][1] + dleta[ii] * ul[a2][i] * ul[a2][j];
                        gFuu[i][j][2] = gFuu[i][j][2] + dlzeta[ii] * ul[a3][i] * ul[a3][j];
                    }
                }
            }
            double gradRho[3] = {0}, gradP[3] = {0}, gradE[3] = {0}, gradRE[3] = {0};
            double gradk[3] = {0}, gradRk[3] = {0}, gradU[3][3] = {{0}}, gradQ[3][3] = {{0}};
            double divF[3] = {0}, divFuu[3] = {0};
            double divQ = 0.0, divFe = 0.0, divFue = 0.0, divFk = 0.0, divFuk = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double h = HE(i, j, g, e);
                    gradRho[i] = gradRho[i] + h * gRho[j];
                    gradP[i] = gradP[i] + h * gP[j];
                    gradE[i] = gradE[i] + h * gE[j];
                    gradRE[i] = gradRE[i] + h * gRE[j];
                    divFe = divFe + h * gFe[i][j];
                    divFue = divFue + h * gFue[i][j];
                    gradk[i] = gradk[i] + h * gk[j];
                    gradRk[i] = gradRk[i] + h * gRk[j];
                    divFk = divFk + h * gFk[i][j];
                    divFuk = divFuk + h * gFuk[i][j];
                    for (int k = 0; k < 3; ++k) {
                        double hjk = HE(j, k, g, e);
                        gradU[i][j] = gradU[i][j] + hjk * gU[i][k];
                        gradQ[i][j] = gradQ[i][j] + hjk * gQ[i][k];
                        divF[i] = divF[i] + hjk * gF[i][j][k];
                        divFuu[i] = divFuu[i] + hjk * gFuu[i][j][k];
                    }
                }
            double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            divQ = gradQ[0][0] + gradQ[1][1] + gradQ[2][2];
            Re_mass[g] = 0.5 * (divQ + rhol[g] * divU);
            Re_ener[g] = 0.25 * (divFe + El[g] * divQ + REl[g] * divU + rhol[g] * divFue) +
                         0.25 * (divFk + kl[g] * divQ + Rkl[g] * divU + rhol[g] * divFuk);
            for (int i = 0; i < 3; ++i) {
                Re_mom[g][i] = 0.25 * (divF[i] + ul[g][i] * divQ + ql[g][i] * divU +
                                       rhol[g] * divFuu[i]) +
                               gradP[i];
                Re_mass[g] = Re_mass[g] + 0.5 * (ul[g][i] * gradRho[i]);
                Re_ener[g] = Re_ener[g] +
                             0.25 * (ql[g][i] * gradE[i] + ul[g][i] * gradRE[i] +
                                     El[g] * ul[g][i] * gradRho[i]) +
                             0.25 * (ql[g][i] * gradk[i] + ul[g][i] * gradRk[i] +
                                     kl[g] * ul[g][i] * gradRho[i]);
                for (int j = 0; j < 3; ++j) {
                    double aux = ul[g][i] * ul[g][j];
                    Re_mom[g][i] = Re_mom[g][i] + 0.25 * (aux * gradRho[j] + ql[g][j] * gradU[i][j] +
                                                          ul[g][j] * gradQ[i][j]);
                }
                Re_mom[g][i] = GPVOL(g, e) * Re_mom[g][i];
            }
            Re_mass[g] = GPVOL(g, e) * Re_mass[g];
            Re_ener[g] = GPVOL(g, e) * Re_ener[g];
        }
        for (int a = 0; a < nnode; ++a) {
            atomic_add(&Rmass[ip[a]], aux_fact * Re_mass[a], par);
            atomic_add(&Rener[ip[a]], aux_fact * Re_ener[a], par);
            for (int d = 0; d < 3; ++d) atomic_add(&NODAL(Rmom, ip[a], d), aux_fact * Re_mom[a][d], par);
        }
    }
}

/* full_diffusion_ijk (elem_diffu.f90:12-223) */
void or_full_diffusion_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                           const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                           const int *gI, const int *gJ, const int *gK, double Cp, double Pr,
                           const double *rho_n, const double *rho, const double *u,
                           const double *Tem, const double *mu_fluid, const double *mu_e,
                           const double *mu_sgs, double c_rho, double c_ener, double *Rmass,
                           double *Rmom, double *Rener, int initialize, double fact, int nthreads)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    const double twoThirds = 2.0 / 3.0;
    if (initialize) {
        for (int n = 0; n < npoin; ++n) {
            Rmass[n] = 0.0;
            Rener[n] = 0.0;
            for (int d = 0; d < 3; ++d) NODAL(Rmom, n, d) = 0.0;
        }
    }
    const double aux_fact = fact;
    const int par = nthreads > 1;
    if (nnode > 128) return;
#pragma omp parallel for schedule(static) num_threads(par ? nthreads : 1) if (par)
    for (int e = 0; e < nelem; ++e) {
        int ip[128];
        double ul[128][3], rhol[128], Teml[128], mufluidl[128], rhonl[128];
        double tauXl[128][3], tauYl[128][3], tauZl[128][3], gradTl[128][3], gradRhol[128][3],
            tauUl[128][3];
        for (int a = 0; a < nnode; ++a) ip[a] = CONNEC(e, a);
        for (int a = 0; a < nnode; ++a) {
            rhol[a] = rho[ip[a]];
            rhonl[a] = rho_n[ip[a]];
            Teml[a] = Tem[ip[a]];
            mufluidl[a] = mu_fluid[ip[a]];
        }
        for (int a = 0; a < nnode; ++a)
            for (int d = 0; d < 3; ++d) ul[a][d] = NODAL(u, ip[a], d);
        for (int g = 0; g < ngaus; ++g) {
            double mu_fgp = mufluidl[g] + rhonl[g] * ELGP(mu_sgs, e, g);
            double mu_egp = ELGP(mu_e, e, g);
            int I = gI[g] - 1, J = gJ[g] - 1, K = gK[g] - 1;
            double gRho[3] = {0}, gT[3] = {0}, gU[3][3] = {{0}};
            for (int ii = 0; ii < n1; ++ii) {
                int a1 = INV(ii, J, K), a2 = INV(I, ii, K), a3 = INV(I, J, ii);
                gRho[0] = gRho[0] + DL(g, 0, ii) * rhol[a1];
                gRho[1] = gRho[1] + DL(g, 1, ii) * rhol[a2];
                gRho[2] = gRho[2] + DL(g, 2, ii) * rhol[a3];
                gT[0] = gT[0] + DL(g, 0, ii) * Teml[a1];
                gT[1] = gT[1] + DL(g, 1, ii) * Teml[a2];
                gT[2] = gT[2] + DL(g, 2, ii) * Teml[a3];
                for (int i = 0; i < 3; ++i) {
                    gU[i][0] = gU[i][0] + DL(g, 0, ii) * ul[a1][i];
                    gU[i][1] = gU[i][1] + DL(g, 1, ii) * ul[a2][i];
                    gU[i][2] = gU[i][2] + DL(g, 2, ii) * ul[a3][i];
                }
            }
            double gradRho[3] = {0}, gradT[3] = {0}, gradU[3][3] = {{0}};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    gradRho[i] = gradRho[i] + HE(i, j, g, e) * gRho[j];
                    gradT[i] = gradT[i] + HE(i, j, g, e) * gT[j];
                    for (int k = 0; k < 3; ++k) gradU[i][j] = gradU[i][j] + HE(j, k, g, e) * gU[i][k];
                }
            double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            double tauU[3] = {0}, tau[3][3];
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) {
                    tauU[i] = tauU[i] + (mu_fgp + mu_egp) * (gradU[i][j] + gradU[j][i]) * ul[g][j];
                    tau[i][j] = (mu_fgp + mu_egp) * (gradU[i][j] + gradU[j][i]);
                }
                tauU[i] = tauU[i] - (mu_fgp)*twoThirds * divU * ul[g][i];
                tau[i][i] = tau[i][i] - (mu_fgp)*twoThirds * divU;
            }
            for (int d = 0; d < 3; ++d) {
                tauXl[g][d] = tau[0][d];
                tauYl[g][d] = tau[1][d];
                tauZl[g][d] = tau[2][d];
                gradTl[g][d] = gradT[d];
                gradRhol[g][d] = gradRho[d];
                tauUl[g][d] = tauU[d];
            }
        }
        for (int g = 0; g < ngaus; ++g) {
            double nu_e = c_rho * ELGP(mu_e, e, g) / rhonl[g];
            double kappa_e = mufluidl[g] * Cp / Pr + c_ener * ELGP(mu_e, e, g) * Cp / Pr +
                             Cp * rhonl[g] * ELGP(mu_sgs, e, g) / 0.9;
            int I = gI[g] - 1, J = gJ[g] - 1, K = gK[g] - 1;
            double divDe = 0.0, divDr = 0.0, divDm[3] = {0, 0, 0};
            for (int ii = 0; ii < n1; ++ii) {
                int b1 = INV(ii, J, K), b2 = INV(I, ii, K), b3 = INV(I, J, ii);
                for (int d = 0; d < 3; ++d) {
                    double w1 = HE(d, 0, b1, e) * GPVOL(b1, e) * DL(b1, 0, I);
                    double w2 = HE(d, 1, b2, e) * GPVOL(b2, e) * DL(b2, 1, J);
                    double w3 = HE(d, 2, b3, e) * GPVOL(b3, e) * DL(b3, 2, K);
                    divDr = divDr + w1 * gradRhol[b1][d];
                    divDr = divDr + w2 * gradRhol[b2][d];
                    divDr = divDr + w3 * gradRhol[b3][d];
                    divDe = divDe + w1 * (tauUl[b1][d] + kappa_e * gradTl[b1][d]);
                    divDe = divDe + w2 * (tauUl[b2][d] + kappa_e * gradTl[b2][d]);
                    divDe = divDe + w3 * (tauUl[b3][d] + kappa_e * gradTl[b3][d]);
                    divDm[0] = divDm[0] + w1 * tauXl[b1][d];
                    divDm[0] = divDm[0] + w2 * tauXl[b2][d];
                    divDm[0] = divDm[0] + w3 * tauXl[b3][d];
                    divDm[1] = divDm[1] + w1 * tauYl[b1][d];
                    divDm[1] = divDm[1] + w2 * tauYl[b2][d];
                    divDm[1] = divDm[1] + w3 * tauYl[b3][d];
                    divDm[2] = divDm[2] + w1 * tauZl[b1][d];
                    divDm[2] = divDm[2] + w2 * tauZl[b2][d];
                    divDm[2] = divDm[2] + w3 * tauZl[b3][d];
                }
            }
            atomic_add(&Rmass[ip[g]], aux_fact * nu_e * divDr, par);
            atomic_add(&Rener[ip[g]], aux_fact * divDe, par);
            for (int d = 0; d < 3; ++d) atomic_add(&NODAL(Rmom, ip[g], d), aux_fact * divDm[d], par);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* LPS stabilisation                                                         */
/* ------------------------------------------------------------------------ */

/* Reference-space gradients of rho, T and u at GP g of element e from the
 * gathered nodal values (the first loops of elem_stab.f90:300-335 and
 * :497-531), then physical ones through He. */
static void lps_gradients(int p, int nelem, int e, int g, const double *He,
                          const double *dlxigp_ip, const int *invAtoIJK, const int *gI,
                          const int *gJ, const int *gK, const double *rhol, const double *Teml,
                          const double (*ul)[3], double gradRho[3], double gradT[3],
                          double gradU[3][3])
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1;
    int I = gI[g] - 1, J = gJ[g] - 1, K = gK[g] - 1;
    double gRho[3] = {0}, gT[3] = {0}, gU[3][3] = {{0}};
    for (int ii = 0; ii < n1; ++ii) {
        int a1 = INV(ii, J, K), a2 = INV(I, ii, K), a3 = INV(I, J, ii);
        gRho[0] = gRho[0] + DL(g, 0, ii) * rhol[a1];
        gRho[1] = gRho[1] + DL(g, 1, ii) * rhol[a2];
        gRho[2] = gRho[2] + DL(g, 2, ii) * rhol[a3];
        gT[0] = gT[0] + DL(g, 0, ii) * Teml[a1];
        gT[1] = gT[1] + DL(g, 1, ii) * Teml[a2];
        gT[2] = gT[2] + DL(g, 2, ii) * Teml[a3];
        for (int i = 0; i < 3; ++i) {
            gU[i][0] = gU[i][0] + DL(g, 0, ii) * ul[a1][i];
            gU[i][1] = gU[i][1] + DL(g, 1, ii) * ul[a2][i];
            gU[i][2] = gU[i][2] + DL(g, 2, ii) * ul[a3][i];
        }
    }
    for (int i = 0; i < 3; ++i) {
        gradRho[i] = 0.0;
        gradT[i] = 0.0;
        for (int j = 0; j < 3; ++j) gradU[i][j] = 0.0;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            gradRho[i] = gradRho[i] + HE(i, j, g, e) * gRho[j];
            gradT[i] = gradT[i] + HE(i, j, g, e) * gT[j];
            for (int k = 0; k < 3; ++k) gradU[i][j] = gradU[i][j] + HE(j, k, g, e) * gU[i][k];
        }
}

/* comp_tau (elem_stab.f90:589-616): tau(e) = max over the element's nodes of
 * helem_k(e) * (c_species_stab / p) * (f * c + |u|), f = 0 when the sound
 * speed is dropped (flag_drop_c_in_envit), else 1. */
void or_comp_tau(int p, int nelem, int npoin, const int *connec, const double *csound,
                 const double *u, const double *helem_k, double c_species_stab, int drop_c,
                 double *tau)
{
    const int nnode = (p + 1) * (p + 1) * (p + 1);
    const double fact_low_mach = drop_c ? 0.0 : 1.0;
    for (int e = 0; e < nelem; ++e) {
        double taul = 0.0;
        for (int a = 0; a < nnode; ++a) {
            int n = CONNEC(e, a);
            double uu = 0.0;
            for (int d = 0; d < 3; ++d) uu = uu + NODAL(u, n, d) * NODAL(u, n, d);
            double aux1 = fact_low_mach * csound[n] + sqrt(uu);
            taul = fmax(taul, helem_k[e] * (c_species_stab / (double)p) * aux1);
        }
        tau[e] = taul;
    }
}

/* full_proj_ijk (elem_stab.f90:441-586): lumped L2 projections of grad rho
 * (ProjMass), grad T (ProjEner) and the rows of the traceless strain
 * 2 S - 2/3 div u I (ProjMX/MY/MZ), each (npoin,3). */
void or_full_proj_ijk(int p, int nelem, int npoin, int npoin_w, const int *connec,
                      const int *lpoin_w, const double *He, const double *gpvol,
                      const double *dlxigp_ip, const int *invAtoIJK, const int *gI, const int *gJ,
                      const int *gK, const double *rho, const double *u, const double *Tem,
                      const double *invMl, double *ProjMass, double *ProjEner, double *ProjMX,
                      double *ProjMY, double *ProjMZ, int nthreads)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    const double twoThirds = 2.0 / 3.0;
    for (size_t n = 0; n < (size_t)npoin * 3; ++n) {
        ProjMass[n] = 0.0;
        ProjEner[n] = 0.0;
        ProjMX[n] = 0.0;
        ProjMY[n] = 0.0;
        ProjMZ[n] = 0.0;
    }
    if (nnode > 128) return;
    const int par = nthreads > 1;
#pragma omp parallel for schedule(static) num_threads(par ? nthreads : 1) if (par)
    for (int e = 0; e < nelem; ++e) {
        int ip[128];
        double ul[128][3], rhol[128], Teml[128];
        for (int a = 0; a < nnode; ++a) {
            ip[a] = CONNEC(e, a);
            rhol[a] = rho[ip[a]];
            Teml[a] = Tem[ip[a]];
            for (int d = 0; d < 3; ++d) ul[a][d] = NODAL(u, ip[a], d);
        }
        for (int g = 0; g < ngaus; ++g) {
            double gradRho[3], gradT[3], gradU[3][3], tau[3][3];
            lps_gradients(p, nelem, e, g, He, dlxigp_ip, invAtoIJK, gI, gJ, gK, rhol, Teml,
                          (const double(*)[3])ul, gradRho, gradT, gradU);
            double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) tau[i][j] = (gradU[i][j] + gradU[j][i]);
                tau[i][i] = tau[i][i] - twoThirds * divU;
            }
            double w = GPVOL(g, e);
            for (int d = 0; d < 3; ++d) {
                atomic_add(&NODAL(ProjMX, ip[g], d), w * tau[0][d], par);
                atomic_add(&NODAL(ProjMY, ip[g], d), w * tau[1][d], par);
                atomic_add(&NODAL(ProjMZ, ip[g], d), w * tau[2][d], par);
                atomic_add(&NODAL(ProjMass, ip[g], d), w * gradRho[d], par);
                atomic_add(&NODAL(ProjEner, ip[g], d), w * gradT[d], par);
            }
        }
    }
    or_lumped_solver_vect(npoin, npoin_w, lpoin_w, invMl, ProjMX);
    or_lumped_solver_vect(npoin, npoin_w, lpoin_w, invMl, ProjMY);
    or_lumped_solver_vect(npoin, npoin_w, lpoin_w, invMl, ProjMZ);
    or_lumped_solver_vect(npoin, npoin_w, lpoin_w, invMl, ProjMass);
    or_lumped_solver_vect(npoin, npoin_w, lpoin_w, invMl, ProjEner);
}

/* full_diff_stab_ijk (elem_stab.f90:207-438): the viscous fluxes of
 * full_diffusion_ijk with the entropy-viscosity coefficients applied at each
 * GP (nu_e grad rho, kappa_e grad T), plus the LPS fluxes
 * -c_lps tau_e (Proj - grad) scaled by rho_n (momentum), rho_n Cp/Pr
 * (energy) and 1 (mass); c_lps is c_lps_comp, a default-real constant. */
void or_full_diff_stab_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                           const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                           const int *gI, const int *gJ, const int *gK, double Cp, double Pr,
                           const double *rho_n, const double *rho, const double *u,
                           const double *Tem, const double *mu_fluid, const double *mu_e,
                           const double *mu_sgs, const double *ProjMass, const double *ProjEner,
                           const double *ProjMX, const double *ProjMY, const double *ProjMZ,
                           const double *tau_stab, double c_rho, double c_ener, double c_lps,
                           double *Rmass, double *Rmom, double *Rener, int initialize,
                           double fact, int nthreads)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    const double twoThirds = 2.0 / 3.0;
    if (initialize) {
        for (int n = 0; n < npoin; ++n) {
            Rmass[n] = 0.0;
            Rener[n] = 0.0;
            for (int d = 0; d < 3; ++d) NODAL(Rmom, n, d) = 0.0;
        }
    }
    const double aux_fact = fact;
    if (nnode > 128) return;
    const int par = nthreads > 1;
#pragma omp parallel for schedule(static) num_threads(par ? nthreads : 1) if (par)
    for (int e = 0; e < nelem; ++e) {
        int ip[128];
        double ul[128][3], rhol[128], Teml[128], mufluidl[128], rhonl[128];
        double tauXl[128][3], tauYl[128][3], tauZl[128][3], gradTl[128][3], gradRhol[128][3];
        for (int a = 0; a < nnode; ++a) {
            ip[a] = CONNEC(e, a);
            rhol[a] = rho[ip[a]];
            rhonl[a] = rho_n[ip[a]];
            Teml[a] = Tem[ip[a]];
            mufluidl[a] = mu_fluid[ip[a]];
            for (int d = 0; d < 3; ++d) ul[a][d] = NODAL(u, ip[a], d);
        }
        const double taustabl = tau_stab[e];
        for (int g = 0; g < ngaus; ++g) {
            double nu_e = c_rho * ELGP(mu_e, e, g) / rhonl[g];
            double mu_fgp = mufluidl[g] + rhonl[g] * ELGP(mu_sgs, e, g);
            double mu_egp = ELGP(mu_e, e, g);
            double kappa_e = mufluidl[g] * Cp / Pr + c_ener * ELGP(mu_e, e, g) * Cp / Pr +
                             Cp * rhonl[g] * ELGP(mu_sgs, e, g) / 0.9;
            double gradRho[3], gradT[3], gradU[3][3], tau[3][3];
            lps_gradients(p, nelem, e, g, He, dlxigp_ip, invAtoIJK, gI, gJ, gK, rhol, Teml,
                          (const double(*)[3])ul, gradRho, gradT, gradU);
            double divU = gradU[0][0] + gradU[1][1] + gradU[2][2];
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) tau[i][j] = (gradU[i][j] + gradU[j][i]);
                tau[i][i] = tau[i][i] - twoThirds * divU;
            }
            const int n = ip[g];
            for (int d = 0; d < 3; ++d) {
                tauXl[g][d] = -c_lps * taustabl * (NODAL(ProjMX, n, d) - tau[0][d]) * rhonl[g];
                tauYl[g][d] = -c_lps * taustabl * (NODAL(ProjMY, n, d) - tau[1][d]) * rhonl[g];
                tauZl[g][d] = -c_lps * taustabl * (NODAL(ProjMZ, n, d) - tau[2][d]) * rhonl[g];
                gradTl[g][d] =
                    -c_lps * taustabl * (NODAL(ProjEner, n, d) - gradT[d]) * rhonl[g] * Cp / Pr;
                gradRhol[g][d] = -c_lps * taustabl * (NODAL(ProjMass, n, d) - gradRho[d]);
            }
            double tauU[3] = {0, 0, 0};
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 3; ++j) {
                    tauU[i] = tauU[i] + (mu_fgp + mu_egp) * (gradU[i][j] + gradU[j][i]) * ul[g][j];
                    tau[i][j] = (mu_fgp + mu_egp) * (gradU[i][j] + gradU[j][i]);
                }
                tauU[i] = tauU[i] - (mu_fgp)*twoThirds * divU * ul[g][i];
                tau[i][i] = tau[i][i] - (mu_fgp)*twoThirds * divU;
            }
            for (int d = 0; d < 3; ++d) {
                tauXl[g][d] = tauXl[g][d] + tau[0][d];
                tauYl[g][d] = tauYl[g][d] + tau[1][d];
                tauZl[g][d] = tauZl[g][d] + tau[2][d];
                gradTl[g][d] = gradTl[g][d] + kappa_e * gradT[d] + tauU[d];
                gradRhol[g][d] = gradRhol[g][d] + nu_e * gradRho[d];
            }
        }
        for (int g = 0; g < ngaus; ++g) {
            int I = gI[g] - 1, J = gJ[g] - 1, K = gK[g] - 1;
            double divDe = 0.0, divDr = 0.0, divDm[3] = {0, 0, 0};
            for (int ii = 0; ii < n1; ++ii) {
                int b1 = INV(ii, J, K), b2 = INV(I, ii, K), b3 = INV(I, J, ii);
                for (int d = 0; d < 3; ++d) {
                    double w1 = HE(d, 0, b1, e) * GPVOL(b1, e) * DL(b1, 0, I);
                    double w2 = HE(d, 1, b2, e) * GPVOL(b2, e) * DL(b2, 1, J);
                    double w3 = HE(d, 2, b3, e) * GPVOL(b3, e) * DL(b3, 2, K);
                    divDr = divDr + w1 * gradRhol[b1][d];
                    divDr = divDr + w2 * gradRhol[b2][d];
                    divDr = divDr + w3 * gradRhol[b3][d];
                    divDe = divDe + w1 * gradTl[b1][d];
                    divDe = divDe + w2 * gradTl[b2][d];
                    divDe = divDe + w3 * gradTl[b3][d];
                    divDm[0] = divDm[0] + w1 * tauXl[b1][d];
                    divDm[0] = divDm[0] + w2 * tauXl[b2][d];
                    divDm[0] = divDm[0] + w3 * tauXl[b3][d];
                    divDm[1] = divDm[1] + w1 * tauYl[b1][d];
                    divDm[1] = divDm[1] + w2 * tauYl[b2][d];
                    divDm[1] = divDm[1] + w3 * tauYl[b3][d];
                    divDm[2] = divDm[2] + w1 * tauZl[b1][d];
                    divDm[2] = divDm[2] + w2 * tauZl[b2][d];
                    divDm[2] = divDm[2] + w3 * tauZl[b3][d];
                }
            }
            atomic_add(&Rmass[ip[g]], aux_fact * divDr, par);
            atomic_add(&Rener[ip[g]], aux_fact * divDe, par);
            for (int d = 0; d < 3; ++d) atomic_add(&NODAL(Rmom, ip[g], d), aux_fact * divDm[d], par);
        }
    }
}

/* generic_scalar_convec_ijk (elem_convec.f90:472-594) */
void or_generic_scalar_convec_ijk(int p, int nelem, int npoin, const int *connec, const double *He,
                                  const double *gpvol, const double *dlxigp_ip,
                                  const int *invAtoIJK, const int *gI, const int *gJ,
                                  const int *gK, const double *q, const double *eta,
                                  const double *u, double *Rconvec, int initialize, double fact,
                                  int nthreads)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    if (initialize)
        for (int n = 0; n < npoin; ++n) Rconvec[n] = 0.0;
    if (nnode > 128) return;
    const int par = nthreads > 1;
#pragma omp parallel for schedule(static) num_threads(par ? nthreads : 1) if (par)
    for (int e = 0; e < nelem; ++e) {
        int ip[128];
        double ul[128][3], fel[128][3], etal[128], Re[128];
        for (int a = 0; a < nnode; ++a) ip[a] = CONNEC(e, a);
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < nnode; ++a) {
                ul[a][d] = NODAL(u, ip[a], d);
                fel[a][d] = NODAL(q, ip[a], d);
            }
        for (int a = 0; a < nnode; ++a) etal[a] = eta[ip[a]];
        for (int g = 0; g < ngaus; ++g) {
            int I = gI[g] - 1, J = gJ[g] - 1, K = gK[g] - 1;
            double gE[3] = {0}, gU[3][3] = {{0}}, gFe[3][3] = {{0}};
            for (int ii = 0; ii < n1; ++ii) {
                int a1 = INV(ii, J, K), a2 = INV(I, ii, K), a3 = INV(I, J, ii);
                gE[0] = gE[0] + DL(g, 0, ii) * etal[a1];
                gE[1] = gE[1] + DL(g, 1, ii) * etal[a2];
                gE[2] = gE[2] + DL(g, 2, ii) * etal[a3];
                for (int i = 0; i < 3; ++i) {
                    gU[i][0] = gU[i][0] + DL(g, 0, ii) * ul[a1][i];
                    gU[i][1] = gU[i][1] + DL(g, 1, ii) * ul[a2][i];
                    gU[i][2] = gU[i][2] + DL(g, 2, ii) * ul[a3][i];
                    gFe[i][0] = gFe[i][0] + DL(g, 0, ii) * fel[a1][i];
                    gFe[i][1] = gFe[i][1] + DL(g, 1, ii) * fel[a2][i];
                    gFe[i][2] = gFe[i][2] + DL(g, 2, ii) * fel[a3][i];
                }
            }
            double gradE[3] = {0}, divU = 0.0, divFe = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    gradE[i] = gradE[i] + HE(i, j, g, e) * gE[j];
                    divU = divU + HE(i, j, g, e) * gU[i][j];
                    divFe = divFe + HE(i, j, g, e) * gFe[i][j];
                }
            Re[g] = 0.5 * (divFe + etal[g] * divU);
            for (int i = 0; i < 3; ++i) Re[g] = Re[g] + 0.5 * (ul[g][i] * gradE[i]);
            Re[g] = GPVOL(g, e) * Re[g];
        }
        for (int a = 0; a < nnode; ++a) atomic_add(&Rconvec[ip[a]], fact * Re[a], par);
    }
}

/* ------------------------------------------------------------------------ */
/* Entropy viscosity                                                         */
/* ------------------------------------------------------------------------ */

/* init_filters: convertIJK(0:p+2) (mod_filters.f90:15-55) */
void or_convert_ijk(int p, int *c)
{
    for (int ii = 3; ii <= p + 1; ++ii) c[ii - 1] = ii;
    c[0] = 3;
    c[1] = 1;
    c[p + 1] = 2;
    c[p + 2] = p;
}

/* smart_visc_spectral (mod_entropy_viscosity.f90:14-111); helem is (nelem,nnode),
 * helem_k is (nelem) -- the caller passes helem_l and helem respectively
 * (time_integ.f90:223-224). */
void or_smart_visc_spectral(int p, int nelem, int npoin, int npoin_w, const int *connec,
                            const int *lpoin_w, const double *Reta, double gamma_gas,
                            const double *rho, const double *u, const double *Tem,
                            const double *eta, const double *helem, const double *helem_k,
                            const int *invAtoIJK, const or_visc_params *prm, double *mu_e,
                            double *mue_l)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    const double Rgas = prm->Rgas;
    const double fact_low_mach = prm->flag_drop_c_in_envit ? 0.0 : 1.0;
    const double w[3] = {1.0 / 4.0, 2.0 / 4.0, 1.0 / 4.0};
    int cijk[40];
    or_convert_ijk(p, cijk);
    for (size_t i = 0; i < (size_t)nelem * nnode; ++i) mue_l[i] = mu_e[i];
    double norm;
    if (prm->flag_normalise_entropy == 1) {
        double maxEta_r = 0.0;
        for (int i = 0; i < npoin_w; ++i) maxEta_r = maxEta_r + eta[lpoin_w[i] - 1];
        double maxEta = maxEta_r / (double)npoin_w;
        double norm_r = 0.0;
        for (int i = 0; i < npoin_w; ++i) norm_r = fmax(norm_r, fabs(eta[lpoin_w[i] - 1] - maxEta));
        norm = norm_r;
    } else {
        norm = 1.0;
    }
    for (int e = 0; e < nelem; ++e) {
        double ced = prm->ce;
        double betae = 0.0;
        for (int a = 0; a < nnode; ++a) {
            int n = CONNEC(e, a);
            double u0 = NODAL(u, n, 0), u1 = NODAL(u, n, 1), u2 = NODAL(u, n, 2);
            double aux2 = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            double aux3 = fact_low_mach * sqrt(Rgas * gamma_gas * Tem[n]);
            double aux1 = aux2 + aux3;
            betae = fmax(betae, (rho[n] * helem_k[e]) * (prm->cmax / (double)p) * aux1);
        }
        for (int a = 0; a < nnode; ++a) {
            int n = CONNEC(e, a);
            double R1 = rho[n] * fabs(Reta[n]) / norm;
            double h = helem[e + (size_t)nelem * a];
            double Ve = ced * R1 * (h * h);
            mue_l[e + (size_t)nelem * a] = prm->cglob * fmin(Ve, betae);
        }
        for (int ii = 1; ii <= n1; ++ii)
            for (int jj = 1; jj <= n1; ++jj)
                for (int kk = 1; kk <= n1; ++kk) {
                    double aux1 = 0.0;
                    for (int ll = -1; ll <= 1; ++ll)
                        for (int mm = -1; mm <= 1; ++mm)
                            for (int nn = -1; nn <= 1; ++nn) {
                                int a = INV(cijk[ii + ll] - 1, cijk[jj + mm] - 1, cijk[kk + nn] - 1);
                                aux1 = aux1 + w[ll + 1] * w[mm + 1] * w[nn + 1] *
                                                  mue_l[e + (size_t)nelem * a];
                            }
                    int a = INV(cijk[ii] - 1, cijk[jj] - 1, cijk[kk] - 1);
                    mu_e[e + (size_t)nelem * a] = aux1;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Live-path sensor: smart_visc_spectral_imex, nodal divergence and         */
/* vorticity, Ducros limiter                                                 */
/* ------------------------------------------------------------------------ */

/* smart_visc_spectral_imex (mod_entropy_viscosity.f90:113-227).  helem_k is the
 * caller's helem(nelem) (the helem(nelem,nnode) argument is not read).  mue_l
 * receives the incoming mu_e.  With the normalisation on and flag_high_mach the
 * norm is the minimum density over the working nodes and prm->ce becomes
 * prm->ce_comp, as the reference assigns its module global (:148-166). */
void or_smart_visc_spectral_imex(int p, int nelem, int npoin, int npoin_w, const int *connec,
                                 const int *lpoin_w, const double *Reta, const double *rho,
                                 const double *u, const double *csound, const double *eta,
                                 const double *helem_k, or_visc_params *prm, int flag_high_mach,
                                 double *mu_e, double *mue_l)
{
    const int n1 = p + 1, nnode = n1 * n1 * n1;
    const double fact_low_mach = prm->flag_drop_c_in_envit ? 0.0 : 1.0;
    for (size_t i = 0; i < (size_t)nelem * nnode; ++i) mue_l[i] = mu_e[i];
    double norm = 1.0;
    if (prm->flag_normalise_entropy == 1) {
        if (flag_high_mach) {
            double rmin = 100000.0;
            for (int i = 0; i < npoin_w; ++i) rmin = fmin(rmin, rho[lpoin_w[i] - 1]);
            norm = rmin;
            prm->ce = prm->ce_comp;
        } else {
            double sum = 0.0;
            for (int i = 0; i < npoin_w; ++i) sum = sum + eta[lpoin_w[i] - 1];
            const double mean = sum / (double)npoin_w;
            double dev = 0.0;
            for (int i = 0; i < npoin_w; ++i) dev = fmax(dev, fabs(eta[lpoin_w[i] - 1] - mean));
            norm = dev;
        }
    }
    for (int e = 0; e < nelem; ++e) {
        const double ced = prm->ce;
        double betae = 0.0;
        for (int a = 0; a < nnode; ++a) {
            const int n = CONNEC(e, a);
            const double u0 = NODAL(u, n, 0), u1 = NODAL(u, n, 1), u2 = NODAL(u, n, 2);
            const double speed = sqrt(u0 * u0 + u1 * u1 + u2 * u2) + fact_low_mach * csound[n];
            betae = fmax(betae, (rho[n] * helem_k[e]) * (prm->cmax / (double)p) * speed);
        }
        const double hp = helem_k[e] / (double)p;
        for (int a = 0; a < nnode; ++a) {
            const int n = CONNEC(e, a);
            const double R1 = rho[n] * fabs(Reta[n]) / norm;
            const double Ve = ced * R1 * (hp * hp);
            ELGP(mu_e, e, a) = prm->cglob * fmin(Ve, betae);
        }
    }
}

/* line derivatives of the velocity at GLL point g of an element whose nodal
 * velocities are ul: gI[d][k] = d u_d / d xi_k, as the reference's gradIsoU
 * (summed over the line index ii in order) */
static void or_grad_iso(int n1, int ngaus, const double *dlxigp_ip, const int *invAtoIJK, int g,
                        int I, int J, int K, double ul[][3], double gI[3][3])
{
    for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 3; ++k) gI[d][k] = 0.0;
    for (int ii = 0; ii < n1; ++ii) {
        const int a1 = INV(ii, J, K), a2 = INV(I, ii, K), a3 = INV(I, J, ii);
        for (int d = 0; d < 3; ++d) {
            gI[d][0] = gI[d][0] + DL(g, 0, ii) * ul[a1][d];
            gI[d][1] = gI[d][1] + DL(g, 1, ii) * ul[a2][d];
            gI[d][2] = gI[d][2] + DL(g, 2, ii) * ul[a3][d];
        }
    }
}

/* eval_divergence (mod_operators.f90:485-570): divU(n) = sum over the elements
 * of n of gpvol * div u at n's GLL point, in element order */
void or_eval_divergence(int p, int nelem, int npoin, const int *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int *invAtoIJK,
                        const int *gI, const int *gJ, const int *gK, const double *u,
                        double *divU)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    for (int n = 0; n < npoin; ++n) divU[n] = 0.0;
    if (nnode > 128) return;
    for (int e = 0; e < nelem; ++e) {
        int ip[128];
        double ul[128][3], Re[128];
        for (int a = 0; a < nnode; ++a) {
            ip[a] = CONNEC(e, a);
            for (int d = 0; d < 3; ++d) ul[a][d] = NODAL(u, ip[a], d);
        }
        for (int g = 0; g < ngaus; ++g) {
            double gi[3][3];
            or_grad_iso(n1, ngaus, dlxigp_ip, invAtoIJK, g, gI[g] - 1, gJ[g] - 1, gK[g] - 1, ul,
                        gi);
            double dv = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) dv = dv + HE(i, j, g, e) * gi[i][j];
            Re[g] = GPVOL(g, e) * dv;
        }
        for (int a = 0; a < nnode; ++a) divU[ip[a]] = divU[ip[a]] + Re[a];
    }
}

/* compute_vorticity with makeConti (mod_operators.f90:684-775):
 * curlU(n,i) = sum over the elements of n of eps_ijk du_j/dx_k (minus the
 * curl) at n's GLL point (leviCivi of CFDSolverBase.f90:1605-1611: eps_123 =
 * +1), each term
 * added in (i,j,k) order, then divided by lelpn(n) on the working nodes */
void or_compute_vorticity(int p, int nelem, int npoin, int npoin_w, const int *lpoin_w,
                          const int *connec, const int *lelpn, const double *He,
                          const double *dlxigp_ip, const int *invAtoIJK, const int *gI,
                          const int *gJ, const int *gK, const double *u, double *curlU)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    double lc[3][3][3] = {{{0}}};
    lc[1][2][0] = 1.0;
    lc[2][1][0] = -1.0;
    lc[0][2][1] = -1.0;
    lc[2][0][1] = 1.0;
    lc[0][1][2] = 1.0;
    lc[1][0][2] = -1.0;
    for (size_t n = 0; n < 3 * (size_t)npoin; ++n) curlU[n] = 0.0;
    if (nnode > 128) return;
    for (int e = 0; e < nelem; ++e) {
        double ul[128][3];
        for (int a = 0; a < nnode; ++a)
            for (int d = 0; d < 3; ++d) ul[a][d] = NODAL(u, CONNEC(e, a), d);
        for (int g = 0; g < ngaus; ++g) {
            double gi[3][3], gu[3][3];
            or_grad_iso(n1, ngaus, dlxigp_ip, invAtoIJK, g, gI[g] - 1, gJ[g] - 1, gK[g] - 1, ul,
                        gi);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    gu[i][j] = 0.0;
                    for (int k = 0; k < 3; ++k) gu[i][j] = gu[i][j] + HE(j, k, g, e) * gi[i][k];
                }
            const int n = CONNEC(e, g);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k)
                        NODAL(curlU, n, i) = NODAL(curlU, n, i) + lc[i][j][k] * gu[j][k];
        }
    }
    for (int w = 0; w < npoin_w; ++w) {
        const int n = lpoin_w[w] - 1;
        for (int d = 0; d < 3; ++d) NODAL(curlU, n, d) = NODAL(curlU, n, d) / (double)lelpn[n];
    }
}

/* ducros_visc_limit (mod_entropy_viscosity.f90:277-314): mu_e scaled in place by
 * max(phi - ducros_min_val, 0) phi2 + ducros_min_val, with divUl = divU/Ml,
 * phi = min(4/3 divUl^2 / (divUl^2 + |vorti|^2 + eps), 1) and
 * phi2 = divUl^2 / (divUl^2 + (|u|/helem)^2 + eps), eps the machine epsilon.
 * helem(nelem,nnode) is the caller's helem_l. */
void or_ducros_visc_limit(int p, int nelem, int npoin, const int *connec, const double *u,
                          const double *divU, const double *vorti, const double *Ml,
                          const double *helem, double ducros_min_val, double *mu_e)
{
    const int n1 = p + 1, nnode = n1 * n1 * n1;
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nnode; ++a) {
            const int n = CONNEC(e, a);
            const double dv = divU[n] / Ml[n];
            const double u0 = NODAL(u, n, 0), u1 = NODAL(u, n, 1), u2 = NODAL(u, n, 2);
            const double w0 = NODAL(vorti, n, 0), w1 = NODAL(vorti, n, 1), w2 = NODAL(vorti, n, 2);
            const double umag = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
            const double phi =
                fmin((4.0 / 3.0) * (dv * dv) / (dv * dv + (w0 * w0 + w1 * w1 + w2 * w2) + DBL_EPSILON),
                     1.0);
            const double t = 1.0 * umag / ELGP(helem, e, a);
            const double phi2 = (dv * dv) / (dv * dv + t * t + DBL_EPSILON);
            ELGP(mu_e, e, a) = ELGP(mu_e, e, a) * (fmax(phi - ducros_min_val, 0.0) * phi2 + ducros_min_val);
        }
}

/* elemPerNode (mod_geom.f90:68-99): lelpn(n) = incidences of n in connec */
void or_elem_per_node(int nelem, int nnode, int npoin, const int *connec, int *lelpn)
{
    for (int n = 0; n < npoin; ++n) lelpn[n] = 0;
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nnode; ++a) lelpn[CONNEC(e, a)] += 1;
}

/* thermo_entropy (mod_entropy_viscosity.f90:411-433); eta is (npoin,2) */
void or_thermo_entropy(int npoin, int npoin_w, const int *lpoin_w, double gamma_gas,
                       double p_inf, double rho_inf, const double *rho, const double *pr,
                       const double *u, double *eta, double *feta)
{
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        eta[n] = eta[n + npoin];
        eta[n + npoin] = (rho[n] / (gamma_gas - 1.0)) *
                         log(fmax(pr[n] / p_inf, DBL_EPSILON) / pow(rho[n] / rho_inf, gamma_gas));
        for (int d = 0; d < 3; ++d) NODAL(feta, n, d) = NODAL(u, n, d) * eta[n];
    }
}

/* ------------------------------------------------------------------------ */
/* Nodal updates                                                             */
/* ------------------------------------------------------------------------ */

/* lumped_solver_scal_opt (mod_solver.f90:40-59) */
void or_lumped_solver_scal(int npoin, int npoin_w, const int *lpoin_w, const double *invMl,
                           double *R)
{
    (void)npoin;
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        R[n] = R[n] * invMl[n];
    }
}

/* lumped_solver_vect_opt (mod_solver.f90:81-104) */
void or_lumped_solver_vect(int npoin, int npoin_w, const int *lpoin_w, const double *invMl,
                           double *R)
{
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        for (int d = 0; d < 3; ++d) NODAL(R, n, d) = NODAL(R, n, d) * invMl[n];
    }
}

/* sutherland_viscosity (mod_fluid_viscosity.f90:21-41) */
void or_sutherland(int npoin, const double *Tem, const double *mu_factor, double *mu_fluid)
{
    const double C1 = 0.000001458, S = 110.4;
    for (int n = 0; n < npoin; ++n)
        mu_fluid[n] = mu_factor[n] * C1 * pow(Tem[n], 1.5) / (Tem[n] + S);
}

/* updateF "Update u and EOS" (time_integ_ls.f90:574-587) */
void or_eos_stage(int npoin, int npoin_w, const int *lpoin_w, double gamma_gas, double Rgas,
                  const double *rho, const double *q, const double *E, double *u, double *e_int,
                  double *pr, double *aux_h, double *Tem)
{
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        double umag = 0.0;
        for (int d = 0; d < 3; ++d) {
            NODAL(u, n, d) = NODAL(q, n, d) / rho[n];
            umag = umag + (NODAL(u, n, d) * NODAL(u, n, d));
        }
        e_int[n] = (E[n] / rho[n]) - 0.5 * umag;
        pr[n] = rho[n] * (gamma_gas - 1.0) * e_int[n];
        aux_h[n] = (gamma_gas / (gamma_gas - 1.0)) * pr[n] / rho[n];
        Tem[n] = pr[n] / (rho[n] * Rgas);
    }
}

/* rk_4_ls_main "Update velocity and equations of state" (time_integ_ls.f90:391-407) */
void or_eos_step(int npoin, int npoin_w, const int *lpoin_w, double gamma_gas, double Rgas,
                 const double *rho, const double *q, const double *E, double *u, double *e_int,
                 double *pr, double *csound, double *machno, double *Tem)
{
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        double umag = 0.0;
        for (int d = 0; d < 3; ++d) {
            NODAL(u, n, d) = NODAL(q, n, d) / rho[n];
            umag = umag + NODAL(u, n, d) * NODAL(u, n, d);
        }
        e_int[n] = (E[n] / rho[n]) - 0.5 * umag;
        pr[n] = rho[n] * (gamma_gas - 1.0) * e_int[n];
        csound[n] = sqrt(gamma_gas * pr[n] / rho[n]);
        umag = sqrt(umag);
        machno[n] = umag / csound[n];
        Tem[n] = pr[n] / (rho[n] * Rgas);
    }
}

/* init_rk4_ls_solver coefficient tables (time_integ_ls.f90:87-171).  Constants
 * the reference writes without a d0 exponent are default-real (FP32) literals
 * converted to real(rp); they are reproduced through (double)(float). */
int or_rk_ls_coeffs(int stages, int variant, double *L, double *G)
{
#define LIJ(i, j) L[((i)-1) + 6 * ((j)-1)]
#define GIJ(i, j) G[((i)-1) + 6 * ((j)-1)]
#define F32(x) ((double)(float)(x))
    for (int i = 0; i < 36; ++i) L[i] = G[i] = 0.0;
    if (stages == 3) {
        LIJ(2, 1) = 1.0;
        LIJ(3, 1) = 3.0 / 4.0;
        LIJ(3, 2) = 1.0 / 4.0;
        LIJ(4, 1) = 1.0 / 3.0;
        LIJ(4, 3) = 2.0 / 3.0;
        GIJ(2, 1) = 1.0;
        GIJ(3, 2) = 0.25;
        GIJ(4, 3) = 2.0 / 3.0;
    } else if (stages == 4) {
        LIJ(2, 1) = 1.0;
        LIJ(3, 2) = 1.0;
        LIJ(4, 1) = 2.0 / 3.0;
        LIJ(4, 3) = 1.0 / 3.0;
        LIJ(5, 4) = 1.0;
        GIJ(2, 1) = 0.5;
        GIJ(3, 2) = 0.5;
        GIJ(4, 3) = 1.0 / 6.0;
        GIJ(5, 4) = 0.5;
    } else if (stages == 5) {
        LIJ(2, 1) = 1.0;
        LIJ(3, 2) = 1.0;
        LIJ(4, 3) = 1.0;
        LIJ(5, 4) = 1.0;
        LIJ(6, 5) = 1.0;
        switch (variant) {
        case 1:
            LIJ(5, 1) = 0.571403511494104;
            LIJ(5, 4) = 0.428596488505896;
            GIJ(2, 1) = 0.443568244942995;
            GIJ(3, 2) = 0.291111420073766;
            GIJ(4, 3) = 0.270612601278217;
            GIJ(5, 4) = 0.110577759392786;
            GIJ(6, 5) = 0.458557505351052;
            break;
        case 2:
            LIJ(4, 1) = 0.682342861037239;
            LIJ(4, 3) = 0.317657138962761;
            LIJ(6, 1) = 0.045230974482400;
            LIJ(6, 5) = 0.954769025517600;
            GIJ(2, 1) = F32(0.465388589249323);
            GIJ(3, 2) = F32(0.465388589249323);
            GIJ(4, 3) = F32(0.124745797313998);
            GIJ(5, 4) = F32(0.465388589249323);
            GIJ(6, 5) = F32(0.154263303748666);
            break;
        case 3:
            LIJ(5, 1) = F32(0.592032910942121);
            LIJ(5, 4) = 1.0 - LIJ(5, 1);
            GIJ(2, 1) = F32(0.266541020678955);
            GIJ(3, 2) = F32(0.548560709048532);
            GIJ(4, 3) = F32(0.289517014154401);
            GIJ(5, 4) = F32(0.086408328057923);
            GIJ(6, 5) = F32(0.462943578481813);
            break;
        case 4:
            LIJ(4, 1) = F32(0.707858560931430);
            LIJ(4, 3) = 1.0 - LIJ(4, 1);
            LIJ(6, 1) = F32(0.222853615080669);
            LIJ(6, 5) = 1.0 - LIJ(6, 1);
            GIJ(2, 1) = F32(0.292845746913355);
            GIJ(3, 2) = F32(0.339532793976408);
            GIJ(4, 3) = F32(0.200532330324672);
            GIJ(5, 4) = F32(0.701676169006879);
            GIJ(6, 5) = F32(0.155278812461877);
            break;
        default:
            return -1;
        }
    } else {
        return -1;
    }
    return 0;
#undef LIJ
#undef GIJ
#undef F32
}

/* init_rk4_solver, flag_rk_order == 4 (time_integ.f90:84-87) */
void or_rk4_coeffs(double *a, double *b, double *c)
{
    a[0] = 0.0; a[1] = 0.5; a[2] = 0.5; a[3] = 1.0;
    c[0] = 0.0; c[1] = 0.5; c[2] = 0.5; c[3] = 1.0;
    b[0] = 1.0 / 6.0; b[1] = 1.0 / 3.0; b[2] = 1.0 / 3.0; b[3] = 1.0 / 6.0;
}

/* Shu-Osher 2N stage combine (time_integ_ls.f90:323-331 with l1=0, ls=1; :350-358) */
void or_rk_ls_update(int npoin, int npoin_w, const int *lpoin_w, double l1, double ls, double gdt,
                     const double *rho1, const double *q1, const double *E1, double *rho,
                     double *q, double *E, const double *Rmass, const double *Rmom,
                     const double *Rener)
{
    for (int i = 0; i < npoin_w; ++i) {
        int n = lpoin_w[i] - 1;
        rho[n] = l1 * rho1[n] + ls * rho[n] + gdt * Rmass[n];
        E[n] = l1 * E1[n] + ls * E[n] + gdt * Rener[n];
        for (int d = 0; d < 3; ++d)
            NODAL(q, n, d) = l1 * NODAL(q1, n, d) + ls * NODAL(q, n, d) + gdt * NODAL(Rmom, n, d);
    }
}

/* ------------------------------------------------------------------------ */
/* Diagnostics and dt                                                        */
/* ------------------------------------------------------------------------ */

/* volAvg_EK (mod_analysis.f90:10-44) */
double or_volavg_ek(int p, int nelem, int npoin, const int *connec, const double *gpvol,
                    const double *Ngp, double rho0, const double *rho, const double *u)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    double EK_l = 0.0;
    for (int e = 0; e < nelem; ++e) {
        double R1 = 0.0;
        for (int g = 0; g < ngaus; ++g)
            for (int a = 0; a < nnode; ++a) {
                int n = CONNEC(e, a);
                double uu = NODAL(u, n, 0) * NODAL(u, n, 0) + NODAL(u, n, 1) * NODAL(u, n, 1) +
                            NODAL(u, n, 2) * NODAL(u, n, 2);
                R1 = R1 + GPVOL(g, e) * Ngp[g + (size_t)ngaus * a] * 0.5 * rho[n] * uu;
            }
        EK_l = EK_l + R1;
    }
    EK_l = EK_l / (rho0 * pow(2.0 * 3.14159, 3.0));
    return EK_l;
}

/* visc_dissipationRate (mod_analysis.f90:46-144) */
void or_visc_dissipation(int p, int nelem, int npoin, const int *connec, double rho0,
                         const double *mu_fluid, const double *mu_e, const double *u, double volT,
                         const double *gpvol, const double *He, const double *dNgp,
                         double *eps_S, double *eps_D, double *eps_T)
{
    const int n1 = p + 1, ngaus = n1 * n1 * n1, nnode = ngaus;
    double lc[3][3][3] = {{{0}}};
    lc[1][2][0] = 1.0;
    lc[2][1][0] = -1.0;
    lc[0][2][1] = -1.0;
    lc[2][0][1] = 1.0;
    lc[0][1][2] = 1.0;
    lc[1][0][2] = -1.0;
    double eS = 0.0, eD = 0.0;
    double *gpcar = (double *)malloc(sizeof(double) * 3 * nnode);
    for (int e = 0; e < nelem; ++e) {
        double R1 = 0.0, R2 = 0.0;
        for (int g = 0; g < ngaus; ++g) {
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < nnode; ++a) {
                    const double *dn = &dNgp[3 * (a + (size_t)nnode * g)];
                    gpcar[i + 3 * a] = HE(i, 0, g, e) * dn[0] + HE(i, 1, g, e) * dn[1] +
                                       HE(i, 2, g, e) * dn[2];
                }
            double gradU[3][3];
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    double aux = 0.0;
                    for (int a = 0; a < nnode; ++a)
                        aux = aux + gpcar[j + 3 * a] * NODAL(u, CONNEC(e, a), k);
                    gradU[j][k] = aux;
                }
            double div2U = gradU[0][0] + gradU[1][1] + gradU[2][2];
            div2U = div2U * div2U;
            double curl2U = 0.0;
            for (int i = 0; i < 3; ++i) {
                double aux2 = 0.0;
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) aux2 = aux2 + (lc[i][j][k] * gradU[j][k]);
                curl2U = curl2U + aux2 * aux2;
            }
            int ng = CONNEC(e, g);
            R1 = R1 + GPVOL(g, e) * (mu_fluid[ng] + ELGP(mu_e, e, g)) * curl2U;
            R2 = R2 + GPVOL(g, e) * mu_fluid[ng] * div2U;
        }
        eS = eS + R1;
        eD = eD + R2;
    }
    free(gpcar);
    double alpha = 1.0 / (rho0 * volT);
    eS = eS * alpha;
    eD = (double)(4.0f / 3.0f) * eD * alpha; /* default-real literal (mod_analysis.f90:139) */
    *eps_T = eS + eD;
    *eps_S = eS;
    *eps_D = eD;
}

/* maxMach (mod_analysis.f90:146-168) */
double or_max_mach(int npoin, int npoin_w, const int *lpoin_w, const double *machno)
{
    (void)npoin;
    double m = 0.0;
    for (int i = 0; i < npoin_w; ++i) m = fmax(m, machno[lpoin_w[i] - 1]);
    return m;
}

/* adapt_dt_cfl (mod_time_ops.f90:11-63); mu_sgs and mu_e are (nelem,nnode) */
double or_adapt_dt_cfl(int p, int nelem, int npoin, const int *connec, const double *helem,
                       const double *u, const double *csound, double cfl_conv, double cfl_diff,
                       const double *mu_fluid, const double *mu_sgs, const double *mu_e,
                       const double *rho)
{
    const int n1 = p + 1, nnode = n1 * n1 * n1;
    double dt_conv = 100000000000000.0, dt_diff = 100000000000000.0, dt_l = 100000000000000.0;
    for (int e = 0; e < nelem; ++e) {
        double L3 = 0.0;
        for (int a = 0; a < nnode; ++a) {
            int n = CONNEC(e, a);
            double umag = fabs(NODAL(u, n, 0));
            umag = fmax(umag, fabs(NODAL(u, n, 1)));
            umag = fmax(umag, fabs(NODAL(u, n, 2)));
            double aux1 = umag + csound[n];
            L3 = fmax(L3, aux1);
        }
        double aux2 = cfl_conv * (helem[e] / (double)(2.0 * p + 1)) / L3;
        dt_conv = fmin(dt_conv, aux2);
        if (mu_fluid && mu_sgs && rho) {
            double max_MU = 0.0;
            for (int a = 0; a < nnode; ++a) {
                int n = CONNEC(e, a);
                max_MU = fmax(max_MU, (ELGP(mu_sgs, e, a) + mu_fluid[n] + ELGP(mu_e, e, a)) / rho[n]);
            }
            double hh = helem[e] / (double)(2.0 * p + 1);
            double aux4 = cfl_diff * (hh * hh) / max_MU;
            dt_diff = fmin(dt_diff, aux4);
        }
        dt_l = fmin(dt_conv, dt_diff);
    }
    return dt_l;
}

/* ------------------------------------------------------------------------ */
/* TGV case                                                                  */
/* ------------------------------------------------------------------------ */

/* TGVSolver_initializeParameters derived constants (TGVSolver.f90:134-139) */
void or_tgv_params_derive(or_tgv_params *t)
{
    double mul = (t->rho0 * 1.0 * 1.0) / t->Re;
    t->Rgas = t->Cp * (t->gamma_gas - 1.0) / t->gamma_gas;
    t->to = 1.0 * 1.0 / (t->gamma_gas * t->Rgas * t->M * t->M);
    t->po = t->rho0 * t->Rgas * t->to;
    double mur = 0.000001458 * pow(t->to, 1.50) / (t->to + 110.40);
    t->mu_factor = mul / mur;
}

/* TGVSolver_evalInitialConditions (TGVSolver.f90:154-221); coordinates are
 * real(4) and the velocity is evaluated in single precision. */
void or_tgv_ic(int npoin, const double *coord, const or_tgv_params *t, double *u, double *q,
               double *rho, double *pr, double *E, double *e_int, double *Tem, double *csound,
               double *eta, double *machno)
{
    const float V0 = 1.0f, L = 1.0f;
    const double g = t->gamma_gas;
    for (int n = 0; n < npoin; ++n) {
        float x = (float)NODAL(coord, n, 0), y = (float)NODAL(coord, n, 1),
              z = (float)NODAL(coord, n, 2);
        NODAL(u, n, 0) = (double)(V0 * sinf(x / L) * cosf(y / L) * cosf(z / L));
        NODAL(u, n, 1) = (double)(-(V0 * cosf(x / L) * sinf(y / L) * cosf(z / L)));
        NODAL(u, n, 2) = 0.0;
        pr[n] = t->po + ((1.0 * (double)V0 * (double)V0) / (16.0)) *
                            (cos(2.0 * (double)x / (double)L) + cos(2.0 * (double)y / (double)L)) *
                            (cos(2.0 * (double)z / (double)L) + 2.0);
        rho[n] = pr[n] / t->Rgas / t->to;
    }
    for (int n = 0; n < npoin; ++n) {
        double uu = NODAL(u, n, 0) * NODAL(u, n, 0) + NODAL(u, n, 1) * NODAL(u, n, 1) +
                    NODAL(u, n, 2) * NODAL(u, n, 2);
        e_int[n] = pr[n] / (rho[n] * (g - 1.0));
        Tem[n] = pr[n] / (rho[n] * t->Rgas);
        E[n] = rho[n] * (0.5 * uu + e_int[n]);
        for (int d = 0; d < 3; ++d) NODAL(q, n, d) = rho[n] * NODAL(u, n, d);
        csound[n] = sqrt(g * pr[n] / rho[n]);
        eta[n] = (rho[n] / (g - 1.0)) * log(pr[n] / pow(rho[n], g));
        machno[n] = uu / csound[n];
    }
}
