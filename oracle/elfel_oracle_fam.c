/*
 * elfel_oracle_fam.c -- CPU restatement of the remaining element families of
 * Elfel.jl and the weak forms that use them, in plain C (FP64, unfused).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library's element-family
 * kernels.  The product path never calls it.
 *
 * Families (src/FElements.jl):
 *   FEH1_L2         :203-225  two-node segment, manifold dimension 1
 *   FEH1_T3_BUBBLE  :322-354  T3 + cubic bubble on the cell (4 basis functions,
 *                             3 geometry nodes, one dof on the cell)
 *   FEH1_T4         :356-384  four-node tetrahedron, manifold dimension 3
 *   FEL2_Q4         :386-413  one constant basis function on the cell; the
 *                             nodal Q4 is its geometry carrier (:393)
 * The nodal T3 / Q4 / T6 come from elfel_oracle.c (orc_bfun).
 *
 * Rules (src/RefShapes.jl): interval Gauss :269-290 (orc_gauss1), triangle
 * :301-323, square :333-366, tetrahedron :232-258,376-398.
 *
 * Geometry (src/FElements.jl:113-168): J = sum_k x_k (outer) dN_k/dxi over the
 * geometry nodes, first term assigned then added; det J: |J11| for a curve in
 * one space dimension (norm, :105-110), J11 J22 - J21 J12 (:120-129), the
 * unrolled 3x3 expansion (:138-146).  Spatial gradients (src/QPIterators.jl:
 * 132-140, g / J): g / J11 in 1D; in 2D and 3D the row vector g times the
 * adjugate of J, divided by det J, each component summed in index order.
 * (StaticArrays' own small solves are not in the reference tree, so the last
 * bit of that rounding is this restatement's choice; the HIP kernels use the
 * same expressions.)
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "elfel_oracle.h"

int orc_fam_nbf(int fe)
{
    switch (fe) {
    case ORC_T3: return 3;
    case ORC_T6: return 6;
    case ORC_Q4: return 4;
    case ORC_T3B: return 4;
    case ORC_Q4L2: return 1;
    case ORC_T4: return 4;
    case ORC_L2: return 2;
    }
    return -1;
}

int orc_fam_ngeo(int fe)
{
    switch (fe) {
    case ORC_T3: case ORC_T3B: return 3;
    case ORC_T6: return 6;
    case ORC_Q4: case ORC_Q4L2: return 4;
    case ORC_T4: return 4;
    case ORC_L2: return 2;
    }
    return -1;
}

int orc_fam_mdim(int fe)
{
    if (fe == ORC_L2) return 1;
    if (fe == ORC_T4) return 3;
    return orc_fam_nbf(fe) < 0 ? -1 : 2;
}

/* the rule a family integrates with: triangle npts, square / interval Gauss
 * order, tetrahedron npts.  pc holds 3 coordinates per point. */
int orc_fam_quad(int fe, int rule, double *pc, double *w)
{
    double p2[32];
    int n = -1;
    memset(pc, 0, sizeof(double) * 3 * 16);
    if (fe == ORC_L2) {
        double p1[8];
        n = orc_gauss1(rule, p1, w);
        for (int q = 0; q < n; q++) pc[3 * q] = p1[q];
        return n;
    }
    if (fe == ORC_T4) {
        if (rule == 1) {
            pc[0] = pc[1] = pc[2] = 0.25;
            w[0] = 1.0 / 6.0;
            return 1;
        }
        if (rule == 4) {
            static const double a[4][3] = {{0.13819660, 0.13819660, 0.13819660},
                                           {0.58541020, 0.13819660, 0.13819660},
                                           {0.13819660, 0.58541020, 0.13819660},
                                           {0.13819660, 0.13819660, 0.58541020}};
            for (int q = 0; q < 4; q++) {
                pc[3 * q] = a[q][0]; pc[3 * q + 1] = a[q][1]; pc[3 * q + 2] = a[q][2];
                w[q] = 0.041666666666666666667;
            }
            return 4;
        }
        return -1;
    }
    if (fe == ORC_Q4 || fe == ORC_Q4L2) n = orc_quad_square(rule, p2, w);
    else n = orc_quad_triangle(rule, p2, w);
    for (int q = 0; q < n; q++) { pc[3 * q] = p2[2 * q]; pc[3 * q + 1] = p2[2 * q + 1]; }
    return n;
}

void orc_fam_bfun(int fe, const double *pc, double *N)
{
    double r = pc[0], s = pc[1], t = pc[2];
    switch (fe) {
    case ORC_T3B:
        N[0] = (1 - r - s); N[1] = r; N[2] = s;
        N[3] = (1 - r - s) * r * s;
        return;
    case ORC_Q4L2: N[0] = 1.0; return;
    case ORC_T4: N[0] = (1 - r - s - t); N[1] = r; N[2] = s; N[3] = t; return;
    case ORC_L2: N[0] = (1. - r) / 2.0; N[1] = (1. + r) / 2.0; return;
    default: orc_bfun(fe, pc, N);
    }
}

/* g[k*mdim + d] = dN_k / dxi_d */
void orc_fam_bfungradpar(int fe, const double *pc, double *g)
{
    double r = pc[0], s = pc[1];
    switch (fe) {
    case ORC_T3B:
        g[0] = -1.; g[1] = -1.; g[2] = +1.; g[3] = 0.; g[4] = 0.; g[5] = +1.;
        g[6] = (-r * s + (1 - r - s) * s);
        g[7] = (-r * s + (1 - r - s) * r);
        return;
    case ORC_Q4L2: g[0] = 0.0; g[1] = 0.0; return;
    case ORC_T4: {
        static const double G[12] = {-1.0, -1.0, -1.0, +1.0, 0.0, 0.0, 0.0, +1.0, 0.0, 0.0, 0.0, +1.0};
        memcpy(g, G, sizeof(G));
        return;
    }
    case ORC_L2: g[0] = -1.0 / 2.0; g[1] = +1.0 / 2.0; return;
    default: orc_bfungradpar(fe, pc, g);
    }
}

/* ---- geometry ---------------------------------------------------------- */
typedef struct { double j[3][3], det; int m; } jacn_t;

static void jacn(int m, int ngeo, const double *X /* ngeo*m */, const double *G /* ngeo*m */, jacn_t *J)
{
    J->m = m;
    for (int s = 0; s < m; s++)
        for (int d = 0; d < m; d++) J->j[s][d] = X[s] * G[d];
    for (int k = 1; k < ngeo; k++)
        for (int s = 0; s < m; s++)
            for (int d = 0; d < m; d++) J->j[s][d] += X[k * m + s] * G[k * m + d];
    const double (*a)[3] = J->j;
    if (m == 1) J->det = fabs(a[0][0]);
    else if (m == 2) J->det = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    else
        J->det = (+a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2])
                  - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
                  + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]));
}

/* gx[k*m + s] = (g_k / J)_s */
static void gradn(const jacn_t *J, int nbf, const double *G, double *gx)
{
    const int m = J->m;
    const double (*a)[3] = J->j;
    if (m == 1) {
        for (int k = 0; k < nbf; k++) gx[k] = G[k] / a[0][0];
        return;
    }
    if (m == 2) {
        for (int k = 0; k < nbf; k++) {
            const double g1 = G[2 * k], g2 = G[2 * k + 1];
            gx[2 * k] = (a[1][1] * g1 - a[1][0] * g2) / J->det;
            gx[2 * k + 1] = (a[0][0] * g2 - a[0][1] * g1) / J->det;
        }
        return;
    }
    /* cofactors C[s][d] of J; (g J^-1)_s = sum_d g_d C[s][d] / det */
    double C[3][3];
    C[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    C[0][1] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    C[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    C[1][0] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    C[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    C[1][2] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    C[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    C[2][1] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    C[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    for (int k = 0; k < nbf; k++) {
        const double *g = G + 3 * k;
        for (int s = 0; s < 3; s++) gx[3 * k + s] = (g[0] * C[s][0] + g[1] * C[s][1] + g[2] * C[s][2]) / J->det;
    }
}

static void gather_x(const int64_t *conn, int nen, int64_t e, int m, const double *x, double *X)
{
    for (int k = 0; k < nen; k++) {
        const int64_t n = conn[(int64_t)nen * e + k] - 1;
        for (int s = 0; s < m; s++) X[k * m + s] = x[m * n + s];
    }
}

typedef struct {
    int nq, nbf, m;
    double w[16];
    double N[16][8];
    double G[16][24];
} ftab_t;

static int make_ftab(int fe, int rule, ftab_t *t)
{
    double pc[48];
    int n = orc_fam_quad(fe, rule, pc, t->w);
    if (n < 1 || n > 16) return -1;
    t->nq = n;
    t->nbf = orc_fam_nbf(fe);
    t->m = orc_fam_mdim(fe);
    for (int q = 0; q < n; q++) {
        orc_fam_bfun(fe, pc + 3 * q, t->N[q]);
        orc_fam_bfungradpar(fe, pc + 3 * q, t->G[q]);
    }
    return 0;
}

static int nodal(int fe) { return fe == ORC_T3 || fe == ORC_T6 || fe == ORC_Q4 || fe == ORC_T4 || fe == ORC_L2; }

/* ---- scalar forms on a nodal family ---------------------------------------
 * heat: ke[i,j] += dot(gradN_i, gradN_j) * (kappa JxW), fe[j] += N_j Q JxW
 *       (examples/heat/poisson/t4.jl:33-55; the same loop as test/test_heat.jl:36-60);
 * mass: me[i,j] += (N_i N_j) * (kappa JxW)
 *       (examples/klein_gordon/klein_gordon.jl:34-51 with kappa = 1, :94-112 with m^2).
 * dot sums the components in index order (StaticArrays' unrolled dot).
 * COO: per element ke column-major, fe in dof order. */
int64_t orc_fam_scalar_coo_len(int fe, int64_t E, int64_t *nvec)
{
    const int n = orc_fam_nbf(fe);
    if (nvec) *nvec = E * n;
    return E * n * n;
}

int orc_fam_scalar_coo(int kind, int fe, int rule, int64_t E, const int64_t *conn, const double *x,
                       const int64_t *dofnums, double kappa, double Q, int64_t *I, int64_t *J, double *V,
                       int64_t *FI, double *FV)
{
    if (!nodal(fe) || (kind != ORC_HEAT && kind != ORC_MASS)) return -1;
    ftab_t t;
    if (make_ftab(fe, rule, &t)) return -2;
    const int n = t.nbf, m = t.m;
    for (int64_t e = 0; e < E; e++) {
        double X[24], ke[64] = {0}, fe_[8] = {0}, gx[24];
        gather_x(conn, n, e, m, x, X);
        for (int q = 0; q < t.nq; q++) {
            jacn_t Jc;
            jacn(m, n, X, t.G[q], &Jc);
            gradn(&Jc, n, t.G[q], gx);
            const double JxW = Jc.det * t.w[q];
            for (int j = 0; j < n; j++) {
                for (int i = 0; i < n; i++) {
                    if (kind == ORC_HEAT) {
                        double d = gx[i * m] * gx[j * m];
                        for (int s = 1; s < m; s++) d += gx[i * m + s] * gx[j * m + s];
                        ke[i + n * j] += d * (kappa * JxW);
                    } else {
                        ke[i + n * j] += (t.N[q][i] * t.N[q][j]) * (kappa * JxW);
                    }
                }
                if (kind == ORC_HEAT) fe_[j] += t.N[q][j] * Q * JxW;
            }
        }
        int64_t d[8];
        for (int k = 0; k < n; k++) d[k] = dofnums[conn[(int64_t)n * e + k] - 1];
        int64_t o = e * n * n;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) { I[o] = d[i]; J[o] = d[j]; V[o] = ke[i + n * j]; o++; }
        if (kind == ORC_HEAT && FI)
            for (int j = 0; j < n; j++) { FI[e * n + j] = d[j]; FV[e * n + j] = fe_[j]; }
    }
    return 0;
}

/* ---- Reddy-form Stokes over three scalar spaces -----------------------------
 * test/test_stokes.jl:373-431 (T6 velocity, T3 pressure),
 * examples/stokes/colliding_flow/p1b_p1.jl:50-108 (T3 + bubble, T3),
 * examples/stokes/colliding_flow/q1_q0.jl:52-110 (Q4, FEL2_Q4).
 *   kxx += (mu JxW)(2 gx_i gx_j + gy_i gy_j),  kyy += (mu JxW)(gx_i gx_j + 2 gy_i gy_j),
 *   kxy += (mu JxW)(gx_i gy_j),  kxp += (-JxW)(gx_i Np_j),  kyp += (-JxW)(gy_i Np_j)
 * The Jacobian comes from the pressure element where it is nodal (T3: the
 * first two drivers call jacjac(pel, pqp)) and from the velocity element for
 * the L2 pressure (q1_q0.jl calls jacjac(uxel, uxqp)).  Element dofs: the
 * nodes' dofs, then the cell's (src/FEIterators.jl:161-171, _storedofs! over
 * the incidence relations of dimension 0, then 2).  Blocks are appended
 * kxx, kxy, kxy', kyy, kxp, kxp', kyp, kyp' (column-major; a transpose
 * row-major of its parent). */
int64_t orc_fam_reddy_coo_len(int vfe, int pfe, int64_t E)
{
    const int nu = orc_fam_nbf(vfe), np = orc_fam_nbf(pfe);
    return E * (4 * nu * nu + 4 * nu * np);
}

int orc_fam_reddy_coo(int vfe, int pfe, int rule, int64_t E, const int64_t *conn_v, const double *xy_v,
                      const int64_t *conn_p, const double *xy_p, const int64_t *dn_ux, const int64_t *dn_uy,
                      const int64_t *de_ux, const int64_t *de_uy, const int64_t *dn_p, const int64_t *de_p,
                      double mu, int64_t *I, int64_t *J, double *V)
{
    const int ok = (vfe == ORC_T6 && pfe == ORC_T3) || (vfe == ORC_T3B && pfe == ORC_T3) ||
                   (vfe == ORC_Q4 && pfe == ORC_Q4L2);
    if (!ok) return -1;
    if (vfe == ORC_T3B && (!de_ux || !de_uy)) return -1;
    if (pfe == ORC_Q4L2 ? !de_p : (!conn_p || !xy_p || !dn_p)) return -1;
    ftab_t tv, tp, tg;
    const int gfe = pfe == ORC_T3 ? ORC_T3 : vfe;   /* geometry element */
    if (make_ftab(vfe, rule, &tv) || make_ftab(pfe, rule, &tp) || make_ftab(gfe, rule, &tg)) return -2;
    const int nu = tv.nbf, np = tp.nbf, nv = orc_fam_ngeo(vfe), ng = orc_fam_ngeo(gfe);
    int64_t o = 0;
    for (int64_t e = 0; e < E; e++) {
        double X[12];
        if (pfe == ORC_T3) gather_x(conn_p, 3, e, 2, xy_p, X);
        else gather_x(conn_v, nv, e, 2, xy_v, X);
        double kxx[36] = {0}, kyy[36] = {0}, kxy[36] = {0}, kxp[18] = {0}, kyp[18] = {0};
        for (int q = 0; q < tv.nq; q++) {
            jacn_t Jc;
            jacn(2, ng, X, tg.G[q], &Jc);
            const double JxW = Jc.det * tv.w[q];
            double g[16];
            gradn(&Jc, nu, tv.G[q], g);
            const double *Np = tp.N[q];
            for (int j = 0; j < nu; j++)
                for (int i = 0; i < nu; i++) {
                    const double gxi = g[2 * i], gyi = g[2 * i + 1], gxj = g[2 * j], gyj = g[2 * j + 1];
                    kxx[i + nu * j] += (mu * JxW) * (2 * gxi * gxj + gyi * gyj);
                    kyy[i + nu * j] += (mu * JxW) * (gxi * gxj + 2 * gyi * gyj);
                    kxy[i + nu * j] += (mu * JxW) * (gxi * gyj);
                }
            for (int j = 0; j < np; j++)
                for (int i = 0; i < nu; i++) {
                    kxp[i + nu * j] += (-JxW) * (g[2 * i] * Np[j]);
                    kyp[i + nu * j] += (-JxW) * (g[2 * i + 1] * Np[j]);
                }
        }
        int64_t dx[6], dy[6], dp[3];
        for (int k = 0; k < nv; k++) {
            const int64_t nd = conn_v[(int64_t)nv * e + k] - 1;
            dx[k] = dn_ux[nd]; dy[k] = dn_uy[nd];
        }
        if (vfe == ORC_T3B) { dx[3] = de_ux[e]; dy[3] = de_uy[e]; }
        if (pfe == ORC_T3)
            for (int k = 0; k < 3; k++) dp[k] = dn_p[conn_p[3 * e + k] - 1];
        else
            dp[0] = de_p[e];
#define EMIT(M, R, C, NR, NC)                                                    \
        for (int j = 0; j < NC; j++)                                             \
            for (int i = 0; i < NR; i++) { I[o] = R[i]; J[o] = C[j]; V[o] = M[i + NR * j]; o++; }
#define EMITT(M, R, C, NR, NC)                                                   \
        for (int b = 0; b < NR; b++)                                             \
            for (int a = 0; a < NC; a++) { I[o] = C[a]; J[o] = R[b]; V[o] = M[b + NR * a]; o++; }
        EMIT(kxx, dx, dx, nu, nu);
        EMIT(kxy, dx, dy, nu, nu);
        EMITT(kxy, dx, dy, nu, nu);
        EMIT(kyy, dy, dy, nu, nu);
        EMIT(kxp, dx, dp, nu, np);
        EMITT(kxp, dx, dp, nu, np);
        EMIT(kyp, dy, dp, nu, np);
        EMITT(kyp, dy, dp, nu, np);
#undef EMIT
#undef EMITT
    }
    return 0;
}

/* ---- structured meshes for the families (stand-ins for MeshSteward) --------
 * L2block(L, n): nodes x_i = L i / n, elements (i, i+1).
 * T4block(A, B, C, nA, nB, nC): nodes x-fastest, then y, then z; each cell
 * (x-index fastest, then y, then z) split into six tetrahedra around the
 * diagonal from its corner (i, j, k) to (i+1, j+1, k+1), all positively
 * oriented.  MeshSteward's own T4block is not in the reference tree, so the
 * element order and split are this restatement's; the example's check of the
 * exact nodal temperature (examples/heat/poisson/t4.jl:82-90) holds on it. */
int64_t orc_l2block(double L, int n, int64_t *conn, double *x)
{
    for (int i = 0; i <= n; i++) x[i] = L * (double)i / (double)n;
    for (int i = 0; i < n; i++) { conn[2 * i] = i + 1; conn[2 * i + 1] = i + 2; }
    return n;
}

int64_t orc_t4block(double A, double B, double C, int nA, int nB, int nC, int64_t *conn, double *xyz)
{
    const int64_t W = nA + 1, S = (int64_t)(nA + 1) * (nB + 1);
    for (int k = 0; k <= nC; k++)
        for (int j = 0; j <= nB; j++)
            for (int i = 0; i <= nA; i++) {
                const int64_t n = k * S + j * W + i;
                xyz[3 * n] = A * (double)i / (double)nA;
                xyz[3 * n + 1] = B * (double)j / (double)nB;
                xyz[3 * n + 2] = C * (double)k / (double)nC;
            }
    /* corners of the unit cube, bit 0 = x, bit 1 = y, bit 2 = z; the six
     * monotone paths 0 -> 7 give the tetrahedra (0, a, a|b, 7) */
    static const int T[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
    int64_t e = 0;
    for (int k = 0; k < nC; k++)
        for (int j = 0; j < nB; j++)
            for (int i = 0; i < nA; i++) {
                int64_t v[8];
                for (int b = 0; b < 8; b++) v[b] = (k + ((b >> 2) & 1)) * S + (j + ((b >> 1) & 1)) * W + (i + (b & 1)) + 1;
                for (int t = 0; t < 6; t++) {
                    int64_t *c = conn + 4 * e;
                    c[0] = v[T[t][0]]; c[1] = v[T[t][1]]; c[2] = v[T[t][2]]; c[3] = v[T[t][3]];
                    /* orientation: (x1-x0) . ((x2-x0) x (x3-x0)) > 0, else swap the last two */
                    const int a1 = T[t][1], a2 = T[t][2];
                    const int dx1 = a1 & 1, dy1 = (a1 >> 1) & 1, dz1 = (a1 >> 2) & 1;
                    const int dx2 = a2 & 1, dy2 = (a2 >> 1) & 1, dz2 = (a2 >> 2) & 1;
                    const int det = dx1 * (dy2 * 1 - dz2 * 1) - dy1 * (dx2 * 1 - dz2 * 1) + dz1 * (dx2 * 1 - dy2 * 1);
                    if (det < 0) { const int64_t s = c[2]; c[2] = c[3]; c[3] = s; }
                    e++;
                }
            }
    return e;
}
