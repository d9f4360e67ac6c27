/*
 * elfel_oracle.c -- CPU restatement of Elfel.jl's element-integration and
 * global-assembly path (FP64, built with -ffp-contract=off so no product is
 * fused into an FMA, matching Julia's unfused IEEE arithmetic).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline ("port") timed beside it.  Never called by the product path.
 *
 * Each function names the reference lines it restates.  Pins (what checks
 * each function against the reference's own tests or against mathematics)
 * are listed in tests/test_oracle_pins.py.  Functions whose exact rounding
 * depends on code absent from the reference tree (StaticArrays' 2x2 solve and
 * 3x3 mat-vec, Julia's SparseArrays.sparse summation order) follow the
 * published algorithm and are pinned only through end-to-end golden values;
 * the tests say so where they apply.
 */
#include "elfel_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Quadrature: src/RefShapes.jl:85-110 (_gauss1), :113-230 (_triangle),      */
/* :333-366 (square tensor rule, i outer / j inner).  Literal constants are  */
/* the reference's own (15 significant digits for Gauss order 2).            */
/* ------------------------------------------------------------------------- */
int orc_gauss1(int order, double *pc, double *w)
{
    switch (order) {
    case 1: pc[0] = 0.0; w[0] = 2.0; return 1;
    case 2:
        pc[0] = -0.577350269189626; pc[1] = 0.577350269189626;
        w[0] = 1.0; w[1] = 1.0; return 2;
    case 3:
        pc[0] = -0.774596669241483; pc[1] = 0.0; pc[2] = 0.774596669241483;
        w[0] = 0.5555555555555556; w[1] = 0.8888888888888889; w[2] = 0.5555555555555556;
        return 3;
    case 4:
        pc[0] = -0.86113631159405; pc[1] = -0.33998104358486;
        pc[2] = 0.33998104358486; pc[3] = 0.86113631159405;
        w[0] = 0.34785484513745; w[1] = 0.65214515486255;
        w[2] = 0.65214515486255; w[3] = 0.34785484513745;
        return 4;
    case 5:
        pc[0] = -0.906179845938664; pc[1] = -0.538469310105683; pc[2] = 0.0;
        pc[3] = 0.538469310105683; pc[4] = 0.906179845938664;
        w[0] = 0.236926885056189; w[1] = 0.478628670499367; w[2] = 0.568888888888889;
        w[3] = 0.478628670499367; w[4] = 0.236926885056189;
        return 5;
    default: return -1; /* higher orders use an eigen-solver in the reference */
    }
}

int orc_quad_triangle(int npts, double *pc, double *w)
{
    if (npts == 1) {
        pc[0] = 1.0 / 3.; pc[1] = 1.0 / 3.;
        w[0] = 1.0 / 2.0;
        return 1;
    }
    if (npts == 3) {
        pc[0] = 2.0 / 3; pc[1] = 1.0 / 6;
        pc[2] = 1.0 / 6; pc[3] = 2.0 / 3;
        pc[4] = 1.0 / 6; pc[5] = 1.0 / 6;
        for (int i = 0; i < 3; i++) w[i] = (1.0 / 3) / 2;
        return 3;
    }
    if (npts == 6) {
        static const double a[6][2] = {{0.816847572980459, 0.091576213509771},
                                       {0.091576213509771, 0.816847572980459},
                                       {0.091576213509771, 0.091576213509771},
                                       {0.108103018168070, 0.445948490915965},
                                       {0.445948490915965, 0.108103018168070},
                                       {0.445948490915965, 0.445948490915965}};
        for (int i = 0; i < 6; i++) {
            pc[2 * i] = a[i][0]; pc[2 * i + 1] = a[i][1];
            w[i] = (i < 3 ? 0.109951743655322 : 0.223381589678011) / 2;
        }
        return 6;
    }
    return -1;
}

int orc_quad_square(int order, double *pc, double *w)
{
    double p1[8], w1[8];
    int n = orc_gauss1(order, p1, w1);
    if (n < 0) return -1;
    int r = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            pc[2 * r] = p1[i]; pc[2 * r + 1] = p1[j];
            w[r] = w1[i] * w1[j];
            r++;
        }
    return r;
}

/* ------------------------------------------------------------------------- */
/* Basis functions: T3 src/FElements.jl:239-246, T6 :264-288, Q4 :306-320.   */
/* ------------------------------------------------------------------------- */
int orc_nen(int fe) { return fe == ORC_T3 ? 3 : (fe == ORC_T6 ? 6 : (fe == ORC_Q4 ? 4 : -1)); }

void orc_bfun(int fe, const double *pc, double *N)
{
    double r = pc[0], s = pc[1];
    if (fe == ORC_T3) {
        N[0] = (1 - r - s); N[1] = r; N[2] = s;
    } else if (fe == ORC_T6) {
        double t = 1. - r - s;
        N[0] = t * (t + t - 1);
        N[1] = r * (r + r - 1);
        N[2] = s * (s + s - 1);
        N[3] = 4 * r * t;
        N[4] = 4 * r * s;
        N[5] = 4 * s * t;
    } else {
        N[0] = 0.25 * (1. - r) * (1. - s);
        N[1] = 0.25 * (1. + r) * (1. - s);
        N[2] = 0.25 * (1. + r) * (1. + s);
        N[3] = 0.25 * (1. - r) * (1. + s);
    }
}

void orc_bfungradpar(int fe, const double *pc, double *g)
{
    double r = pc[0], s = pc[1];
    if (fe == ORC_T3) {
        g[0] = -1.; g[1] = -1.; g[2] = +1.; g[3] = 0.; g[4] = 0.; g[5] = +1.;
    } else if (fe == ORC_T6) {
        g[0] = -3 + 4 * r + 4 * s; g[1] = -3 + 4 * r + 4 * s;
        g[2] = 4 * r - 1;          g[3] = 0.0;
        g[4] = 0.0;                g[5] = 4 * s - 1;
        g[6] = 4 - 8 * r - 4 * s;  g[7] = -4 * r;
        g[8] = 4 * s;              g[9] = 4 * r;
        g[10] = -4 * s;            g[11] = 4 - 4 * r - 8 * s;
    } else {
        g[0] = -(1. - s) * 0.25; g[1] = -(1. - r) * 0.25;
        g[2] = (1. - s) * 0.25;  g[3] = -(1. + r) * 0.25;
        g[4] = (1. + s) * 0.25;  g[5] = (1. + r) * 0.25;
        g[6] = -(1. + s) * 0.25; g[7] = (1. - r) * 0.25;
    }
}

/* ------------------------------------------------------------------------- */
/* Structured meshes.  MeshSteward is not in the reference tree; its T3 node  */
/* and element order is pinned by test/qmesh-conn.dat (x-fastest nodes,      */
/* cells x-index outer / y-index inner, diagonal (i,j)-(i+1,j+1)).  Q4 uses   */
/* the same cell order.  T6 midside nodes follow the vertices and are         */
/* numbered by first appearance (element order, local edges 1-2, 2-3, 3-1);   */
/* that numbering is this build's own and is shared with the host package.    */
/* ------------------------------------------------------------------------- */
static double lin(double L, int n, int i) { return L * (double)i / (double)n; }

int64_t orc_t3block(double A, double B, int nA, int nB, int64_t *conn, double *xy)
{
    int64_t Nn = (int64_t)(nA + 1) * (nB + 1);
    if (xy) {
        int64_t f = 0;
        for (int j = 0; j <= nB; j++)
            for (int i = 0; i <= nA; i++) {
                xy[2 * f] = lin(A, nA, i); xy[2 * f + 1] = lin(B, nB, j); f++;
            }
    }
    if (conn) {
        int64_t gc = 0;
        for (int i = 0; i < nA; i++)
            for (int j = 0; j < nB; j++) {
                int64_t f = (int64_t)j * (nA + 1) + i + 1;
                conn[3 * gc] = f; conn[3 * gc + 1] = f + 1; conn[3 * gc + 2] = f + (nA + 1) + 1; gc++;
                conn[3 * gc] = f; conn[3 * gc + 1] = f + (nA + 1) + 1; conn[3 * gc + 2] = f + (nA + 1); gc++;
            }
    }
    return Nn;
}

int64_t orc_q4block(double A, double B, int nA, int nB, int64_t *conn, double *xy)
{
    int64_t Nn = orc_t3block(A, B, nA, nB, NULL, xy);
    if (conn) {
        int64_t gc = 0;
        for (int i = 0; i < nA; i++)
            for (int j = 0; j < nB; j++) {
                int64_t f = (int64_t)j * (nA + 1) + i + 1;
                conn[4 * gc] = f; conn[4 * gc + 1] = f + 1;
                conn[4 * gc + 2] = f + (nA + 1) + 1; conn[4 * gc + 3] = f + (nA + 1);
                gc++;
            }
    }
    return Nn;
}

/* T6 block: T3 block plus midside nodes numbered by first appearance.       */
/* Edge lookup uses the structured layout: every edge of the T3 block is      */
/* horizontal, vertical or a (i,j)-(i+1,j+1) diagonal.                        */
int64_t orc_t6block(double A, double B, int nA, int nB, int64_t *conn, double *xy)
{
    int64_t Nv = (int64_t)(nA + 1) * (nB + 1);
    int64_t E = 2 * (int64_t)nA * nB;
    int64_t nh = (int64_t)nA * (nB + 1), nv = (int64_t)(nA + 1) * nB, nd = (int64_t)nA * nB;
    int64_t *edgenum = (int64_t *)malloc(sizeof(int64_t) * (nh + nv + nd));
    for (int64_t k = 0; k < nh + nv + nd; k++) edgenum[k] = 0;
    int64_t *c3 = (int64_t *)malloc(sizeof(int64_t) * 3 * E);
    double *xyv = (double *)malloc(sizeof(double) * 2 * Nv);
    orc_t3block(A, B, nA, nB, c3, xyv);
    int64_t next = Nv;
    int64_t Nn = Nv + nh + nv + nd;
    if (xy) memcpy(xy, xyv, sizeof(double) * 2 * Nv);
    for (int64_t e = 0; e < E; e++) {
        int64_t v[3] = {c3[3 * e] - 1, c3[3 * e + 1] - 1, c3[3 * e + 2] - 1};
        if (conn) for (int k = 0; k < 3; k++) conn[6 * e + k] = v[k] + 1;
        for (int k = 0; k < 3; k++) {
            int64_t a = v[k], b = v[(k + 1) % 3];
            int64_t lo = a < b ? a : b, hi = a < b ? b : a;
            int64_t ilo = lo % (nA + 1), jlo = lo / (nA + 1);
            int64_t ihi = hi % (nA + 1), jhi = hi / (nA + 1);
            int64_t key;
            if (jlo == jhi) key = jlo * nA + ilo;                          /* horizontal */
            else if (ilo == ihi) key = nh + jlo * (nA + 1) + ilo;          /* vertical */
            else key = nh + nv + jlo * nA + ilo;                           /* diagonal */
            if (edgenum[key] == 0) {
                edgenum[key] = ++next; /* 1-based node number */
                if (xy) {
                    xy[2 * (next - 1)] = (xyv[2 * a] + xyv[2 * b]) / 2;
                    xy[2 * (next - 1) + 1] = (xyv[2 * a + 1] + xyv[2 * b + 1]) / 2;
                }
            }
            if (conn) conn[6 * e + 3 + k] = edgenum[key];
        }
    }
    free(edgenum); free(c3); free(xyv);
    (void)Nn;
    return next;
}

/* T6 -> T3 on the corner nodes, renumbered compactly in increasing order of */
/* the T6 node number (the pressure mesh of test/test_stokes.jl:32,37).      */
int64_t orc_t6tot3(int64_t E, const int64_t *conn6, int64_t Nn6, const double *xy6,
                   int64_t *conn3, double *xy3)
{
    int64_t *map = (int64_t *)calloc((size_t)Nn6, sizeof(int64_t));
    for (int64_t e = 0; e < E; e++)
        for (int k = 0; k < 3; k++) map[conn6[6 * e + k] - 1] = 1;
    int64_t n = 0;
    for (int64_t i = 0; i < Nn6; i++)
        if (map[i]) {
            map[i] = ++n;
            if (xy3) { xy3[2 * (n - 1)] = xy6[2 * i]; xy3[2 * (n - 1) + 1] = xy6[2 * i + 1]; }
        }
    if (conn3)
        for (int64_t e = 0; e < E; e++)
            for (int k = 0; k < 3; k++) conn3[3 * e + k] = map[conn6[6 * e + k] - 1];
    free(map);
    return n;
}

/* ------------------------------------------------------------------------- */
/* Dof numbering.  numberdofs!(spaces) (src/FESpaces.jl:250-259): free dofs  */
/* of each space in turn (numberfreedofs!, src/FEFields.jl:137-154: term-     */
/* major, component-minor, data dofs zeroed), then data dofs of each space   */
/* (numberdatadofs!, :164-177).  Each space starts one past the highest      */
/* number of the previous space (highestfreedofnum / highestdatadofnum).     */
/* ------------------------------------------------------------------------- */
int orc_numberdofs(int nspaces, const int64_t *nterms, const int32_t *ncomp,
                   const uint8_t *isdatum, int64_t *dofnums)
{
    int64_t off = 0, hi_free_prev = 0;
    int64_t *base = (int64_t *)malloc(sizeof(int64_t) * (nspaces + 1));
    for (int s = 0; s < nspaces; s++) { base[s] = off; off += nterms[s] * ncomp[s]; }
    base[nspaces] = off;
    /* free dofs */
    for (int s = 0; s < nspaces; s++) {
        int64_t num = (s == 0) ? 1 : hi_free_prev + 1;
        int64_t hi = 0;
        for (int64_t k = base[s]; k < base[s + 1]; k++) {
            if (!isdatum[k]) { dofnums[k] = num; if (num > hi) hi = num; num++; }
            else dofnums[k] = 0;
        }
        hi_free_prev = hi;
    }
    /* data dofs */
    int64_t hi_data_prev = 0;
    for (int s = 0; s < nspaces; s++) {
        int64_t num = (s == 0) ? hi_free_prev + 1 : hi_data_prev + 1;
        int64_t hi = 0;
        for (int64_t k = base[s]; k < base[s + 1]; k++)
            if (isdatum[k]) { dofnums[k] = num; if (num > hi) hi = num; num++; }
        hi_data_prev = hi;
    }
    free(base);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Jacobian and spatial gradients.                                            */
/* jacjac: J = sum_k x_k (outer) gradNpar_k, first term assigned then added   */
/* (src/FElements.jl:148-156,165-168); det = J11*J22 - J21*J12 (:120-129).    */
/* bfungrad: gradNpar / J, i.e. the 2x2 solve J' \ g by Cramer's rule with    */
/* the same determinant (src/QPIterators.jl:132-140, StaticArrays' 2x2 solve).*/
/* ------------------------------------------------------------------------- */
typedef struct { double j11, j12, j21, j22, det; } jac_t;

static jac_t jacjac(int nen, const double *X, const double *gp /* nen*2 */)
{
    jac_t J;
    J.j11 = X[0] * gp[0]; J.j21 = X[1] * gp[0];
    J.j12 = X[0] * gp[1]; J.j22 = X[1] * gp[1];
    for (int k = 1; k < nen; k++) {
        J.j11 += X[2 * k] * gp[2 * k];
        J.j21 += X[2 * k + 1] * gp[2 * k];
        J.j12 += X[2 * k] * gp[2 * k + 1];
        J.j22 += X[2 * k + 1] * gp[2 * k + 1];
    }
    J.det = J.j11 * J.j22 - J.j21 * J.j12;
    return J;
}

static void bfungrad(const jac_t *J, int nbf, const double *gp, double *gx, double *gy)
{
    for (int k = 0; k < nbf; k++) {
        double a = gp[2 * k], b = gp[2 * k + 1];
        gx[k] = (J->j22 * a - J->j21 * b) / J->det;
        gy[k] = (J->j11 * b - J->j12 * a) / J->det;
    }
}

/* Quadrature + tables for an element/rule pair (src/QPIterators.jl:14-50). */
typedef struct {
    int npts, nbf;
    double w[16];
    double N[16][8];      /* N[q][k] */
    double G[16][16];     /* G[q][2k+d] parametric gradients */
} qtab_t;

static int make_qtab(int fe, int rule, qtab_t *t)
{
    double pc[32];
    int n = (fe == ORC_Q4) ? orc_quad_square(rule, pc, t->w) : orc_quad_triangle(rule, pc, t->w);
    if (n < 0 || n > 16) return -1;
    t->npts = n;
    t->nbf = orc_nen(fe);
    for (int q = 0; q < n; q++) {
        orc_bfun(fe, pc + 2 * q, t->N[q]);
        orc_bfungradpar(fe, pc + 2 * q, t->G[q]);
    }
    return 0;
}

static void gather_xy(const int64_t *conn, int nen, int64_t e, const double *xy, double *X)
{
    for (int k = 0; k < nen; k++) {
        int64_t n = conn[(int64_t)nen * e + k] - 1;
        X[2 * k] = xy[2 * n]; X[2 * k + 1] = xy[2 * n + 1];
    }
}

/* B(g, c) of test/test_elasticity.jl:151: c=0 -> (g1, 0, g2), c=1 -> (0, g2, g1). */
static void Bvec(double gx, double gy, int c, double *b)
{
    if (c == 0) { b[0] = gx; b[1] = 0.0; b[2] = gy; }
    else { b[0] = 0.0; b[1] = gy; b[2] = gx; }
}

/* D * b with D column-major: row r = D[r,1]*b1 + D[r,2]*b2 + D[r,3]*b3. */
static void Dmul(const double *D, const double *b, double *o)
{
    for (int r = 0; r < 3; r++) o[r] = D[r] * b[0] + D[r + 3] * b[1] + D[r + 6] * b[2];
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ------------------------------------------------------------------------- */
/* Per-element matrices.                                                     */
/* heat:   test/test_heat.jl:36-60 (also examples/heat/poisson/q4.jl:31-54)   */
/* elast:  test/test_elasticity.jl:150-172                                    */
/* stokes: examples/stokes/colliding_flow/ht_p2_p1_gen.jl:47-81 (gen),        */
/*         test/test_stokes.jl:630-659 (veclap)                               */
/* ------------------------------------------------------------------------- */
#define MAXDOF 12

static int nedof_of(const orc_problem *p) { return orc_nen(p->fe) * p->ncomp; }

static int64_t block_entries(const orc_problem *p)
{
    int n = nedof_of(p);
    if (p->kind == ORC_STOKES_GEN || p->kind == ORC_STOKES_VECLAP) return (int64_t)n * n + 2 * n * 3;
    return (int64_t)n * n;
}

static void element_compute(const orc_problem *p, const qtab_t *tu, const qtab_t *tp, int64_t e,
                            double *ke /* n*n col-major */, double *fe, double *kup /* n*3 */)
{
    int nen = p->nen, nc = p->ncomp, n = nen * nc;
    double X[16];
    gather_xy(p->conn, nen, e, p->xy, X);
    memset(ke, 0, sizeof(double) * n * n);
    if (fe) memset(fe, 0, sizeof(double) * n);
    if (kup) memset(kup, 0, sizeof(double) * n * 3);
    double kap = p->kappa_e ? p->kappa_e[e] : p->kappa;
    double Qe = p->Q_e ? p->Q_e[e] : p->Q;
    double D[9];
    double sc = p->scale_e ? p->scale_e[e] : 1.0;
    for (int i = 0; i < 9; i++) D[i] = p->scale_e ? sc * p->D[i] : p->D[i];
    for (int q = 0; q < tu->npts; q++) {
        jac_t J = jacjac(nen, X, tu->G[q]);
        double sgx[8], sgy[8];
        bfungrad(&J, nen, tu->G[q], sgx, sgy);
        double JxW = J.det * tu->w[q];
        if (p->kind == ORC_HEAT) {
            for (int j = 0; j < n; j++) {
                for (int i = 0; i < n; i++)
                    ke[i + n * j] += (sgx[i] * sgx[j] + sgy[i] * sgy[j]) * (kap * JxW);
                if (fe) fe[j] += tu->N[q][j] * Qe * JxW;
            }
        } else if (p->kind == ORC_ELAST || p->kind == ORC_STOKES_GEN) {
            for (int j = 0; j < n; j++) {
                double Bj[3], DBj[3];
                Bvec(sgx[j / nc], sgy[j / nc], j % nc, Bj);
                Dmul(D, Bj, DBj);
                for (int i = 0; i < n; i++) {
                    double Bi[3];
                    Bvec(sgx[i / nc], sgy[i / nc], i % nc, Bi);
                    ke[i + n * j] += dot3(DBj, Bi) * JxW;
                }
            }
        } else { /* veclap */
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++)
                    if ((i % nc) == (j % nc))
                        ke[i + n * j] += (p->mu * JxW) * (sgx[i / nc] * sgx[j / nc] + sgy[i / nc] * sgy[j / nc]);
        }
        if (kup) {
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < n; i++) {
                    double g = (i % nc) == 0 ? sgx[i / nc] : sgy[i / nc];
                    kup[i + n * j] += (-JxW * tp->N[q][j]) * g;
                }
        }
    }
}

int64_t orc_coo_len(const orc_problem *p, int64_t *nvec)
{
    if (nvec) *nvec = (p->kind == ORC_HEAT) ? p->E * nedof_of(p) : 0;
    return p->E * block_entries(p);
}

static int check_problem(const orc_problem *p)
{
    if (orc_nen(p->fe) != p->nen) return -1;
    if (nedof_of(p) > MAXDOF) return -1;
    if (p->kind == ORC_STOKES_GEN || p->kind == ORC_STOKES_VECLAP)
        if (p->fe != ORC_T6 || p->ncomp != 2 || !p->conn_p || !p->dofnums_p) return -1;
    return 0;
}

/* Element dofs: entity-major, component-minor (src/FEIterators.jl:161-171). */
static void eldofs(const orc_problem *p, int64_t e, int64_t *d)
{
    for (int k = 0; k < p->nen; k++) {
        int64_t node = p->conn[(int64_t)p->nen * e + k] - 1;
        for (int c = 0; c < p->ncomp; c++) d[k * p->ncomp + c] = p->dofnums[node * p->ncomp + c];
    }
}

/* One element's COO entries in Elfel stream order: LocalMatrixAssembler is  */
/* appended column-major (src/LocalAssemblers.jl:68-83, Assemblers.jl:97-102),*/
/* a transposed block row-major of the parent (Assemblers.jl:109-114).       */
static void emit_element(const orc_problem *p, const qtab_t *tu, const qtab_t *tp, int64_t e,
                         int64_t *I, int64_t *J, double *V, int64_t *FI, double *FV)
{
    int n = nedof_of(p);
    double ke[MAXDOF * MAXDOF], fe[MAXDOF], kup[MAXDOF * 3];
    int stokes = (p->kind == ORC_STOKES_GEN || p->kind == ORC_STOKES_VECLAP);
    element_compute(p, tu, tp, e, ke, p->kind == ORC_HEAT ? fe : NULL, stokes ? kup : NULL);
    int64_t d[MAXDOF];
    eldofs(p, e, d);
    int64_t o = e * block_entries(p);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) { I[o] = d[i]; J[o] = d[j]; V[o] = ke[i + n * j]; o++; }
    if (stokes) {
        int64_t dp[3];
        for (int k = 0; k < 3; k++) dp[k] = p->dofnums_p[p->conn_p[3 * e + k] - 1];
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < n; i++) { I[o] = d[i]; J[o] = dp[j]; V[o] = kup[i + n * j]; o++; }
        for (int b = 0; b < n; b++)
            for (int a = 0; a < 3; a++) { I[o] = dp[a]; J[o] = d[b]; V[o] = kup[b + n * a]; o++; }
    }
    if (p->kind == ORC_HEAT && FI) {
        int64_t fo = e * n;
        for (int j = 0; j < n; j++) { FI[fo + j] = d[j]; FV[fo + j] = fe[j]; }
    }
}

int orc_coo(const orc_problem *p, int64_t *I, int64_t *J, double *V, int64_t *FI, double *FV)
{
    return orc_coo_parallel(p, I, J, V, FI, FV, 1);
}

int orc_coo_parallel(const orc_problem *p, int64_t *I, int64_t *J, double *V,
                     int64_t *FI, double *FV, int nthreads)
{
    if (check_problem(p)) return -1;
    qtab_t tu, tp;
    if (make_qtab(p->fe, p->rule, &tu)) return -2;
    if (p->kind == ORC_STOKES_GEN || p->kind == ORC_STOKES_VECLAP)
        if (make_qtab(ORC_T3, p->rule, &tp)) return -2;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t e = 0; e < p->E; e++) emit_element(p, &tu, &tp, e, I, J, V, FI, FV);
    (void)nthreads;
    return 0;
}

/* Reddy form over three scalar spaces, Jacobian from the pressure (T3)      */
/* element: test/test_stokes.jl:373-431.                                      */
int64_t orc_stokes_reddy_coo_len(int64_t E) { return E * (36 * 4 + 18 * 4); }

int orc_stokes_reddy_coo(int64_t E, const int64_t *conn6, const double *xy6,
                         const int64_t *conn3, const double *xy3,
                         const int64_t *dn_ux, const int64_t *dn_uy, const int64_t *dn_p,
                         double mu, int64_t *I, int64_t *J, double *V)
{
    qtab_t t6, t3;
    if (make_qtab(ORC_T6, 3, &t6) || make_qtab(ORC_T3, 3, &t3)) return -2;
    (void)xy6;
    int64_t o = 0;
    for (int64_t e = 0; e < E; e++) {
        double X[6];
        gather_xy(conn3, 3, e, xy3, X);
        double kxx[36] = {0}, kyy[36] = {0}, kxy[36] = {0}, kxp[18] = {0}, kyp[18] = {0};
        for (int q = 0; q < 3; q++) {
            jac_t Jc = jacjac(3, X, t3.G[q]);
            double JxW = Jc.det * t3.w[q];
            double gx[6], gy[6];
            bfungrad(&Jc, 6, t6.G[q], gx, gy);
            const double *Np = t3.N[q];
            for (int j = 0; j < 6; j++)
                for (int i = 0; i < 6; i++) {
                    kxx[i + 6 * j] += (mu * JxW) * (2 * gx[i] * gx[j] + gy[i] * gy[j]);
                    kyy[i + 6 * j] += (mu * JxW) * (gx[i] * gx[j] + 2 * gy[i] * gy[j]);
                    kxy[i + 6 * j] += (mu * JxW) * (gx[i] * gy[j]);
                }
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < 6; i++) {
                    kxp[i + 6 * j] += (-JxW) * (gx[i] * Np[j]);
                    kyp[i + 6 * j] += (-JxW) * (gy[i] * Np[j]);
                }
        }
        int64_t dx[6], dy[6], dp[3];
        for (int k = 0; k < 6; k++) {
            int64_t nd = conn6[6 * e + k] - 1;
            dx[k] = dn_ux[nd]; dy[k] = dn_uy[nd];
        }
        for (int k = 0; k < 3; k++) dp[k] = dn_p[conn3[3 * e + k] - 1];
#define EMIT(M, R, C, NR, NC)                                                    \
        for (int j = 0; j < NC; j++)                                             \
            for (int i = 0; i < NR; i++) { I[o] = R[i]; J[o] = C[j]; V[o] = M[i + NR * j]; o++; }
#define EMITT(M, R, C, NR, NC)                                                   \
        for (int b = 0; b < NR; b++)                                             \
            for (int a = 0; a < NC; a++) { I[o] = C[a]; J[o] = R[b]; V[o] = M[b + NR * a]; o++; }
        EMIT(kxx, dx, dx, 6, 6);
        EMIT(kxy, dx, dy, 6, 6);
        EMITT(kxy, dx, dy, 6, 6);
        EMIT(kyy, dy, dy, 6, 6);
        EMIT(kxp, dx, dp, 6, 3);
        EMITT(kxp, dx, dp, 6, 3);
        EMIT(kyp, dy, dp, 6, 3);
        EMITT(kyp, dy, dp, 6, 3);
#undef EMIT
#undef EMITT
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* sparse(I, J, V, nrow, ncol) (src/Assemblers.jl:121-123): duplicates are   */
/* combined with + in stream order (left fold: first value, then each later  */
/* duplicate added), explicit zeros stay structural, indices sorted within   */
/* each column (csc=1) or row (csc=0).  Algorithm: stable counting sort by    */
/* the major index, first-appearance dedup with a marker array, then an      */
/* insertion sort of each (short) segment by the minor index.  Out-of-range  */
/* indices return -1 (sparse throws ArgumentError).  ptr/ind/val must hold    */
/* n (>= nnz) entries; returns nnz.                                           */
/* ------------------------------------------------------------------------- */
int64_t orc_sparse(int64_t n, const int64_t *I, const int64_t *J, const double *V,
                   int64_t nrow, int64_t ncol, int csc,
                   int64_t *ptr, int64_t *ind, double *val)
{
    const int64_t *maj = csc ? J : I, *mnr = csc ? I : J;
    int64_t nmaj = csc ? ncol : nrow, nmnr = csc ? nrow : ncol;
    for (int64_t k = 0; k < n; k++)
        if (I[k] < 1 || I[k] > nrow || J[k] < 1 || J[k] > ncol) return -1;
    int64_t *cnt = (int64_t *)calloc((size_t)nmaj + 1, sizeof(int64_t));
    for (int64_t k = 0; k < n; k++) cnt[maj[k]]++;
    for (int64_t m = 0; m < nmaj; m++) cnt[m + 1] += cnt[m];
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nmaj + 1));
    memcpy(fill, cnt, sizeof(int64_t) * (size_t)(nmaj + 1));
    for (int64_t k = 0; k < n; k++) perm[fill[maj[k] - 1]++] = k;
    int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nmnr + 1));
    for (int64_t i = 0; i <= nmnr; i++) mark[i] = -1;
    int64_t nnz = 0;
    ptr[0] = 0;
    for (int64_t m = 0; m < nmaj; m++) {
        int64_t start = nnz;
        for (int64_t t = cnt[m]; t < cnt[m + 1]; t++) {
            int64_t k = perm[t];
            int64_t r = mnr[k];
            if (mark[r] >= start) val[mark[r]] += V[k];
            else { mark[r] = nnz; ind[nnz] = r - 1; val[nnz] = V[k]; nnz++; }
        }
        for (int64_t a = start + 1; a < nnz; a++) {
            int64_t ki = ind[a]; double kv = val[a]; int64_t b = a - 1;
            while (b >= start && ind[b] > ki) { ind[b + 1] = ind[b]; val[b + 1] = val[b]; b--; }
            ind[b + 1] = ki; val[b + 1] = kv;
        }
        ptr[m + 1] = nnz;
    }
    free(cnt); free(perm); free(fill); free(mark);
    return nnz;
}

/* SysvecAssembler: start! zeros, assemble! adds in stream order, finish!    */
/* copies (src/Assemblers.jl:194-230).  Out-of-range index returns -1.       */
int orc_vec_assemble(int64_t n, const int64_t *FI, const double *FV, int64_t ndofs, double *out)
{
    for (int64_t i = 0; i < ndofs; i++) out[i] = 0.0;
    for (int64_t k = 0; k < n; k++) {
        if (FI[k] < 1 || FI[k] > ndofs) return -1;
        out[FI[k] - 1] = out[FI[k] - 1] + FV[k];
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Post-processing integrals for the colliding-flow manufactured solution    */
/* (test/test_stokes.jl:354-356 exact fields; :438-496 and :675-731 norms).  */
/* location(): sum_k x_k N_k, first term assigned (src/FEIterators.jl:227-235)*/
/* ------------------------------------------------------------------------- */
static double truep(double x, double y) { return 60 * (x * x) * y - 20 * (y * y * y); }
static double trueux(double x, double y) { return 20 * x * (y * y * y); }
static double trueuy(double x, double y) { return 5 * (x * x * x * x) - 5 * (y * y * y * y); }

double orc_colliding_p_error(int64_t E, const int64_t *conn3, const double *xy3,
                             const int64_t *dn_p, const double *U)
{
    qtab_t t;
    make_qtab(ORC_T3, 3, &t);
    double Err = 0.0;
    for (int64_t e = 0; e < E; e++) {
        double X[6], dv[3];
        gather_xy(conn3, 3, e, xy3, X);
        for (int k = 0; k < 3; k++) dv[k] = U[dn_p[conn3[3 * e + k] - 1] - 1];
        for (int q = 0; q < 3; q++) {
            jac_t J = jacjac(3, X, t.G[q]);
            double JxW = J.det * t.w[q];
            double lx = X[0] * t.N[q][0], ly = X[1] * t.N[q][0];
            for (int k = 1; k < 3; k++) { lx += X[2 * k] * t.N[q][k]; ly += X[2 * k + 1] * t.N[q][k]; }
            double pt = truep(lx, ly), pa = 0.0;
            for (int j = 0; j < 3; j++) pa += dv[j] * t.N[q][j];
            Err += JxW * (pa - pt) * (pa - pt);
        }
    }
    return sqrt(Err);
}

static double v_error(int64_t E, const int64_t *conn6, const double *xy6,
                      const int64_t *dx, const int64_t *dy, int stride, const double *U)
{
    qtab_t t;
    make_qtab(ORC_T6, 3, &t);
    double Err = 0.0;
    for (int64_t e = 0; e < E; e++) {
        double X[12], vx[6], vy[6];
        gather_xy(conn6, 6, e, xy6, X);
        for (int k = 0; k < 6; k++) {
            int64_t nd = conn6[6 * e + k] - 1;
            vx[k] = U[dx[stride * nd] - 1]; vy[k] = U[dy[stride * nd] - 1];
        }
        for (int q = 0; q < 3; q++) {
            jac_t J = jacjac(6, X, t.G[q]);
            double JxW = J.det * t.w[q];
            double lx = X[0] * t.N[q][0], ly = X[1] * t.N[q][0];
            for (int k = 1; k < 6; k++) { lx += X[2 * k] * t.N[q][k]; ly += X[2 * k + 1] * t.N[q][k]; }
            double ua = 0.0, va = 0.0;
            for (int j = 0; j < 6; j++) { ua += vx[j] * t.N[q][j]; va += vy[j] * t.N[q][j]; }
            double ux = trueux(lx, ly), uy = trueuy(lx, ly);
            Err += JxW * ((ua - ux) * (ua - ux) + (va - uy) * (va - uy));
        }
    }
    return sqrt(Err);
}

double orc_colliding_v_error_scalar(int64_t E, const int64_t *conn6, const double *xy6,
                                    const int64_t *dn_ux, const int64_t *dn_uy, const double *U)
{
    return v_error(E, conn6, xy6, dn_ux, dn_uy, 1, U);
}

double orc_colliding_v_error_vector(int64_t E, const int64_t *conn6, const double *xy6,
                                    const int64_t *dn_u, const double *U)
{
    return v_error(E, conn6, xy6, dn_u, dn_u + 1, 2, U);
}

/* ------------------------------------------------------------------------- */
/* FP64 Jacobi-PCG on the free block (columns < nfree) of a CSR matrix.  The */
/* reference solves with SuiteSparse `\` (test/test_heat.jl:73-76); this is   */
/* the textbook preconditioned conjugate-gradient method (Saad, Iterative    */
/* Methods for Sparse Linear Systems, Alg. 9.1) used as the FP64 comparator.  */
/* ------------------------------------------------------------------------- */
static void spmv_free(int64_t n, const int64_t *ip, const int64_t *ix, const double *v,
                      const double *x, double *y)
{
    for (int64_t r = 0; r < n; r++) {
        double s = 0.0;
        for (int64_t k = ip[r]; k < ip[r + 1]; k++)
            if (ix[k] < n) s += v[k] * x[ix[k]];
        y[r] = s;
    }
}

int orc_pcg(int64_t n, const int64_t *ip, const int64_t *ix, const double *v,
            const double *b, double *x, double tol, int maxit, double *relres)
{
    double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
    double *pv = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
    double *dinv = (double *)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; i++) {
        double d = 0.0;
        for (int64_t k = ip[i]; k < ip[i + 1]; k++) if (ix[k] == i) d = v[k];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
    spmv_free(n, ip, ix, v, x, q);
    double bn = 0.0, rz = 0.0, rn;
    for (int64_t i = 0; i < n; i++) {
        r[i] = b[i] - q[i]; z[i] = dinv[i] * r[i]; pv[i] = z[i];
        bn += b[i] * b[i]; rz += r[i] * z[i];
    }
    bn = sqrt(bn);
    if (bn == 0.0) bn = 1.0;
    int it = 0;
    for (; it < maxit; it++) {
        rn = 0.0;
        for (int64_t i = 0; i < n; i++) rn += r[i] * r[i];
        if (sqrt(rn) / bn <= tol) break;
        spmv_free(n, ip, ix, v, pv, q);
        double pq = 0.0;
        for (int64_t i = 0; i < n; i++) pq += pv[i] * q[i];
        double alpha = rz / pq, rz2 = 0.0;
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * pv[i]; r[i] -= alpha * q[i]; z[i] = dinv[i] * r[i]; rz2 += r[i] * z[i];
        }
        double beta = rz2 / rz;
        rz = rz2;
        for (int64_t i = 0; i < n; i++) pv[i] = z[i] + beta * pv[i];
    }
    /* true residual */
    spmv_free(n, ip, ix, v, x, q);
    rn = 0.0;
    for (int64_t i = 0; i < n; i++) rn += (b[i] - q[i]) * (b[i] - q[i]);
    if (relres) *relres = sqrt(rn) / bn;
    free(r); free(z); free(pv); free(q); free(dinv);
    return it;
}

/* ------------------------------------------------------------------------- */
/* CPU baseline: element loop (COO in stream order) and sparse() timed       */
/* separately, as the reference's `@time integrate!` split does              */
/* (examples/heat/poisson/t3.jl:72).                                          */
/* ------------------------------------------------------------------------- */
static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_time_assembly(const orc_problem *p, int reps, int nthreads,
                      double *t_elem, double *t_sparse, int64_t *nnz_out)
{
    int64_t nv = 0, n = orc_coo_len(p, &nv);
    int64_t ndofs = 0;
    for (int64_t i = 0; i < p->Nn * p->ncomp; i++) if (p->dofnums[i] > ndofs) ndofs = p->dofnums[i];
    if (p->dofnums_p)
        for (int64_t i = 0; i < p->Np; i++) if (p->dofnums_p[i] > ndofs) ndofs = p->dofnums_p[i];
    int64_t *I = (int64_t *)malloc(sizeof(int64_t) * n), *J = (int64_t *)malloc(sizeof(int64_t) * n);
    double *V = (double *)malloc(sizeof(double) * n);
    int64_t *FI = nv ? (int64_t *)malloc(sizeof(int64_t) * nv) : NULL;
    double *FV = nv ? (double *)malloc(sizeof(double) * nv) : NULL;
    int64_t *ptr = (int64_t *)malloc(sizeof(int64_t) * (ndofs + 1));
    int64_t *ind = (int64_t *)malloc(sizeof(int64_t) * n);
    double *val = (double *)malloc(sizeof(double) * n);
    double *F = (double *)malloc(sizeof(double) * (ndofs + 1));
    double te = 0, ts = 0;
    int rc = 0;
    for (int r = 0; r < reps && !rc; r++) {
        double t0 = now_s();
        rc = orc_coo_parallel(p, I, J, V, FI, FV, nthreads);
        double t1 = now_s();
        int64_t nnz = orc_sparse(n, I, J, V, ndofs, ndofs, 1, ptr, ind, val);
        if (nv) orc_vec_assemble(nv, FI, FV, ndofs, F);
        double t2 = now_s();
        if (nnz < 0) rc = -3;
        if (nnz_out) *nnz_out = nnz;
        te += t1 - t0; ts += t2 - t1;
    }
    *t_elem = te / reps; *t_sparse = ts / reps;
    free(I); free(J); free(V); free(FI); free(FV); free(ptr); free(ind); free(val); free(F);
    return rc;
}
