// qtable.h -- quadrature rules and basis tables, constexpr on host and device.
//
// Rules: src/RefShapes.jl:113-119 (triangle 1- and 3-point), :85-96 and
// :333-366 (Gauss order 1/2, tensor order i outer, j inner).  Bases:
// src/FElements.jl:239-246 (T3), :264-288 (T6), :306-320 (Q4).  The literal
// constants and the expression forms are the reference's, so the FP64 tables
// are bit-identical to what QPIterator tabulates (src/QPIterators.jl:14-50).
// The host builds the tables handed to the FP64 kernels with qt_make(); the
// mixed kernels evaluate the same function at compile time (qt_fixed), so
// their table entries are literals and structural zeros drop out.
#pragma once
#include <hip/hip_runtime.h>

#define ELFEL_MAXQ 4     // quadrature points per element (triangle 1/3, square 1x1/2x2)
#define ELFEL_MAXBF 6    // basis functions of the scalar element (T3 3, Q4 4, T6 6)

// element codes as in elfel_hip.h (ELFEL_T3, ELFEL_T6, ELFEL_Q4)
enum { QT_T3 = 0, QT_T6 = 1, QT_Q4 = 2 };

struct QTable {
    int nq, nbf;
    double w[ELFEL_MAXQ];
    double N[ELFEL_MAXQ][ELFEL_MAXBF];
    double G[ELFEL_MAXQ][ELFEL_MAXBF][2];
    double Np[ELFEL_MAXQ][3];  // T3 basis at the same points (Taylor-Hood pressure)
    float wf[ELFEL_MAXQ];
    float Nf[ELFEL_MAXQ][ELFEL_MAXBF];
    float Gf[ELFEL_MAXQ][ELFEL_MAXBF][2];
    float Npf[ELFEL_MAXQ][3];
};

__host__ __device__ constexpr void qt_bfun(int fe, double r, double s, double *N)
{
    if (fe == QT_T3) {
        N[0] = (1 - r - s); N[1] = r; N[2] = s;
    } else if (fe == QT_T6) {
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

__host__ __device__ constexpr void qt_bgrad(int fe, double r, double s, double (*g)[2])
{
    if (fe == QT_T3) {
        g[0][0] = -1.; g[0][1] = -1.; g[1][0] = +1.; g[1][1] = 0.; g[2][0] = 0.; g[2][1] = +1.;
    } else if (fe == QT_T6) {
        g[0][0] = -3 + 4 * r + 4 * s; g[0][1] = -3 + 4 * r + 4 * s;
        g[1][0] = 4 * r - 1;          g[1][1] = 0.0;
        g[2][0] = 0.0;                g[2][1] = 4 * s - 1;
        g[3][0] = 4 - 8 * r - 4 * s;  g[3][1] = -4 * r;
        g[4][0] = 4 * s;              g[4][1] = 4 * r;
        g[5][0] = -4 * s;             g[5][1] = 4 - 4 * r - 8 * s;
    } else {
        g[0][0] = -(1. - s) * 0.25; g[0][1] = -(1. - r) * 0.25;
        g[1][0] = (1. - s) * 0.25;  g[1][1] = -(1. + r) * 0.25;
        g[2][0] = (1. + s) * 0.25;  g[2][1] = (1. + r) * 0.25;
        g[3][0] = -(1. + s) * 0.25; g[3][1] = (1. - r) * 0.25;
    }
}

// Fills *t for (fe, rule); returns -1 for an unsupported pair.
__host__ __device__ constexpr int qt_make(int fe, int rule, QTable *t)
{
    double pc[ELFEL_MAXQ][2] = {}, w[ELFEL_MAXQ] = {};
    int nq = 0;
    if (fe == QT_T3 || fe == QT_T6) {
        if (rule == 1) {
            pc[0][0] = 1.0 / 3.; pc[0][1] = 1.0 / 3.; w[0] = 1.0 / 2.0; nq = 1;
        } else if (rule == 3) {
            pc[0][0] = 2.0 / 3; pc[0][1] = 1.0 / 6;
            pc[1][0] = 1.0 / 6; pc[1][1] = 2.0 / 3;
            pc[2][0] = 1.0 / 6; pc[2][1] = 1.0 / 6;
            for (int i = 0; i < 3; i++) w[i] = (1.0 / 3) / 2;
            nq = 3;
        } else {
            return -1;
        }
    } else if (fe == QT_Q4) {
        double p1[2] = {}, w1[2] = {};
        int n1 = 0;
        if (rule == 1) { p1[0] = 0.0; w1[0] = 2.0; n1 = 1; }
        else if (rule == 2) { p1[0] = -0.577350269189626; p1[1] = 0.577350269189626; w1[0] = 1.0; w1[1] = 1.0; n1 = 2; }
        else return -1;
        for (int i = 0; i < n1; i++)
            for (int j = 0; j < n1; j++) {
                pc[nq][0] = p1[i]; pc[nq][1] = p1[j]; w[nq] = w1[i] * w1[j]; nq++;
            }
    } else {
        return -1;
    }
    *t = QTable{};
    t->nq = nq;
    t->nbf = fe == QT_T3 ? 3 : (fe == QT_T6 ? 6 : 4);
    for (int q = 0; q < nq; q++) {
        t->w[q] = w[q];
        qt_bfun(fe, pc[q][0], pc[q][1], t->N[q]);
        qt_bgrad(fe, pc[q][0], pc[q][1], t->G[q]);
        if (fe != QT_Q4) qt_bfun(QT_T3, pc[q][0], pc[q][1], t->Np[q]);
        t->wf[q] = (float)t->w[q];
        for (int k = 0; k < ELFEL_MAXBF; k++) {
            t->Nf[q][k] = (float)t->N[q][k];
            t->Gf[q][k][0] = (float)t->G[q][k][0];
            t->Gf[q][k][1] = (float)t->G[q][k][1];
        }
        for (int k = 0; k < 3; k++) t->Npf[q][k] = (float)t->Np[q][k];
    }
    return 0;
}

// The table a kernel instantiation implies: (nodes, points) fixes the element
// and the rule (T3/T6 with 1 or 3 points, Q4 with 1 or 4).
template <int NEN, int NQ>
__host__ __device__ constexpr QTable qt_fixed()
{
    QTable t{};
    const int fe = NEN == 3 ? QT_T3 : (NEN == 6 ? QT_T6 : QT_Q4);
    const int rule = fe == QT_Q4 ? (NQ == 1 ? 1 : 2) : NQ;
    qt_make(fe, rule, &t);
    return t;
}
