// tables.cpp -- host tabulation of quadrature rules and basis functions.
//
// Rules: src/RefShapes.jl:113-119 (triangle 1- and 3-point), :85-96 and
// :333-366 (Gauss order 1/2, tensor order i outer, j inner).  Bases:
// src/FElements.jl:239-246 (T3), :264-288 (T6), :306-320 (Q4).  The literal
// constants and the expression forms are the reference's, so the FP64 tables
// are bit-identical to what QPIterator tabulates (src/QPIterators.jl:14-50).
#include "common.h"

static void bfun(int fe, double r, double s, double *N)
{
    if (fe == ELFEL_T3) {
        N[0] = (1 - r - s); N[1] = r; N[2] = s;
    } else if (fe == ELFEL_T6) {
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

static void bgrad(int fe, double r, double s, double g[][2])
{
    if (fe == ELFEL_T3) {
        g[0][0] = -1.; g[0][1] = -1.; g[1][0] = +1.; g[1][1] = 0.; g[2][0] = 0.; g[2][1] = +1.;
    } else if (fe == ELFEL_T6) {
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

int make_qtable(int fe, int rule, QTable *t)
{
    double pc[ELFEL_MAXQ][2], w[ELFEL_MAXQ];
    int nq = 0;
    if (fe == ELFEL_T3 || fe == ELFEL_T6) {
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
    } else if (fe == ELFEL_Q4) {
        double p1[2], w1[2];
        int n1;
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
    t->nbf = fe == ELFEL_T3 ? 3 : (fe == ELFEL_T6 ? 6 : 4);
    for (int q = 0; q < nq; q++) {
        t->w[q] = w[q];
        bfun(fe, pc[q][0], pc[q][1], t->N[q]);
        bgrad(fe, pc[q][0], pc[q][1], t->G[q]);
        if (fe != ELFEL_Q4) bfun(ELFEL_T3, pc[q][0], pc[q][1], t->Np[q]);
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
