/*
 * elfel_oracle.h -- CPU restatement of Elfel.jl's element-integration and
 * global-assembly path, in plain C (FP64, no FMA contraction).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library and
 * the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never calls it.
 *
 * Conventions: connectivity is element-major and 1-based (conn[e*nen + k]),
 * exactly the memory layout of MeshCore's Vector{SVector{nen,Int64}}; node
 * coordinates are interleaved xy[2*n + d] (Vector{SVector{2,Float64}});
 * dof numbers are interleaved dofnums[ncomp*n + c], 1-based, as in
 * FEField.dofnums (src/FEFields.jl:15).  COO streams are 1-based like the
 * row/col vectors of SysmatAssemblerSparse (src/Assemblers.jl:19-24).
 * Sparse outputs are 0-based (scipy layout) with int64 indices.
 */
#ifndef ELFEL_ORACLE_H
#define ELFEL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* finite element codes: FEH1_T3, FEH1_T6, FEH1_Q4, FEH1_T3_BUBBLE, FEL2_Q4,
 * FEH1_T4, FEH1_L2 */
enum { ORC_T3 = 0, ORC_T6 = 1, ORC_Q4 = 2, ORC_T3B = 3, ORC_Q4L2 = 4, ORC_T4 = 5, ORC_L2 = 6 };
/* weak forms */
enum { ORC_HEAT = 0, ORC_ELAST = 1, ORC_STOKES_GEN = 2, ORC_STOKES_VECLAP = 3, ORC_STOKES_REDDY = 4,
       ORC_MASS = 5 };

typedef struct {
    int kind;          /* ORC_HEAT / ORC_ELAST / ORC_STOKES_* */
    int fe;            /* element of the primary (velocity) field */
    int rule;          /* triangle: number of points; square: Gauss order */
    int64_t E;         /* number of elements */
    int nen;           /* nodes per element of the primary mesh */
    const int64_t *conn;     /* E*nen, 1-based */
    const double *xy;        /* 2*Nn */
    int64_t Nn;
    int ncomp;               /* copies of the scalar element (1 heat, 2 elasticity/velocity) */
    const int64_t *dofnums;  /* Nn*ncomp, 1-based */
    /* Stokes pressure field (T3 on the corner nodes) */
    const int64_t *conn_p;   /* E*3, 1-based, pressure-mesh node numbers */
    const int64_t *dofnums_p;/* Np*1, 1-based */
    int64_t Np;
    /* coefficients */
    double kappa, Q;               /* heat: uniform conductivity and source */
    const double *kappa_e, *Q_e;   /* optional per-element values (NULL: uniform) */
    double D[9];                   /* material matrix, column-major 3x3 */
    const double *scale_e;         /* optional per-element scale of D (NULL: 1) */
    double mu;                     /* viscosity for the vector-Laplacian Stokes form */
} orc_problem;

/* ---- reference shapes and quadrature (src/RefShapes.jl) ---- */
int orc_gauss1(int order, double *pc, double *w);
int orc_quad_triangle(int npts, double *pc, double *w);
int orc_quad_square(int order, double *pc, double *w);

/* ---- basis functions (src/FElements.jl) ---- */
int orc_nen(int fe);
void orc_bfun(int fe, const double *pc, double *N);
void orc_bfungradpar(int fe, const double *pc, double *g);

/* ---- structured meshes (stand-ins for MeshSteward's block generators) ---- */
int64_t orc_t3block(double A, double B, int nA, int nB, int64_t *conn, double *xy);
int64_t orc_q4block(double A, double B, int nA, int nB, int64_t *conn, double *xy);
int64_t orc_t6block(double A, double B, int nA, int nB, int64_t *conn, double *xy);
int64_t orc_t6tot3(int64_t E, const int64_t *conn6, int64_t Nn6, const double *xy6,
                   int64_t *conn3, double *xy3);

/* ---- dof numbering (src/FEFields.jl:137-177, src/FESpaces.jl:141-173,250-259) ---- */
int orc_numberdofs(int nspaces, const int64_t *nterms, const int32_t *ncomp,
                   const uint8_t *isdatum_concat, int64_t *dofnums_concat);

/* ---- element loops producing the COO stream in Elfel's order ---- */
int64_t orc_coo_len(const orc_problem *p, int64_t *nvec);
int orc_coo(const orc_problem *p, int64_t *I, int64_t *J, double *V,
            int64_t *FI, double *FV);
int orc_coo_parallel(const orc_problem *p, int64_t *I, int64_t *J, double *V,
                     int64_t *FI, double *FV, int nthreads);
/* Reddy-form Stokes over three scalar spaces (test/test_stokes.jl:373-431) */
int64_t orc_stokes_reddy_coo_len(int64_t E);
int orc_stokes_reddy_coo(int64_t E, const int64_t *conn6, const double *xy6,
                         const int64_t *conn3, const double *xy3,
                         const int64_t *dn_ux, const int64_t *dn_uy, const int64_t *dn_p,
                         double mu, int64_t *I, int64_t *J, double *V);

/* ---- element families (elfel_oracle_fam.c) ---- */
int orc_fam_nbf(int fe);
int orc_fam_ngeo(int fe);
int orc_fam_mdim(int fe);
int orc_fam_quad(int fe, int rule, double *pc /* 3 per point */, double *w);
void orc_fam_bfun(int fe, const double *pc, double *N);
void orc_fam_bfungradpar(int fe, const double *pc, double *g /* nbf*mdim */);
/* heat / mass on a nodal family (T3, T6, Q4, T4, L2); x holds mdim
 * coordinates per node; dofnums one per node */
int64_t orc_fam_scalar_coo_len(int fe, int64_t E, int64_t *nvec);
int orc_fam_scalar_coo(int kind, int fe, int rule, int64_t E, const int64_t *conn, const double *x,
                       const int64_t *dofnums, double kappa, double Q, int64_t *I, int64_t *J, double *V,
                       int64_t *FI, double *FV);
/* Reddy-form Stokes: (T6, T3), (T3B, T3), (Q4, Q4L2) */
int64_t orc_fam_reddy_coo_len(int vfe, int pfe, int64_t E);
int orc_fam_reddy_coo(int vfe, int pfe, int rule, int64_t E, const int64_t *conn_v, const double *xy_v,
                      const int64_t *conn_p, const double *xy_p, const int64_t *dn_ux, const int64_t *dn_uy,
                      const int64_t *de_ux, const int64_t *de_uy, const int64_t *dn_p, const int64_t *de_p,
                      double mu, int64_t *I, int64_t *J, double *V);
int64_t orc_l2block(double L, int n, int64_t *conn, double *x);
int64_t orc_t4block(double A, double B, double C, int nA, int nB, int nC, int64_t *conn, double *xyz);

/* ---- SparseArrays.sparse equivalent (src/Assemblers.jl:121-123) ---- */
int64_t orc_sparse(int64_t n, const int64_t *I, const int64_t *J, const double *V,
                   int64_t nrow, int64_t ncol, int csc,
                   int64_t *ptr, int64_t *ind, double *val);
int orc_vec_assemble(int64_t n, const int64_t *FI, const double *FV, int64_t ndofs,
                     double *out);

/* ---- post-processing integrals (test/test_stokes.jl:438-496,675-731) ---- */
double orc_colliding_p_error(int64_t E, const int64_t *conn3, const double *xy3,
                             const int64_t *dn_p, const double *U);
double orc_colliding_v_error_scalar(int64_t E, const int64_t *conn6, const double *xy6,
                                    const int64_t *dn_ux, const int64_t *dn_uy,
                                    const double *U);
double orc_colliding_v_error_vector(int64_t E, const int64_t *conn6, const double *xy6,
                                    const int64_t *dn_u, const double *U);

/* ---- FP64 Jacobi-PCG on the free block of a CSR matrix ---- */
int orc_pcg(int64_t nfree, const int64_t *indptr, const int64_t *indices,
            const double *vals, const double *b, double *x, double tol, int maxit,
            double *relres);

/* ---- timing helpers for the CPU baseline ---- */
int orc_time_assembly(const orc_problem *p, int reps, int nthreads,
                      double *t_elem, double *t_sparse, int64_t *nnz);

#ifdef __cplusplus
}
#endif
#endif
