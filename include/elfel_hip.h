/*
 * elfel_hip.h -- C ABI of libelfel_hip.so, the MI355X (gfx950) element
 * integration + global assembly library behind Elfel.jl's assembler API.
 *
 * Plain C: no C++ or torch types cross this boundary.  Every function returns
 * an int status (0 = ELFEL_OK, < 0 = error class); elfel_hip_last_error()
 * gives the text.  No C++ exception crosses the ABI.  One context = one HIP
 * stream on one device; calls on a context are serialised by the caller
 * (Elfel's assemblers are mutable and single-threaded, src/Assemblers.jl).
 *
 * Index conventions follow Julia: connectivity, dof numbers and COO indices
 * are 1-based Int64, laid out exactly as the Julia arrays are in memory
 * (Vector{SVector{nen,Int64}} -> conn[e*nen + k]; Vector{SVector{2,Float64}}
 * -> xy[2*n + d]; FEField.dofnums -> dofnums[ncomp*n + c]).  Sparse outputs
 * are copied as 1-based Int64 (ready for SparseMatrixCSC) or 0-based.
 */
#ifndef ELFEL_HIP_H
#define ELFEL_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct elfel_hip_ctx elfel_hip_ctx;

/* status codes */
enum {
    ELFEL_OK = 0,
    ELFEL_EINVAL = -1,   /* bad argument (null pointer, bad size, bad enum) */
    ELFEL_ERANGE = -2,   /* index out of range: sparse() ArgumentError / BoundsError */
    ELFEL_ENOMEM = -3,   /* device or host allocation failed */
    ELFEL_EHIP = -4,     /* HIP runtime error */
    ELFEL_ESTATE = -5,   /* call out of order (no mesh, no dofs, no result) */
    ELFEL_EGEOM = -6,    /* element with non-positive Jacobian determinant */
    ELFEL_ENOTSUP = -7   /* combination not supported */
};

/* weak forms: heat (test/test_heat.jl:36-60, examples/heat/poisson/t4.jl:33-55),
 * plane elasticity, Taylor-Hood Stokes over one vector space (generic B'DB or
 * vector Laplacian), Reddy-form Stokes over three scalar spaces ux, uy, p
 * (test/test_stokes.jl:373-431, examples/stokes/colliding_flow/p1b_p1.jl:50-108,
 * q1_q0.jl:52-110), and the mass matrix N_i N_j (kappa JxW)
 * (examples/klein_gordon/klein_gordon.jl:34-51,94-112) */
enum { ELFEL_HEAT = 0, ELFEL_ELAST = 1, ELFEL_STOKES_GEN = 2, ELFEL_STOKES_VECLAP = 3, ELFEL_STOKES_REDDY = 4,
       ELFEL_MASS = 5 };
/* scalar finite elements: FEH1_T3, FEH1_T6, FEH1_Q4, FEH1_T3_BUBBLE (T3 + a
 * cubic bubble with one dof on the cell), FEL2_Q4 (one dof on the cell, the
 * Q4 its geometry carrier), FEH1_T4 (3D), FEH1_L2 (1D); src/FElements.jl */
enum { ELFEL_T3 = 0, ELFEL_T6 = 1, ELFEL_Q4 = 2, ELFEL_T3B = 3, ELFEL_Q4L2 = 4, ELFEL_T4 = 5, ELFEL_L2 = 6 };
/* arithmetic of the element kernels */
enum { ELFEL_FP64 = 0, ELFEL_MIXED = 1 };
/* global build method */
enum { ELFEL_SORT = 0,    /* COO in stream order -> stable radix sort -> in-order segmented reduce */
       ELFEL_REUSE = 1 }; /* cached pattern + fused row-owner kernel; for the element families the
                             cached sort permutation + in-order run sums (no sort, no atomics) */
/* sparse layout of the result */
enum { ELFEL_CSR = 0, ELFEL_CSC = 1 };
/* timing phases (elfel_hip_timing_read) */
enum { ELFEL_PH_GATHER = 0, ELFEL_PH_ELEMENT = 1, ELFEL_PH_SORT = 2, ELFEL_PH_REDUCE = 3,
       ELFEL_PH_NUMERIC = 4, ELFEL_PH_SYMBOLIC = 5, ELFEL_PH_TOTAL = 6, ELFEL_PH_SOLVE = 7,
       ELFEL_NPHASES = 8 };

/* ---- context -------------------------------------------------------------
 * No counterpart in the reference (Julia needs no device context); the Julia
 * assembler object that owns it is SysmatAssemblerHIP. */
int elfel_hip_init(int device, elfel_hip_ctx **ctx);
int elfel_hip_free(elfel_hip_ctx *ctx);
const char *elfel_hip_last_error(const elfel_hip_ctx *ctx);
int elfel_hip_device_name(elfel_hip_ctx *ctx, char *buf, int len, int *num_cus);
/* Bandwidth of a plain device-to-device copy of `bytes` (a multiple of 16)
 * through HBM, timed over `reps` launches after one warm-up: *gbs = bytes read
 * + bytes written per second / 1e9.  The measured ceiling beside which the
 * benchmark reads its roofline fractions (the published HBM peak is 8 TB/s).
 * Allocates and frees two buffers.  No counterpart in the reference. */
int elfel_hip_copy_bandwidth(elfel_hip_ctx *ctx, int64_t bytes, int reps, double *gbs);

/* ---- generic drop-in for SysmatAssemblerSparse ----------------------------
 * Replaces finish!(self::SysmatAssemblerSparse) = sparse(row, col, val, nrow,
 * ncol) at src/Assemblers.jl:121-123, for a COO stream gathered on the host by
 * assemble!(a, r, c, v) (:85-90), assemble!(a, lma) (:97-102) and
 * assemble!(a, transpose(lma)) (:109-114).  Duplicates are summed in stream
 * order; explicit zeros stay structural; out-of-range indices -> ELFEL_ERANGE. */
int elfel_hip_sparse(elfel_hip_ctx *ctx, int64_t nrow, int64_t ncol, int64_t n,
                     const int64_t *row, const int64_t *col, const double *val,
                     int layout, int64_t *nnz);

/* Allocation-free variant of elfel_hip_sparse.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc): the call makes no
 * hipMalloc / hipFree and no allocating library call.  row, col (1-based
 * Int64), val are DEVICE arrays of n COO entries; the result goes to the
 * caller's DEVICE arrays ptr (nmajor + 1 int32), ind and out (room for n
 * entries: nnz <= n), 0-based; *nnz (host) receives the entry count.  The
 * sort's scratch lives in one static device array per GPU, sized for
 * n <= ELFEL_NOALLOC_MAX_COO entries (about 41 MiB, held until the process
 * ends); a larger n returns ELFEL_EINVAL.  All calls on one GPU share that
 * array, so calls that overlap on one GPU (two contexts, two threads) race;
 * the context's own calls are serialised on its stream.  Same order and
 * semantics as elfel_hip_sparse (stable order, structural zeros kept,
 * ELFEL_ERANGE for indices out of range); the call waits for the result. */
#define ELFEL_NOALLOC_MAX_COO (1 << 20)
int elfel_hip_sparse_noalloc(elfel_hip_ctx *ctx, int64_t nrow, int64_t ncol, int64_t n, const int64_t *row,
                             const int64_t *col, const double *val, int layout, int32_t *ptr, int32_t *ind,
                             double *out, int64_t *nnz);

/* Copies the matrix built by the last elfel_hip_sparse / elfel_hip_assemble.
 * base = 1: 1-based Int64 (SparseMatrixCSC colptr/rowval when layout = CSC);
 * base = 0: 0-based.  ptr has nmajor+1 entries, ind and val nnz entries. */
int elfel_hip_copy_sparse(elfel_hip_ctx *ctx, int base, int64_t *ptr, int64_t *ind, double *val);
int elfel_hip_result_info(elfel_hip_ctx *ctx, int64_t *nrow, int64_t *ncol, int64_t *nnz, int *layout);

/* ---- generic drop-in for SysvecAssembler ----------------------------------
 * Replaces start!/assemble!/finish! of SysvecAssembler (src/Assemblers.jl:
 * 194-230): out[i-1] = sum of val[k] with idx[k] == i, added in stream order
 * from 0.0.  Index outside 1..ndofs -> ELFEL_ERANGE (BoundsError at :218). */
int elfel_hip_vector(elfel_hip_ctx *ctx, int64_t ndofs, int64_t n, const int64_t *idx,
                     const double *val, double *out);

/* ---- device-resident problem description ----------------------------------
 * field 0: the primary (velocity / temperature / displacement) space;
 * field 1: the pressure space of a Taylor-Hood pair (T3 on the corners).
 * Replaces what FEIterator gathers per element (src/FEIterators.jl:54-81,
 * 161-194): base incidence relation (conn), "geom" attribute (xy) and
 * FEField.dofnums (src/FEFields.jl:15).  Dof 0 (unnumbered) -> ELFEL_ERANGE. */
int elfel_hip_set_mesh(elfel_hip_ctx *ctx, int field, int32_t nen, int64_t E, const int64_t *conn,
                       int64_t Nn, const double *xy);
int elfel_hip_set_dofs(elfel_hip_ctx *ctx, int field, int32_t ncomp, int64_t Nn,
                       const int64_t *dofnums, int64_t ndofs);
/* A mesh in sdim = 1, 2 or 3 space dimensions (x[sdim*n + d], the memory of
 * Vector{SVector{sdim,Float64}}): L2 meshes (sdim 1) and T4 meshes (sdim 3)
 * for the element families; elfel_hip_set_mesh is sdim = 2. */
int elfel_hip_set_mesh_sdim(elfel_hip_ctx *ctx, int field, int32_t nen, int64_t E, const int64_t *conn,
                            int64_t Nn, int32_t sdim, const double *x);
/* Dof numbers of the cells (the incidence relation of the element's own
 * dimension, FESpace._irsfields[2], src/FESpaces.jl:76-86): dofnums[ncomp*e +
 * c], 1-based.  The bubble of FEH1_T3_BUBBLE and the FEL2_Q4 pressure; an
 * element's dofs are its nodes' dofs followed by its cell dofs, the order of
 * eldofs (src/FEIterators.jl:161-171).  A field may hold cell dofs only. */
int elfel_hip_set_element_dofs(elfel_hip_ctx *ctx, int field, int32_t ncomp, int64_t E, const int64_t *dofnums,
                               int64_t ndofs);
/* Per-element coefficients (any pointer may be NULL = uniform value of the
 * problem): conductivity, heat source, scale of the material matrix D. */
int elfel_hip_set_coefficients(elfel_hip_ctx *ctx, int64_t E, const double *kappa_e,
                               const double *Q_e, const double *scale_e);

typedef struct {
    int32_t kind;       /* ELFEL_HEAT / ELFEL_ELAST / ELFEL_STOKES_GEN / ELFEL_STOKES_VECLAP */
    int32_t fe;         /* ELFEL_T3 / ELFEL_T6 / ELFEL_Q4 (velocity element for Stokes) */
    int32_t rule;       /* triangle: npts (1 or 3); square: Gauss order (1 or 2) */
    int32_t precision;  /* ELFEL_FP64 / ELFEL_MIXED */
    int32_t method;     /* ELFEL_SORT / ELFEL_REUSE */
    int32_t layout;     /* ELFEL_CSR / ELFEL_CSC */
    double kappa, Q;    /* heat: uniform conductivity and source */
    double mu;          /* Stokes vector-Laplacian viscosity */
    double D[9];        /* elasticity / Stokes-gen material matrix, column-major */
    int32_t fe_p;       /* Reddy Stokes: pressure element, ELFEL_T3 (field 1's nodes) or ELFEL_Q4L2
                           (field 1's cell dofs); pairs (T6, T3), (T3B, T3), (Q4, Q4L2) */
    int32_t flags;      /* ELFEL_CHECK_GEOM or 0 */
} elfel_hip_problem;

/* elfel_hip_problem.flags: ELFEL_CHECK_GEOM makes elfel_hip_assemble first
 * evaluate det J of field 0's elements (T3, T6, Q4 in 2D) at the points of
 * the problem's rule, in FP64, and return ELFEL_EGEOM without assembling if
 * any is <= 0.  The reference does no such check (jacjac, src/FElements.jl:
 * 148-168, returns whatever determinant the nodes give and the element loop
 * goes on), so without the flag an inverted element is assembled as the
 * reference would assemble it. */
enum { ELFEL_CHECK_GEOM = 1 };

/* Integrates every element and builds the global matrix (and the load vector
 * for heat) on the device.  Replaces the user integrate! loops with
 * LocalMatrixAssembler + SysmatAssemblerSparse (test/test_heat.jl:36-71,
 * test/test_elasticity.jl:149-181, examples/stokes/colliding_flow/
 * ht_p2_p1_gen.jl:46-90, test/test_stokes.jl:629-668).  nnz may be NULL:
 * then the call does not wait for the device.
 * Element families (T3B, Q4L2, T4, L2) and the Reddy and mass forms run on
 * the plain FP64 method (ELFEL_SORT, ELFEL_FP64) or on ELFEL_REUSE: the sort
 * of their COO stream is done once per mesh + dof map + weak form and kept as
 * a permutation, and each assembly runs the element kernel and sums every
 * matrix entry's run through it, in the sort's order (ELFEL_FP64: the plain
 * method's bits; ELFEL_MIXED: FP64 geometry, FP32 products and element
 * values, FP64 sums).
 * Heat on T4 (rule = tetrahedron points 1 or 4) and L2 (Gauss order 1-3, sdim
 * 1); mass on T3, T6, Q4, T4, L2;
 * Reddy Stokes with field 0 the velocity mesh, its node dofs ncomp = 2
 * (component 1 the ux space's numbers, 2 the uy space's) and for T3B its cell
 * dofs ncomp = 2, field 1 the pressure (T3 mesh + node dofs, or Q4L2 cell
 * dofs).  Blocks are appended as the drivers do: kxx, kxy, kxy', kyy, kxp,
 * kxp', kyp, kyp'. */
int elfel_hip_assemble(elfel_hip_ctx *ctx, const elfel_hip_problem *pb, int64_t *nnz);
/* What the last elfel_hip_assemble ran.  The REUSE method needs every node in
 * at most 16 elements, rows of at most 128 entries, each dof owned by one
 * (node, component) and elements whose dofs are distinct; a mesh or dof map
 * that breaks one of these (valid input: the reference's sparse() assembles any
 * COO stream, src/Assemblers.jl:97-123) is assembled by summing the plain
 * method's stream in its sort order, cached on the first call, in FP64: the
 * same pattern and the reference's values.  method and
 * precision receive the method and arithmetic used; note (len bytes, may be
 * null) the reason when they differ from the request.  No counterpart in the
 * reference, which has one method. */
int elfel_hip_assembly_info(elfel_hip_ctx *ctx, int *method, int *precision, char *note, int len);
/* Element matrices: the ke block of every element of field 0 (what
 * LocalMatrixAssembler holds after the element loop of test/test_elasticity.jl:
 * 150-172, src/LocalAssemblers.jl:68-83), plane elasticity on T6 with the
 * 3-point rule in mixed arithmetic, as ke[e * 144 + i + 12 j] (column-major per
 * element, FP32).  method ELFEL_EM_VALU: vector FMAs, one work-item per
 * element; ELFEL_EM_MFMA: the B^T D B contraction on the matrix cores
 * (v_mfma_f32_4x4x1f32, 16 elements per wavefront).  *ms receives the kernel
 * time.  n = 144 E; ke NULL with n 0 runs and times the kernel and leaves the
 * blocks on the device.  The assembly kernels use the vector form; this entry
 * point measures the matrix cores on the path's one dense contraction. */
enum { ELFEL_EM_VALU = 0, ELFEL_EM_MFMA = 1 };
int elfel_hip_element_matrices(elfel_hip_ctx *ctx, const elfel_hip_problem *pb, int method, float *ke, int64_t n,
                               double *ms);
/* Load vector of the last heat assembly (finish!(SysvecAssembler), :228-230). */
int elfel_hip_copy_vector(elfel_hip_ctx *ctx, double *F, int64_t n);
/* Blocks until the context's stream is idle. */
int elfel_hip_synchronize(elfel_hip_ctx *ctx);
/* Mesh quality of field 0: the smallest Jacobian determinant at any
 * quadrature point (FP64), and the number of elements with det J <= 0 or a
 * shape-quality proxy sum_k |grad N_k|^2 det J >= 1024 at some point.
 * Informational: every kernel treats every element alike. */
int elfel_hip_geometry_stats(elfel_hip_ctx *ctx, double *min_det, int64_t *n_bad_shape);

/* ---- solve --------------------------------------------------------------
 * Replaces the driver's solve! (test/test_heat.jl:73-76: T[1:nu] =
 * K[1:nu,1:nu] \ (F - K*T)[1:nu]) for SPD problems with Jacobi-PCG on the
 * matrix of the last assembly (CSR).  x holds the data values on entry
 * (x[nfree:]) and the solution on return.  precision = ELFEL_FP64: FP64 PCG;
 * ELFEL_MIXED: PCG with an FP32 matrix copy and FP32 r, z, p, q, the iterate
 * kept in FP64 (its steps gathered in FP32 over 8 iterations), and reliable
 * updates (the FP64 residual replaces the FP32 one
 * at every 100x drop).  Convergence: ||b - A x|| / ||b|| <= tol in FP64 on
 * the free block.  *iters = PCG iterations, *outer = FP64 residual
 * replacements, *relres = the final FP64 relative residual. */
int elfel_hip_pcg(elfel_hip_ctx *ctx, int64_t nfree, const double *F, double *x, int precision,
                  double tol, int maxit, int *iters, int *outer, double *relres);

/* ---- dof numbering and system vectors ----------------------------------------
 * numberdofs! over nspaces spaces (src/FESpaces.jl:250-273 with
 * src/FEFields.jl:137-177): isdatum holds, space after space, nterms[s] x
 * ncomp[s] flags (entity-major, component-minor, nonzero = prescribed);
 * dofnums receives the 1-based numbers in the same layout: the free dofs of
 * each space consecutively (first space from 1, later spaces from the
 * previous space's highest free number + 1), then the data dofs likewise
 * (first from the last space's highest free number + 1). */
int elfel_hip_number_dofs(elfel_hip_ctx *ctx, int32_t nspaces, const int64_t *nterms, const int32_t *ncomp,
                          const uint8_t *isdatum, int64_t *dofnums);
/* gathersysvec! of one field (src/FEFields.jl:247-257): v[dofnums[k] - 1] =
 * dofvals[k] for every dof of the field bound by set_dofs; nvals = Nn * ncomp;
 * entries of v that no dof maps to keep their values.  A dof number outside
 * 1..ndofs -> ELFEL_ERANGE (BoundsError). */
int elfel_hip_gather_sysvec(elfel_hip_ctx *ctx, int field, int64_t nvals, const double *dofvals, int64_t ndofs,
                            double *v);
/* scattersysvec! of one field (src/FEFields.jl:264-279): dofvals[k] =
 * v[dofnums[k] - 1] for the dofs numbered within v (others keep theirs); an
 * unnumbered dof -> ELFEL_ERANGE. */
int elfel_hip_scatter_sysvec(elfel_hip_ctx *ctx, int field, int64_t ndofs, const double *v, int64_t nvals,
                             double *dofvals);

/* ---- meshes generated on the device -----------------------------------------
 * The structured generators of the reference's drivers, written straight into
 * field `field` (as elfel_hip_set_mesh would, with the same element and node
 * order and the same coordinate bits as the generators): T3block
 * (test/test_heat.jl:29), Q4block (:144) and T6block (test/test_stokes.jl:32)
 * of an A x B rectangle with nA x nB cells; fe = ELFEL_T3, ELFEL_Q4 or
 * ELFEL_T6.  *E and *Nn receive the element and node counts. */
int elfel_hip_block_mesh(elfel_hip_ctx *ctx, int field, int fe, double A, double B, int64_t nA, int64_t nB,
                         int64_t *E, int64_t *Nn);
/* T6toT3 (test/test_stokes.jl:37): field dst receives the T3 mesh on the
 * corner nodes of field src's T6 mesh, renumbered compactly in increasing T6
 * node order; *Nn receives its node count. */
int elfel_hip_t6_to_t3(elfel_hip_ctx *ctx, int src, int dst, int64_t *Nn);
/* connectedv(boundary(mesh)) of field `field` (test/test_heat.jl:31): flags[n]
 * = 1 for the nodes of the edges that belong to one element (T6: with the
 * edge's midside node), else 0; Nn must equal the field's node count. */
int elfel_hip_boundary_nodes(elfel_hip_ctx *ctx, int field, int64_t Nn, uint8_t *flags, int64_t *count);
/* load(Mesh(), "name.mesh") of MeshSteward, as the reference's tests read
 * test/qmesh.mesh and test/mt3gen3.mesh: a manifest of three lines (coordinate
 * file, shape T3/Q4/T6, connectivity file, both text tables next to the
 * manifest) read on the host and set as field `field`'s mesh. */
int elfel_hip_load_mesh(elfel_hip_ctx *ctx, int field, const char *path, int32_t *nen, int64_t *E, int64_t *Nn);
/* Copies field `field`'s mesh to the host: conn (E x nen, 1-based, row per
 * element) and xy (Nn x sdim, sdim of the mesh); either may be null. */
int elfel_hip_copy_mesh(elfel_hip_ctx *ctx, int field, int64_t E, int64_t *conn, int64_t Nn, double *xy);

/* ---- post-processing -------------------------------------------------------
 * L2 errors of a Taylor-Hood solution against the colliding-flow exact
 * fields p = 60x^2y - 20y^3, ux = 20xy^3, uy = 5x^4 - 5y^4
 * (test/test_stokes.jl:354-356): ep = sqrt(sum_e sum_q JxW (p_h - p)^2) over
 * field 1 (scalar T3) and ev = sqrt(sum_e sum_q JxW |u_h - u|^2) over field 0
 * (vector T6), 3-point rules, U the full system vector (ndofs entries,
 * 1-based dof numbers of set_dofs).  Replaces evaluate_pressure_error and
 * evaluate_velocity_error (test/test_stokes.jl:438-496, :675-731); the
 * element terms are summed in a fixed tree order, so the norms agree with
 * the reference's left-to-right sum to rounding. */
int elfel_hip_colliding_errors(elfel_hip_ctx *ctx, int64_t ndofs, const double *U, double *ep, double *ev);

/* ---- batched problems ------------------------------------------------------
 * Many independent problems of uneven size (heat T3 1-point, heat Q4 2x2 and
 * plane elasticity T6 3-point) assembled by one persistent fused launch
 * replayed from a hipGraph.  Each problem is one reference driver's
 * integrate! + finish! (test/test_heat.jl:35-71, examples/heat/poisson/q4.jl:
 * 31-54, test/test_elasticity.jl:149-181); the problems of a batch share their
 * precision (ELFEL_MIXED, or ELFEL_FP64: the FP64 REUSE method's bits) and
 * layout, problems of one kind their uniform coefficients (kappa, Q, D);
 * coef_e is the per-element conductivity (heat) or scale of D (elasticity),
 * or NULL.  batch_build runs the symbolic phase of every problem (a mesh the
 * REUSE tiles do not hold is refused with ELFEL_ENOTSUP; the build uses field
 * 0 of the context, whose mesh and dof map it replaces), picks the launch
 * shape (autotune = 1 times every variant on the batch; 0 uses the on-disk
 * cache when it has an entry) and captures the graph.  No counterpart in the
 * reference beyond the per-problem loops. */
int elfel_hip_batch_begin(elfel_hip_ctx *ctx);
int elfel_hip_batch_add(elfel_hip_ctx *ctx, const elfel_hip_problem *pb, int32_t nen, int64_t E, const int64_t *conn,
                        int64_t Nn, const double *xy, int32_t ncomp, const int64_t *dofnums, int64_t ndofs,
                        const double *coef_e);
int elfel_hip_batch_build(elfel_hip_ctx *ctx, int autotune);
int elfel_hip_batch_assemble(elfel_hip_ctx *ctx, int wait);
int elfel_hip_batch_info(elfel_hip_ctx *ctx, int32_t k, int64_t *ndofs, int64_t *nnz, int64_t *nelem, int *block,
                         int *grid);
int elfel_hip_batch_copy(elfel_hip_ctx *ctx, int32_t k, int base, int64_t *ptr, int64_t *ind, double *val, double *F);
int elfel_hip_batch_stats(elfel_hip_ctx *ctx, int64_t *nprob, int64_t *nelem, int64_t *nnz, int64_t *ntiles,
                          double *tuned_ms);

/* ---- timing ---------------------------------------------------------------
 * enable = 2 (region mode): one hipEvent pair on the context's stream around
 * everything submitted until elfel_hip_timing_read, which then returns the
 * elapsed time in ms[ELFEL_PH_TOTAL] and the number of assemblies (single or
 * batched) in counts[ELFEL_PH_TOTAL]; no per-phase events in between.
 * enable = 1: hipEvent timing of each phase on the context's stream, accumulated while
 * enabled.  ms[ELFEL_NPHASES], counts[ELFEL_NPHASES] (launch counts). */
int elfel_hip_timing(elfel_hip_ctx *ctx, int enable);
int elfel_hip_timing_read(elfel_hip_ctx *ctx, double *ms, int64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
