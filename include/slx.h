/* slx — MI355X-native weak-form assembly: C ABI.
 *
 * Two layers, both plain C (no C++ or torch types cross this boundary):
 *
 *  1. Assembly plan layer (the hot path).  A plan holds the symbolic state
 *     of one formulation on one mesh: element buckets with their Gauss
 *     tables, shape-function tables, DOF address tables, term tables and
 *     coefficient programs, plus the cached CSR pattern.  slx_assemble runs
 *     the numeric phase on the GPU: Jacobians, Gauss-point coefficients,
 *     element contractions and the deterministic reduction into the CSR.
 *       replaces contribution::generate (src/formulation/contribution.cpp:20-307)
 *       and rawmat::process (src/formulation/rawmat.cpp:127-298).
 *
 *  2. Formulation layer, mirroring the reference objects a user drives:
 *     mesh, field, formulation (+= integral), generate(), A(), b().
 *       replaces formulation::operator+= (src/formulation/formulation.cpp:51-149),
 *       formulation::generate() (formulation.cpp:211-218),
 *       formulation::A()/b() (formulation.cpp:358-422).
 *
 * Every function returns SLX_OK (0) or a negative status; the message of
 * the last failure on the calling thread is returned by slx_last_error().
 * Errors never propagate as C++ exceptions (the reference throws
 * std::runtime_error from logs::error(), src/logs.cpp:9-12).
 */
#ifndef SLX_H
#define SLX_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t slx_status;
#define SLX_OK 0
#define SLX_ERR_INVALID -1     /* bad argument or malformed description */
#define SLX_ERR_DEVICE -2      /* HIP runtime failure */
#define SLX_ERR_UNSUPPORTED -3 /* feature outside the implemented subset */
#define SLX_ERR_NOMEM -4       /* device or host allocation failed */

/* Precision of the numeric phase.  The three MFMA modes run the matrix
 * slices of every bucket that has an MFMA kernel variant (batched kernel for
 * element matrices up to 256 entries with coefficients constant over the
 * Gauss points, factorized kernel for larger ones; the checked-in kernel
 * table may send a class to the FP32 VALU kernel instead) on MFMA with FP32
 * accumulation, and everything else on the FP32 VALU element kernel; the
 * CSR reduction is FP64 in every mode.  The FP16 split also takes the batched
 * kernel for Gauss-point-varying coefficients when the table says so. */
#define SLX_PREC_FP64 0        /* plain FP64 element kernel (FP64 everywhere) */
#define SLX_PREC_F16_SPLIT 1   /* FP16 hi/lo split MFMA (3 products), FP32 accumulation, FP64 reduction */
#define SLX_PREC_BF16_SPLIT 2  /* BF16 hi/mid/lo split MFMA (6 products), FP32 accumulation, FP64 reduction */
#define SLX_PREC_FP32 3        /* FP32 VALU element kernel, FP64 reduction */
#define SLX_PREC_FP32_MFMA 4   /* FP32 MFMA (v_mfma_f32_16x16x4_f32), FP64 reduction */

/* Coefficient program opcodes (one instruction = {op, dst, a, b, imm}). */
enum slx_opcode {
    SLX_OP_CONST = 0,   /* dst = consts[imm] */
    SLX_OP_PARAM = 1,   /* dst = elem_params[imm][element] */
    SLX_OP_JAC = 2,     /* dst = J[a][b]  (a: reference derivative, b: physical coordinate) */
    SLX_OP_INVJAC = 3,  /* dst = inverse(J)[a][b] */
    SLX_OP_DETJ = 4,    /* dst = det(J) (signed) */
    SLX_OP_COORD = 5,   /* dst = physical coordinate a at the Gauss point */
    SLX_OP_FIELD = 6,   /* dst = field imm, space derivative a (0 value, 1..3 d/dx,y,z) */
    SLX_OP_ADD = 7,     /* dst = r[a] + r[b] */
    SLX_OP_MUL = 8,     /* dst = r[a] * r[b] */
    SLX_OP_INV = 9,     /* dst = 1 / r[a] */
    SLX_OP_POW = 10,    /* dst = pow(r[a], r[b]) */
    SLX_OP_ABS = 11,
    SLX_OP_EXP = 12,
    SLX_OP_LOG = 13,
    SLX_OP_SIN = 14,
    SLX_OP_COS = 15,
    SLX_OP_TAN = 16,
    SLX_OP_SQRT = 17,
    SLX_OP_OUT = 18,    /* coefficient slot imm = r[a] (the kernel multiplies by |det J|) */
    SLX_OP_COUNT = 19
};

typedef struct slx_instr {
    int16_t op, dst, a, b;
    int32_t imm;
} slx_instr;

/* One term of a slice: tf derivative x dof derivative with a coefficient
 * register (extractdoftf entries, src/expression/expression.cpp:2253-2281). */
typedef struct slx_term {
    int32_t tf_deriv;   /* 0 value, 1 d/dki, 2 d/deta, 3 d/dphi */
    int32_t dof_deriv;  /* same; ignored for right-hand-side slices */
    int32_t coef;       /* coefficient slot written by SLX_OP_OUT */
} slx_term;

/* A slice = one (tf field, dof field) pair of a contribution inside a bucket. */
typedef struct slx_slice_desc {
    int32_t frag_rank;        /* position of this fragment in the reference
                                 accumulation order (rawmat fragment order) */
    int32_t is_rhs;           /* 1: no dof, contributes -coef*tf to the rhs */
    int32_t n_tf, n_dof;      /* local form functions (rows / columns) */
    const int32_t* tf_addr;   /* [n_tf][n_elem] row-major, -1 = skipped   */
    const int32_t* dof_addr;  /* [n_dof][n_elem], NULL for rhs slices    */
    int32_t n_terms;
    const slx_term* terms;
    int32_t tf_table, dof_table; /* indices into the bucket's shape tables */
    /* Orientation folding (NULL = every sign +1).  A bucket may hold elements of
     * several total orientations when each orientation's basis is a signed
     * renumbering of orientation 0 (hexahedra, quadrangles): the tables are those
     * of orientation 0, the address tables are already renumbered, and
     * sign[i][e] = +1/-1 multiplies row i (tf) or column i (dof) of element e's
     * matrix.  No reference counterpart: the reference keeps one batch per total
     * orientation (elementselector.cpp:4-133); folding gives the same sums. */
    const int8_t* tf_sign;    /* [n_tf][n_elem] or NULL */
    const int8_t* dof_sign;   /* [n_dof][n_elem] or NULL */
} slx_slice_desc;

/* A bucket: elements of one type sharing Gauss rule, orders and (when the
 * basis depends on it) total orientation (elementselector.cpp:4-133). */
typedef struct slx_bucket_desc {
    int32_t elem_type;          /* 1 line, 2 tri, 3 quad, 4 tet, 5 hex, 6 prism */
    int32_t n_elem;
    int32_t n_geo;              /* corner (order-1 geometry) nodes per element */
    const int32_t* geo_nodes;   /* [n_elem][n_geo] node numbers */
    const int32_t* elem_ids;    /* [n_elem] element numbers (index into per-element parameters) */
    int32_t n_gp;
    const double* gp_coords;    /* [n_gp][3] reference coordinates */
    const double* gp_weights;   /* [n_gp] */
    int32_t n_tables;
    const int32_t* table_nff;   /* [n_tables] functions per table */
    const double* const* tables;/* [n_tables] -> [4 derivs][nff][n_gp] */
    int32_t n_slices;
    const slx_slice_desc* slices;
    int32_t n_instr;
    const slx_instr* prog;      /* coefficient program, run per (element, Gauss point) */
    int32_t n_regs;             /* registers the program uses */
    int32_t n_coefs;            /* coefficient slots written by SLX_OP_OUT */
    int32_t q_invariant;        /* 1: program value identical at every Gauss point */
} slx_bucket_desc;

typedef struct slx_plan_desc {
    int32_t mesh_dim;           /* problem (mesh) dimension, 1..3 */
    int64_t n_nodes;
    int64_t n_dofs;             /* dofmanager::countdofs */
    const uint8_t* constrained; /* [n_dofs] 1 = Dirichlet-constrained DOF */
    int32_t n_consts;
    const double* consts;
    int32_t n_buckets;
    const slx_bucket_desc* buckets;
    int32_t n_fields;           /* fields read by SLX_OP_FIELD */
    const int32_t* field_nff;   /* per field: form functions per element, per bucket [n_fields*n_buckets] */
    const int32_t* const* field_addr;   /* [n_fields*n_buckets] -> [nff][n_elem] addresses into field data */
    const double* const* field_tables;  /* [n_fields*n_buckets] -> [4][nff][n_gp] */
    int32_t has_rhs, has_matrix;
    const double* dirichlet;    /* [n_dofs] or NULL: value of every constrained DOF, written into
                                   the constrained rhs entries like formulation::b()
                                   (vec::updateconstraints, src/formulation/vec.cpp:43-61) */
} slx_plan_desc;

/* Numeric inputs of one assembly (host pointers; copied to the device). */
typedef struct slx_numeric_desc {
    const double* node_xyz;       /* [n_nodes][3] */
    int32_t n_params;
    int64_t param_len;            /* entries per parameter array */
    const double* params;         /* [n_params][param_len] per-element values */
    const double* const* field_data;  /* [n_fields] coefficient arrays */
    const int64_t* field_len;     /* [n_fields] */
    int32_t precision;            /* SLX_PREC_* */
} slx_numeric_desc;

typedef struct slx_ctx_s* slx_ctx;
typedef struct slx_plan_s* slx_plan;

const char* slx_last_error(void);

slx_status slx_create(int device, slx_ctx* out);
slx_status slx_destroy(slx_ctx ctx);

/* Symbolic phase: reads the description (which must stay valid for the life
 * of the plan: the numeric phase reads its node count and programs), builds
 * the element-row lists, the
 * CSR pattern (per-row sorted unique columns = sorted COO keys), the A/D
 * split with local renumbering and the reduction maps.  Pattern equals
 * rawmat::process (rawmat.cpp:127-298) on the same address tables.
 * With ctx == NULL only the symbolic phase runs (host memory, no device):
 * the pattern can be read with slx_csr_copyout (value pointers NULL).
 * With a context the phase runs on the device (sorts and per-row column
 * kernels; the same arrays as the host phase, bit for bit); a row of more
 * than 8192 element-matrix entries or 1024 element rows makes it fall back
 * to the host phase. */
slx_status slx_plan_create(slx_ctx ctx, const slx_plan_desc* desc, slx_plan* out);
/* slx_plan_create with flags: SLX_PLAN_HOST_SYMBOLIC runs the symbolic phase
 * on the host even with a context (the reference path for the device phase). */
#define SLX_PLAN_HOST_SYMBOLIC 1
slx_status slx_plan_create_ex(slx_ctx ctx, const slx_plan_desc* desc, int32_t flags, slx_plan* out);
/* Where the plan's symbolic phase ran (1 device, 0 host) and its wall time
 * in seconds (layout, uploads of the bucket data and the phase itself). */
slx_status slx_plan_symbolic_info(slx_plan plan, int32_t* on_device, double* seconds);
slx_status slx_plan_shape(slx_plan plan, int64_t* nfree, int64_t* ncons, int64_t* nnzA, int64_t* nnzD);
/* Numeric phase: uploads the numeric inputs and assembles A, D and b on the
 * plan's stream (asynchronous; slx_plan_sync waits). */
slx_status slx_assemble(slx_plan plan, const slx_numeric_desc* in);
/* Numeric phase on inputs already resident on the device (no host copies,
 * no allocation): device pointers laid out like slx_numeric_desc.  This is
 * the entry the timed loops and graph capture use. */
slx_status slx_assemble_device(slx_plan plan, const double* d_node_xyz, const double* d_params,
                               int64_t param_len, const double* const* d_field_data, int32_t precision,
                               void* hip_stream);
/* slx_assemble on a caller stream (SURVEY's slx_assemble(plan, desc,
 * stream)): the host inputs are copied to the device before it returns, the
 * numeric phase is enqueued on hip_stream (NULL: the plan's stream) and runs
 * asynchronously to the host. */
slx_status slx_assemble_async(slx_plan plan, const slx_numeric_desc* in, void* hip_stream);
slx_status slx_plan_sync(slx_plan plan);
/* mat::getainds / mat::getdinds (src/formulation/mat.cpp:105-106): the DOF
 * numbers of the rows and columns of A (the free DOFs, ascending: nfree
 * entries) and of the columns of D (the constrained DOFs, ascending: ncons
 * entries).  Host memory; either pointer may be NULL.  Pattern-only plans too. */
slx_status slx_plan_dof_maps(slx_plan plan, int64_t* ainds, int64_t* dinds);
/* Captures one numeric phase on these device inputs into a HIP graph owned by
 * the plan (for Newton loops: update the field data in place, relaunch).  The
 * captured work is kernels only (no memset, no allocation, no host sync). */
slx_status slx_plan_graph_capture(slx_plan plan, const double* d_node_xyz, const double* d_params, int64_t param_len,
                                  const double* const* d_field_data, int32_t precision);
/* Launches the captured graph on hip_stream (NULL: the plan's stream). */
slx_status slx_plan_graph_launch(slx_plan plan, void* hip_stream);
/* Kernel timing on the plan's stream with HIP events around every launch:
 * runs `iters` numeric phases on device-resident inputs and returns the
 * average milliseconds per phase of the Gauss-point coefficient kernels
 * (ms[0]), the MFMA element contractions (ms[1]), the VALU element kernels
 * (ms[2]), the CSR value reduction (ms[3]) and the rhs reduction (ms[4]). */
slx_status slx_plan_profile(slx_plan plan, const double* d_node_xyz, const double* d_params, int64_t param_len,
                            const double* const* d_field_data, int32_t precision, int32_t iters, double* ms);
/* Device views of the assembled CSR (valid until slx_plan_destroy). */
slx_status slx_csr_device(slx_plan plan, const int64_t** Arow, const int32_t** Acol, const double** Aval,
                          const int64_t** Drow, const int32_t** Dcol, const double** Dval, const double** rhs);
/* Copies to caller-owned host arrays (any pointer may be NULL). */
slx_status slx_csr_copyout(slx_plan plan, int64_t* Arow, int32_t* Acol, double* Aval, int64_t* Drow,
                           int32_t* Dcol, double* Dval, double* rhs);
/* Segment counts of the reduction: element-row segments, element-matrix
 * entries (including right-hand-side values) and COO entries. */
slx_status slx_plan_stats(slx_plan plan, int64_t* n_elem_rows, int64_t* n_ke, int64_t* n_coo);
/* Buckets whose matrix slices run on the MFMA split kernel (the others and
 * every right-hand-side slice use the VALU element kernel). */
slx_status slx_plan_kernels(slx_plan plan, int32_t* n_mfma_buckets, int32_t* n_buckets);
/* For one precision mode: buckets whose matrix slices run on the batched
 * MFMA kernel, and slices on the factorized MFMA kernel (0 and 0 for the
 * VALU modes). */
slx_status slx_plan_kernel_counts(slx_plan plan, int32_t precision, int32_t* n_batched, int32_t* n_factorized);
/* For one precision mode: matrix slices whose element matrices run on the
 * one-tile-per-element MFMA kernel (small slices with Gauss-point-varying
 * coefficients; the FP16 split, BF16 split and FP32-MFMA modes, 0 for the FP32
 * VALU and FP64 modes). */
slx_status slx_plan_tile_count(slx_plan plan, int32_t precision, int32_t* n_tile);
/* The fused launch of a precision mode: factorized slices of every bucket in
 * one persistent launch (contribution.cpp:61-133 loops over the element
 * groups one after the other; here their elements share one work queue,
 * most expensive first).  n_slices = 0 when the mode runs per-slice launches. */
slx_status slx_plan_fused_info(slx_plan plan, int32_t precision, int32_t* n_slices, int64_t* n_items);
/* Solve of the last assembled system on the device, replacing
 * sl::solve(A, b) (src/expression/operation/sl.cpp:1565-1627): the
 * Dirichlet elimination b_A - D x_D of mat::eliminate
 * (src/formulation/mat.cpp:78-103), a solve with A, and the merge of the
 * free and constrained values of mat::xbmerge (x_out: n_dofs entries, the
 * constrained ones equal to the Dirichlet values).  The reference factorizes
 * A (MUMPS LU / Cholesky); here the solve is iterative: FP64 iterative
 * refinement (FP64 residuals r = b_A - D x_D - A x) around an FP32 inner
 * solve with a Jacobi preconditioner, either conjugate gradients
 * (SLX_SOLVE_CG: symmetric positive definite A) or one GMRES cycle of at most
 * min(max_inner, 100) Krylov vectors per refinement (SLX_SOLVE_GMRES: any
 * nonsingular A).  SLX_SOLVE_AUTO checks A for symmetry (per entry: 1e-5
 * relative plus 1e-6 of the two diagonal entries, so assembly rounding does
 * not count) and picks CG or GMRES.  Stops when ||r|| <= rel_tol ||b_A - D x_D||
 * or after max_outer refinements; max_inner bounds the inner iterations of
 * one refinement.  Returns the counts, the final relative FP64 residual and
 * the method used; not reaching rel_tol is reported through rel_residual,
 * not as an error.  Allocates its device workspace per call.  x_out: host
 * (d_x_out NULL) or device (x_out NULL) memory. */
#define SLX_SOLVE_AUTO 0
#define SLX_SOLVE_CG 1
#define SLX_SOLVE_GMRES 2
slx_status slx_plan_solve(slx_plan plan, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner, double* x_out,
                          double* d_x_out, int32_t* outer_its, int32_t* inner_its, double* rel_residual,
                          int32_t* method_used);
/* slx_plan_solve without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid allocation).  Same arguments, method choice,
 * arithmetic and results as slx_plan_solve; the workspace (256 bytes, then
 * 8 (3 nfree + ncons + n_dofs + 2056) bytes, then 4 (6 nfree + nnzA + 4)
 * bytes, plus for GMRES 4 (restart + 1) nfree + 8 (restart + 1) 1025 bytes)
 * lives in one fixed device array of SLX_SOLVE_SCRATCH_BYTES per GPU, held
 * until the process ends; a system whose workspace exceeds it is rejected
 * with SLX_ERR_INVALID before any device work.  Every call on a GPU shares
 * that array: calls that overlap on one GPU race. */
#define SLX_SOLVE_SCRATCH_BYTES ((size_t)128 << 20)
slx_status slx_plan_solve_noalloc(slx_plan plan, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner,
                                  double* x_out, double* d_x_out, int32_t* outer_its, int32_t* inner_its,
                                  double* rel_residual, int32_t* method_used);
/* mat::eliminate(b) (src/formulation/mat.cpp:78-103): b_A - D b_D of a
 * vector in DOF numbering (n_dofs entries; b NULL: the rhs of the last
 * assembly, whose constrained entries are the Dirichlet values), on the
 * device.  out: nfree entries in the free-DOF order of A.  Host memory. */
slx_status slx_plan_eliminate(slx_plan plan, const double* b, double* out);
/* mat * vec (src/formulation/mat.cpp:182-205): y_A = A x_A + D x_D on the
 * device, and the constrained entries of y are those of x (mat::xbmerge).
 * x, y: n_dofs entries in DOF numbering, host memory. */
slx_status slx_plan_matvec(slx_plan plan, const double* x, double* y);
/* Coefficient programs: 0 (auto, the default) compiles the program of every
 * bucket of at least 4096 elements at plan creation into its own kernel
 * (hipRTC; values in registers instead of the interpreter's scratch register
 * file; the same arithmetic, results equal to rounding), 1 uses the
 * interpreter everywhere, 2
 * compiles every bucket's program.  Returns the number of compiled
 * (bucket, precision) kernels in use; mode 2 fails if any compilation fails.
 * No reference counterpart (the reference evaluates expression trees on the
 * host, operation::interpolate). */
slx_status slx_plan_coef_mode(slx_plan plan, int32_t mode, int32_t* n_compiled);
slx_status slx_plan_destroy(slx_plan plan);
/* Element matrices of the last assembly (n_ke entries, as FP64) in the
 * layout given by slx_plan_reduction_maps: for the element-matrix error of
 * the mixed-precision paths. */
slx_status slx_plan_element_matrices(slx_plan plan, double* out, int64_t cap);
/* Gauss-point coefficient table of the last bucket of the last assembly
 * (as FP64): out[(slot * nq + q) * n_elem + e] = coefficient * |det J|, with
 * nq = 1 when the bucket's coefficients do not vary over the Gauss points.
 * *n_out receives the entry count; out may be NULL to query it. */
slx_status slx_plan_coefficients(slx_plan plan, double* out, int64_t cap, int64_t* n_out);
/* Diagnostic: the reduction maps (pattern-only plans, ctx == NULL) and the
 * element-matrix layout (every plan), so the reduction can be replayed on
 * the host.
 * sizes[0..3] = element-row segments, slot entries, rhs entries, slices.
 * erow_ptr/slot_ptr/rhs_ptr have nfree + 1 entries; erows and rhs_erows are
 * packed (length << 48 | offset); slots are column positions in the row
 * (A columns first, then D), -1 = skipped; ke_base/ke_ld list every
 * (bucket, slice): entry (e, i, j) at ke_base + (e*n_tf + i)*ke_ld + j.
 * Any output pointer may be NULL. */
slx_status slx_plan_reduction_maps(slx_plan plan, int64_t* sizes, int64_t* erow_ptr, uint64_t* erows, int64_t* slot_ptr,
                                   int32_t* slots, int64_t* rhs_ptr, uint64_t* rhs_erows, int64_t* ke_base, int64_t* ke_ld);
/* Diagnostic: a copy of one device array of the symbolic phase, to compare
 * the device and host phases.  which: 0 reduction descriptors (4 words per
 * work position), 1 element-row segments in work order, 2 slots in work
 * order (slot_bytes each), 3 rhs segment pointers, 4 rhs segments, 5 free
 * DOFs, 6 constrained DOFs, 7 row blocks of the solve's product, 8 reduction
 * classes (begin, end, longest row: int64 triples), 9 scalars (int64: slot
 * bytes, longest row, longest segment, segments, COO entries).  *n_bytes
 * receives the size; out may be NULL to query it. */
slx_status slx_plan_symbolic_array(slx_plan plan, int32_t which, void* out, int64_t cap_bytes, int64_t* n_bytes);

/* Diagnostic: runs one v_mfma_f32_16x16x32_f16 on fixed operands and returns
 * the 16x16 result (row-major) so the operand lane layout the kernels assume
 * can be checked against a host product. */
slx_status slx_selftest_mfma_layout(slx_ctx ctx, float* out256);
/* The same for kind 1 (v_mfma_f32_16x16x32_bf16), 2 (v_mfma_f32_16x16x4_f32)
 * and 3 (v_mfma_f64_16x16x4_f64) on A[m][k] = (m + 2k) % 7 - 3,
 * B[k][n] = (3k + n) % 5 - 2 (kind 0 is slx_selftest_mfma_layout). */
slx_status slx_selftest_mfma_layout_kind(slx_ctx ctx, int32_t kind, float* out256);
/* Test hook: the streaming FP32 CSR product of slx_plan_solve's inner
 * iteration on a host CSR (square, nrows rows; values and x rounded to FP32):
 * y = A x and xy = x . y (FP64 sum of the FP32 products). */
slx_status slx_selftest_spmv_stream(slx_ctx ctx, int64_t nrows, const int64_t* rowp, const int32_t* col,
                                    const double* val, const double* x, double* y, double* xy);

/* ---- Reference tables ------------------------------------------------------
 * gausspoints(type, order) (src/gausspoint/gausspoints.cpp:4-54): writes up to
 * `cap` points (3 coordinates each) and weights, *npts = count. */
slx_status slx_gauss_rule(int32_t elem_type, int32_t order, int32_t cap, int32_t* npts, double* coords, double* weights);
/* Number of H1 form functions of order <= `order` on the element
 * (h1*::count, e.g. src/shapefunction/hierarchical/h1/h1triangle.cpp:6-12). */
slx_status slx_h1_count(int32_t elem_type, int32_t order, int32_t* nff);
/* hierarchicalformfunctioncontainer::tomatrix (hierarchicalformfunctioncontainer.cpp:78-112)
 * for all four derivative slots: out[deriv][ff][pt], deriv 0 value, 1..3 d/dki,eta,phi. */
slx_status slx_h1_tables(int32_t elem_type, int32_t order, int32_t total_orientation, int32_t npts, const double* pts,
                         double* out);
/* Number of H(curl) form functions of order <= `order` (order >= 0) on a
 * line, triangle, quadrangle, tetrahedron or hexahedron
 * (hcurl*::count, e.g. src/shapefunction/hierarchical/hcurl/hcurltriangle.cpp:6-14). */
slx_status slx_hcurl_count(int32_t elem_type, int32_t order, int32_t* nff);
/* H(curl) tables (hcurl*::evalat + tomatrix per component): family 1 the
 * function values, family 2 their curls, both in reference coordinates:
 * out[slot][ff][pt], slot 0 zero, slots 1..3 the ki, eta, phi components. */
slx_status slx_hcurl_tables(int32_t elem_type, int32_t order, int32_t total_orientation, int32_t family, int32_t npts,
                            const double* pts, double* out);
/* Geometry (Lagrange) functions of an element type with n_geo nodes per
 * element: the corner nodes (order 1) or every node of a curved order-2
 * element in gmsh order (3 line, 6 triangle, 9 quadrangle, 10 tetrahedron,
 * 27 hexahedron, 18 prism; lagrange*.cpp getnodecoordinates / evalat):
 * out[deriv][node][pt], deriv 0 value, 1..3 d/dki,eta,phi. */
slx_status slx_geometry_tables(int32_t elem_type, int32_t n_geo, int32_t npts, const double* pts, double* out);

/* ---- Formulation layer ---------------------------------------------------- */
typedef struct slx_problem_s* slx_problem;

/* Synthetic workloads of the benchmark suite and small test problems.
 * kind: "poisson2d" (P1 tris, unit square grid n x n, jitter, seed),
 *       "elasticity3d" (P2 tets, h1xyz, n^3 cubes x 6 tets),
 *       "hppoisson3d" (hex, per-element order 1..5),
 *       "quadbatch" (independent quad instances, K + k^2 M),
 *       "newton2d" (P2 tris, nonlinear reluctivity, Newton re-assembly),
 *       "msh" (a gmsh 2.2 file), "electrostatics1d" / "hoprojection2d" (the
 *       reference examples of the same names on their mesh files).
 * The spec string is "key=value;key=value" (n, seed, jitter, order, ...). */
slx_status slx_problem_create(slx_ctx ctx, const char* kind, const char* spec, slx_problem* out);
slx_status slx_problem_destroy(slx_problem p);
/* Build (or reuse) the assembly plan of the problem's formulation. */
slx_status slx_problem_plan(slx_problem p, slx_plan* plan_out);
/* The numeric inputs of the problem (host pointers owned by the problem). */
slx_status slx_problem_numeric(slx_problem p, slx_numeric_desc* out);
/* formulation::generate() + A()/b(): assemble with the given precision. */
slx_status slx_problem_generate(slx_problem p, int32_t precision);
/* Sizes: dofs, free dofs, elements, buckets. */
slx_status slx_problem_info(slx_problem p, int64_t* ndofs, int64_t* nfree, int64_t* nelem, int64_t* nbuckets);
/* The plan description (for the CPU oracle), valid while the problem lives. */
slx_status slx_problem_desc(slx_problem p, const slx_plan_desc** desc);
/* Newton: set the field data used by SLX_OP_FIELD (setdata). */
slx_status slx_problem_setdata(slx_problem p, int32_t field, const double* values, int64_t n);
/* sl::setdata(vec) (src/expression/operation/sl.cpp): the fields of the
 * formulation take their data from a solution vector in DOF numbering. */
slx_status slx_problem_setdata_vec(slx_problem p, const double* x, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* SLX_H */
