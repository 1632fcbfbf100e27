/*
 * cfdgpu.h — C-ABI of the MI355X (gfx950) linear-solver backend for the unstructured
 * finite-volume solver Yipexnn/CFD.
 *
 * The library replaces the device side of the reference's accelerator library
 * (namespace cfd::gpu, src/gpu/include/cfd/gpu/krylov.hpp). Every entry point below names the
 * reference declaration it stands in for. All functions return a cfdg_status; on failure the
 * message is kept per thread and can be read with cfdg_last_error(). Non-convergence is not an
 * error: it is reported in cfdg_solve_stats.converged, as the reference reports it in
 * SolveStats (krylov.hpp:26-31).
 *
 * Arrays are plain host pointers unless the name says "device"; the caller owns them and they
 * are only borrowed for the duration of the call (krylov.hpp:79-103, device_solver.cpp:58-73).
 * Indices are int32 (core/types.hpp:15), values are IEEE binary64 (core/types.hpp:11).
 * A handle may be used by one host thread at a time; different handles are independent.
 * There is no host fallback: without a HIP device every solve returns CFDG_ERR_NO_DEVICE.
 */
#ifndef CFDGPU_H
#define CFDGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFDG_VERSION_MAJOR 0
#define CFDG_VERSION_MINOR 1

typedef enum cfdg_status {
    CFDG_OK = 0,
    CFDG_ERR_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument */
    CFDG_ERR_NO_DEVICE = 2,        /* no HIP device visible */
    CFDG_ERR_HIP = 3,              /* reference: std::runtime_error("CUDA ...") */
    CFDG_ERR_OUT_OF_MEMORY = 4,
    CFDG_ERR_INTERNAL = 5,
    CFDG_ERR_CAPACITY = 6          /* request larger than the admitted footprint */
} cfdg_status;

/* Method (krylov.hpp:19). */
typedef enum cfdg_method { CFDG_METHOD_CG = 0, CFDG_METHOD_BICGSTAB = 1 } cfdg_method;

/* Arithmetic of the AMG-PCG solve. FP64 is the reference algorithm (damped-Jacobi V-cycle,
 * standard PCG; krylov_kernels.inc:182-301). The mixed modes run the V-cycle in FP32
 * (level-0 off-diagonals optionally stored as per-row-scaled BF16) inside an FP64 flexible-CG
 * outer loop; they have no counterpart in the reference. */
typedef enum cfdg_precision {
    CFDG_PREC_FP64 = 0,
    CFDG_PREC_MIXED_FP32 = 1,
    CFDG_PREC_MIXED_BF16 = 2
} cfdg_precision;

/* CsrView (krylov.hpp:10-16): square (or, for transfer operators, rectangular) CSR matrix. */
typedef struct cfdg_csr {
    int n_rows;
    int n_cols; /* 0 means n_cols = n_rows */
    int nnz;
    const int* row_ptr; /* n_rows + 1 */
    const int* col;     /* nnz */
    const double* val;  /* nnz */
} cfdg_csr;

/* SolveControl (krylov.hpp:21-24). */
typedef struct cfdg_solve_control {
    double abs_tolerance; /* stop when the recurrence |r|_2 <= abs_tolerance */
    int max_iterations;
} cfdg_solve_control;

/* SolveStats (krylov.hpp:26-31), extended. */
typedef struct cfdg_solve_stats {
    int iterations;
    int converged;
    int on_device;
    int exit_reason;      /* 0 running, 1 converged, 2 breakdown, 3 iteration limit */
    double residual;      /* recurrence |r|_2 at exit (reference: SolveStats.residual) */
    double true_residual; /* |b - A x|_2 recomputed in FP64 at exit */
    double rhs_norm;      /* |b|_2 */
    double solve_ms;      /* device time of the solve, HIP events on the solve stream */
} cfdg_solve_stats;

/* AmgLevelView / AmgView (krylov.hpp:48-58): per level the operator a and, except on the
 * coarsest level, the prolongation p (n_l x n_{l+1}) and restriction r (n_{l+1} x n_l);
 * coarse_inverse is the dense row-major inverse of the coarsest operator. */
typedef struct cfdg_amg_level_view {
    cfdg_csr a;
    cfdg_csr p;
    cfdg_csr r;
} cfdg_amg_level_view;

typedef struct cfdg_amg_view {
    int n_levels;
    const cfdg_amg_level_view* levels;
    const double* coarse_inverse;
    int smooth_sweeps;    /* reference default 2 */
    double jacobi_weight; /* reference default 0.7 */
} cfdg_amg_view;

/* LduPattern (krylov.hpp:68-77). */
typedef struct cfdg_ldu_pattern {
    int n;
    int n_faces;
    int nnz;
    const int* row_ptr;
    const int* col;
    const int* diag_pos;
    const int* upper_pos;
    const int* lower_pos;
} cfdg_ldu_pattern;

/* ------------------------------------------------------------------------------------------
 * Library / device information
 * ---------------------------------------------------------------------------------------- */

/* Replaces backend_name() (krylov.hpp:119): writes "hip". */
cfdg_status cfdg_backend_name(char* buf, size_t len);
/* Replaces device_available() (krylov.hpp:121). Never fails; *available = 0 without a GPU. */
cfdg_status cfdg_device_available(int* available);
/* Replaces device_description() (krylov.hpp:123), e.g. "hip (device: AMD Instinct MI355X)". */
cfdg_status cfdg_device_description(char* buf, size_t len);
/* Message of the last failed call on this thread (reference: exception what()). */
cfdg_status cfdg_last_error(char* buf, size_t len);
/* Library version as major * 1000 + minor. */
int cfdg_version(void);
/* No counterpart in the reference. Device buffers come from a pool of large slabs that keeps
 * released memory for reuse; this returns the slabs no buffer uses to the device and reports
 * the bytes freed (may be NULL) and the bytes the pool still holds (may be NULL). */
cfdg_status cfdg_release_cached_memory(size_t* freed_bytes, size_t* reserved_bytes);

/* ------------------------------------------------------------------------------------------
 * One-shot solves (device)
 * ---------------------------------------------------------------------------------------- */

/* Replaces solve() (krylov.hpp:37-38; dispatch.cpp:5-13): Jacobi-preconditioned CG or
 * BiCGStab, x is the initial guess on entry and the solution on exit. */
cfdg_status cfdg_solve_csr(const cfdg_csr* a, const double* b, double* x, cfdg_method method,
                           const cfdg_solve_control* control, cfdg_solve_stats* stats);

/* Replaces solve_amg_cg() (krylov.hpp:61-62; dispatch.cpp:15-23): CG preconditioned by one
 * V-cycle of the given hierarchy per iteration, A = levels[0].a. */
cfdg_status cfdg_solve_amg_cg(const cfdg_amg_view* amg, const double* b, double* x,
                              const cfdg_solve_control* control, cfdg_solve_stats* stats);

/* Same, with the V-cycle precision selectable (no counterpart in the reference). */
cfdg_status cfdg_solve_amg_cg_prec(const cfdg_amg_view* amg, const double* b, double* x,
                                   cfdg_precision precision, const cfdg_solve_control* control,
                                   cfdg_solve_stats* stats);

/* Allocation-free variant of cfdg_solve_csr. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Every array is a DEVICE pointer of the current device: the
 * matrix (a->row_ptr[n_rows + 1], a->col[nnz], a->val[nnz]), b[n_rows] and x[n_rows] (initial
 * guess in, solution out); the structs themselves are host memory. The matrix is checked on the
 * device before use (CFDG_ERR_INVALID_ARGUMENT). Work vectors, partial sums, the solve state and
 * the launch table live in a static device array of the library, CFDG_NOALLOC_SCRATCH_BYTES on
 * every device the library runs on, held until the process ends; more than CFDG_NOALLOC_MAX_ROWS
 * rows are rejected with CFDG_ERR_CAPACITY. Same arithmetic as cfdg_solve_csr (bit-identical
 * results). Runs on the null stream of the current device and returns when it has finished;
 * stats->solve_ms is host wall time. Every call on a device shares that device's array, which
 * keeps the last call's contents: calls that overlapped on one device would race on it, so the
 * library serialises these calls within a process (one lock); each process has its own array. */
#define CFDG_NOALLOC_MAX_ROWS 524288
#define CFDG_NOALLOC_SCRATCH_BYTES 51380224
cfdg_status cfdg_solve_csr_noalloc(const cfdg_csr* a, const double* b, double* x, cfdg_method method,
                                   const cfdg_solve_control* control, cfdg_solve_stats* stats);

/* Replaces assemble_convection_diffusion() (krylov.hpp:105-116; host loop
 * krylov_openmp.cpp:289-300). upper/lower are written, diag/divergence are accumulated.
 * The device kernel gathers per cell in ascending face order, so every cell sees its faces in
 * the order of the reference's serial loop and the result equals it exactly. */
cfdg_status cfdg_assemble_convection_diffusion(int n_cells, int n_faces, const int* owner,
                                               const int* neighbour, const double* flux,
                                               const double* gamma, const double* delta_coeff,
                                               double convection, double* upper, double* lower,
                                               double* diag, double* divergence);

/* ------------------------------------------------------------------------------------------
 * Persistent system: replaces class DeviceSolver (krylov.hpp:79-103, device_solver.cpp)
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_system cfdg_system;

/* DeviceSolver(const LduPattern&) (device_solver.cpp:58-73). */
cfdg_status cfdg_system_create(const cfdg_ldu_pattern* pattern, cfdg_system** out);
/* ~DeviceSolver() (device_solver.cpp:75-81). NULL is accepted. */
cfdg_status cfdg_system_destroy(cfdg_system* sys);
/* DeviceSolver::on_device() (device_solver.cpp:83-85). */
cfdg_status cfdg_system_on_device(const cfdg_system* sys, int* on_device);
/* DeviceSolver::set_amg() (device_solver.cpp:87-115): levels[0].a is ignored (the assembled
 * matrix is used), everything else is copied to the device now. An empty hierarchy is
 * CFDG_ERR_INVALID_ARGUMENT, as the reference throws std::invalid_argument. */
cfdg_status cfdg_system_set_amg(cfdg_system* sys, const cfdg_amg_view* amg);
/* DeviceSolver::solve() (device_solver.cpp:117-150): uploads diag (n), upper / lower
 * (n_faces), b and x, assembles A on the device and solves; x receives the solution. */
cfdg_status cfdg_system_solve(cfdg_system* sys, const double* diag, const double* upper,
                              const double* lower, const double* b, double* x,
                              cfdg_method method, int use_amg, const cfdg_solve_control* control,
                              cfdg_solve_stats* stats);

/* Resident use (no counterpart in the reference): the matrix, right-hand side and solution
 * stay in HBM between calls, so repeated solves move no data over PCIe. */
cfdg_status cfdg_system_assemble(cfdg_system* sys, const double* diag, const double* upper,
                                 const double* lower);
cfdg_status cfdg_system_set_rhs(cfdg_system* sys, const double* b);
cfdg_status cfdg_system_set_x(cfdg_system* sys, const double* x); /* NULL: x = 0 */
cfdg_status cfdg_system_get_x(cfdg_system* sys, double* x);
cfdg_status cfdg_system_solve_resident(cfdg_system* sys, cfdg_method method, int use_amg,
                                       cfdg_precision precision,
                                       const cfdg_solve_control* control,
                                       cfdg_solve_stats* stats);
/* Replaces LduMatrix::normalised_residual / residual_norms (ldu_matrix.cpp:65-103) for the
 * resident matrix, right-hand side and iterate, computed on the device: res = |b - A x|_1,
 * norm = sum_c |(A x)_c - (A xref)_c| + |b_c - (A xref)_c| with A xref = rowsum(A) mean(x),
 * normalised = res / (norm + 1e-20). Sums in a fixed order (reproducible run to run; they agree
 * with the reference's serial face-order sums to rounding). Output pointers may be NULL. */
cfdg_status cfdg_system_normalised_residual(cfdg_system* sys, double* normalised, double* res,
                                            double* norm);
/* Kernel timing for benchmarks (no counterpart in the reference). kind 1 brackets every launch
 * of the last level-0 Jacobi sweep of the V-cycle, kind 2 every outer FP64 q = A p, with HIP
 * events on the system's stream during the following single-system solves; 0 turns it off.
 * Setting a kind resets the accumulators. */
cfdg_status cfdg_system_probe(cfdg_system* sys, int kind);
/* Average duration of the probed launches that did work, their number, and the algorithmic
 * bytes of one launch (stored matrix entries x (value + 4-byte index) + vectors read and written
 * once) for the given precision. */
cfdg_status cfdg_system_probe_read(const cfdg_system* sys, int precision, double* avg_ms,
                                   long long* launches, double* bytes_per_launch);
/* Device bytes held by the system (matrices, hierarchy, work vectors). */
cfdg_status cfdg_system_footprint(const cfdg_system* sys, size_t* bytes);

/* ------------------------------------------------------------------------------------------
 * Batched solves (no counterpart in the reference): many resident systems of different sizes
 * solved together, one launch per solver stage for all of them. Each system keeps its own
 * reductions, scalars and stopping test, so its result does not depend on the other members.
 * The systems must stay alive and unchanged in structure (set_amg) while the batch exists;
 * their matrices, right-hand sides and initial guesses may be updated between solves.
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_batch cfdg_batch;

cfdg_status cfdg_batch_create(cfdg_system* const* systems, int count, cfdg_method method,
                              int use_amg, cfdg_precision precision, cfdg_batch** out);
cfdg_status cfdg_batch_destroy(cfdg_batch* batch);
/* controls and stats hold one entry per system, in creation order. */
cfdg_status cfdg_batch_solve(cfdg_batch* batch, const cfdg_solve_control* controls,
                             cfdg_solve_stats* stats);
/* Sets every member's initial guess to zero with one launch on the batch's stream (what
 * cfdg_system_set_x(member, NULL) does member by member, without a launch per member). */
cfdg_status cfdg_batch_zero_x(cfdg_batch* batch);

/* ------------------------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference): a request is one mesh's system with a run of
 * time-step right-hand sides. Requests are admitted in arrival order while their estimated
 * device footprint fits the memory budget; an admitted request's system and hierarchy stay
 * resident for all of its steps, each step starting from the previous step's solution.
 * Solves of different requests share every kernel launch (continuous batching: a solve joins
 * or leaves the running set at any iteration boundary) and keep their own state, so every
 * result is bit-identical to the one the request gets when served alone.
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_server cfdg_server;

typedef struct cfdg_server_options {
    int max_active_solves;       /* solves sharing the launches (slots), default 64 */
    cfdg_method method;          /* default CFDG_METHOD_CG */
    int use_amg;                 /* default 1 */
    cfdg_precision precision;    /* default CFDG_PREC_MIXED_FP32 */
    int smooth_sweeps;           /* V-cycle sweeps every hierarchy must use, default 2 */
    size_t memory_budget_bytes;  /* resident requests' device bytes; 0: 90% of device memory */
    int staging_threads;         /* host threads uploading and retiring requests, default 4 */
} cfdg_server_options;

typedef struct cfdg_server_request {
    cfdg_ldu_pattern pattern;
    const double* diag;          /* operator coefficients (n, n_faces, n_faces) */
    const double* upper;
    const double* lower;
    const cfdg_amg_view* amg;    /* hierarchy (level 0's matrix unused); NULL without AMG */
    int n_steps;
    const double* const* rhs;    /* n_steps right-hand sides of n entries */
    const double* x0;            /* initial guess of the first step; NULL: zero */
    double rel_tolerance;        /* each step stops at |r| <= rel_tolerance |b| */
    int max_iterations;
    /* Outputs, written before cfdg_server_wait returns (any may be NULL): */
    double* const* x_out;        /* n_steps solutions (entries may be NULL) */
    cfdg_solve_stats* stats_out; /* n_steps */
    double* step_latency_ms;     /* n_steps: from the step being ready (the system resident in
                                    HBM for the first, the previous step's result after) to its
                                    own result; waiting for a free slot is included, staging is
                                    not (cfdg_server_wait's latency runs from submission) */
} cfdg_server_request;

typedef struct cfdg_server_counters {
    unsigned long long submitted;
    unsigned long long completed;
    unsigned long long solves;
    size_t resident_bytes;
    size_t peak_resident_bytes;
    size_t budget_bytes;
    int active_solves;
    int peak_active_solves;
    int waiting;                 /* submitted requests not yet resident */
    double staging_ms;           /* summed over staging threads: making requests resident */
    double retiring_ms;          /* summed over staging threads: returning results */
    double engine_busy_ms;       /* engine thread time with at least one solve in flight */
} cfdg_server_counters;

/* ------------------------------------------------------------------------------------------
 * Finite-volume cell kernels (replaces gpu::DeviceFv, fv_device.hpp:14-71): cell gradients
 * (Green-Gauss, weighted least squares, skewness-corrected Gauss), the Barth-Jespersen /
 * Venkatakrishnan limiter and the explicit face corrections of the scalar and momentum
 * equations. The mesh is copied to the device once; each call moves only the fields. Every cell
 * gathers over its faces in ascending face order, so results equal the reference's serial face
 * loops (no atomics, deterministic).
 * ---------------------------------------------------------------------------------------- */

/* FvMeshView (fv_device.hpp:16-32). Faces [0, n_internal_faces) are internal, boundary faces
 * follow; vectors are 3 doubles per entry. */
typedef struct cfdg_fv_mesh {
    int n_cells;
    int n_internal_faces;
    int n_faces;
    const int* owner;          /* n_faces */
    const int* neighbour;      /* n_internal_faces */
    const int* cell_offsets;   /* n_cells + 1: faces of each cell (CSR) ... */
    const int* cell_faces;     /* ... in ascending face order */
    const double* weight;      /* n_internal_faces: phi_f = w phi_P + (1 - w) phi_N */
    const double* sf;          /* 3 n_faces: face area vectors */
    const double* face_centre; /* 3 n_faces */
    const double* cell_centre; /* 3 n_cells */
    const double* volume;      /* n_cells */
    const double* correction;  /* 3 n_faces: non-orthogonal correction vectors */
    const double* lsq_d;       /* 3 n_faces: d / |d|^2 (d = C_N - C_P or C_f - C_P) */
    const double* lsq_inverse; /* 9 n_cells: (sum d d^T / |d|^2)^-1, column-major */
} cfdg_fv_mesh;

typedef struct cfdg_fv cfdg_fv;

/* Replaces DeviceFv::DeviceFv (fv_device.hpp:36, fv_device.cpp:24-59): checks the sizes,
 * copies the mesh to the device. Errors as the reference: CFDG_ERR_INVALID_ARGUMENT for invalid
 * sizes or a missing array. */
cfdg_status cfdg_fv_create(const cfdg_fv_mesh* mesh, cfdg_fv** out);
cfdg_status cfdg_fv_destroy(cfdg_fv* fv);
/* Replaces DeviceFv::gradient (fv_device.hpp:43-45): scheme 0 Green-Gauss, 1 least squares,
 * 2 skewness-corrected Gauss; phi n_cells, bv n_boundary, grad 3 n_cells. */
cfdg_status cfdg_fv_gradient(cfdg_fv* fv, const double* phi, const double* bv, int scheme,
                             double* grad);
/* Replaces DeviceFv::vector_gradient (fv_device.hpp:46-50): u 3 n_cells, ub 3 n_boundary, grad
 * 9 n_cells, column-major 3x3 per cell with grad(j, i) = du_i / dx_j. u, ub and the gradient
 * stay on the device for cfdg_fv_momentum_corrections. */
cfdg_status cfdg_fv_vector_gradient(cfdg_fv* fv, const double* u, const double* ub, int scheme,
                                    double* grad);
/* Replaces DeviceFv::scalar_corrections (fv_device.hpp:51-56): source (n_cells) += explicit
 * internal-face fluxes; flux and gamma per internal face. */
cfdg_status cfdg_fv_scalar_corrections(cfdg_fv* fv, const double* phi, const double* bv, int scheme,
                                       const double* flux, const double* gamma, int linear_upwind,
                                       double* source);
/* Replaces DeviceFv::momentum_corrections (fv_device.hpp:57-66) with the field and gradient of
 * the last cfdg_fv_vector_gradient (CFDG_ERR_INVALID_ARGUMENT before one, as the reference's
 * logic_error); eps2: 3 values, read only when venkat. source 3 n_cells. */
cfdg_status cfdg_fv_momentum_corrections(cfdg_fv* fv, const double* flux, const double* nu_face,
                                         int linear_upwind, int limited, int venkat,
                                         const double* eps2, double central, int transpose_term,
                                         double* source);
/* Pressure equation of a SIMPLE / PISO pressure correction, assembled on the device: replaces
 * the host loops of FlowSolver::pressure_correction (flow_solver.cpp:688-796): H(U)/A and
 * rAU = V / mean(A) from the momentum matrix (LduMatrix::off_diagonal_product,
 * ldu_matrix.cpp:39-49), the predicted face flux phiHbyA and rAU_f with the boundary kinds, the
 * optional relaxation-independent flux correction (Majumdar), the Laplacian coefficients and
 * div(phiHbyA) of gpu::assemble_convection_diffusion with convection 0 (krylov_openmp.cpp:257-301,
 * in face order), the boundary terms and the pin LduMatrix::set_reference(0, 0)
 * (ldu_matrix.cpp:105-109). Boundary kinds follow FlowBoundary (flow_solver.hpp:16). */
enum {
    CFDG_BC_VELOCITY_INLET = 0,
    CFDG_BC_WALL = 1,
    CFDG_BC_PRESSURE_OUTLET = 2,
    CFDG_BC_SYMMETRY = 3,
    CFDG_BC_EMPTY = 4
};

typedef struct cfdg_pressure_inputs {
    const double* upper;           /* n_internal_faces: momentum matrix off-diagonals (sys.base) */
    const double* lower;           /* n_internal_faces */
    const double* diag;            /* 3 n_cells: per-component diagonal (sys.diag), x y z per cell */
    const double* source;          /* 3 n_cells: per-component source without grad p (sys.source) */
    const double* u;               /* 3 n_cells: velocity */
    const double* delta_coeffs;    /* n_faces */
    const int* boundary_kind;      /* n_boundary faces: CFDG_BC_* */
    const double* fixed_velocity;  /* 3 n_boundary: inlet / wall velocity (fixed_velocity_) */
    const double* outlet_pressure; /* n_boundary: outlet pressure (BoundaryCondition::pressure) */
    const double* u_prev;          /* 3 n_cells, or NULL: with phi_prev and alpha_u < 1 the */
    const double* phi_prev;        /* n_faces, or NULL:   relaxation-independent face flux */
    double alpha_u;
    int pin_reference;             /* 1: set_reference(0, 0.0) (no fixed-pressure boundary) */
} cfdg_pressure_inputs;

typedef struct cfdg_pressure_outputs { /* every pointer may be NULL: not copied back */
    double* diag;     /* n_cells: pressure matrix */
    double* upper;    /* n_internal_faces */
    double* lower;    /* n_internal_faces */
    double* source;   /* n_cells */
    double* phi_hbya; /* n_faces */
    double* rau_f;    /* n_faces */
    double* hbya;     /* 3 n_cells */
    double* rau;      /* n_cells */
} cfdg_pressure_outputs;

/* Assembles the pressure equation on the device; the results stay resident in the object (for
 * cfdg_system_assemble_pressure) and are copied to the non-NULL outputs. Results equal the
 * serial face loops bit for bit. CFDG_ERR_INVALID_ARGUMENT for a missing array or a boundary
 * kind out of range. */
cfdg_status cfdg_fv_pressure_equation(cfdg_fv* fv, const cfdg_pressure_inputs* in,
                                      cfdg_pressure_outputs* out);
/* No counterpart in the reference: the system's matrix and right-hand side become the pressure
 * equation last assembled by fv, copied within HBM (no PCIe transfer). The system must have been
 * created from the pattern of fv's mesh (same cells and internal faces, in the same order). */
cfdg_status cfdg_system_assemble_pressure(cfdg_system* sys, const cfdg_fv* fv);
/* Replaces DeviceFv::on_device (fv_device.hpp:68): 1. */
cfdg_status cfdg_fv_on_device(const cfdg_fv* fv, int* on_device);
/* No counterpart in the reference: device time (ms) of the kernels of the last call on this
 * object when timing is on (HIP events around each launch; off by default). */
cfdg_status cfdg_fv_timing(cfdg_fv* fv, int enable);
cfdg_status cfdg_fv_last_kernel_ms(const cfdg_fv* fv, double* ms);

/* ------------------------------------------------------------------------------------------
 * Layout tuning (no counterpart in the reference). Coarse operators and transfer matrices are
 * stored as sliced ELL or CSR with `lanes` lanes per row. The library's built-in rule picks
 * them from the mean row length; a tuned table (written by python -m cfd_amd.autotune from
 * timings of cfdg_spmv_time on this GPU) overrides it for the matrices it covers. The table is
 * process-wide and applies to hierarchies uploaded after it is set.
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_layout_rule {
    int min_rows;
    int max_rows;
    double min_mean;  /* mean entries per row */
    double max_mean;
    int lanes;        /* power of two <= 64 */
    int csr;          /* 0 sliced ELL, 1 CSR */
} cfdg_layout_rule;

/* count 0 restores the built-in rule. */
cfdg_status cfdg_set_layout_rules(const cfdg_layout_rule* rules, int count);
/* The layout the library picks for a matrix with these row offsets (rules, then built-in). */
cfdg_status cfdg_layout_of(int n_rows, const int* row_ptr, int* lanes, int* csr);
/* Smoother of the mixed-precision V-cycle: damped-Jacobi sweeps per side (default 1) and the
 * ratio of the damping interval [h / ratio, h] (default 6), h from a per-level estimate of the
 * largest eigenvalue of D^-1 A. Ratio 0 selects the hierarchy's own damping weight instead
 * (jacobi_weight, the reference smoother), so that with 2 sweeps the mixed V-cycle is the FP64
 * path's V-cycle in FP32. The ratio applies to hierarchies set afterwards. */
cfdg_status cfdg_set_mixed_smoother(int sweeps, double ratio);
/* Average device time of y = A x (fp32 != 0: FP32 values, else FP64) over reps launches with
 * the given layout. */
cfdg_status cfdg_spmv_time(const cfdg_csr* a, int lanes, int csr, int fp32, int reps, double* avg_ms);

/* Upper bound of the device bytes a system of this pattern and hierarchy (may be NULL) holds;
 * host-only, usable for capacity planning without a device. */
cfdg_status cfdg_estimate_footprint(const cfdg_ldu_pattern* pattern, const cfdg_amg_view* amg,
                                    size_t* bytes);
cfdg_status cfdg_server_default_options(cfdg_server_options* out);
cfdg_status cfdg_server_create(const cfdg_server_options* options, cfdg_server** out);
/* Every array the request points at must stay valid until cfdg_server_wait returns for it. */
cfdg_status cfdg_server_submit(cfdg_server* server, const cfdg_server_request* request,
                               unsigned long long* id);
/* Blocks until the request has completed; latency_ms (may be NULL) receives arrival to last
 * result. Returns the request's own status (e.g. CFDG_ERR_CAPACITY when it can never fit). */
cfdg_status cfdg_server_wait(cfdg_server* server, unsigned long long id, double* latency_ms);
cfdg_status cfdg_server_query(cfdg_server* server, cfdg_server_counters* out);
/* Completes every submitted request, then stops the server. */
cfdg_status cfdg_server_destroy(cfdg_server* server);

/* ------------------------------------------------------------------------------------------
 * AMG hierarchy setup on the host (replaces fv::AmgPreconditioner::setup, amg.cpp:122-190,
 * and coarse_inverse, amg.cpp:227-240, which the reference runs through Eigen in its caller)
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_amg cfdg_amg;

/* AmgSettings (fv/include/cfd/fv/amg.hpp:17-23) plus the V-cycle parameters of AmgView. */
typedef struct cfdg_amg_settings {
    double strength_threshold; /* 0.08 */
    int max_levels;            /* 25 */
    int coarse_size;           /* 400 */
    int smooth_sweeps;         /* 2 */
    double jacobi_weight;      /* 0.7 */
    double strength_decay;     /* threshold factor per level: level l uses
                                  strength_threshold * strength_decay^l. 1.0 is the reference
                                  (fixed threshold); 0.5 is the halving per level of Vanek,
                                  Mandel & Brezina (1996), which the reference cites. */
} cfdg_amg_settings;

cfdg_status cfdg_amg_default_settings(cfdg_amg_settings* s);
cfdg_status cfdg_amg_build(const cfdg_csr* a, const cfdg_amg_settings* settings,
                           cfdg_amg** out);
/* Replaces the same setup (amg.cpp:122-190) with the sparse products of every level (smoothed
 * prolongation, transpose, Galerkin product, pruning) on the GPU and the greedy aggregation
 * (amg.cpp:12-78, sequential by definition) on the host. The hierarchy is bit-identical to
 * cfdg_amg_build's; CFDG_ERR_NO_DEVICE without a GPU. */
cfdg_status cfdg_amg_build_device(const cfdg_csr* a, const cfdg_amg_settings* settings,
                                  cfdg_amg** out);
cfdg_status cfdg_amg_destroy(cfdg_amg* amg);
cfdg_status cfdg_amg_n_levels(const cfdg_amg* amg, int* n_levels);
/* which: 0 = a, 1 = p, 2 = r. Sizes of a level's matrix (nnz = 0 on the coarsest p / r). */
cfdg_status cfdg_amg_level_size(const cfdg_amg* amg, int level, int which, int* n_rows,
                                int* n_cols, int* nnz);
/* Copies a level's CSR arrays out (any of the three pointers may be NULL). */
cfdg_status cfdg_amg_level_copy(const cfdg_amg* amg, int level, int which, int* row_ptr,
                                int* col, double* val);
/* Aggregate index per row of a level (the aggregation that produced level + 1). */
cfdg_status cfdg_amg_aggregates(const cfdg_amg* amg, int level, int* agg);
/* Dense row-major inverse of the coarsest operator (n_c * n_c doubles). */
cfdg_status cfdg_amg_coarse_inverse(const cfdg_amg* amg, double* out);
/* Borrowed view of the hierarchy (valid while amg lives); levels must hold n_levels slots. */
cfdg_status cfdg_amg_view_of(const cfdg_amg* amg, cfdg_amg_level_view* levels,
                             cfdg_amg_view* view);
/* Setup wall time of the last build in milliseconds. */
cfdg_status cfdg_amg_setup_ms(const cfdg_amg* amg, double* ms);
/* 1 when the hierarchy's coarsest inverse was computed on the matrix cores (cfdg_amg_build_device
 * inverts it with cfdg_dense_inverse_batch's kernel), 0 when on the host. */
cfdg_status cfdg_amg_coarse_on_device(const cfdg_amg* amg, int* on_device);

/* Batched dense inversion of coarsest-level operators (no single counterpart in the reference,
 * which inverts its one coarsest operator on the host: n_c LDL^T solves, amg.cpp:185-189,
 * 227-240). count row-major host matrices a[i] of order order[i] are inverted into inv[i]; on the
 * device, one workgroup per matrix runs block Gauss-Jordan elimination in 16 x 16 tiles on the
 * FP64 matrix cores, for orders up to 512. Matrices the device does not take (larger orders, or a
 * zero or non-finite pivot, where LDL^T fails too) are inverted by the host path, which falls back
 * to the pseudo-inverse as the reference does. on_device[i] (may be NULL): 1 when the device
 * produced inv[i]. */
cfdg_status cfdg_dense_inverse_batch(int count, const int* order, const double* const* a,
                                     double* const* inv, int* on_device);

/* ------------------------------------------------------------------------------------------
 * The reference's caller policy for the device solver: LduMatrix::solve with a CSR pattern and
 * an AmgCache (ldu_matrix.cpp:111-131 and solve_persistent 287-366; AmgCache
 * ldu_matrix.hpp:61-75), so that a caller without the reference's fv library gets the same
 * behaviour: the initial normalised residual and the early return below the tolerance, the
 * tolerance mapping target = max(rel_tolerance, tolerance / initial) * |b - A x0|_2, the
 * hierarchy rebuilt from the current matrix when missing or max_age solves old, AMG-PCG only
 * when the hierarchy has more than one level (else Jacobi CG), BiCGStab without AMG, the
 * hierarchy otherwise stale (built from an older matrix). Optionally the host AmgCg path's
 * stale-hierarchy rule (ldu_matrix.cpp:225-258): a solve that needs more than twice the
 * iterations of the first solve after the last setup, plus 2, rebuilds the hierarchy and
 * continues from the current iterate.
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_amg_cache cfdg_amg_cache;

typedef struct cfdg_amg_cache_options {
    int max_age;                /* rebuild after this many solves (ldu_matrix.hpp:74: 20) */
    int stale_rebuild;          /* 0: solve_persistent (the device path); 1: also the rule above */
    cfdg_amg_settings amg;      /* hierarchy settings (cfdg_amg_default_settings) */
    int device_setup;           /* 1 (default): rebuilds run the setup's sparse products and the
                                   coarsest inverse on the GPU (cfdg_amg_build_device: the same
                                   hierarchy, the inverse equal to rounding); 0: the host setup */
} cfdg_amg_cache_options;

/* LinearSolverSettings (ldu_matrix.hpp:23-28) for the device solver types. */
typedef struct cfdg_linear_settings {
    cfdg_method method;         /* CG: GpuCg, BICGSTAB: GpuBiCgStab */
    double tolerance;           /* absolute, on the normalised residual (reference default 1e-10) */
    double rel_tolerance;       /* relative to the initial normalised residual (default 0.01) */
    int max_iterations;         /* default 1000 */
    cfdg_precision precision;   /* V-cycle precision of AMG-PCG; FP64 is the reference */
} cfdg_linear_settings;

/* LinearSolveResult (ldu_matrix.hpp:30-34) plus what the policy did. */
typedef struct cfdg_linear_result {
    double initial_residual;    /* normalised, before the solve */
    double final_residual;      /* normalised, after the solve */
    int iterations;             /* Krylov iterations (both solves when the stale rule rebuilt) */
    int solved;                 /* 0: returned before solving (initial residual < tolerance) */
    int used_amg;               /* AMG-PCG (1) or the method's Jacobi-preconditioned solver (0) */
    int setups;                 /* hierarchies built during this call */
    double target;              /* absolute tolerance on |r|_2 handed to the device solver */
    cfdg_solve_stats stats;     /* of the last device solve of the call */
} cfdg_linear_result;

cfdg_status cfdg_amg_cache_default_options(cfdg_amg_cache_options* out);
/* The cache owns a device system on the pattern (created here) and the current hierarchy. */
cfdg_status cfdg_amg_cache_create(const cfdg_ldu_pattern* pattern, const cfdg_amg_cache_options* options,
                                  cfdg_amg_cache** out);
cfdg_status cfdg_amg_cache_destroy(cfdg_amg_cache* cache);
/* AmgCache::age, the levels of the current hierarchy (0: none yet) and hierarchies built so far. */
cfdg_status cfdg_amg_cache_state(const cfdg_amg_cache* cache, int* age, int* n_levels, int* setups);
/* LduMatrix::solve (ldu_matrix.cpp:111-131, 287-366): host LDU coefficients and source, x in/out. */
cfdg_status cfdg_ldu_solve(cfdg_amg_cache* cache, const double* diag, const double* upper,
                           const double* lower, const double* source, double* x,
                           const cfdg_linear_settings* settings, cfdg_linear_result* result);

/* ------------------------------------------------------------------------------------------
 * Case generators (restate the reference's box mesher and FV geometry so that the benchmark
 * systems can be synthesised; make_box_mesh box_mesher.cpp:73-145, Mesh::compute_geometry
 * mesh.cpp:100-186, face weights fv_geometry.cpp:52-58)
 * ---------------------------------------------------------------------------------------- */

typedef struct cfdg_box_spec {
    int nx, ny, nz;
    double lo[3];
    double hi[3];
    double jitter;  /* interior vertices moved by U[-jitter, jitter] * spacing per axis */
    uint64_t seed;  /* mt19937_64 seed of the jitter */
} cfdg_box_spec;

cfdg_status cfdg_box_sizes(const cfdg_box_spec* spec, int* n_cells, int* n_internal_faces);
/* Internal faces in the reference order (owner ascending, then neighbour); per face the
 * coefficient |S_f| / |C_N - C_P| and the interpolation weight; per cell volume and centre
 * (3 doubles). Any output pointer except owner / neighbour may be NULL. */
cfdg_status cfdg_box_build(const cfdg_box_spec* spec, int* owner, int* neighbour,
                           double* area_over_distance, double* weight, double* cell_volume,
                           double* cell_centre);
/* n draws of mt19937_64(seed) mapped to lo + (hi - lo) * (u >> 11) * 2^-53. */
cfdg_status cfdg_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out);
/* CSR pattern of the cell graph with value positions (build_csr_pattern,
 * ldu_matrix.cpp:572-612). row_ptr n+1, col n+2*n_faces. */
cfdg_status cfdg_csr_pattern(int n, int n_faces, const int* owner, const int* neighbour,
                             int* row_ptr, int* col, int* diag_pos, int* upper_pos,
                             int* lower_pos);

#ifdef __cplusplus
}
#endif

#endif /* CFDGPU_H */
