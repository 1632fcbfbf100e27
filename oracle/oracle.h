/*
 * oracle.h — CPU restatement of the reference's pressure / momentum linear-solver path.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker or the timed CPU baseline; the product path
 * (cfd_amd/, libcfdgpu) never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to the Yipexnn/CFD
 * repository). Arithmetic is written in the reference's order and the library is built with
 * -ffp-contract=off, so a serial run reproduces the reference's host results bit for bit where
 * the reference itself is serial (n <= 20000, src/gpu/src/krylov_openmp.cpp:21-25).
 *
 * How the Eigen-based parts are pinned (Eigen, and so the reference's fv library, is absent):
 * orc_sgs_pcg's V-cycle equals the dense two-level symmetric Gauss-Seidel operator and its outer
 * loop equals scipy's preconditioned CG with that operator (tests/test_oracle_pins.py); the
 * setup's sparse product follows hand-derived known answers in which any other summation order
 * keeps or drops an exact zero, and the whole setup equals dense brute force. Still parity
 * unpinned: bit-exact iteration counts against Eigen's vectorised norms in SGS-PCG, and
 * orc_pressure_equation's Vec3 dot / mean order against Eigen's evaluation (pinned only by the
 * mathematical checks of tests/test_pressure_oracle.py).
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Thread count used by the OpenMP loops (the reference threads loops with n > 20000). */
void orc_set_threads(int n);

/* ---- LDU matrix (src/fv/src/ldu_matrix.cpp) ---- */
void orc_ldu_multiply(int n, int nf, const int* own, const int* nei, const double* diag,
                      const double* upper, const double* lower, const double* x, double* y);
void orc_ldu_residual(int n, int nf, const int* own, const int* nei, const double* diag,
                      const double* upper, const double* lower, const double* source,
                      const double* x, double* r);
double orc_ldu_normalised_residual(int n, int nf, const int* own, const int* nei,
                                   const double* diag, const double* upper, const double* lower,
                                   const double* source, const double* x);
void orc_ldu_set_reference(double* diag, double* source, int cell, double value);
int orc_csr_pattern(int n, int nf, const int* own, const int* nei, int* rp, int* ci,
                    int* diag_pos, int* upper_pos, int* lower_pos);
void orc_ldu_to_csr(int n, int nf, const int* diag_pos, const int* upper_pos,
                    const int* lower_pos, const double* diag, const double* upper,
                    const double* lower, double* val);
void orc_assemble(int n_cells, int n_faces, const int* own, const int* nei, const double* flux,
                  const double* gamma, const double* delta, double convection, double* upper,
                  double* lower, double* diag, double* divergence);

/* ---- Krylov kernels (src/gpu/src/krylov_kernels.inc) ---- */
typedef struct orc_stats {
    int iterations;
    int converged;
    double residual;
} orc_stats;

void orc_spmv(int n, const int* rp, const int* ci, const double* va, const double* x, double* y);
double orc_dot(int n, const double* x, const double* y);
/* method 0 = Jacobi CG, 1 = Jacobi BiCGStab (solve_openmp, krylov_openmp.cpp:93-138). */
orc_stats orc_solve(int n, const int* rp, const int* ci, const double* va, const double* b,
                    double* x, int method, double abs_tol, int max_it);

/* ---- AMG hierarchy (src/fv/src/amg.cpp) ---- */
typedef struct orc_amg orc_amg;
int orc_aggregate(int n, const int* rp, const int* ci, const double* va, const double* diag,
                  double theta, int* agg);
orc_amg* orc_amg_setup(int n, const int* rp, const int* ci, const double* va, double theta,
                       int max_levels, int coarse_size);
/* Same with the threshold of level l set to theta * decay^l (decay 1: the reference). */
orc_amg* orc_amg_setup_decay(int n, const int* rp, const int* ci, const double* va, double theta,
                             double decay, int max_levels, int coarse_size);
void orc_amg_free(orc_amg* h);
int orc_amg_levels(const orc_amg* h);
/* which: 0 a, 1 p, 2 r. */
void orc_amg_size(const orc_amg* h, int level, int which, int* rows, int* cols, int* nnz);
void orc_amg_copy(const orc_amg* h, int level, int which, int* rp, int* ci, double* va);
void orc_amg_aggregates(const orc_amg* h, int level, int* agg);
/* Dense row-major inverse of the coarsest operator; returns 1 if the LDL^T succeeded. */
int orc_amg_coarse_inverse(const orc_amg* h, double* out);

/* Jacobi-V-cycle AMG-PCG on a hierarchy given as CSR arrays per level (the reference's
 * solve_amg_cg_host: krylov_openmp.cpp:187-255 -> krylov_kernels.inc:182-301).
 * Level l has a (n_l rows), and for l < L-1 p (n_l x n_{l+1}) and r (n_{l+1} x n_l). */
orc_stats orc_amg_cg(int n_levels, const int* n, const int* const* a_rp, const int* const* a_ci,
                     const double* const* a_va, const int* const* p_rp, const int* const* p_ci,
                     const double* const* p_va, const int* const* r_rp, const int* const* r_ci,
                     const double* const* r_va, const double* coarse_inverse, int sweeps,
                     double weight, const double* b, double* x, double abs_tol, int max_it);

/* The same, taking the hierarchy built by orc_amg_setup with level 0 replaced by (rp, ci, va)
 * when rp != NULL (DeviceSolver's stale-hierarchy semantics, device_solver.cpp:138-147). */
orc_stats orc_amg_cg_h(const orc_amg* h, const int* rp, const int* ci, const double* va,
                       int sweeps, double weight, const double* b, double* x, double abs_tol,
                       int max_it);

/* CPU AmgCg: PCG with the symmetric Gauss-Seidel V-cycle and a direct coarsest solve
 * (fv::pcg amg.cpp:253-285, cycle amg.cpp:192-216). Breakdown when p.Ap <= 0. */
orc_stats orc_sgs_pcg(const orc_amg* h, const double* b, double* x, double tol, int max_it);
/* apply (amg.cpp:218-225): z = M^-1 r, one SGS V-cycle from z = 0. */
void orc_sgs_apply(const orc_amg* h, const double* r, double* z);
/* The setup's conservative sparse product C = A B (amg.cpp:170, 176-177), explicit zeros kept;
 * returns nnz(C), or -1 when cap is too small. */
int orc_spgemm(int rows, int inner, int cols, const int* a_rp, const int* a_ci, const double* a_va,
               const int* b_rp, const int* b_ci, const double* b_va, int cap, int* c_rp, int* c_ci,
               double* c_va);

/* ---- Finite-volume cell kernels (gpu::DeviceFv, src/gpu/src/fv_cell.inc, fv_openmp.cpp) ----
 * The mesh struct has FvMeshView's fields in its order (src/gpu/include/cfd/gpu/fv_device.hpp:16-32). */
typedef struct orc_fv_mesh {
    int nc, nif, nf;
    const int* own;
    const int* nei;
    const int* start;
    const int* faces;
    const double* w;
    const double* sf;
    const double* cf;
    const double* cc;
    const double* vol;
    const double* kv;
    const double* lsq_d;
    const double* lsq;
} orc_fv_mesh;

/* Pressure equation of FlowSolver::pressure_correction (flow_solver.cpp:688-796) with the
 * accelerator assembly path (gpu::assemble_convection_diffusion, convection 0), as face loops.
 * Inputs: momentum off-diagonals upper / lower (nif), per-component diag / source / u (3 per
 * cell), delta coefficients (nf), boundary kinds 0 inlet, 1 wall, 2 outlet, 3 symmetry, 4 empty,
 * fixed velocity (3 per boundary face), outlet pressure, optional u_prev / phi_prev (NULL: off),
 * alpha_u and the pin flag. Outputs: pressure diag / upper / lower / source, phi_hbya and rau_f
 * (nf), hbya (3 per cell) and rau. */
void orc_pressure_equation(const orc_fv_mesh* m, const double* upper, const double* lower, const double* diag3,
                           const double* source3, const double* u, const double* dc, const int* kind,
                           const double* fixed_velocity, const double* outlet_p, const double* u_prev,
                           const double* phi_prev, double alpha_u, int pin, double* p_diag, double* p_upper,
                           double* p_lower, double* p_source, double* phi_hbya, double* rau_f, double* hbya,
                           double* rau);

/* DeviceFv::gradient of a scalar field: scheme 0 Green-Gauss, 1 least squares, 2 corrected. */
void orc_fv_gradient(const orc_fv_mesh* m, const double* phi, const double* bv, int scheme,
                     double* grad);
/* DeviceFv::vector_gradient: u 3 per cell, grad 9 per cell (column i = d u_i / d x). */
void orc_fv_vector_gradient(const orc_fv_mesh* m, const double* u, const double* ub, int scheme,
                            double* grad);
/* DeviceFv::scalar_corrections: source += explicit internal-face fluxes. */
void orc_fv_scalar_corrections(const orc_fv_mesh* m, const double* phi, const double* bv,
                               int scheme, const double* flux, const double* gamma,
                               int linear_upwind, double* source);
/* Limiter of DeviceFv::momentum_corrections (lim 3 per cell). */
void orc_fv_limiter(const orc_fv_mesh* m, const double* u, const double* ub, const double* grad,
                    int venkat, const double* eps2, double* lim);
/* DeviceFv::momentum_corrections with the field u / ub and gradient grad given explicitly. */
void orc_fv_momentum_corrections(const orc_fv_mesh* m, const double* u, const double* ub,
                                 const double* grad, const double* flux, const double* nu_face,
                                 int linear_upwind, int limited, int venkat, const double* eps2,
                                 double central, int transpose_term, double* source);

#ifdef __cplusplus
}
#endif

#endif
