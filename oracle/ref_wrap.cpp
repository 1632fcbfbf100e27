// ref_wrap.cpp — C entry points over the reference's own host solver code.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile, in this container only, together with
// the reference's Eigen-free accelerator sources (src/gpu/src/krylov_openmp.cpp,
// dispatch.cpp, device_solver.cpp, fv_device.cpp, fv_openmp.cpp, compiled where they lie under
// the reference checkout with no
// backend defined, i.e. the "host" build of cfd::gpu) into oracle/_ref/libcfdref.so. Tests use
// it to check the oracle's restatement against the reference's actual host algorithms and
// bench.py times it as the CPU baseline. Nothing here is part of the product.
#include <cfd/gpu/fv_device.hpp>
#include <cfd/gpu/krylov.hpp>

#include <cstring>
#include <memory>
#include <vector>

using cfd::gpu::AmgLevelView;
using cfd::gpu::AmgView;
using cfd::gpu::CsrView;
using cfd::gpu::Method;
using cfd::gpu::SolveControl;
using cfd::gpu::SolveStats;

namespace {

void put(const SolveStats& st, int* iterations, int* converged, double* residual) {
    if (iterations != nullptr) {
        *iterations = st.iterations;
    }
    if (converged != nullptr) {
        *converged = st.converged ? 1 : 0;
    }
    if (residual != nullptr) {
        *residual = st.residual;
    }
}

AmgView make_view(int n_levels, const int* n, const int* nnz_a, const int* const* a_rp,
                  const int* const* a_ci, const double* const* a_va, const int* nnz_p,
                  const int* const* p_rp, const int* const* p_ci, const double* const* p_va,
                  const int* nnz_r, const int* const* r_rp, const int* const* r_ci,
                  const double* const* r_va, const double* coarse_inverse, int sweeps,
                  double weight) {
    AmgView v;
    for (int l = 0; l < n_levels; ++l) {
        AmgLevelView lv;
        if (a_rp[l] != nullptr) {
            lv.a = CsrView{n[l], nnz_a[l], a_rp[l], a_ci[l], a_va[l]};
        }
        if (l + 1 < n_levels) {
            lv.p = CsrView{n[l], nnz_p[l], p_rp[l], p_ci[l], p_va[l]};
            lv.r = CsrView{n[l + 1], nnz_r[l], r_rp[l], r_ci[l], r_va[l]};
        }
        v.levels.push_back(lv);
    }
    v.coarse_inverse = coarse_inverse;
    v.smooth_sweeps = sweeps;
    v.jacobi_weight = weight;
    return v;
}

} // namespace

extern "C" {

int ref_backend_name(char* buf, int len) {
    const std::string s = cfd::gpu::backend_name();
    std::strncpy(buf, s.c_str(), static_cast<std::size_t>(len));
    return static_cast<int>(s.size());
}

// cfd::gpu::solve_host (krylov.hpp:41-42).
void ref_solve_host(int n, int nnz, const int* rp, const int* ci, const double* va,
                    const double* b, double* x, int method, double tol, int max_it,
                    int* iterations, int* converged, double* residual) {
    const CsrView a{n, nnz, rp, ci, va};
    const SolveStats st = cfd::gpu::solve_host(a, b, x, method == 0 ? Method::Cg : Method::BiCgStab,
                                               SolveControl{tol, max_it});
    put(st, iterations, converged, residual);
}

// cfd::gpu::solve_amg_cg_host (krylov.hpp:63-64).
void ref_solve_amg_cg_host(int n_levels, const int* n, const int* nnz_a, const int* const* a_rp,
                           const int* const* a_ci, const double* const* a_va, const int* nnz_p,
                           const int* const* p_rp, const int* const* p_ci,
                           const double* const* p_va, const int* nnz_r, const int* const* r_rp,
                           const int* const* r_ci, const double* const* r_va,
                           const double* coarse_inverse, int sweeps, double weight,
                           const double* b, double* x, double tol, int max_it, int* iterations,
                           int* converged, double* residual) {
    const AmgView v = make_view(n_levels, n, nnz_a, a_rp, a_ci, a_va, nnz_p, p_rp, p_ci, p_va,
                                nnz_r, r_rp, r_ci, r_va, coarse_inverse, sweeps, weight);
    const SolveStats st = cfd::gpu::solve_amg_cg_host(v, b, x, SolveControl{tol, max_it});
    put(st, iterations, converged, residual);
}

// cfd::gpu::assemble_convection_diffusion (krylov.hpp:113-116), host build.
void ref_assemble(int n_cells, int n_faces, const int* owner, const int* neighbour,
                  const double* flux, const double* gamma, const double* delta,
                  double convection, double* upper, double* lower, double* diag,
                  double* divergence) {
    cfd::gpu::assemble_convection_diffusion(n_cells, n_faces, owner, neighbour, flux, gamma, delta,
                                            convection, upper, lower, diag, divergence);
}

// cfd::gpu::DeviceSolver host path (device_solver.cpp).
void* ref_device_solver_create(int n, int n_faces, int nnz, const int* rp, const int* ci,
                               const int* diag_pos, const int* upper_pos, const int* lower_pos) {
    const cfd::gpu::LduPattern p{n, n_faces, nnz, rp, ci, diag_pos, upper_pos, lower_pos};
    return new cfd::gpu::DeviceSolver(p);
}

void ref_device_solver_destroy(void* h) {
    delete static_cast<cfd::gpu::DeviceSolver*>(h);
}

void ref_device_solver_set_amg(void* h, int n_levels, const int* n, const int* nnz_a,
                               const int* const* a_rp, const int* const* a_ci,
                               const double* const* a_va, const int* nnz_p,
                               const int* const* p_rp, const int* const* p_ci,
                               const double* const* p_va, const int* nnz_r,
                               const int* const* r_rp, const int* const* r_ci,
                               const double* const* r_va, const double* coarse_inverse,
                               int sweeps, double weight) {
    const AmgView v = make_view(n_levels, n, nnz_a, a_rp, a_ci, a_va, nnz_p, p_rp, p_ci, p_va,
                                nnz_r, r_rp, r_ci, r_va, coarse_inverse, sweeps, weight);
    static_cast<cfd::gpu::DeviceSolver*>(h)->set_amg(v);
}

void ref_device_solver_solve(void* h, const double* diag, const double* upper,
                             const double* lower, const double* b, double* x, int method,
                             int use_amg, double tol, int max_it, int* iterations,
                             int* converged, double* residual) {
    const SolveStats st = static_cast<cfd::gpu::DeviceSolver*>(h)->solve(
        diag, upper, lower, b, x, method == 0 ? Method::Cg : Method::BiCgStab, use_amg != 0,
        SolveControl{tol, max_it});
    put(st, iterations, converged, residual);
}

// gpu::DeviceFv on host threads (fv_device.cpp:51-59 -> fv_openmp.cpp). The mesh arrays follow
// FvMeshView's field order (fv_device.hpp:16-32).
void* ref_fv_create(int n_cells, int n_internal_faces, int n_faces, const int* owner,
                    const int* neighbour, const int* cell_offsets, const int* cell_faces,
                    const double* weight, const double* sf, const double* face_centre,
                    const double* cell_centre, const double* volume, const double* correction,
                    const double* lsq_d, const double* lsq_inverse) {
    cfd::gpu::FvMeshView v;
    v.n_cells = n_cells;
    v.n_internal_faces = n_internal_faces;
    v.n_faces = n_faces;
    v.owner = owner;
    v.neighbour = neighbour;
    v.cell_offsets = cell_offsets;
    v.cell_faces = cell_faces;
    v.weight = weight;
    v.sf = sf;
    v.face_centre = face_centre;
    v.cell_centre = cell_centre;
    v.volume = volume;
    v.correction = correction;
    v.lsq_d = lsq_d;
    v.lsq_inverse = lsq_inverse;
    return new cfd::gpu::DeviceFv(v);
}

void ref_fv_destroy(void* h) { delete static_cast<cfd::gpu::DeviceFv*>(h); }

void ref_fv_gradient(void* h, const double* phi, const double* bv, int scheme, double* grad) {
    static_cast<cfd::gpu::DeviceFv*>(h)->gradient(phi, bv, scheme, grad);
}

void ref_fv_vector_gradient(void* h, const double* u, const double* ub, int scheme, double* grad) {
    static_cast<cfd::gpu::DeviceFv*>(h)->vector_gradient(u, ub, scheme, grad);
}

void ref_fv_scalar_corrections(void* h, const double* phi, const double* bv, int scheme,
                               const double* flux, const double* gamma, int linear_upwind,
                               double* source) {
    static_cast<cfd::gpu::DeviceFv*>(h)->scalar_corrections(phi, bv, scheme, flux, gamma,
                                                            linear_upwind != 0, source);
}

void ref_fv_momentum_corrections(void* h, const double* flux, const double* nu_face,
                                 int linear_upwind, int limited, int venkat, const double* eps2,
                                 double central, int transpose_term, double* source) {
    static_cast<cfd::gpu::DeviceFv*>(h)->momentum_corrections(
        flux, nu_face, linear_upwind != 0, limited != 0, venkat != 0, eps2, central,
        transpose_term != 0, source);
}

} // extern "C"
