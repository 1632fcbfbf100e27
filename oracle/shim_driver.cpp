// Test infrastructure: C entry points over the reference-side binding of the integration guide
// (extracted to _ref/krylov_hip.cpp by build_shim.py and included
// here). Each entry point builds the reference's own cfd::gpu types (CsrView, AmgLevelView,
// AmgView, LduPattern, SolveControl, SolveStats, Method of src/gpu/include/cfd/gpu/krylov.hpp)
// from plain arrays and calls the binding's cfd::gpu::detail functions, which call libcfdgpu;
// the binding's exceptions come back as codes (1 std::invalid_argument, 2 std::runtime_error).
// Built against the reference's headers only; nothing here is part of the library.
#include <cfd/gpu/krylov.hpp>

#include "krylov_hip.cpp"

#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

namespace {

using namespace cfd::gpu;

struct ShimStats {
    int iterations;
    int converged;
    int on_device;
    double residual;
};

template <class F>
int run(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (const std::exception&) {
        return 2;
    }
}

void put(const SolveStats& s, ShimStats* out) {
    out->iterations = s.iterations;
    out->converged = s.converged ? 1 : 0;
    out->on_device = s.on_device ? 1 : 0;
    out->residual = s.residual;
}

// Per level l: a, p, r as (rows, nnz, row_ptr, col, val) arrays; p and r are empty on the
// coarsest level.
AmgView hierarchy(int n_levels, const int* rows, const int* nnz, const int* const* rp,
                  const int* const* ci, const double* const* va, const double* coarse_inverse,
                  int sweeps, double weight) {
    AmgView v;
    for (int l = 0; l < n_levels; ++l) {
        AmgLevelView lv;
        for (int w = 0; w < 3; ++w) {
            const int k = 3 * l + w;
            const CsrView m{rows[k], nnz[k], rp[k], ci[k], va[k]};
            (w == 0 ? lv.a : w == 1 ? lv.p : lv.r) = m;
        }
        v.levels.push_back(lv);
    }
    v.coarse_inverse = coarse_inverse;
    v.smooth_sweeps = sweeps;
    v.jacobi_weight = weight;
    return v;
}

} // namespace

extern "C" int shim_device_name(char* buf, int len) {
    return run([&] {
        const std::string s = detail::hip_device_name();
        std::strncpy(buf, s.c_str(), static_cast<std::size_t>(len) - 1);
        buf[len - 1] = '\0';
    });
}

extern "C" int shim_solve(int n, int nnz, const int* rp, const int* ci, const double* va, const double* b,
                          double* x, int method, double tol, int max_it, ShimStats* out) {
    return run([&] {
        const CsrView a{n, nnz, rp, ci, va};
        put(detail::solve_hip(a, b, x, method == 0 ? Method::Cg : Method::BiCgStab, SolveControl{tol, max_it}),
            out);
    });
}

extern "C" int shim_solve_amg(int n_levels, const int* rows, const int* nnz, const int* const* rp,
                              const int* const* ci, const double* const* va, const double* coarse_inverse,
                              int sweeps, double weight, const double* b, double* x, double tol, int max_it,
                              ShimStats* out) {
    return run([&] {
        const AmgView v = hierarchy(n_levels, rows, nnz, rp, ci, va, coarse_inverse, sweeps, weight);
        put(detail::solve_amg_hip(v, b, x, SolveControl{tol, max_it}), out);
    });
}

extern "C" int shim_assemble(int n_cells, int n_faces, const int* owner, const int* neighbour, const double* flux,
                             const double* gamma, const double* delta, double convection, double* upper,
                             double* lower, double* diag, double* divergence) {
    return run([&] {
        detail::assemble_hip(n_cells, n_faces, owner, neighbour, flux, gamma, delta, convection, upper, lower, diag,
                             divergence);
    });
}

extern "C" int shim_system_create(int n, int n_faces, int nnz, const int* rp, const int* ci, const int* diag_pos,
                                  const int* upper_pos, const int* lower_pos, void** out) {
    return run([&] {
        const LduPattern p{n, n_faces, nnz, rp, ci, diag_pos, upper_pos, lower_pos};
        *out = detail::hip_system_create(p);
    });
}

extern "C" int shim_system_set_amg(void* h, int n_levels, const int* rows, const int* nnz, const int* const* rp,
                                   const int* const* ci, const double* const* va, const double* coarse_inverse,
                                   int sweeps, double weight) {
    return run([&] {
        detail::hip_system_set_amg(h, hierarchy(n_levels, rows, nnz, rp, ci, va, coarse_inverse, sweeps, weight));
    });
}

extern "C" int shim_system_solve(void* h, const double* diag, const double* upper, const double* lower,
                                 const double* b, double* x, int method, int use_amg, double tol, int max_it,
                                 ShimStats* out) {
    return run([&] {
        put(detail::hip_system_solve(h, diag, upper, lower, b, x, method == 0 ? Method::Cg : Method::BiCgStab,
                                     use_amg != 0, SolveControl{tol, max_it}),
            out);
    });
}

extern "C" void shim_system_destroy(void* h) { detail::hip_system_destroy(h); }
