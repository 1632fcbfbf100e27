// The reference's caller policy for the device solver: LduMatrix::solve with a CSR pattern and
// an AmgCache (src/fv/src/ldu_matrix.cpp:111-131, solve_persistent 287-366; AmgCache
// src/fv/include/cfd/fv/ldu_matrix.hpp:61-75), plus, as an option, the host AmgCg path's
// stale-hierarchy rebuild (src/fv/src/ldu_matrix.cpp:225-258). The cache owns one resident
// system on the pattern and the current hierarchy; every call uploads the LDU coefficients, the
// source and the iterate, as DeviceSolver::solve does.
#include "common.hpp"
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

using namespace cfdg;

struct cfdg_amg_cache {
    int n = 0;
    int nf = 0;
    std::vector<int> row_ptr, col, diag_pos, upper_pos, lower_pos;
    cfdg_amg_cache_options opt{};
    std::unique_ptr<System> sys;
    cfdg_amg* amg = nullptr; // current hierarchy (host copy; the device holds it when levels > 1)
    int levels = 0;
    int age = 0;              // solves since the last setup
    int setup_iterations = 0; // iterations of the first solve after the last setup (stale rule)
    int setups = 0;

    ~cfdg_amg_cache() {
        if (amg != nullptr) {
            (void)cfdg_amg_destroy(amg);
        }
    }

    // A new hierarchy from the current coefficients (ldu_matrix.cpp:315-326): CSR values placed
    // through the pattern's positions, setup, and the upload of levels > 1
    // (ldu_matrix.cpp:327-353; a one-level hierarchy is kept on the host only).
    void rebuild(const double* diag, const double* upper, const double* lower) {
        std::vector<double> val(col.size(), 0.0);
        for (int c = 0; c < n; ++c) {
            val[static_cast<std::size_t>(diag_pos[static_cast<std::size_t>(c)])] = diag[c];
        }
        for (int f = 0; f < nf; ++f) {
            val[static_cast<std::size_t>(upper_pos[static_cast<std::size_t>(f)])] = upper[f];
            val[static_cast<std::size_t>(lower_pos[static_cast<std::size_t>(f)])] = lower[f];
        }
        cfdg_csr a{};
        a.n_rows = n;
        a.n_cols = n;
        a.nnz = static_cast<int>(col.size());
        a.row_ptr = row_ptr.data();
        a.col = col.data();
        a.val = val.data();
        cfdg_amg* h = nullptr;
        rethrow(opt.device_setup != 0 ? cfdg_amg_build_device(&a, &opt.amg, &h) : cfdg_amg_build(&a, &opt.amg, &h));
        std::unique_ptr<cfdg_amg, cfdg_status (*)(cfdg_amg*)> keep(h, cfdg_amg_destroy);
        int nl = 0;
        rethrow(cfdg_amg_n_levels(h, &nl));
        if (nl > 1) {
            std::vector<cfdg_amg_level_view> lv(static_cast<std::size_t>(nl));
            cfdg_amg_view v{};
            rethrow(cfdg_amg_view_of(h, lv.data(), &v));
            sys->set_amg(v, false); // rebuilt every max_age solves: no fused products
        }
        if (amg != nullptr) {
            (void)cfdg_amg_destroy(amg);
        }
        amg = keep.release();
        levels = nl;
        ++setups;
    }

    // res / (norm + k_small) and |b - A x|_2 of the resident system (ldu_matrix.cpp:65-103).
    void norms(double& normalised, double& r2) {
        double o[3];
        sys->residual_norms(o);
        normalised = o[0] / (o[1] + 1e-20);
        r2 = std::sqrt(o[2]);
    }

    // A status of a nested C-ABI call back into the exception it stood for.
    static void rethrow(cfdg_status s) {
        switch (s) {
        case CFDG_OK:
            return;
        case CFDG_ERR_INVALID_ARGUMENT:
            throw InvalidArgument(last_error());
        case CFDG_ERR_NO_DEVICE:
            throw NoDevice(last_error());
        case CFDG_ERR_HIP:
            throw HipError(last_error());
        case CFDG_ERR_OUT_OF_MEMORY:
            throw OutOfMemory(last_error());
        case CFDG_ERR_CAPACITY:
            throw CapacityError(last_error());
        default:
            throw std::runtime_error(last_error());
        }
    }
};

extern "C" cfdg_status cfdg_amg_cache_default_options(cfdg_amg_cache_options* out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = cfdg_amg_cache_options{};
        out->max_age = 20;
        out->stale_rebuild = 0;
        out->device_setup = 1;
        cfdg_amg_cache::rethrow(cfdg_amg_default_settings(&out->amg));
    });
}

extern "C" cfdg_status cfdg_amg_cache_create(const cfdg_ldu_pattern* pattern, const cfdg_amg_cache_options* options,
                                             cfdg_amg_cache** out) {
    return guarded([&] {
        require(pattern != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        auto c = std::make_unique<cfdg_amg_cache>();
        if (options != nullptr) {
            c->opt = *options;
        } else {
            cfdg_amg_cache::rethrow(cfdg_amg_cache_default_options(&c->opt));
        }
        require(c->opt.max_age >= 1, "AmgCache: max_age must be >= 1");
        c->sys = std::make_unique<System>(*pattern); // validates the pattern
        c->n = pattern->n;
        c->nf = pattern->n_faces;
        const auto n = static_cast<std::size_t>(pattern->n);
        const auto nf = static_cast<std::size_t>(pattern->n_faces);
        c->row_ptr.assign(pattern->row_ptr, pattern->row_ptr + n + 1);
        c->col.assign(pattern->col, pattern->col + pattern->nnz);
        c->diag_pos.assign(pattern->diag_pos, pattern->diag_pos + n);
        c->upper_pos.assign(pattern->upper_pos, pattern->upper_pos + nf);
        c->lower_pos.assign(pattern->lower_pos, pattern->lower_pos + nf);
        *out = c.release();
    });
}

extern "C" cfdg_status cfdg_amg_cache_destroy(cfdg_amg_cache* cache) {
    return guarded([&] { delete cache; });
}

extern "C" cfdg_status cfdg_amg_cache_state(const cfdg_amg_cache* cache, int* age, int* n_levels, int* setups) {
    return guarded([&] {
        require(cache != nullptr, "null argument");
        if (age != nullptr) {
            *age = cache->age;
        }
        if (n_levels != nullptr) {
            *n_levels = cache->levels;
        }
        if (setups != nullptr) {
            *setups = cache->setups;
        }
    });
}

extern "C" cfdg_status cfdg_ldu_solve(cfdg_amg_cache* c, const double* diag, const double* upper,
                                      const double* lower, const double* source, double* x,
                                      const cfdg_linear_settings* s, cfdg_linear_result* result) {
    return guarded([&] {
        require(c != nullptr && s != nullptr && result != nullptr && x != nullptr && source != nullptr &&
                    diag != nullptr && (c->nf == 0 || (upper != nullptr && lower != nullptr)),
                "null argument");
        require(s->method == CFDG_METHOD_CG || s->method == CFDG_METHOD_BICGSTAB, "unknown method");
        require(s->precision == CFDG_PREC_FP64 || s->precision == CFDG_PREC_MIXED_FP32 ||
                    s->precision == CFDG_PREC_MIXED_BF16,
                "unknown precision");
        require(s->max_iterations >= 0, "max_iterations must be >= 0");
        *result = cfdg_linear_result{};
        System& sys = *c->sys;
        sys.assemble(diag, upper, lower);
        sys.set_rhs(source);
        sys.set_x(x);
        // LduMatrix::solve (ldu_matrix.cpp:117-124): nothing to do below the tolerance.
        double initial = 0.0;
        double r0 = 0.0;
        c->norms(initial, r0);
        result->initial_residual = initial;
        result->final_residual = initial;
        if (initial < s->tolerance || c->n == 0) {
            return;
        }
        // solve_persistent (ldu_matrix.cpp:291-366).
        const double target = std::max(s->rel_tolerance, s->tolerance / std::max(initial, 1e-20)) * r0;
        const bool cg = s->method == CFDG_METHOD_CG;
        bool fresh = false;
        if (cg && (c->amg == nullptr || c->age >= c->opt.max_age)) {
            c->rebuild(diag, upper, lower);
            c->age = 0;
            fresh = true;
            ++result->setups;
        }
        if (cg) {
            ++c->age;
        }
        const bool use_amg = cg && c->amg != nullptr && c->levels > 1;
        const cfdg_precision prec = use_amg ? s->precision : CFDG_PREC_FP64;
        const cfdg_solve_control control{std::max(target, 1e-300), s->max_iterations};
        result->target = control.abs_tolerance;
        result->used_amg = use_amg ? 1 : 0;
        result->stats = solve_system(sys, cg ? CFDG_METHOD_CG : CFDG_METHOD_BICGSTAB, use_amg, prec, control);
        result->iterations = result->stats.iterations;
        result->solved = 1;
        if (c->opt.stale_rebuild != 0 && cg) {
            // The host AmgCg path's rule (ldu_matrix.cpp:242-256): remember the iterations of the
            // first solve after a setup; a later solve needing more than twice as many plus 2
            // rebuilds from the current matrix and continues from the current iterate.
            if (fresh) {
                c->setup_iterations = std::max(result->iterations, 1);
            } else if (result->iterations > 2 * c->setup_iterations + 2) {
                c->rebuild(diag, upper, lower);
                ++result->setups;
                const bool amg2 = c->levels > 1;
                result->used_amg = amg2 ? 1 : 0;
                result->stats = solve_system(sys, CFDG_METHOD_CG, amg2, amg2 ? s->precision : CFDG_PREC_FP64,
                                             control);
                result->iterations += result->stats.iterations;
                c->age = 1; // reset to 0 by the setup, then counted for this solve
                c->setup_iterations = std::max(result->stats.iterations, 1);
            }
        }
        sys.get_x(x);
        double final_res = 0.0;
        double r_end = 0.0;
        c->norms(final_res, r_end);
        result->final_residual = final_res;
    });
}
