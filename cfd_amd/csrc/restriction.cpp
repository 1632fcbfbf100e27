// Fused restriction of the mixed-precision V-cycle: with one damped-Jacobi sweep from zero per
// side, a level's down leg is x1 = W b, r = b - A x1, b_next = R r (krylov_kernels.inc:182-238
// with x = 0 on entry, W = diag(weight / a_ii)). The three steps are the single linear map
// b_next = R (I - A W) b, formed here once per hierarchy so that the down leg of a level is one
// sparse product instead of a residual pass and a restriction pass.
#include "common.hpp"

#include <algorithm>
#include <cstddef>
#include <vector>

namespace cfdg {

// Row j of R (I - A W): R_jc minus, for every entry R_ji of the row in storage order, R_ji A_ic w_c
// over row i of A in storage order; columns sorted, computed in FP64.
Csr fused_restriction(const cfdg_csr& r, const cfdg_csr& a, const float* wd) {
    const int n = r.n_rows;
    const int cols = a.n_rows;
    require(r.row_ptr != nullptr && a.row_ptr != nullptr && wd != nullptr, "fused restriction: null input");
    require((r.n_cols == 0 ? r.n_rows : r.n_cols) == cols, "fused restriction: R and A do not conform");
    Csr c;
    c.n_rows = n;
    c.n_cols = cols;
    std::vector<int> count(static_cast<std::size_t>(n) + 1, 0);
#pragma omp parallel if (host_parallel())
    {
        std::vector<int> mark(static_cast<std::size_t>(cols), -1);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < n; ++j) {
            int cnt = 0;
            auto touch = [&](int col) {
                if (mark[static_cast<std::size_t>(col)] != j) {
                    mark[static_cast<std::size_t>(col)] = j;
                    ++cnt;
                }
            };
            for (int k = r.row_ptr[j]; k < r.row_ptr[j + 1]; ++k) {
                const int i = r.col[k];
                touch(i);
                for (int e = a.row_ptr[i]; e < a.row_ptr[i + 1]; ++e) {
                    touch(a.col[e]);
                }
            }
            count[static_cast<std::size_t>(j) + 1] = cnt;
        }
    }
    c.rp.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int j = 0; j < n; ++j) {
        c.rp[static_cast<std::size_t>(j) + 1] = c.rp[static_cast<std::size_t>(j)] + count[static_cast<std::size_t>(j) + 1];
    }
    const std::size_t nnz = static_cast<std::size_t>(c.rp[static_cast<std::size_t>(n)]);
    c.ci.resize(nnz);
    c.va.resize(nnz);
#pragma omp parallel if (host_parallel())
    {
        std::vector<int> mark(static_cast<std::size_t>(cols), -1);
        std::vector<double> acc(static_cast<std::size_t>(cols), 0.0);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < n; ++j) {
            int* cj = c.ci.data() + c.rp[static_cast<std::size_t>(j)];
            int cnt = 0;
            auto add = [&](int col, double v) {
                if (mark[static_cast<std::size_t>(col)] != j) {
                    mark[static_cast<std::size_t>(col)] = j;
                    acc[static_cast<std::size_t>(col)] = v;
                    cj[cnt++] = col;
                } else {
                    acc[static_cast<std::size_t>(col)] += v;
                }
            };
            for (int k = r.row_ptr[j]; k < r.row_ptr[j + 1]; ++k) {
                add(r.col[k], r.val[k]);
            }
            for (int k = r.row_ptr[j]; k < r.row_ptr[j + 1]; ++k) {
                const int i = r.col[k];
                const double rji = r.val[k];
                for (int e = a.row_ptr[i]; e < a.row_ptr[i + 1]; ++e) {
                    const int col = a.col[e];
                    add(col, -(rji * a.val[e]) * static_cast<double>(wd[col]));
                }
            }
            std::sort(cj, cj + cnt);
            double* vj = c.va.data() + c.rp[static_cast<std::size_t>(j)];
            for (int q = 0; q < cnt; ++q) {
                vj[q] = acc[static_cast<std::size_t>(cj[q])];
            }
        }
    }
    return c;
}

} // namespace cfdg
