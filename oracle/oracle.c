/*
 * oracle.c — CPU restatement of the reference's linear-solver path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py as the checker; never part of the product.
 *
 * Citations are into the reference repository (Yipexnn/CFD). Built with
 * -O2 -fopenmp -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PAR_THRESHOLD 20000 /* src/gpu/src/krylov_openmp.cpp:21-25 */

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

static void* xmalloc(size_t bytes) {
    void* p = malloc(bytes > 0 ? bytes : 1);
    if (!p) {
        abort();
    }
    return p;
}

/* ============================ LDU matrix ============================ */

/* off_diagonal_product + multiply (ldu_matrix.cpp:39-56). */
void orc_ldu_multiply(int n, int nf, const int* own, const int* nei, const double* diag,
                      const double* upper, const double* lower, const double* x, double* y) {
    for (int c = 0; c < n; ++c) {
        y[c] = 0.0;
    }
    for (int f = 0; f < nf; ++f) {
        const int p = own[f];
        const int q = nei[f];
        y[p] += upper[f] * x[q];
        y[q] += lower[f] * x[p];
    }
    for (int c = 0; c < n; ++c) {
        y[c] += diag[c] * x[c];
    }
}

/* residual (ldu_matrix.cpp:58-63). */
void orc_ldu_residual(int n, int nf, const int* own, const int* nei, const double* diag,
                      const double* upper, const double* lower, const double* source,
                      const double* x, double* r) {
    orc_ldu_multiply(n, nf, own, nei, diag, upper, lower, x, r);
    for (int c = 0; c < n; ++c) {
        r[c] = source[c] - r[c];
    }
}

/* normalised_residual / residual_norms for one process (ldu_matrix.cpp:65-103). */
double orc_ldu_normalised_residual(int n, int nf, const int* own, const int* nei,
                                   const double* diag, const double* upper, const double* lower,
                                   const double* source, const double* x) {
    if (n == 0) {
        return 0.0 / (0.0 + 1e-20);
    }
    double* ax = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* row_sum = (double*)xmalloc(sizeof(double) * (size_t)n);
    orc_ldu_multiply(n, nf, own, nei, diag, upper, lower, x, ax);
    double sx = 0.0;
    for (int c = 0; c < n; ++c) {
        sx += x[c];
    }
    const double x_ref = sx / (double)n;
    for (int c = 0; c < n; ++c) {
        row_sum[c] = diag[c];
    }
    for (int f = 0; f < nf; ++f) {
        row_sum[own[f]] += upper[f];
        row_sum[nei[f]] += lower[f];
    }
    double res = 0.0;
    double norm = 0.0;
    for (int c = 0; c < n; ++c) {
        const double ax_ref = row_sum[c] * x_ref;
        res += fabs(source[c] - ax[c]);
        norm += fabs(ax[c] - ax_ref) + fabs(source[c] - ax_ref);
    }
    free(ax);
    free(row_sum);
    return res / (norm + 1e-20); /* k_small, ldu_matrix.cpp:22 */
}

/* set_reference (ldu_matrix.cpp:105-109). */
void orc_ldu_set_reference(double* diag, double* source, int cell, double value) {
    source[cell] += diag[cell] * value;
    diag[cell] += diag[cell];
}

static int cmp_int(const void* a, const void* b) {
    const int x = *(const int*)a;
    const int y = *(const int*)b;
    return (x > y) - (x < y);
}

/* build_csr_pattern (ldu_matrix.cpp:572-612): triplets (c,c), (own,nei), (nei,own) assembled
 * with sorted columns; positions found by a linear scan of the row. Returns nnz or -1 if an
 * entry is missing. */
int orc_csr_pattern(int n, int nf, const int* own, const int* nei, int* rp, int* ci,
                    int* diag_pos, int* upper_pos, int* lower_pos) {
    int* cnt = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int c = 0; c < n; ++c) {
        cnt[c + 1] += 1;
    }
    for (int f = 0; f < nf; ++f) {
        cnt[own[f] + 1] += 1;
        cnt[nei[f] + 1] += 1;
    }
    rp[0] = 0;
    for (int c = 0; c < n; ++c) {
        rp[c + 1] = rp[c] + cnt[c + 1];
    }
    int* fill = (int*)xmalloc(sizeof(int) * ((size_t)n + 1));
    memcpy(fill, rp, sizeof(int) * (size_t)n);
    for (int c = 0; c < n; ++c) {
        ci[fill[c]++] = c;
    }
    for (int f = 0; f < nf; ++f) {
        ci[fill[own[f]]++] = nei[f];
        ci[fill[nei[f]]++] = own[f];
    }
    for (int c = 0; c < n; ++c) {
        qsort(ci + rp[c], (size_t)(rp[c + 1] - rp[c]), sizeof(int), cmp_int);
    }
    free(cnt);
    free(fill);
    /* find(): first k in the row with ci[k] == col (ldu_matrix.cpp:593-600). */
#define ORC_FIND(row, col, out)                                                                   \
    do {                                                                                          \
        int found_ = -1;                                                                          \
        for (int k_ = rp[(row)]; k_ < rp[(row) + 1]; ++k_) {                                      \
            if (ci[k_] == (col)) {                                                                \
                found_ = k_;                                                                      \
                break;                                                                            \
            }                                                                                     \
        }                                                                                         \
        if (found_ < 0) {                                                                         \
            return -1;                                                                            \
        }                                                                                         \
        (out) = found_;                                                                           \
    } while (0)
    for (int c = 0; c < n; ++c) {
        ORC_FIND(c, c, diag_pos[c]);
    }
    for (int f = 0; f < nf; ++f) {
        ORC_FIND(own[f], nei[f], upper_pos[f]);
        ORC_FIND(nei[f], own[f], lower_pos[f]);
    }
#undef ORC_FIND
    return rp[n];
}

/* Host assembly of CSR values from LDU coefficients (device_solver.cpp:127-135). */
void orc_ldu_to_csr(int n, int nf, const int* diag_pos, const int* upper_pos,
                    const int* lower_pos, const double* diag, const double* upper,
                    const double* lower, double* val) {
    for (int i = 0; i < n; ++i) {
        val[diag_pos[i]] = diag[i];
    }
    for (int f = 0; f < nf; ++f) {
        val[upper_pos[f]] = upper[f];
        val[lower_pos[f]] = lower[f];
    }
}

/* Serial face loop of assemble_openmp (krylov_openmp.cpp:288-300). */
void orc_assemble(int n_cells, int n_faces, const int* own, const int* nei, const double* flux,
                  const double* gamma, const double* delta, double convection, double* upper,
                  double* lower, double* diag, double* divergence) {
    (void)n_cells;
    for (int f = 0; f < n_faces; ++f) {
        const double fl = flux[f];
        const double g = gamma[f] * delta[f];
        const double pos = fl > 0.0 ? convection * fl : 0.0;
        const double neg = fl < 0.0 ? convection * fl : 0.0;
        upper[f] = neg - g;
        lower[f] = -pos - g;
        diag[own[f]] += pos + g;
        diag[nei[f]] += -neg + g;
        divergence[own[f]] += fl;
        divergence[nei[f]] -= fl;
    }
}

/* ============================ Krylov kernels ============================ */

/* spmv (krylov_kernels.inc:12-25). */
void orc_spmv(int n, const int* rp, const int* ci, const double* va, const double* x, double* y) {
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            s += va[k] * x[ci[k]];
        }
        y[i] = s;
    }
}

/* dot (krylov_kernels.inc:27-34). */
double orc_dot(int n, const double* x, const double* y) {
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        sum += x[i] * y[i];
    }
    return sum;
}

/* scale (krylov_kernels.inc:36-42). */
static void orc_scale(int n, const double* x, const double* d, double* y) {
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        y[i] = x[i] * d[i];
    }
}

/* inverse_diagonal (krylov_openmp.cpp:147-157): 1 / a_ii, or 1.0 when a_ii is 0. */
static void orc_inverse_diagonal(int n, const int* rp, const int* ci, const double* va,
                                 double* d) {
    for (int i = 0; i < n; ++i) {
        d[i] = 1.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ci[k] == i && va[k] != 0.0) {
                d[i] = 1.0 / va[k];
            }
        }
    }
}

/* cg (krylov_kernels.inc:44-94). */
static orc_stats orc_cg(int n, const int* rp, const int* ci, const double* va, const double* b,
                        double* x, const double* dinv, double* r, double* z, double* p,
                        double* q, double tol, int max_it) {
    orc_stats st = {0, 0, 0.0};
    orc_spmv(n, rp, ci, va, x, q);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
    }
    double res = sqrt(orc_dot(n, r, r));
    if (res <= tol) {
        st.residual = res;
        st.converged = 1;
        return st;
    }
    orc_scale(n, r, dinv, z);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        p[i] = z[i];
    }
    double rz = orc_dot(n, r, z);
    for (st.iterations = 1; st.iterations <= max_it; ++st.iterations) {
        orc_spmv(n, rp, ci, va, p, q);
        const double pq = orc_dot(n, p, q);
        if (pq == 0.0) {
            break;
        }
        const double alpha = rz / pq;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        res = sqrt(orc_dot(n, r, r));
        if (res <= tol) {
            st.converged = 1;
            break;
        }
        orc_scale(n, r, dinv, z);
        const double rz_new = orc_dot(n, r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            p[i] = z[i] + beta * p[i];
        }
    }
    if (st.iterations > max_it) {
        st.iterations = max_it;
    }
    st.residual = res;
    return st;
}

/* bicgstab (krylov_kernels.inc:96-169). */
static orc_stats orc_bicgstab(int n, const int* rp, const int* ci, const double* va,
                              const double* b, double* x, const double* dinv, double* r,
                              double* r0, double* p, double* v, double* s, double* t, double* ph,
                              double* sh, double tol, int max_it) {
    orc_stats st = {0, 0, 0.0};
    orc_spmv(n, rp, ci, va, x, v);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - v[i];
        r0[i] = r[i];
        p[i] = 0.0;
        v[i] = 0.0;
    }
    double res = sqrt(orc_dot(n, r, r));
    if (res <= tol) {
        st.residual = res;
        st.converged = 1;
        return st;
    }
    double rho = 1.0;
    double alpha = 1.0;
    double omega = 1.0;
    for (st.iterations = 1; st.iterations <= max_it; ++st.iterations) {
        const double rho_new = orc_dot(n, r0, r);
        if (rho_new == 0.0 || omega == 0.0) {
            break;
        }
        const double beta = (rho_new / rho) * (alpha / omega);
        rho = rho_new;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        orc_scale(n, p, dinv, ph);
        orc_spmv(n, rp, ci, va, ph, v);
        const double r0v = orc_dot(n, r0, v);
        if (r0v == 0.0) {
            break;
        }
        alpha = rho / r0v;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            s[i] = r[i] - alpha * v[i];
        }
        const double s_norm = sqrt(orc_dot(n, s, s));
        if (s_norm <= tol) {
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * ph[i];
            }
            res = s_norm;
            st.converged = 1;
            break;
        }
        orc_scale(n, s, dinv, sh);
        orc_spmv(n, rp, ci, va, sh, t);
        const double tt = orc_dot(n, t, t);
        omega = tt > 0.0 ? orc_dot(n, t, s) / tt : 0.0;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = s[i] - omega * t[i];
        }
        res = sqrt(orc_dot(n, r, r));
        if (res <= tol) {
            st.converged = 1;
            break;
        }
    }
    if (st.iterations > max_it) {
        st.iterations = max_it;
    }
    st.residual = res;
    return st;
}

/* solve_openmp host path (krylov_openmp.cpp:93-138). */
orc_stats orc_solve(int n, const int* rp, const int* ci, const double* va, const double* b,
                    double* x, int method, double abs_tol, int max_it) {
    double* dinv = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_inverse_diagonal(n, rp, ci, va, dinv);
    const int nvec = method == 0 ? 4 : 8;
    double* w = (double*)calloc((size_t)nvec * (size_t)(n > 0 ? n : 1), sizeof(double));
    orc_stats st;
    if (method == 0) {
        st = orc_cg(n, rp, ci, va, b, x, dinv, w, w + n, w + 2 * n, w + 3 * n, abs_tol, max_it);
    } else {
        st = orc_bicgstab(n, rp, ci, va, b, x, dinv, w, w + n, w + 2 * n, w + 3 * n, w + 4 * n,
                          w + 5 * n, w + 6 * n, w + 7 * n, abs_tol, max_it);
    }
    free(dinv);
    free(w);
    return st;
}

/* ============================ AMG setup ============================ */

typedef struct {
    int rows, cols, nnz;
    int* rp;
    int* ci;
    double* va;
} ocsr;

static ocsr ocsr_alloc(int rows, int cols, int nnz) {
    ocsr m;
    m.rows = rows;
    m.cols = cols;
    m.nnz = nnz;
    m.rp = (int*)calloc((size_t)rows + 1, sizeof(int));
    m.ci = (int*)xmalloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    m.va = (double*)xmalloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    return m;
}

static void ocsr_free(ocsr* m) {
    free(m->rp);
    free(m->ci);
    free(m->va);
    memset(m, 0, sizeof(*m));
}

static ocsr ocsr_copy(int rows, int cols, const int* rp, const int* ci, const double* va) {
    ocsr m = ocsr_alloc(rows, cols, rp[rows]);
    memcpy(m.rp, rp, sizeof(int) * ((size_t)rows + 1));
    memcpy(m.ci, ci, sizeof(int) * (size_t)rp[rows]);
    memcpy(m.va, va, sizeof(double) * (size_t)rp[rows]);
    return m;
}

/* Conservative sparse product C = A * B: entry (i, j) sums a_ik * b_kj over the columns k of
 * row i of A in storage order, starting from the first product; explicit zeros are kept
 * (Eigen's conservative_sparse_sparse_product as used at amg.cpp:170, 176-177).
 * One dense accumulator row; the touched columns are insertion-sorted before output. */
static ocsr ocsr_mul(const ocsr* a, const ocsr* b) {
    double* acc = (double*)calloc((size_t)b->cols + 1, sizeof(double));
    int* stamp = (int*)xmalloc(sizeof(int) * ((size_t)b->cols + 1));
    for (int j = 0; j < b->cols; ++j) {
        stamp[j] = -1;
    }
    int cap = 64;
    int* touched = (int*)xmalloc(sizeof(int) * (size_t)cap);
    int out_cap = a->nnz + 16;
    ocsr c = ocsr_alloc(a->rows, b->cols, out_cap);
    int nnz = 0;
    for (int i = 0; i < a->rows; ++i) {
        int m = 0;
        for (int ka = a->rp[i]; ka < a->rp[i + 1]; ++ka) {
            const int k = a->ci[ka];
            const double y = a->va[ka];
            for (int kb = b->rp[k]; kb < b->rp[k + 1]; ++kb) {
                const int j = b->ci[kb];
                const double v = b->va[kb] * y;
                if (stamp[j] != i) {
                    stamp[j] = i;
                    acc[j] = v;
                    if (m == cap) {
                        cap *= 2;
                        touched = (int*)realloc(touched, sizeof(int) * (size_t)cap);
                    }
                    /* insertion into the sorted list of touched columns */
                    int q = m++;
                    while (q > 0 && touched[q - 1] > j) {
                        touched[q] = touched[q - 1];
                        --q;
                    }
                    touched[q] = j;
                } else {
                    acc[j] += v;
                }
            }
        }
        if (nnz + m > out_cap) {
            while (nnz + m > out_cap) {
                out_cap *= 2;
            }
            c.ci = (int*)realloc(c.ci, sizeof(int) * (size_t)out_cap);
            c.va = (double*)realloc(c.va, sizeof(double) * (size_t)out_cap);
        }
        for (int q = 0; q < m; ++q) {
            c.ci[nnz] = touched[q];
            c.va[nnz] = acc[touched[q]];
            ++nnz;
        }
        c.rp[i + 1] = nnz;
    }
    c.nnz = nnz;
    free(acc);
    free(stamp);
    free(touched);
    return c;
}

static ocsr ocsr_transpose(const ocsr* a) {
    ocsr t = ocsr_alloc(a->cols, a->rows, a->nnz);
    for (int k = 0; k < a->nnz; ++k) {
        t.rp[a->ci[k] + 1] += 1;
    }
    for (int i = 0; i < t.rows; ++i) {
        t.rp[i + 1] += t.rp[i];
    }
    int* pos = (int*)xmalloc(sizeof(int) * ((size_t)t.rows + 1));
    memcpy(pos, t.rp, sizeof(int) * ((size_t)t.rows + 1));
    for (int i = 0; i < a->rows; ++i) {
        for (int k = a->rp[i]; k < a->rp[i + 1]; ++k) {
            const int d = pos[a->ci[k]]++;
            t.ci[d] = i;
            t.va[d] = a->va[k];
        }
    }
    free(pos);
    return t;
}

/* prune(0.0): keeps entries with !(|v| <= 0) (amg.cpp:172, 178). */
static void ocsr_prune(ocsr* a) {
    int out = 0;
    int start = 0;
    for (int i = 0; i < a->rows; ++i) {
        const int end = a->rp[i + 1];
        for (int k = start; k < end; ++k) {
            if (!(fabs(a->va[k]) <= 0.0)) {
                a->ci[out] = a->ci[k];
                a->va[out] = a->va[k];
                ++out;
            }
        }
        start = end;
        a->rp[i + 1] = out;
    }
    a->nnz = out;
}

/* diagonal (amg.cpp:82-92). */
static void orc_diagonal(const ocsr* a, double* d) {
    for (int i = 0; i < a->rows; ++i) {
        d[i] = 0.0;
        for (int k = a->rp[i]; k < a->rp[i + 1]; ++k) {
            if (a->ci[k] == i) {
                d[i] = a->va[k];
            }
        }
    }
}

/* amg_aggregate (amg.cpp:12-78). */
int orc_aggregate(int n, const int* rp, const int* ci, const double* va, const double* diag,
                  double theta, int* agg) {
#define ORC_STRONG(i, k) ((i) != ci[k] && fabs(va[k]) > theta * sqrt(fabs(diag[(i)] * diag[ci[k]])))
    for (int i = 0; i < n; ++i) {
        agg[i] = -1;
    }
    int n_agg = 0;
    /* Pass 1: seed aggregates from nodes whose strong neighbourhood is still free. */
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) {
            continue;
        }
        int free_ = 1;
        int has_strong = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ORC_STRONG(i, k)) {
                has_strong = 1;
                if (agg[ci[k]] >= 0) {
                    free_ = 0;
                    break;
                }
            }
        }
        if (!free_ || !has_strong) {
            continue;
        }
        agg[i] = n_agg;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ORC_STRONG(i, k)) {
                agg[ci[k]] = n_agg;
            }
        }
        ++n_agg;
    }
    /* Pass 2: attach to the aggregate of the strongest pass-1 neighbour (strict >). */
    int* pass1 = (int*)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(pass1, agg, sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) {
            continue;
        }
        double best = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (ORC_STRONG(i, k) && pass1[j] >= 0 && fabs(va[k]) > best) {
                best = fabs(va[k]);
                agg[i] = pass1[j];
            }
        }
    }
    free(pass1);
    /* Pass 3: leftovers form their own aggregates with their still-free strong neighbours. */
    for (int i = 0; i < n; ++i) {
        if (agg[i] < 0) {
            agg[i] = n_agg;
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                if (ORC_STRONG(i, k) && agg[ci[k]] < 0) {
                    agg[ci[k]] = n_agg;
                }
            }
            ++n_agg;
        }
    }
#undef ORC_STRONG
    return n_agg;
}

typedef struct {
    ocsr a, p, r;
    int* agg;
} olevel;

struct orc_amg {
    int n_levels;
    olevel* levels;
    /* LDL^T factor of the coarsest operator (dense, unit lower L below the diagonal, D on it). */
    int coarse_ok;
    double* coarse_factor;
};

/* Dense LDL^T (no pivoting) in place on a row-major n x n matrix; 0 on a zero pivot. */
static int orc_ldlt(int n, double* m) {
    for (int j = 0; j < n; ++j) {
        double dj = m[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) {
            const double ljk = m[(size_t)j * n + k];
            dj -= ljk * ljk * m[(size_t)k * n + k];
        }
        if (dj == 0.0 || !isfinite(dj)) {
            return 0;
        }
        m[(size_t)j * n + j] = dj;
        for (int i = j + 1; i < n; ++i) {
            double s = m[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) {
                s -= m[(size_t)i * n + k] * m[(size_t)j * n + k] * m[(size_t)k * n + k];
            }
            m[(size_t)i * n + j] = s / dj;
        }
    }
    return 1;
}

static void orc_ldlt_solve(int n, const double* f, const double* b, double* x) {
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) {
            s -= f[(size_t)i * n + k] * x[k];
        }
        x[i] = s;
    }
    for (int i = 0; i < n; ++i) {
        x[i] /= f[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) {
            s -= f[(size_t)k * n + i] * x[k];
        }
        x[i] = s;
    }
}

/* AmgPreconditioner::setup (amg.cpp:122-190). */
orc_amg* orc_amg_setup(int n, const int* rp, const int* ci, const double* va, double theta,
                       int max_levels, int coarse_size) {
    return orc_amg_setup_decay(n, rp, ci, va, theta, 1.0, max_levels, coarse_size);
}

orc_amg* orc_amg_setup_decay(int n, const int* rp, const int* ci, const double* va, double theta,
                             double decay, int max_levels, int coarse_size) {
    double theta_l = theta;
    orc_amg* h = (orc_amg*)calloc(1, sizeof(orc_amg));
    h->levels = (olevel*)calloc((size_t)(max_levels > 0 ? max_levels : 1), sizeof(olevel));
    ocsr current = ocsr_copy(n, n, rp, ci, va);
    for (int l = 0; l < max_levels; ++l) {
        olevel* lv = &h->levels[h->n_levels];
        lv->a = current;
        const int nr = current.rows;
        double* d = (double*)xmalloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
        double* inv_diag = (double*)xmalloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
        orc_diagonal(&current, d);
        for (int i = 0; i < nr; ++i) {
            inv_diag[i] = d[i] != 0.0 ? 1.0 / d[i] : 0.0;
        }
        const int last = nr <= coarse_size || l + 1 == max_levels;
        if (last) {
            h->n_levels++;
            free(d);
            free(inv_diag);
            break;
        }
        int* agg = (int*)xmalloc(sizeof(int) * (size_t)(nr > 0 ? nr : 1));
        if (l > 0) {
            theta_l = theta * pow(decay, (double)l);
        }
        const int n_agg = orc_aggregate(nr, current.rp, current.ci, current.va, d, theta_l, agg);
        if (n_agg >= nr || n_agg == 0) {
            free(agg);
            h->n_levels++;
            free(d);
            free(inv_diag);
            break;
        }
        /* Tentative prolongation P_tent(i, agg_i) = 1 (amg.cpp:152-158). */
        ocsr pt = ocsr_alloc(nr, n_agg, nr);
        for (int i = 0; i < nr; ++i) {
            pt.rp[i + 1] = i + 1;
            pt.ci[i] = agg[i];
            pt.va[i] = 1.0;
        }
        /* rho = max_i (sum_j |a_ij|) |inv_diag_i|, omega = 4 / (3 rho) (amg.cpp:160-168). */
        double rho = 0.0;
        for (int i = 0; i < nr; ++i) {
            double row = 0.0;
            for (int k = current.rp[i]; k < current.rp[i + 1]; ++k) {
                row += fabs(current.va[k]);
            }
            const double v = row * fabs(inv_diag[i]);
            rho = rho > v ? rho : v;
        }
        const double omega = rho > 0.0 ? 4.0 / (3.0 * rho) : 0.0;
        /* dinv_a = diag(inv_diag) * A (amg.cpp:169). */
        ocsr dinv_a = ocsr_copy(nr, nr, current.rp, current.ci, current.va);
        for (int i = 0; i < nr; ++i) {
            for (int k = dinv_a.rp[i]; k < dinv_a.rp[i + 1]; ++k) {
                dinv_a.va[k] = dinv_a.va[k] * inv_diag[i];
            }
        }
        ocsr smoothed = ocsr_mul(&dinv_a, &pt);
        /* P = P_tent - omega * smoothed: union of the patterns in column order; an entry
         * present in only one operand takes 0 for the other (amg.cpp:171). */
        ocsr p = ocsr_alloc(nr, n_agg, smoothed.nnz + nr);
        int np = 0;
        for (int i = 0; i < nr; ++i) {
            int ka = pt.rp[i];
            const int ea = pt.rp[i + 1];
            int kb = smoothed.rp[i];
            const int eb = smoothed.rp[i + 1];
            while (ka < ea || kb < eb) {
                int col;
                double v;
                if (ka < ea && kb < eb && pt.ci[ka] == smoothed.ci[kb]) {
                    col = pt.ci[ka];
                    v = pt.va[ka] - omega * smoothed.va[kb];
                    ++ka;
                    ++kb;
                } else if (ka < ea && (kb >= eb || pt.ci[ka] < smoothed.ci[kb])) {
                    col = pt.ci[ka];
                    v = pt.va[ka] - 0.0;
                    ++ka;
                } else {
                    col = smoothed.ci[kb];
                    v = 0.0 - omega * smoothed.va[kb];
                    ++kb;
                }
                p.ci[np] = col;
                p.va[np] = v;
                ++np;
            }
            p.rp[i + 1] = np;
        }
        p.nnz = np;
        ocsr_prune(&p);
        ocsr r = ocsr_transpose(&p);
        ocsr ap = ocsr_mul(&current, &p);
        ocsr coarse = ocsr_mul(&r, &ap);
        ocsr_prune(&coarse);
        lv->p = p;
        lv->r = r;
        lv->agg = agg;
        h->n_levels++;
        ocsr_free(&pt);
        ocsr_free(&dinv_a);
        ocsr_free(&smoothed);
        ocsr_free(&ap);
        free(d);
        free(inv_diag);
        current = coarse;
    }
    /* Coarsest-level factorisation (amg.cpp:183-189): our own dense LDL^T in natural order. */
    const ocsr* c = &h->levels[h->n_levels - 1].a;
    const int nc = c->rows;
    h->coarse_factor = (double*)calloc((size_t)nc * (size_t)nc + 1, sizeof(double));
    for (int i = 0; i < nc; ++i) {
        for (int k = c->rp[i]; k < c->rp[i + 1]; ++k) {
            h->coarse_factor[(size_t)i * nc + c->ci[k]] = c->va[k];
        }
    }
    h->coarse_ok = orc_ldlt(nc, h->coarse_factor);
    return h;
}

void orc_amg_free(orc_amg* h) {
    if (!h) {
        return;
    }
    for (int l = 0; l < h->n_levels; ++l) {
        ocsr_free(&h->levels[l].a);
        ocsr_free(&h->levels[l].p);
        ocsr_free(&h->levels[l].r);
        free(h->levels[l].agg);
    }
    free(h->levels);
    free(h->coarse_factor);
    free(h);
}

int orc_amg_levels(const orc_amg* h) { return h->n_levels; }

static const ocsr* orc_pick(const orc_amg* h, int level, int which) {
    const olevel* l = &h->levels[level];
    return which == 0 ? &l->a : (which == 1 ? &l->p : &l->r);
}

void orc_amg_size(const orc_amg* h, int level, int which, int* rows, int* cols, int* nnz) {
    const ocsr* m = orc_pick(h, level, which);
    *rows = m->rows;
    *cols = m->cols;
    *nnz = m->rp ? m->rp[m->rows] : 0;
}

void orc_amg_copy(const orc_amg* h, int level, int which, int* rp, int* ci, double* va) {
    const ocsr* m = orc_pick(h, level, which);
    if (!m->rp) {
        return;
    }
    const int nnz = m->rp[m->rows];
    if (rp) {
        memcpy(rp, m->rp, sizeof(int) * ((size_t)m->rows + 1));
    }
    if (ci) {
        memcpy(ci, m->ci, sizeof(int) * (size_t)nnz);
    }
    if (va) {
        memcpy(va, m->va, sizeof(double) * (size_t)nnz);
    }
}

void orc_amg_aggregates(const orc_amg* h, int level, int* agg) {
    memcpy(agg, h->levels[level].agg, sizeof(int) * (size_t)h->levels[level].a.rows);
}

/* coarse_inverse (amg.cpp:227-240): column j = A_c^-1 e_j, stored row-major. */
int orc_amg_coarse_inverse(const orc_amg* h, double* out) {
    const int nc = h->levels[h->n_levels - 1].a.rows;
    if (!h->coarse_ok) {
        return 0;
    }
    double* e = (double*)calloc((size_t)nc + 1, sizeof(double));
    double* col = (double*)calloc((size_t)nc + 1, sizeof(double));
    for (int j = 0; j < nc; ++j) {
        memset(e, 0, sizeof(double) * (size_t)nc);
        e[j] = 1.0;
        orc_ldlt_solve(nc, h->coarse_factor, e, col);
        for (int i = 0; i < nc; ++i) {
            out[(size_t)i * nc + j] = col[i];
        }
    }
    free(e);
    free(col);
    return 1;
}

/* ============================ Jacobi-V-cycle AMG-PCG ============================ */

typedef struct {
    int n;
    const int *arp, *aci, *prp, *pci, *rrp, *rci;
    const double *ava, *pva, *rva;
    double* dinv;
    double* x;
    double* b;
    double* t;
} ojl;

/* amg_jacobi (krylov_kernels.inc:182-195). */
static void orc_jacobi(const ojl* lv, int sweeps, double w) {
    const int n = lv->n;
    double* x = lv->x;
    const double* b = lv->b;
    double* t = lv->t;
    const double* d = lv->dinv;
    for (int s = 0; s < sweeps; ++s) {
        orc_spmv(n, lv->arp, lv->aci, lv->ava, x, t);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            x[i] += w * d[i] * (b[i] - t[i]);
        }
    }
}

/* dense_mv (krylov_kernels.inc:197-207). */
static void orc_dense_mv(int n, const double* m, const double* b, double* x) {
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            s += m[(long long)i * n + j] * b[j];
        }
        x[i] = s;
    }
}

/* vcycle (krylov_kernels.inc:209-238). */
static void orc_vcycle(const ojl* levels, int n_levels, int l, const double* ci, int sweeps,
                       double w) {
    const ojl* lv = &levels[l];
    const int n = lv->n;
    if (l == n_levels - 1) {
        orc_dense_mv(n, ci, lv->b, lv->x);
        return;
    }
    double* x = lv->x;
    double* t = lv->t;
    const double* b = lv->b;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        x[i] = 0.0;
    }
    orc_jacobi(lv, sweeps, w);
    orc_spmv(n, lv->arp, lv->aci, lv->ava, x, t);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        t[i] = b[i] - t[i];
    }
    orc_spmv(levels[l + 1].n, lv->rrp, lv->rci, lv->rva, t, levels[l + 1].b);
    orc_vcycle(levels, n_levels, l + 1, ci, sweeps, w);
    orc_spmv(n, lv->prp, lv->pci, lv->pva, levels[l + 1].x, t);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        x[i] += t[i];
    }
    orc_jacobi(lv, sweeps, w);
}

/* precondition lambda of amg_cg (krylov_kernels.inc:259-265): b_0 = r, V-cycle into x_0. */
static void orc_precondition(const ojl* levels, int nl, const double* coarse_inverse, int sweeps,
                             double w, const double* r) {
    const int n = levels[0].n;
    double* zb = levels[0].b;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        zb[i] = r[i];
    }
    orc_vcycle(levels, nl, 0, coarse_inverse, sweeps, w);
}

/* amg_cg (krylov_kernels.inc:240-301) with the work layout of solve_amg_openmp
 * (krylov_openmp.cpp:187-255): zero-initialised per-level x / b / t and CG r / p / q. */
static orc_stats orc_amg_cg_levels(ojl* levels, int nl, const double* coarse_inverse,
                                   int sweeps, double w, const double* b, double* x, double tol,
                                   int max_it) {
    const int n = levels[0].n;
    double* cg = (double*)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
    double* r = cg;
    double* p = cg + n;
    double* q = cg + 2 * n;
    orc_stats st = {0, 0, 0.0};
    orc_spmv(n, levels[0].arp, levels[0].aci, levels[0].ava, x, q);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
    }
    double res = sqrt(orc_dot(n, r, r));
    if (res <= tol) {
        st.residual = res;
        st.converged = 1;
        free(cg);
        return st;
    }
    double* z = levels[0].x;
    orc_precondition(levels, nl, coarse_inverse, sweeps, w, r);
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
    for (int i = 0; i < n; ++i) {
        p[i] = z[i];
    }
    double rz = orc_dot(n, r, z);
    for (st.iterations = 1; st.iterations <= max_it; ++st.iterations) {
        orc_spmv(n, levels[0].arp, levels[0].aci, levels[0].ava, p, q);
        const double pq = orc_dot(n, p, q);
        if (pq == 0.0) {
            break;
        }
        const double alpha = rz / pq;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        res = sqrt(orc_dot(n, r, r));
        if (res <= tol) {
            st.converged = 1;
            break;
        }
        orc_precondition(levels, nl, coarse_inverse, sweeps, w, r);
        const double rz_new = orc_dot(n, r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for if (n > ORC_PAR_THRESHOLD)
        for (int i = 0; i < n; ++i) {
            p[i] = z[i] + beta * p[i];
        }
    }
    if (st.iterations > max_it) {
        st.iterations = max_it;
    }
    st.residual = res;
    free(cg);
    return st;
}

static void ojl_alloc(ojl* l, int n) {
    l->n = n;
    l->dinv = (double*)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    l->x = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    l->b = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    l->t = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    orc_inverse_diagonal(n, l->arp, l->aci, l->ava, l->dinv);
}

static void ojl_free(ojl* l) {
    free(l->dinv);
    free(l->x);
    free(l->b);
    free(l->t);
}

orc_stats orc_amg_cg(int n_levels, const int* n, const int* const* a_rp, const int* const* a_ci,
                     const double* const* a_va, const int* const* p_rp, const int* const* p_ci,
                     const double* const* p_va, const int* const* r_rp, const int* const* r_ci,
                     const double* const* r_va, const double* coarse_inverse, int sweeps,
                     double weight, const double* b, double* x, double abs_tol, int max_it) {
    ojl* lv = (ojl*)calloc((size_t)n_levels, sizeof(ojl));
    for (int l = 0; l < n_levels; ++l) {
        lv[l].arp = a_rp[l];
        lv[l].aci = a_ci[l];
        lv[l].ava = a_va[l];
        if (l + 1 < n_levels) {
            lv[l].prp = p_rp[l];
            lv[l].pci = p_ci[l];
            lv[l].pva = p_va[l];
            lv[l].rrp = r_rp[l];
            lv[l].rci = r_ci[l];
            lv[l].rva = r_va[l];
        }
        ojl_alloc(&lv[l], n[l]);
    }
    orc_stats st = orc_amg_cg_levels(lv, n_levels, coarse_inverse, sweeps, weight, b, x, abs_tol,
                                     max_it);
    for (int l = 0; l < n_levels; ++l) {
        ojl_free(&lv[l]);
    }
    free(lv);
    return st;
}

orc_stats orc_amg_cg_h(const orc_amg* h, const int* rp, const int* ci, const double* va,
                       int sweeps, double weight, const double* b, double* x, double abs_tol,
                       int max_it) {
    const int nl = h->n_levels;
    const int nc = h->levels[nl - 1].a.rows;
    double* cinv = (double*)calloc((size_t)nc * (size_t)nc + 1, sizeof(double));
    orc_amg_coarse_inverse(h, cinv);
    ojl* lv = (ojl*)calloc((size_t)nl, sizeof(ojl));
    for (int l = 0; l < nl; ++l) {
        const olevel* s = &h->levels[l];
        if (l == 0 && rp) {
            lv[l].arp = rp;
            lv[l].aci = ci;
            lv[l].ava = va;
        } else {
            lv[l].arp = s->a.rp;
            lv[l].aci = s->a.ci;
            lv[l].ava = s->a.va;
        }
        if (l + 1 < nl) {
            lv[l].prp = s->p.rp;
            lv[l].pci = s->p.ci;
            lv[l].pva = s->p.va;
            lv[l].rrp = s->r.rp;
            lv[l].rci = s->r.ci;
            lv[l].rva = s->r.va;
        }
        ojl_alloc(&lv[l], s->a.rows);
    }
    orc_stats st = orc_amg_cg_levels(lv, nl, cinv, sweeps, weight, b, x, abs_tol, max_it);
    for (int l = 0; l < nl; ++l) {
        ojl_free(&lv[l]);
    }
    free(lv);
    free(cinv);
    return st;
}

/* ============================ CPU SGS AMG-PCG ============================ */

/* gauss_seidel (amg.cpp:94-118). */
static void orc_gauss_seidel(const ocsr* a, const double* inv_diag, const double* b, double* x,
                             int forward) {
    const int n = a->rows;
    for (int q = 0; q < n; ++q) {
        const int i = forward ? q : n - 1 - q;
        double sum = b[i];
        double d = 0.0;
        for (int k = a->rp[i]; k < a->rp[i + 1]; ++k) {
            if (a->ci[k] == i) {
                d = a->va[k];
            } else {
                sum -= a->va[k] * x[a->ci[k]];
            }
        }
        x[i] = d != 0.0 ? sum * inv_diag[i] : x[i];
    }
}

typedef struct {
    double** inv_diag;
    double** res;
    double** rc;
    double** ec;
} sgs_work;

/* cycle (amg.cpp:192-216), 1 pre / 1 post sweep (amg.hpp:21-22). */
static void orc_cycle(const orc_amg* h, const sgs_work* w, int level, const double* b,
                      double* x) {
    const olevel* lv = &h->levels[level];
    const int n = lv->a.rows;
    if (level + 1 == h->n_levels) {
        if (h->coarse_ok) {
            orc_ldlt_solve(n, h->coarse_factor, b, x);
        } else {
            for (int s = 0; s < 20; ++s) {
                orc_gauss_seidel(&lv->a, w->inv_diag[level], b, x, 1);
                orc_gauss_seidel(&lv->a, w->inv_diag[level], b, x, 0);
            }
        }
        return;
    }
    orc_gauss_seidel(&lv->a, w->inv_diag[level], b, x, 1);
    double* res = w->res[level];
    orc_spmv(n, lv->a.rp, lv->a.ci, lv->a.va, x, res);
    for (int i = 0; i < n; ++i) {
        res[i] = b[i] - res[i];
    }
    const int nc = lv->r.rows;
    double* rc = w->rc[level];
    double* ec = w->ec[level];
    orc_spmv(nc, lv->r.rp, lv->r.ci, lv->r.va, res, rc);
    for (int i = 0; i < nc; ++i) {
        ec[i] = 0.0;
    }
    orc_cycle(h, w, level + 1, rc, ec);
    orc_spmv(n, lv->p.rp, lv->p.ci, lv->p.va, ec, res);
    for (int i = 0; i < n; ++i) {
        x[i] += res[i];
    }
    orc_gauss_seidel(&lv->a, w->inv_diag[level], b, x, 0);
}

static sgs_work sgs_work_alloc(const orc_amg* h) {
    const int nl = h->n_levels;
    sgs_work w;
    w.inv_diag = (double**)calloc((size_t)nl, sizeof(double*));
    w.res = (double**)calloc((size_t)nl, sizeof(double*));
    w.rc = (double**)calloc((size_t)nl, sizeof(double*));
    w.ec = (double**)calloc((size_t)nl, sizeof(double*));
    for (int l = 0; l < nl; ++l) {
        const ocsr* m = &h->levels[l].a;
        w.inv_diag[l] = (double*)calloc((size_t)m->rows + 1, sizeof(double));
        double* d = (double*)calloc((size_t)m->rows + 1, sizeof(double));
        orc_diagonal(m, d);
        for (int i = 0; i < m->rows; ++i) {
            w.inv_diag[l][i] = d[i] != 0.0 ? 1.0 / d[i] : 0.0;
        }
        free(d);
        w.res[l] = (double*)calloc((size_t)m->rows + 1, sizeof(double));
        const int nc = l + 1 < nl ? h->levels[l].r.rows : 0;
        w.rc[l] = (double*)calloc((size_t)nc + 1, sizeof(double));
        w.ec[l] = (double*)calloc((size_t)nc + 1, sizeof(double));
    }
    return w;
}

static void sgs_work_free(const orc_amg* h, sgs_work* w) {
    for (int l = 0; l < h->n_levels; ++l) {
        free(w->inv_diag[l]);
        free(w->res[l]);
        free(w->rc[l]);
        free(w->ec[l]);
    }
    free(w->inv_diag);
    free(w->res);
    free(w->rc);
    free(w->ec);
}

/* apply (amg.cpp:218-225): z = M^-1 r, one SGS V-cycle from z = 0. */
void orc_sgs_apply(const orc_amg* h, const double* r, double* z) {
    sgs_work w = sgs_work_alloc(h);
    for (int i = 0; i < h->levels[0].a.rows; ++i) {
        z[i] = 0.0;
    }
    orc_cycle(h, &w, 0, r, z);
    sgs_work_free(h, &w);
}

/* pcg (amg.cpp:253-285) with apply (amg.cpp:218-225). */
orc_stats orc_sgs_pcg(const orc_amg* h, const double* b, double* x, double tol, int max_it) {
    const ocsr* a = &h->levels[0].a;
    const int n = a->rows;
    sgs_work w = sgs_work_alloc(h);
    double* r = (double*)calloc((size_t)n + 1, sizeof(double));
    double* z = (double*)calloc((size_t)n + 1, sizeof(double));
    double* p = (double*)calloc((size_t)n + 1, sizeof(double));
    double* ap = (double*)calloc((size_t)n + 1, sizeof(double));
    orc_stats st = {0, 0, 0.0};
    orc_spmv(n, a->rp, a->ci, a->va, x, r);
    for (int i = 0; i < n; ++i) {
        r[i] = b[i] - r[i];
    }
    st.residual = sqrt(orc_dot(n, r, r));
    if (st.residual <= tol) {
        st.converged = 1;
    } else {
        for (int i = 0; i < n; ++i) {
            z[i] = 0.0;
        }
        orc_cycle(h, &w, 0, r, z);
        for (int i = 0; i < n; ++i) {
            p[i] = z[i];
        }
        double rz = orc_dot(n, r, z);
        for (int it = 1; it <= max_it; ++it) {
            orc_spmv(n, a->rp, a->ci, a->va, p, ap);
            const double pap = orc_dot(n, p, ap);
            if (pap <= 0.0) {
                break;
            }
            const double alpha = rz / pap;
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * ap[i];
            }
            st.iterations = it;
            st.residual = sqrt(orc_dot(n, r, r));
            if (st.residual <= tol) {
                st.converged = 1;
                break;
            }
            for (int i = 0; i < n; ++i) {
                z[i] = 0.0;
            }
            orc_cycle(h, &w, 0, r, z);
            const double rz_new = orc_dot(n, r, z);
            const double beta = rz_new / rz;
            for (int i = 0; i < n; ++i) {
                p[i] = z[i] + beta * p[i];
            }
            rz = rz_new;
        }
    }
    sgs_work_free(h, &w);
    free(r);
    free(z);
    free(p);
    free(ap);
    return st;
}

/* The conservative product C = A B of the setup (ocsr_mul) on caller arrays, explicit zeros
 * kept; c_* must hold cap entries. Returns nnz(C), or -1 when cap is too small. */
int orc_spgemm(int rows, int inner, int cols, const int* a_rp, const int* a_ci, const double* a_va,
               const int* b_rp, const int* b_ci, const double* b_va, int cap, int* c_rp, int* c_ci,
               double* c_va) {
    ocsr a = ocsr_copy(rows, inner, a_rp, a_ci, a_va);
    ocsr b = ocsr_copy(inner, cols, b_rp, b_ci, b_va);
    ocsr c = ocsr_mul(&a, &b);
    int nnz = c.nnz;
    if (nnz <= cap) {
        memcpy(c_rp, c.rp, sizeof(int) * ((size_t)rows + 1));
        memcpy(c_ci, c.ci, sizeof(int) * (size_t)nnz);
        memcpy(c_va, c.va, sizeof(double) * (size_t)nnz);
    } else {
        nnz = -1;
    }
    ocsr_free(&a);
    ocsr_free(&b);
    ocsr_free(&c);
    return nnz;
}


/* ================================================================================================
 * Finite-volume cell kernels (gpu::DeviceFv). Every cell gathers over its faces in ascending
 * face order; arithmetic in the order of src/gpu/src/fv_cell.inc.
 * ============================================================================================== */

/* Gradient of component comp of a field with `stride` values per cell / boundary face into
 * out[c * ostride + ooff + j] (fv_cell.inc:25-73). */
static void orc_fv_cell_gradient(const orc_fv_mesh* m, int c, const double* phi, const double* bv,
                                 int stride, int comp, int least_squares, double* out,
                                 int ostride, int ooff) {
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const double pc = phi[c * stride + comp];
    for (int k = m->start[c]; k < m->start[c + 1]; ++k) {
        const int f = m->faces[k];
        if (least_squares) {
            double dv;
            if (f < m->nif) {
                dv = phi[m->nei[f] * stride + comp] - phi[m->own[f] * stride + comp];
            } else {
                dv = bv[(f - m->nif) * stride + comp] - pc;
            }
            g0 += m->lsq_d[3 * f] * dv;
            g1 += m->lsq_d[3 * f + 1] * dv;
            g2 += m->lsq_d[3 * f + 2] * dv;
        } else if (f < m->nif) {
            const double wf = m->w[f];
            const double pf = wf * phi[m->own[f] * stride + comp] + (1.0 - wf) * phi[m->nei[f] * stride + comp];
            if (m->own[f] == c) {
                g0 += m->sf[3 * f] * pf;
                g1 += m->sf[3 * f + 1] * pf;
                g2 += m->sf[3 * f + 2] * pf;
            } else {
                g0 -= m->sf[3 * f] * pf;
                g1 -= m->sf[3 * f + 1] * pf;
                g2 -= m->sf[3 * f + 2] * pf;
            }
        } else {
            const double b = bv[(f - m->nif) * stride + comp];
            g0 += m->sf[3 * f] * b;
            g1 += m->sf[3 * f + 1] * b;
            g2 += m->sf[3 * f + 2] * b;
        }
    }
    double* o = out + (size_t)c * ostride + ooff;
    if (least_squares) {
        const double* a = m->lsq + 9 * (size_t)c; /* column-major */
        o[0] = a[0] * g0 + a[3] * g1 + a[6] * g2;
        o[1] = a[1] * g0 + a[4] * g1 + a[7] * g2;
        o[2] = a[2] * g0 + a[5] * g1 + a[8] * g2;
    } else {
        const double v = m->vol[c];
        o[0] = g0 / v;
        o[1] = g1 / v;
        o[2] = g2 / v;
    }
}

/* Skewness-corrected Gauss gradient with the least-squares gradient g (3 per cell)
 * (fv_cell.inc:77-118). */
static void orc_fv_cell_gradient_corrected(const orc_fv_mesh* m, int c, const double* phi,
                                           const double* bv, int stride, int comp, const double* g,
                                           double* out, int ostride, int ooff) {
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int k = m->start[c]; k < m->start[c + 1]; ++k) {
        const int f = m->faces[k];
        if (f < m->nif) {
            const int p = m->own[f];
            const int n = m->nei[f];
            const double wf = m->w[f];
            double pf = wf * phi[p * stride + comp] + (1.0 - wf) * phi[n * stride + comp];
            double dot = 0.0;
            for (int j = 0; j < 3; ++j) {
                const double xip = wf * m->cc[3 * p + j] + (1.0 - wf) * m->cc[3 * n + j];
                const double gf = wf * g[3 * p + j] + (1.0 - wf) * g[3 * n + j];
                if (j == 0) {
                    dot = gf * (m->cf[3 * f] - xip);
                } else {
                    dot = dot + gf * (m->cf[3 * f + j] - xip);
                }
            }
            pf += dot;
            if (p == c) {
                g0 += m->sf[3 * f] * pf;
                g1 += m->sf[3 * f + 1] * pf;
                g2 += m->sf[3 * f + 2] * pf;
            } else {
                g0 -= m->sf[3 * f] * pf;
                g1 -= m->sf[3 * f + 1] * pf;
                g2 -= m->sf[3 * f + 2] * pf;
            }
        } else {
            const double b = bv[(f - m->nif) * stride + comp];
            g0 += m->sf[3 * f] * b;
            g1 += m->sf[3 * f + 1] * b;
            g2 += m->sf[3 * f + 2] * b;
        }
    }
    double* o = out + (size_t)c * ostride + ooff;
    const double v = m->vol[c];
    o[0] = g0 / v;
    o[1] = g1 / v;
    o[2] = g2 / v;
}

/* One component's gradient over all cells (fv_openmp.cpp:257-311): scheme 2 first forms the
 * least-squares gradient, then the corrected one. */
static void orc_fv_gradient_loop(const orc_fv_mesh* m, const double* phi, const double* bv, int stride,
                                 int comp, int scheme, double* out, int ostride, int ooff) {
    if (scheme == 2) {
        double* tmp = (double*)malloc(sizeof(double) * 3 * (size_t)m->nc);
        for (int c = 0; c < m->nc; ++c) {
            orc_fv_cell_gradient(m, c, phi, bv, stride, comp, 1, tmp, 3, 0);
        }
        for (int c = 0; c < m->nc; ++c) {
            orc_fv_cell_gradient_corrected(m, c, phi, bv, stride, comp, tmp, out, ostride, ooff);
        }
        free(tmp);
        return;
    }
    for (int c = 0; c < m->nc; ++c) {
        orc_fv_cell_gradient(m, c, phi, bv, stride, comp, scheme == 1, out, ostride, ooff);
    }
}

/* fv_openmp.cpp:130-135 */
void orc_fv_gradient(const orc_fv_mesh* m, const double* phi, const double* bv, int scheme, double* grad) {
    orc_fv_gradient_loop(m, phi, bv, 1, 0, scheme, grad, 3, 0);
}

/* fv_openmp.cpp:137-145: component i of u into columns 3 i .. 3 i + 2. */
void orc_fv_vector_gradient(const orc_fv_mesh* m, const double* u, const double* ub, int scheme,
                            double* grad) {
    for (int i = 0; i < 3; ++i) {
        orc_fv_gradient_loop(m, u, ub, 3, i, scheme, grad, 9, 3 * i);
    }
}

/* fv_cell.inc:167-195 with the gradient of fv_openmp.cpp:147-152. */
void orc_fv_scalar_corrections(const orc_fv_mesh* m, const double* phi, const double* bv, int scheme,
                               const double* flux, const double* gamma, int linear_upwind, double* source) {
    double* g = (double*)malloc(sizeof(double) * 3 * (size_t)m->nc);
    orc_fv_gradient(m, phi, bv, scheme, g);
    for (int c = 0; c < m->nc; ++c) {
        double s = 0.0;
        for (int k = m->start[c]; k < m->start[c + 1]; ++k) {
            const int f = m->faces[k];
            if (f >= m->nif) {
                break;
            }
            const int p = m->own[f];
            const int n = m->nei[f];
            const double fl = flux[f];
            double e = 0.0;
            if (linear_upwind) {
                const int up = fl >= 0.0 ? p : n;
                const double r0 = m->cf[3 * f] - m->cc[3 * up];
                const double r1 = m->cf[3 * f + 1] - m->cc[3 * up + 1];
                const double r2 = m->cf[3 * f + 2] - m->cc[3 * up + 2];
                e -= fl * (g[3 * up] * r0 + g[3 * up + 1] * r1 + g[3 * up + 2] * r2);
            }
            const double wf = m->w[f];
            const double gf0 = wf * g[3 * p] + (1.0 - wf) * g[3 * n];
            const double gf1 = wf * g[3 * p + 1] + (1.0 - wf) * g[3 * n + 1];
            const double gf2 = wf * g[3 * p + 2] + (1.0 - wf) * g[3 * n + 2];
            e += gamma[f] * (m->kv[3 * f] * gf0 + m->kv[3 * f + 1] * gf1 + m->kv[3 * f + 2] * gf2);
            s = p == c ? s + e : s - e;
        }
        source[c] += s;
    }
    free(g);
}

/* Eigen's Vec3 dot and sum (fixed size 3, SSE2 linear-vectorized reduction: the first two
 * terms as a packet, then the third). Parity with Eigen itself is unpinned (Eigen is absent);
 * the order is fixed here and shared with the device kernels. */
static double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
    return (a0 * b0 + a1 * b1) + a2 * b2;
}

/* FlowSolver::pressure_correction, flow_solver.cpp:688-796, GpuCg branch:
 * H(U)/A and rAU (:688-706, off-diagonal product of ldu_matrix.cpp:39-49), face flux and rAU_f
 * (:708-733), Majumdar correction (:735-753), assemble_convection_diffusion with convection 0
 * (krylov_openmp.cpp:286-300), source -= divergence (:760-769), boundary terms (:782-791) and
 * set_reference(0, 0) (:792-794, ldu_matrix.cpp:105-109). */
void orc_pressure_equation(const orc_fv_mesh* m, const double* upper, const double* lower, const double* diag3,
                           const double* source3, const double* u, const double* dc, const int* kind,
                           const double* fixed_velocity, const double* outlet_p, const double* u_prev,
                           const double* phi_prev, double alpha_u, int pin, double* p_diag, double* p_upper,
                           double* p_lower, double* p_source, double* phi_hbya, double* rau_f, double* hbya,
                           double* rau) {
    const int nc = m->nc, nif = m->nif, nf = m->nf;
    double* off = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    double* comp = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    double* divergence = (double*)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    for (int i = 0; i < 3; ++i) {
        for (int c = 0; c < nc; ++c) {
            comp[c] = u[3 * c + i];
            off[c] = 0.0;
        }
        for (int f = 0; f < nif; ++f) {
            const int p = m->own[f], n = m->nei[f];
            off[p] += upper[f] * comp[n];
            off[n] += lower[f] * comp[p];
        }
        for (int c = 0; c < nc; ++c) {
            hbya[3 * c + i] = (source3[3 * c + i] - off[c]) / diag3[3 * c + i];
        }
    }
    for (int c = 0; c < nc; ++c) {
        const double mean = ((diag3[3 * c] + diag3[3 * c + 1]) + diag3[3 * c + 2]) / 3.0;
        rau[c] = m->vol[c] / mean;
    }
    for (int f = 0; f < nif; ++f) {
        const int p = m->own[f], n = m->nei[f];
        const double w = m->w[f];
        const double v0 = w * hbya[3 * p] + (1.0 - w) * hbya[3 * n];
        const double v1 = w * hbya[3 * p + 1] + (1.0 - w) * hbya[3 * n + 1];
        const double v2 = w * hbya[3 * p + 2] + (1.0 - w) * hbya[3 * n + 2];
        phi_hbya[f] = dot3(v0, v1, v2, m->sf[3 * f], m->sf[3 * f + 1], m->sf[3 * f + 2]);
        rau_f[f] = w * rau[p] + (1.0 - w) * rau[n];
    }
    for (int f = nif; f < nf; ++f) {
        const int p = m->own[f], b = f - nif;
        const double* s = m->sf + 3 * f;
        rau_f[f] = rau[p];
        switch (kind[b]) {
        case 0:
        case 1:
            phi_hbya[f] = dot3(fixed_velocity[3 * b], fixed_velocity[3 * b + 1], fixed_velocity[3 * b + 2], s[0], s[1],
                               s[2]);
            break;
        case 2:
            phi_hbya[f] = dot3(hbya[3 * p], hbya[3 * p + 1], hbya[3 * p + 2], s[0], s[1], s[2]);
            break;
        default:
            phi_hbya[f] = 0.0;
            break;
        }
    }
    if (u_prev != NULL && phi_prev != NULL && alpha_u < 1.0) {
        const double beta = 1.0 - alpha_u;
        for (int f = 0; f < nif; ++f) {
            const int p = m->own[f], n = m->nei[f];
            const double w = m->w[f];
            const double* s = m->sf + 3 * f;
            const double uf0 = w * u_prev[3 * p] + (1.0 - w) * u_prev[3 * n];
            const double uf1 = w * u_prev[3 * p + 1] + (1.0 - w) * u_prev[3 * n + 1];
            const double uf2 = w * u_prev[3 * p + 2] + (1.0 - w) * u_prev[3 * n + 2];
            phi_hbya[f] += beta * (phi_prev[f] - dot3(uf0, uf1, uf2, s[0], s[1], s[2]));
        }
        for (int f = nif; f < nf; ++f) {
            if (kind[f - nif] == 2) {
                const int p = m->own[f];
                const double* s = m->sf + 3 * f;
                phi_hbya[f] += beta * (phi_prev[f] - dot3(u_prev[3 * p], u_prev[3 * p + 1], u_prev[3 * p + 2], s[0],
                                                          s[1], s[2]));
            }
        }
    }
    for (int c = 0; c < nc; ++c) {
        p_diag[c] = 0.0;
        p_source[c] = 0.0;
        divergence[c] = 0.0;
    }
    for (int f = 0; f < nif; ++f) {
        const double fl = phi_hbya[f];
        const double g = rau_f[f] * dc[f];
        const double pos = fl > 0.0 ? 0.0 * fl : 0.0;
        const double neg = fl < 0.0 ? 0.0 * fl : 0.0;
        p_upper[f] = neg - g;
        p_lower[f] = -pos - g;
        p_diag[m->own[f]] += pos + g;
        p_diag[m->nei[f]] += -neg + g;
        divergence[m->own[f]] += fl;
        divergence[m->nei[f]] -= fl;
    }
    for (int c = 0; c < nc; ++c) {
        p_source[c] -= divergence[c];
    }
    for (int f = nif; f < nf; ++f) {
        const int p = m->own[f], b = f - nif;
        p_source[p] -= phi_hbya[f];
        if (kind[b] == 2) {
            const double g = rau_f[f] * dc[f];
            p_diag[p] += g;
            p_source[p] += g * outlet_p[b];
        }
    }
    if (pin && nc > 0) {
        p_source[0] += p_diag[0] * 0.0;
        p_diag[0] += p_diag[0];
    }
    free(off);
    free(comp);
    free(divergence);
}

/* Barth-Jespersen / Venkatakrishnan limiter (fv_cell.inc:123-164). */
void orc_fv_limiter(const orc_fv_mesh* m, const double* u, const double* ub, const double* grad, int venkat,
                    const double* eps2, double* lim) {
    const double e[3] = {venkat ? eps2[0] : 0.0, venkat ? eps2[1] : 0.0, venkat ? eps2[2] : 0.0};
    for (int c = 0; c < m->nc; ++c) {
        double lo[3], hi[3];
        for (int i = 0; i < 3; ++i) {
            lo[i] = u[3 * c + i];
            hi[i] = u[3 * c + i];
        }
        for (int k = m->start[c]; k < m->start[c + 1]; ++k) {
            const int f = m->faces[k];
            const double* v = f < m->nif ? u + 3 * (m->own[f] == c ? m->nei[f] : m->own[f]) : ub + 3 * (f - m->nif);
            for (int i = 0; i < 3; ++i) {
                lo[i] = v[i] < lo[i] ? v[i] : lo[i];
                hi[i] = v[i] > hi[i] ? v[i] : hi[i];
            }
        }
        double l[3] = {1.0, 1.0, 1.0};
        const double* gc = grad + 9 * (size_t)c;
        for (int k = m->start[c]; k < m->start[c + 1]; ++k) {
            const int f = m->faces[k];
            const double r0 = m->cf[3 * f] - m->cc[3 * c];
            const double r1 = m->cf[3 * f + 1] - m->cc[3 * c + 1];
            const double r2 = m->cf[3 * f + 2] - m->cc[3 * c + 2];
            for (int i = 0; i < 3; ++i) {
                const double d = gc[3 * i] * r0 + gc[3 * i + 1] * r1 + gc[3 * i + 2] * r2;
                if (d < 1e-300 && d > -1e-300) {
                    continue;
                }
                const double bound = d > 0.0 ? hi[i] - u[3 * c + i] : lo[i] - u[3 * c + i];
                double q;
                if (venkat) {
                    q = (bound * bound + e[i] + 2.0 * d * bound) / (bound * bound + 2.0 * d * d + d * bound + e[i]);
                } else {
                    q = bound / d > 0.0 ? bound / d : 0.0;
                }
                l[i] = q < l[i] ? q : l[i];
            }
        }
        for (int i = 0; i < 3; ++i) {
            lim[3 * c + i] = l[i];
        }
    }
}

/* fv_cell.inc:199-260 with the limiter pass of fv_openmp.cpp:202-246 (only when limited and
 * linear upwind). */
void orc_fv_momentum_corrections(const orc_fv_mesh* m, const double* u, const double* ub, const double* grad,
                                 const double* flux, const double* nu_face, int linear_upwind, int limited,
                                 int venkat, const double* eps2, double central, int transpose_term,
                                 double* source) {
    const int use_lim = limited && linear_upwind;
    double* lim = NULL;
    if (use_lim) {
        lim = (double*)malloc(sizeof(double) * 3 * (size_t)m->nc);
        orc_fv_limiter(m, u, ub, grad, venkat, eps2, lim);
    }
    for (int c = 0; c < m->nc; ++c) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int k = m->start[c]; k < m->start[c + 1]; ++k) {
            const int f = m->faces[k];
            if (f >= m->nif) {
                break;
            }
            const int p = m->own[f];
            const int n = m->nei[f];
            const double fl = flux[f];
            double e[3] = {0.0, 0.0, 0.0};
            if (linear_upwind) {
                const int up = fl >= 0.0 ? p : n;
                const double r0 = m->cf[3 * f] - m->cc[3 * up];
                const double r1 = m->cf[3 * f + 1] - m->cc[3 * up + 1];
                const double r2 = m->cf[3 * f + 2] - m->cc[3 * up + 2];
                const double* gu = grad + 9 * (size_t)up;
                for (int i = 0; i < 3; ++i) {
                    double d = gu[3 * i] * r0 + gu[3 * i + 1] * r1 + gu[3 * i + 2] * r2;
                    if (central > 0.0) {
                        const double wc = m->w[f];
                        const double uf = wc * u[3 * p + i] + (1.0 - wc) * u[3 * n + i];
                        d = (1.0 - central) * d + central * (uf - u[3 * up + i]);
                    }
                    if (use_lim) {
                        d = lim[3 * up + i] * d;
                    }
                    e[i] -= fl * d;
                }
            }
            const double wf = m->w[f];
            double gf[9];
            for (int j = 0; j < 9; ++j) {
                gf[j] = wf * grad[9 * (size_t)p + j] + (1.0 - wf) * grad[9 * (size_t)n + j];
            }
            const double nu = nu_face[f];
            const double k0 = m->kv[3 * f];
            const double k1 = m->kv[3 * f + 1];
            const double k2 = m->kv[3 * f + 2];
            for (int i = 0; i < 3; ++i) {
                e[i] += nu * (gf[3 * i] * k0 + gf[3 * i + 1] * k1 + gf[3 * i + 2] * k2);
            }
            if (transpose_term) {
                const double a0 = m->sf[3 * f];
                const double a1 = m->sf[3 * f + 1];
                const double a2 = m->sf[3 * f + 2];
                for (int i = 0; i < 3; ++i) {
                    e[i] += nu * (gf[i] * a0 + gf[i + 3] * a1 + gf[i + 6] * a2);
                }
            }
            for (int i = 0; i < 3; ++i) {
                s[i] = p == c ? s[i] + e[i] : s[i] - e[i];
            }
        }
        for (int i = 0; i < 3; ++i) {
            source[3 * c + i] += s[i];
        }
    }
    free(lim);
}
