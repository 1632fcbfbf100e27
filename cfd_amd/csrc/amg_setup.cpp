// Smoothed-aggregation AMG setup on the host (product path).
//
// Restates fv::AmgPreconditioner::setup (src/fv/src/amg.cpp:122-190) and amg_aggregate
// (amg.cpp:12-78) without Eigen. The sparse products run row-parallel with OpenMP; within a row
// every entry is accumulated over the left operand's columns in ascending order starting from
// the first product, the order of Eigen's conservative sparse product, so values, exact zeros
// and therefore the pruned patterns do not depend on the thread count. This file is compiled
// with -ffp-contract=off: the x86-64 reference has no fused multiply-add either.
#include "amg_device.hpp"
#include "common.hpp"
#include "dense.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <string>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <memory>

namespace cfdg {

namespace {

struct Level {
    Csr a;
    Csr p; // n_l x n_{l+1}; empty on the coarsest level
    Csr r; // n_{l+1} x n_l
    std::vector<int> agg;
};

} // namespace

void setup_phase(const char* what, int level) {
    static const bool on = std::getenv("CFDG_SETUP_TIMES") != nullptr;
    thread_local std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    const auto now = std::chrono::steady_clock::now();
    if (on) {
        std::fprintf(stderr, "setup %s level %d: %.2f ms\n", what, level,
                     std::chrono::duration<double, std::milli>(now - last).count());
    }
    last = now;
}

struct AmgHierarchy {
    cfdg_amg_settings settings{};
    std::vector<Level> levels;
    std::vector<double> coarse_inverse; // row-major n_c x n_c
    int coarse_on_device = 0;           // 1: inverted on the matrix cores
    double setup_ms = 0.0;
};

namespace {

std::vector<double> diagonal(const Csr& a) { // amg.cpp:82-92
    std::vector<double> d(static_cast<std::size_t>(a.n_rows), 0.0);
#pragma omp parallel for schedule(static) if (a.n_rows > 20000)
    for (int i = 0; i < a.n_rows; ++i) {
        for (int k = a.rp[static_cast<std::size_t>(i)]; k < a.rp[static_cast<std::size_t>(i) + 1]; ++k) {
            if (a.ci[static_cast<std::size_t>(k)] == i) {
                d[static_cast<std::size_t>(i)] = a.va[static_cast<std::size_t>(k)];
            }
        }
    }
    return d;
}

// Greedy three-pass aggregation (amg.cpp:12-78). Sequential in row order by definition.
std::vector<int> aggregate(const Csr& a, const std::vector<double>& d, double theta, int& n_agg) {
    const int n = a.n_rows;
    std::vector<int> agg(static_cast<std::size_t>(n), -1);
    const int* rp = a.rp.data();
    const int* ci = a.ci.data();
    const double* va = a.va.data();
    // Strength flags per entry, evaluated once: |a_ij| > theta sqrt(|d_i d_j|), i != j.
    std::vector<unsigned char> strong(a.ci.size(), 0);
#pragma omp parallel for schedule(static) if (n > 20000)
    for (int i = 0; i < n; ++i) {
        const double di = d[static_cast<std::size_t>(i)];
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            strong[static_cast<std::size_t>(k)] =
                (i != j && std::abs(va[k]) > theta * std::sqrt(std::abs(di * d[static_cast<std::size_t>(j)])))
                    ? 1
                    : 0;
        }
    }
    n_agg = 0;
    for (int i = 0; i < n; ++i) { // pass 1: roots whose strong neighbourhood is free
        if (agg[static_cast<std::size_t>(i)] >= 0) {
            continue;
        }
        bool free = true;
        bool has_strong = false;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (strong[static_cast<std::size_t>(k)] != 0) {
                has_strong = true;
                if (agg[static_cast<std::size_t>(ci[k])] >= 0) {
                    free = false;
                    break;
                }
            }
        }
        if (!free || !has_strong) {
            continue;
        }
        agg[static_cast<std::size_t>(i)] = n_agg;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (strong[static_cast<std::size_t>(k)] != 0) {
                agg[static_cast<std::size_t>(ci[k])] = n_agg;
            }
        }
        ++n_agg;
    }
    const std::vector<int> pass1 = agg;
    // Pass 2 reads only the pass-1 snapshot, so rows are independent.
#pragma omp parallel for schedule(static) if (n > 20000)
    for (int i = 0; i < n; ++i) {
        if (pass1[static_cast<std::size_t>(i)] >= 0) {
            continue;
        }
        double best = 0.0;
        int choice = -1;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (strong[static_cast<std::size_t>(k)] != 0 && pass1[static_cast<std::size_t>(j)] >= 0 &&
                std::abs(va[k]) > best) {
                best = std::abs(va[k]);
                choice = pass1[static_cast<std::size_t>(j)];
            }
        }
        agg[static_cast<std::size_t>(i)] = choice;
    }
    for (int i = 0; i < n; ++i) { // pass 3: leftovers become roots and grab free neighbours
        if (agg[static_cast<std::size_t>(i)] < 0) {
            agg[static_cast<std::size_t>(i)] = n_agg;
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                if (strong[static_cast<std::size_t>(k)] != 0 && agg[static_cast<std::size_t>(ci[k])] < 0) {
                    agg[static_cast<std::size_t>(ci[k])] = n_agg;
                }
            }
            ++n_agg;
        }
    }
    return agg;
}

// Whether the device setup aggregates a level on the device (passes 1 and 2 as parallel root
// rounds, the same aggregates as the sequential passes): CFDG_DEVICE_AGGREGATION = "all" for every
// level, "short" for levels whose rows hold at most 8 entries on average, anything else or unset
// for none. The host passes are the default: on 96^3-160^3 boxes they build the hierarchy faster
// than the rounds, whose two-hop root checks repeat until every node is decided.
bool device_aggregation(const Csr& a) {
    const char* e = std::getenv("CFDG_DEVICE_AGGREGATION");
    const std::string mode = e != nullptr ? e : "";
    if (mode == "all") {
        return true;
    }
    return mode == "short" && a.n_rows > 0 && a.nnz() <= 8LL * a.n_rows;
}

// Pass 3 of the aggregation (amg.cpp:65-76) on the aggregates of passes 1 and 2: the nodes still
// free, in order, found aggregates with their strong neighbours that are still free.
void aggregate_pass3(const Csr& a, const std::vector<double>& d, double theta, std::vector<int>& agg, int& n_agg) {
    const int n = a.n_rows;
    for (int i = 0; i < n; ++i) {
        if (agg[static_cast<std::size_t>(i)] >= 0) {
            continue;
        }
        agg[static_cast<std::size_t>(i)] = n_agg;
        const double di = d[static_cast<std::size_t>(i)];
        for (int k = a.rp[static_cast<std::size_t>(i)]; k < a.rp[static_cast<std::size_t>(i) + 1]; ++k) {
            const int j = a.ci[static_cast<std::size_t>(k)];
            const bool strong = i != j && std::abs(a.va[static_cast<std::size_t>(k)]) >
                                              theta * std::sqrt(std::abs(di * d[static_cast<std::size_t>(j)]));
            if (strong && agg[static_cast<std::size_t>(j)] < 0) {
                agg[static_cast<std::size_t>(j)] = n_agg;
            }
        }
        ++n_agg;
    }
}

// Row-wise sparse product C = A * B (Gustavson). Entry (i, j) is the left-to-right sum over the
// columns k of row i of A in storage order; explicit zeros are kept (conservative product).
// lscale, if given, multiplies row i of A first (diag(lscale) * A), value lscale_i * a_ik.
Csr spgemm(const Csr& a, const Csr& b, const double* lscale = nullptr) {
    Csr c;
    c.n_rows = a.n_rows;
    c.n_cols = b.n_cols;
    const int n = a.n_rows;
    std::vector<int> count(static_cast<std::size_t>(n) + 1, 0);
#pragma omp parallel
    {
        std::vector<int> mark(static_cast<std::size_t>(b.n_cols), -1);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            int cnt = 0;
            for (int ka = a.rp[static_cast<std::size_t>(i)]; ka < a.rp[static_cast<std::size_t>(i) + 1]; ++ka) {
                const int k = a.ci[static_cast<std::size_t>(ka)];
                for (int kb = b.rp[static_cast<std::size_t>(k)]; kb < b.rp[static_cast<std::size_t>(k) + 1]; ++kb) {
                    const int j = b.ci[static_cast<std::size_t>(kb)];
                    if (mark[static_cast<std::size_t>(j)] != i) {
                        mark[static_cast<std::size_t>(j)] = i;
                        ++cnt;
                    }
                }
            }
            count[static_cast<std::size_t>(i) + 1] = cnt;
        }
    }
    c.rp.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) {
        c.rp[static_cast<std::size_t>(i) + 1] = c.rp[static_cast<std::size_t>(i)] + count[static_cast<std::size_t>(i) + 1];
    }
    const std::size_t nnz = static_cast<std::size_t>(c.rp[static_cast<std::size_t>(n)]);
    c.ci.resize(nnz);
    c.va.resize(nnz);
#pragma omp parallel
    {
        std::vector<int> mark(static_cast<std::size_t>(b.n_cols), -1);
        std::vector<double> acc(static_cast<std::size_t>(b.n_cols), 0.0);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            int* cols = c.ci.data() + c.rp[static_cast<std::size_t>(i)];
            int cnt = 0;
            const double s = lscale != nullptr ? lscale[i] : 1.0;
            for (int ka = a.rp[static_cast<std::size_t>(i)]; ka < a.rp[static_cast<std::size_t>(i) + 1]; ++ka) {
                const int k = a.ci[static_cast<std::size_t>(ka)];
                const double y = lscale != nullptr ? s * a.va[static_cast<std::size_t>(ka)] : a.va[static_cast<std::size_t>(ka)];
                for (int kb = b.rp[static_cast<std::size_t>(k)]; kb < b.rp[static_cast<std::size_t>(k) + 1]; ++kb) {
                    const int j = b.ci[static_cast<std::size_t>(kb)];
                    const double prod = b.va[static_cast<std::size_t>(kb)] * y;
                    if (mark[static_cast<std::size_t>(j)] != i) {
                        mark[static_cast<std::size_t>(j)] = i;
                        acc[static_cast<std::size_t>(j)] = prod;
                        cols[cnt++] = j;
                    } else {
                        acc[static_cast<std::size_t>(j)] += prod;
                    }
                }
            }
            std::sort(cols, cols + cnt);
            double* vals = c.va.data() + c.rp[static_cast<std::size_t>(i)];
            for (int q = 0; q < cnt; ++q) {
                vals[q] = acc[static_cast<std::size_t>(cols[q])];
            }
        }
    }
    return c;
}

// S = diag(dinv) * A * P_tent where P_tent(k, agg_k) = 1 (amg.cpp:152-158, 169-170).
Csr smooth_tentative(const Csr& a, const std::vector<double>& dinv, const std::vector<int>& agg,
                     int n_agg) {
    Csr pt;
    pt.n_rows = a.n_rows;
    pt.n_cols = n_agg;
    pt.rp.resize(static_cast<std::size_t>(a.n_rows) + 1);
    pt.ci.resize(static_cast<std::size_t>(a.n_rows));
    pt.va.assign(static_cast<std::size_t>(a.n_rows), 1.0);
    for (int i = 0; i <= a.n_rows; ++i) {
        pt.rp[static_cast<std::size_t>(i)] = i;
    }
    for (int i = 0; i < a.n_rows; ++i) {
        pt.ci[static_cast<std::size_t>(i)] = agg[static_cast<std::size_t>(i)];
    }
    return spgemm(a, pt, dinv.data());
}

// P = P_tent - omega * S over the union pattern, then prune(0.0) (amg.cpp:171-172).
Csr prolongation(const Csr& s, const std::vector<int>& agg, double omega) {
    const int n = s.n_rows;
    Csr p;
    p.n_rows = n;
    p.n_cols = s.n_cols;
    std::vector<int> count(static_cast<std::size_t>(n) + 1, 0);
    // Row i of P_tent is the single entry (i, agg_i) = 1.
    auto merged = [&](int i, int* out_c, double* out_v) {
        const int ja = agg[static_cast<std::size_t>(i)];
        bool a_done = false;
        int m = 0;
        for (int k = s.rp[static_cast<std::size_t>(i)]; k < s.rp[static_cast<std::size_t>(i) + 1]; ++k) {
            const int j = s.ci[static_cast<std::size_t>(k)];
            const double ws = omega * s.va[static_cast<std::size_t>(k)];
            if (!a_done && ja < j) {
                if (out_c != nullptr) {
                    out_c[m] = ja;
                    out_v[m] = 1.0 - 0.0;
                }
                ++m;
                a_done = true;
            }
            double v;
            if (!a_done && ja == j) {
                v = 1.0 - ws;
                a_done = true;
            } else {
                v = 0.0 - ws;
            }
            if (!(std::abs(v) <= 0.0)) {
                if (out_c != nullptr) {
                    out_c[m] = j;
                    out_v[m] = v;
                }
                ++m;
            }
        }
        if (!a_done) {
            if (out_c != nullptr) {
                out_c[m] = ja;
                out_v[m] = 1.0;
            }
            ++m;
        }
        return m;
    };
#pragma omp parallel for schedule(static) if (n > 20000)
    for (int i = 0; i < n; ++i) {
        count[static_cast<std::size_t>(i) + 1] = merged(i, nullptr, nullptr);
    }
    p.rp.assign(static_cast<std::size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) {
        p.rp[static_cast<std::size_t>(i) + 1] = p.rp[static_cast<std::size_t>(i)] + count[static_cast<std::size_t>(i) + 1];
    }
    p.ci.resize(static_cast<std::size_t>(p.rp[static_cast<std::size_t>(n)]));
    p.va.resize(p.ci.size());
#pragma omp parallel for schedule(static) if (n > 20000)
    for (int i = 0; i < n; ++i) {
        merged(i, p.ci.data() + p.rp[static_cast<std::size_t>(i)], p.va.data() + p.rp[static_cast<std::size_t>(i)]);
    }
    return p;
}

Csr transpose(const Csr& a) { // R = P^T with sorted columns (amg.cpp:175)
    Csr t;
    t.n_rows = a.n_cols;
    t.n_cols = a.n_rows;
    t.rp.assign(static_cast<std::size_t>(t.n_rows) + 1, 0);
    for (int k = 0; k < a.nnz(); ++k) {
        ++t.rp[static_cast<std::size_t>(a.ci[static_cast<std::size_t>(k)]) + 1];
    }
    for (int i = 0; i < t.n_rows; ++i) {
        t.rp[static_cast<std::size_t>(i) + 1] += t.rp[static_cast<std::size_t>(i)];
    }
    t.ci.resize(a.ci.size());
    t.va.resize(a.va.size());
    std::vector<int> fill(t.rp.begin(), t.rp.end() - 1);
    for (int i = 0; i < a.n_rows; ++i) {
        for (int k = a.rp[static_cast<std::size_t>(i)]; k < a.rp[static_cast<std::size_t>(i) + 1]; ++k) {
            const int j = a.ci[static_cast<std::size_t>(k)];
            const int dst = fill[static_cast<std::size_t>(j)]++;
            t.ci[static_cast<std::size_t>(dst)] = i;
            t.va[static_cast<std::size_t>(dst)] = a.va[static_cast<std::size_t>(k)];
        }
    }
    return t;
}

void prune_zeros(Csr& a) { // prune(0.0): drops entries with |v| <= 0 (amg.cpp:172, 178)
    int out = 0;
    RawVec<int> rp(static_cast<std::size_t>(a.n_rows) + 1, 0);
    for (int i = 0; i < a.n_rows; ++i) {
        for (int k = a.rp[static_cast<std::size_t>(i)]; k < a.rp[static_cast<std::size_t>(i) + 1]; ++k) {
            const double v = a.va[static_cast<std::size_t>(k)];
            if (!(std::abs(v) <= 0.0)) {
                a.ci[static_cast<std::size_t>(out)] = a.ci[static_cast<std::size_t>(k)];
                a.va[static_cast<std::size_t>(out)] = v;
                ++out;
            }
        }
        rp[static_cast<std::size_t>(i) + 1] = out;
    }
    a.rp = std::move(rp);
    a.ci.resize(static_cast<std::size_t>(out));
    a.va.resize(static_cast<std::size_t>(out));
}

// Dense LDL^T without pivoting of a symmetric matrix; false on an exactly zero pivot.
bool ldlt(int n, std::vector<double>& m, std::vector<double>& d) {
    d.assign(static_cast<std::size_t>(n), 0.0);
    const auto at = [n](int i, int j) { return static_cast<std::size_t>(i) * n + j; };
    std::vector<double> ld(static_cast<std::size_t>(n));
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < j; ++k) {
            ld[static_cast<std::size_t>(k)] = m[at(j, k)] * d[static_cast<std::size_t>(k)];
        }
        double dj = m[at(j, j)];
        for (int k = 0; k < j; ++k) {
            dj -= m[at(j, k)] * ld[static_cast<std::size_t>(k)];
        }
        if (dj == 0.0 || !std::isfinite(dj)) {
            return false;
        }
        d[static_cast<std::size_t>(j)] = dj;
#pragma omp parallel for schedule(static) if (n > 128)
        for (int i = j + 1; i < n; ++i) {
            double s = m[at(i, j)];
            for (int k = 0; k < j; ++k) {
                s -= m[at(i, k)] * ld[static_cast<std::size_t>(k)];
            }
            m[at(i, j)] = s / dj;
        }
    }
    return true;
}

// Moore-Penrose pseudo-inverse of a symmetric matrix by cyclic Jacobi rotations; the
// reference uses a complete orthogonal decomposition when its LDL^T fails (amg.cpp:239).
std::vector<double> pseudo_inverse(int n, const std::vector<double>& a) {
    std::vector<double> m = a;
    std::vector<double> v(static_cast<std::size_t>(n) * n, 0.0);
    const auto at = [n](int i, int j) { return static_cast<std::size_t>(i) * n + j; };
    for (int i = 0; i < n; ++i) {
        v[at(i, i)] = 1.0;
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                off += m[at(p, q)] * m[at(p, q)];
            }
        }
        if (off < 1e-30) {
            break;
        }
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const double apq = m[at(p, q)];
                if (apq == 0.0) {
                    continue;
                }
                const double theta = (m[at(q, q)] - m[at(p, p)]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                                 (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0);
                const double s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double mkp = m[at(k, p)];
                    const double mkq = m[at(k, q)];
                    m[at(k, p)] = c * mkp - s * mkq;
                    m[at(k, q)] = s * mkp + c * mkq;
                }
                for (int k = 0; k < n; ++k) {
                    const double mpk = m[at(p, k)];
                    const double mqk = m[at(q, k)];
                    m[at(p, k)] = c * mpk - s * mqk;
                    m[at(q, k)] = s * mpk + c * mqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[at(k, p)];
                    const double vkq = v[at(k, q)];
                    v[at(k, p)] = c * vkp - s * vkq;
                    v[at(k, q)] = s * vkp + c * vkq;
                }
            }
        }
    }
    double lmax = 0.0;
    for (int i = 0; i < n; ++i) {
        lmax = std::max(lmax, std::abs(m[at(i, i)]));
    }
    const double tol = static_cast<double>(n) * 2.220446049250313e-16 * lmax;
    std::vector<double> inv(static_cast<std::size_t>(n) * n, 0.0);
    for (int k = 0; k < n; ++k) {
        const double lam = m[at(k, k)];
        if (std::abs(lam) <= tol) {
            continue;
        }
        for (int i = 0; i < n; ++i) {
            const double vi = v[at(i, k)] / lam;
            for (int j = 0; j < n; ++j) {
                inv[at(i, j)] += vi * v[at(j, k)];
            }
        }
    }
    return inv;
}

// Dense inverse of the coarsest operator, column j = A^-1 e_j (amg.cpp:227-240).
// Inverse of a dense row-major n x n matrix as the reference computes the coarsest one
// (amg.cpp:185-189, 227-240): LDL^T and one solve per column, the pseudo-inverse when LDL^T meets
// a zero pivot.
std::vector<double> host_dense_inverse(int n, std::vector<double> m) {
    const std::vector<double> dense = m;
    std::vector<double> d;
    if (!ldlt(n, m, d)) {
        return pseudo_inverse(n, dense);
    }
    std::vector<double> inv(static_cast<std::size_t>(n) * n, 0.0);
#pragma omp parallel for schedule(static) if (n > 64)
    for (int j = 0; j < n; ++j) {
        std::vector<double> y(static_cast<std::size_t>(n), 0.0);
        y[static_cast<std::size_t>(j)] = 1.0;
        for (int i = 0; i < n; ++i) { // L y = e_j (unit lower triangle)
            double s = y[static_cast<std::size_t>(i)];
            for (int k = 0; k < i; ++k) {
                s -= m[static_cast<std::size_t>(i) * n + k] * y[static_cast<std::size_t>(k)];
            }
            y[static_cast<std::size_t>(i)] = s;
        }
        for (int i = 0; i < n; ++i) {
            y[static_cast<std::size_t>(i)] /= d[static_cast<std::size_t>(i)];
        }
        for (int i = n - 1; i >= 0; --i) { // L^T x = y
            double s = y[static_cast<std::size_t>(i)];
            for (int k = i + 1; k < n; ++k) {
                s -= m[static_cast<std::size_t>(k) * n + i] * y[static_cast<std::size_t>(k)];
            }
            y[static_cast<std::size_t>(i)] = s;
        }
        for (int i = 0; i < n; ++i) {
            inv[static_cast<std::size_t>(i) * n + j] = y[static_cast<std::size_t>(i)];
        }
    }
    return inv;
}

std::vector<double> dense_of(const Csr& a) {
    const int n = a.n_rows;
    std::vector<double> m(static_cast<std::size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int k = a.rp[static_cast<std::size_t>(i)]; k < a.rp[static_cast<std::size_t>(i) + 1]; ++k) {
            m[static_cast<std::size_t>(i) * n + a.ci[static_cast<std::size_t>(k)]] = a.va[static_cast<std::size_t>(k)];
        }
    }
    return m;
}

// The coarsest inverse: on the host, or (device setup) on the matrix cores with the host path
// where the device reports a zero pivot or the order exceeds its limit.
std::vector<double> coarse_inverse(const Csr& a, bool device, int& on_device) {
    const int n = a.n_rows;
    std::vector<double> m = dense_of(a);
    on_device = 0;
    if (device && n >= 1) {
        std::vector<double> inv(static_cast<std::size_t>(n) * n, 0.0);
        const std::vector<int> done = dense_inverse_batch({n}, {m.data()}, {inv.data()}, nullptr);
        if (done[0] != 0) {
            on_device = 1;
            return inv;
        }
    }
    return host_dense_inverse(n, std::move(m));
}

// The setup loop (amg.cpp:128-190). With `device`, the sparse products and the first two
// aggregation passes of every level run on the GPU (DeviceSetup, bit-identical results); the
// third pass, over the few nodes the first two leave, stays here.
void build(AmgHierarchy& h, Csr a0, bool device = false) {
    const cfdg_amg_settings& s = h.settings;
    h.levels.clear();
    std::unique_ptr<DeviceSetup> dev;
    if (device) {
        dev = std::make_unique<DeviceSetup>(a0);
    }
    Csr next = std::move(a0);
    setup_phase("start (device setup object)", 0);
    for (int l = 0; l < s.max_levels; ++l) {
        Level level;
        level.a = std::move(next); // the level keeps its operator; no copy
        const Csr& current = level.a;
        const std::vector<double> d = diagonal(current);
        setup_phase("diagonal", l);
        std::vector<double> inv_diag(d.size());
        for (std::size_t i = 0; i < d.size(); ++i) {
            inv_diag[i] = d[i] != 0.0 ? 1.0 / d[i] : 0.0; // amg.cpp:135
        }
        const bool last = current.n_rows <= s.coarse_size || l + 1 == s.max_levels;
        if (last) {
            h.levels.push_back(std::move(level));
            break;
        }
        int n_agg = 0;
        const double theta = s.strength_threshold * std::pow(s.strength_decay, static_cast<double>(l));
        std::vector<int> agg;
        if (dev && device_aggregation(current)) { // passes 1 and 2 on the device, pass 3 here
            agg = dev->aggregate12(d, theta, n_agg);
            aggregate_pass3(current, d, theta, agg, n_agg);
        } else {
            agg = aggregate(current, d, theta, n_agg);
        }
        setup_phase("aggregation", l);
        if (n_agg >= current.n_rows || n_agg == 0) { // amg.cpp:145-148
            h.levels.push_back(std::move(level));
            break;
        }
        double rho = 0.0; // Gershgorin bound (amg.cpp:160-168); rows summed in column order, the
                          // maximum taken in any order (exact)
#pragma omp parallel for schedule(static) reduction(max : rho) if (current.n_rows > 20000)
        for (int i = 0; i < current.n_rows; ++i) {
            double row = 0.0;
            for (int k = current.rp[static_cast<std::size_t>(i)]; k < current.rp[static_cast<std::size_t>(i) + 1]; ++k) {
                row += std::abs(current.va[static_cast<std::size_t>(k)]);
            }
            rho = std::max(rho, row * std::abs(inv_diag[static_cast<std::size_t>(i)]));
        }
        const double omega = rho > 0.0 ? 4.0 / (3.0 * rho) : 0.0;
        setup_phase("rho", l);
        Csr coarse;
        if (dev) {
            dev->level(inv_diag, agg, n_agg, omega, level.p, level.r, coarse);
        } else {
            const Csr smoothed = smooth_tentative(current, inv_diag, agg, n_agg);
            level.p = prolongation(smoothed, agg, omega);
            level.r = transpose(level.p);
            const Csr ap = spgemm(current, level.p);
            coarse = spgemm(level.r, ap);
            prune_zeros(coarse);
        }
        setup_phase("products", l);
        level.agg = std::move(agg);
        h.levels.push_back(std::move(level));
        next = std::move(coarse);
    }
    h.coarse_inverse = coarse_inverse(h.levels.back().a, device, h.coarse_on_device);
    setup_phase("coarse inverse", static_cast<int>(h.levels.size()) - 1);
}

} // namespace

Csr copy_csr(const cfdg_csr& v, const char* what) {
    require(v.n_rows >= 0 && v.nnz >= 0, what);
    Csr c;
    c.n_rows = v.n_rows;
    c.n_cols = v.n_cols > 0 ? v.n_cols : v.n_rows;
    if (v.n_rows == 0) {
        return c;
    }
    require(v.row_ptr != nullptr && (v.nnz == 0 || (v.col != nullptr && v.val != nullptr)), what);
    require(v.row_ptr[0] == 0 && v.row_ptr[v.n_rows] == v.nnz, what);
    c.rp.resize(static_cast<std::size_t>(v.n_rows) + 1);
    c.ci.resize(static_cast<std::size_t>(v.nnz));
    c.va.resize(static_cast<std::size_t>(v.nnz));
    parallel_copy(c.rp.data(), v.row_ptr, c.rp.size() * sizeof(int));
    parallel_copy(c.ci.data(), v.col, c.ci.size() * sizeof(int));
    parallel_copy(c.va.data(), v.val, c.va.size() * sizeof(double));
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad) if (c.n_rows > 20000)
    for (int i = 0; i < c.n_rows; ++i) {
        bad |= c.rp[static_cast<std::size_t>(i) + 1] < c.rp[static_cast<std::size_t>(i)] ? 1 : 0;
    }
    const int nnz = c.nnz();
    const int nc = c.n_cols;
#pragma omp parallel for schedule(static) reduction(| : bad) if (nnz > 100000)
    for (int k = 0; k < nnz; ++k) {
        const int j = c.ci[static_cast<std::size_t>(k)];
        bad |= (j < 0 || j >= nc) ? 1 : 0;
    }
    require(bad == 0, what);
    return c;
}

// Sorts the columns of every row (values follow); rejects duplicate columns. Rows are
// independent (host threads over rows).
void sort_rows(Csr& a) {
    int dup = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : dup) if (a.n_rows > 20000)
    for (int i = 0; i < a.n_rows; ++i) {
        const int b = a.rp[static_cast<std::size_t>(i)];
        const int e = a.rp[static_cast<std::size_t>(i) + 1];
        bool sorted = true;
        for (int k = b + 1; k < e; ++k) {
            if (a.ci[static_cast<std::size_t>(k)] <= a.ci[static_cast<std::size_t>(k) - 1]) {
                sorted = false;
            }
        }
        if (sorted) {
            continue;
        }
        std::vector<std::pair<int, double>> row;
        for (int k = b; k < e; ++k) {
            row.emplace_back(a.ci[static_cast<std::size_t>(k)], a.va[static_cast<std::size_t>(k)]);
        }
        std::stable_sort(row.begin(), row.end(),
                         [](const auto& x, const auto& y) { return x.first < y.first; });
        for (int k = b; k < e; ++k) {
            a.ci[static_cast<std::size_t>(k)] = row[static_cast<std::size_t>(k - b)].first;
            a.va[static_cast<std::size_t>(k)] = row[static_cast<std::size_t>(k - b)].second;
        }
        for (int k = b + 1; k < e; ++k) {
            dup |= a.ci[static_cast<std::size_t>(k)] == a.ci[static_cast<std::size_t>(k) - 1] ? 1 : 0;
        }
    }
    require(dup == 0, "amg setup: duplicate column in a row");
}

} // namespace cfdg

using namespace cfdg;

struct cfdg_amg {
    AmgHierarchy h;
};

extern "C" cfdg_status cfdg_amg_default_settings(cfdg_amg_settings* s) {
    return guarded([&] {
        require(s != nullptr, "null settings");
        s->strength_threshold = 0.08;
        s->max_levels = 25;
        s->coarse_size = 400;
        s->smooth_sweeps = 2;
        s->jacobi_weight = 0.7;
        s->strength_decay = 1.0;
    });
}

namespace {
cfdg_status amg_build(const cfdg_csr* a, const cfdg_amg_settings* settings, cfdg_amg** out, bool device) {
    return guarded([&] {
        require(a != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        require(a->n_cols == 0 || a->n_cols == a->n_rows, "AmgPreconditioner: matrix must be square");
        require(a->n_rows > 0, "amg setup: empty matrix");
        auto amg = std::make_unique<cfdg_amg>();
        if (settings != nullptr) {
            amg->h.settings = *settings;
        } else {
            cfdg_amg_default_settings(&amg->h.settings);
        }
        require(amg->h.settings.max_levels >= 1 && amg->h.settings.coarse_size >= 1 &&
                    amg->h.settings.strength_decay > 0.0 && amg->h.settings.strength_decay <= 1.0,
                "invalid AMG settings");
        const auto t0 = std::chrono::steady_clock::now();
        Csr c = copy_csr(*a, "amg setup: invalid CSR matrix");
        sort_rows(c);
        build(amg->h, std::move(c), device);
        const auto t1 = std::chrono::steady_clock::now();
        amg->h.setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        *out = amg.release();
    });
}
} // namespace

extern "C" cfdg_status cfdg_amg_build(const cfdg_csr* a, const cfdg_amg_settings* settings,
                                      cfdg_amg** out) {
    return amg_build(a, settings, out, false);
}

extern "C" cfdg_status cfdg_amg_build_device(const cfdg_csr* a, const cfdg_amg_settings* settings,
                                             cfdg_amg** out) {
    return amg_build(a, settings, out, true);
}

extern "C" cfdg_status cfdg_amg_destroy(cfdg_amg* amg) {
    return guarded([&] { delete amg; });
}

extern "C" cfdg_status cfdg_amg_n_levels(const cfdg_amg* amg, int* n_levels) {
    return guarded([&] {
        require(amg != nullptr && n_levels != nullptr, "null argument");
        *n_levels = static_cast<int>(amg->h.levels.size());
    });
}

namespace {
const Csr& pick(const cfdg_amg* amg, int level, int which) {
    require(amg != nullptr, "null hierarchy");
    require(level >= 0 && level < static_cast<int>(amg->h.levels.size()), "level out of range");
    require(which >= 0 && which <= 2, "which must be 0 (a), 1 (p) or 2 (r)");
    const auto& l = amg->h.levels[static_cast<std::size_t>(level)];
    return which == 0 ? l.a : (which == 1 ? l.p : l.r);
}
} // namespace

extern "C" cfdg_status cfdg_amg_level_size(const cfdg_amg* amg, int level, int which, int* n_rows,
                                           int* n_cols, int* nnz) {
    return guarded([&] {
        const Csr& m = pick(amg, level, which);
        if (n_rows != nullptr) {
            *n_rows = m.n_rows;
        }
        if (n_cols != nullptr) {
            *n_cols = m.n_cols;
        }
        if (nnz != nullptr) {
            *nnz = m.nnz();
        }
    });
}

extern "C" cfdg_status cfdg_amg_level_copy(const cfdg_amg* amg, int level, int which,
                                           int* row_ptr, int* col, double* val) {
    return guarded([&] {
        const Csr& m = pick(amg, level, which);
        if (row_ptr != nullptr) {
            std::copy(m.rp.begin(), m.rp.end(), row_ptr);
        }
        if (col != nullptr) {
            std::copy(m.ci.begin(), m.ci.end(), col);
        }
        if (val != nullptr) {
            std::copy(m.va.begin(), m.va.end(), val);
        }
    });
}

extern "C" cfdg_status cfdg_amg_aggregates(const cfdg_amg* amg, int level, int* agg) {
    return guarded([&] {
        require(amg != nullptr && agg != nullptr, "null argument");
        require(level >= 0 && level + 1 < static_cast<int>(amg->h.levels.size()),
                "aggregates exist on non-coarsest levels only");
        const auto& a = amg->h.levels[static_cast<std::size_t>(level)].agg;
        std::copy(a.begin(), a.end(), agg);
    });
}

extern "C" cfdg_status cfdg_amg_coarse_inverse(const cfdg_amg* amg, double* out) {
    return guarded([&] {
        require(amg != nullptr && out != nullptr, "null argument");
        std::copy(amg->h.coarse_inverse.begin(), amg->h.coarse_inverse.end(), out);
    });
}

extern "C" cfdg_status cfdg_amg_view_of(const cfdg_amg* amg, cfdg_amg_level_view* levels,
                                        cfdg_amg_view* view) {
    return guarded([&] {
        require(amg != nullptr && levels != nullptr && view != nullptr, "null argument");
        const int nl = static_cast<int>(amg->h.levels.size());
        auto mk = [](const Csr& m) {
            cfdg_csr v{};
            v.n_rows = m.n_rows;
            v.n_cols = m.n_cols;
            v.nnz = m.nnz();
            v.row_ptr = m.n_rows > 0 ? m.rp.data() : nullptr;
            v.col = m.ci.data();
            v.val = m.va.data();
            return v;
        };
        for (int l = 0; l < nl; ++l) {
            const auto& lv = amg->h.levels[static_cast<std::size_t>(l)];
            levels[l].a = mk(lv.a);
            levels[l].p = mk(lv.p);
            levels[l].r = mk(lv.r);
        }
        view->n_levels = nl;
        view->levels = levels;
        view->coarse_inverse = amg->h.coarse_inverse.data();
        view->smooth_sweeps = amg->h.settings.smooth_sweeps;
        view->jacobi_weight = amg->h.settings.jacobi_weight;
    });
}

extern "C" cfdg_status cfdg_amg_coarse_on_device(const cfdg_amg* amg, int* on_device) {
    return guarded([&] {
        require(amg != nullptr && on_device != nullptr, "null argument");
        *on_device = amg->h.coarse_on_device;
    });
}

extern "C" cfdg_status cfdg_dense_inverse_batch(int count, const int* order, const double* const* a,
                                                double* const* inv, int* on_device) {
    return guarded([&] {
        require(count >= 0, "dense inverse: negative count");
        require(count == 0 || (order != nullptr && a != nullptr && inv != nullptr), "null argument");
        std::vector<int> n(order, order + count);
        std::vector<const double*> in(a, a + count);
        std::vector<double*> out(inv, inv + count);
        for (int i = 0; i < count; ++i) {
            require(n[static_cast<std::size_t>(i)] >= 0, "dense inverse: negative order");
            require(n[static_cast<std::size_t>(i)] == 0 || (in[static_cast<std::size_t>(i)] != nullptr &&
                                                           out[static_cast<std::size_t>(i)] != nullptr),
                    "null argument");
        }
        require_device();
        const std::vector<int> done = dense_inverse_batch(n, in, out, nullptr);
        for (int i = 0; i < count; ++i) {
            const std::size_t k = static_cast<std::size_t>(i);
            if (done[k] == 0 && n[k] > 0) {
                const std::size_t nn = static_cast<std::size_t>(n[k]) * static_cast<std::size_t>(n[k]);
                const std::vector<double> r = host_dense_inverse(n[k], std::vector<double>(in[k], in[k] + nn));
                std::copy(r.begin(), r.end(), out[k]);
            }
            if (on_device != nullptr) {
                on_device[k] = done[k];
            }
        }
    });
}

extern "C" cfdg_status cfdg_amg_setup_ms(const cfdg_amg* amg, double* ms) {
    return guarded([&] {
        require(amg != nullptr && ms != nullptr, "null argument");
        *ms = amg->h.setup_ms;
    });
}
