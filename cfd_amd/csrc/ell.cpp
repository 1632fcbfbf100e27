#include "ell.hpp"

#include <algorithm>
#include <cstdlib>
#include <limits>
#include <mutex>

namespace cfdg {

namespace {

// Rounds of slice s: the longest row of the slice divided by the lanes per row, rounded up.
int slice_rounds(int s, int n_rows, const int* rp, int lanes) {
    const int rows = k_wave / lanes;
    int w = 0;
    const int r_end = std::min(n_rows, (s + 1) * rows);
    for (int r = s * rows; r < r_end; ++r) {
        w = std::max(w, rp[r + 1] - rp[r]);
    }
    return (w + lanes - 1) / lanes;
}

} // namespace

long long ell_slots(int n_rows, const int* rp, int lanes) {
    const int rows = k_wave / lanes;
    const int n_slices = (n_rows + rows - 1) / rows;
    long long total = 0;
    for (int s = 0; s < n_slices; ++s) {
        total += static_cast<long long>(slice_rounds(s, n_rows, rp, lanes)) * k_wave;
    }
    return total;
}

std::vector<int> ell_slice_offsets(int n_rows, const int* rp, int lanes) {
    const int rows = k_wave / lanes;
    const int n_slices = (n_rows + rows - 1) / rows;
    std::vector<int> sp(static_cast<std::size_t>(n_slices) + 1, 0);
    std::vector<long long> w(static_cast<std::size_t>(n_slices));
    for (int s = 0; s < n_slices; ++s) {
        w[static_cast<std::size_t>(s)] = static_cast<long long>(slice_rounds(s, n_rows, rp, lanes)) * k_wave;
    }
    long long total = 0;
    for (int s = 0; s < n_slices; ++s) {
        total += w[static_cast<std::size_t>(s)];
        if (total > std::numeric_limits<int>::max()) {
            throw InvalidArgument("matrix too large for 32-bit sliced-ELL offsets");
        }
        sp[static_cast<std::size_t>(s) + 1] = static_cast<int>(total);
    }
    return sp;
}

bool short_offsets_fit(int n_rows, int n_cols, const int* rp, const int* ci) {
    if (n_rows != n_cols || std::getenv("CFDG_NO_SHORT_COLS") != nullptr) {
        return false;
    }
    int fits = 1;
    for (int r = 0; r < n_rows; ++r) {
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int off = ci[k] - r;
            if (off < -32768 || off > 32767) {
                fits = 0;
            }
        }
    }
    return fits != 0;
}

HostEll csr_to_ell(int n_rows, int n_cols, const int* rp, const int* ci, int lanes) {
    require(lanes >= 1 && lanes <= k_wave && (lanes & (lanes - 1)) == 0,
            "csr_to_ell: lanes per row must be a power of two <= 64");
    HostEll e;
    e.n_rows = n_rows;
    e.n_cols = n_cols;
    e.lanes = lanes;
    const int rows = k_wave / lanes;
    e.n_slices = (n_rows + rows - 1) / rows;
    e.sp.assign(static_cast<std::size_t>(e.n_slices) + 1, 0);
    long long total = 0;
    for (int s = 0; s < e.n_slices; ++s) {
        total += static_cast<long long>(slice_rounds(s, n_rows, rp, lanes)) * k_wave;
        if (total > std::numeric_limits<int>::max()) {
            throw InvalidArgument("matrix too large for 32-bit sliced-ELL offsets");
        }
        e.sp[static_cast<std::size_t>(s) + 1] = static_cast<int>(total);
    }
    e.col.resize(static_cast<std::size_t>(total));
    e.src.resize(static_cast<std::size_t>(total));
    for (int s = 0; s < e.n_slices; ++s) {
        const int base = e.sp[static_cast<std::size_t>(s)];
        const int w = (e.sp[static_cast<std::size_t>(s) + 1] - base) / k_wave;
        for (int lane = 0; lane < k_wave; ++lane) {
            const int r = s * rows + lane / lanes;
            const int sub = lane % lanes;
            const int len = r < n_rows ? rp[r + 1] - rp[r] : 0;
            // rows past the end point at the last row of a square matrix (a short offset)
            int pad = (n_rows == n_cols && n_rows > 0) ? n_rows - 1 : 0;
            if (r < n_rows) {
                pad = (n_rows == n_cols) ? r : (len > 0 ? ci[rp[r] + len - 1] : 0);
            }
            for (int k = 0; k < w; ++k) {
                const std::size_t slot = static_cast<std::size_t>(base) + static_cast<std::size_t>(k) * k_wave + lane;
                const int j = k * lanes + sub;
                if (j < len) {
                    e.col[slot] = ci[rp[r] + j];
                    e.src[slot] = rp[r] + j;
                } else {
                    e.col[slot] = pad;
                    e.src[slot] = -1;
                }
            }
        }
    }
    return e;
}

namespace {
double env_or(const char* name, double dflt) {
    const char* v = std::getenv(name);
    return v != nullptr ? std::atof(v) : dflt;
}
} // namespace

int pick_tpr(int n_rows, long long nnz) {
    if (n_rows <= 0) {
        return 1;
    }
    // Tuning knobs (environment): rows with at most CFDG_ELL_MAX_AVG entries on average stay
    // thread-per-row ELL; longer rows get lanes so that each lane sums about
    // CFDG_TPR_ENTRIES entries.
    static const double ell_max = env_or("CFDG_ELL_MAX_AVG", 12.0);
    static const double per_lane = env_or("CFDG_TPR_ENTRIES", 4.0);
    const double avg = static_cast<double>(nnz) / static_cast<double>(n_rows);
    if (avg <= ell_max) {
        return 1;
    }
    int t = 2;
    while (t < 64 && per_lane * t * 2 <= avg) {
        t *= 2;
    }
    return t;
}

void DevEll::upload(const HostEll& e, const double* va, bool fp64, bool fp32, hipStream_t s) {
    n_rows = e.n_rows;
    n_cols = e.n_cols;
    n_slices = e.n_slices;
    tpr = e.lanes;
    csr = 0;
    sp.upload(e.sp.data(), e.sp.size(), s);
    // Square matrices whose columns all lie within 32767 of the row store 16-bit offsets.
    bool short_ok = e.n_rows == e.n_cols;
    std::vector<short> d;
    if (short_ok) {
        d.resize(e.col.size());
        int fits = 1;
        for (int sl = 0; sl < e.n_slices; ++sl) {
            const int base = e.sp[static_cast<std::size_t>(sl)];
            const int end = e.sp[static_cast<std::size_t>(sl) + 1];
            for (int k = base; k < end; ++k) {
                const int row = sl * (k_wave / e.lanes) + ((k - base) % k_wave) / e.lanes;
                const int off = e.col[static_cast<std::size_t>(k)] - row;
                if (off < -32768 || off > 32767) {
                    fits = 0;
                }
                d[static_cast<std::size_t>(k)] = static_cast<short>(off);
            }
        }
        short_ok = fits != 0;
    }
    cbase.resize(0); // host conversion: row-relative offsets for square matrices only
    soff.resize(0);
    if (short_ok && !std::getenv("CFDG_NO_SHORT_COLS")) {
        dcol.upload(d.data(), d.size(), s);
        col.resize(0);
    } else {
        dcol.resize(0);
        col.upload(e.col.data(), e.col.size(), s);
    }
    if (va != nullptr) {
        if (fp64) {
            const std::vector<double> v = ell_values<double>(e, va);
            v64.upload(v.data(), v.size(), s);
        }
        if (fp32) {
            const std::vector<float> v = ell_values<float>(e, va);
            v32.upload(v.data(), v.size(), s);
        }
    }
}

namespace {
std::mutex g_rules_mutex;
std::vector<cfdg_layout_rule> g_rules;
} // namespace

void set_layout_rules(const cfdg_layout_rule* rules, int count) {
    std::vector<cfdg_layout_rule> r;
    for (int i = 0; i < count; ++i) {
        const cfdg_layout_rule& x = rules[i];
        require(x.lanes >= 1 && x.lanes <= k_wave && (x.lanes & (x.lanes - 1)) == 0,
                "layout rule: lanes must be a power of two <= 64");
        require(x.csr == 0 || x.csr == 1, "layout rule: csr must be 0 or 1");
        require(x.min_rows <= x.max_rows && x.min_mean <= x.max_mean, "layout rule: empty range");
        r.push_back(x);
    }
    std::lock_guard<std::mutex> lk(g_rules_mutex);
    g_rules = std::move(r);
}

Layout choose_layout(int rows, const int* rp) {
    const double mean = rows > 0 ? static_cast<double>(rp[rows]) / rows : 0.0;
    {
        std::lock_guard<std::mutex> lk(g_rules_mutex);
        for (const cfdg_layout_rule& x : g_rules) {
            if (rows >= x.min_rows && rows <= x.max_rows && mean >= x.min_mean && mean <= x.max_mean) {
                return Layout{x.lanes, x.csr};
            }
        }
    }
    const int t = pick_tpr(rows, rp[rows]);
    // Sliced ELL (coalesced, 16-bit columns) unless its padding would exceed the CSR entries by
    // more than CFDG_ELL_MAX_PAD (default 1.5).
    static const double max_pad = env_or("CFDG_ELL_MAX_PAD", 1.5);
    const bool ell = t == 1 || static_cast<double>(ell_slots(rows, rp, t)) <= max_pad * static_cast<double>(rp[rows]);
    return Layout{t, ell ? 0 : 1};
}

void DevEll::upload_csr(int rows, int cols, const int* rp, const int* ci, const double* va,
                        hipStream_t s, const Layout* forced) {
    const Layout lay = forced != nullptr ? *forced : choose_layout(rows, rp);
    const int t = lay.lanes;
    if (lay.csr == 0) {
        if (std::getenv("CFDG_HOST_ELL") != nullptr) { // host conversion (same layout), for checks
            upload(csr_to_ell(rows, cols, rp, ci, t), va, true, true, s);
        } else {
            build_on_device(rows, cols, rp, ci, va, t, s);
        }
        return;
    }
    n_rows = rows;
    n_cols = cols;
    n_slices = 0;
    tpr = t;
    csr = 1;
    const std::size_t nnz = static_cast<std::size_t>(rp[rows]);
    sp.upload(rp, static_cast<std::size_t>(rows) + 1, s);
    dcol.resize(0); // CSR rows always carry 32-bit columns
    cbase.resize(0);
    soff.resize(0);
    col.upload(ci, nnz, s);
    v64.upload(va, nnz, s);
    std::vector<float> f(nnz);
    const long long m = static_cast<long long>(nnz);
    for (long long k = 0; k < m; ++k) {
        f[static_cast<std::size_t>(k)] = static_cast<float>(va[k]);
    }
    v32.upload(f.data(), nnz, s);
}

} // namespace cfdg
