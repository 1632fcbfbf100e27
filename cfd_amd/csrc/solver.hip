// Device-resident systems, the batched solve driver and the C-ABI of the solver.
//
// Reference counterparts: DeviceSolver (src/gpu/src/device_solver.cpp), the CUDA system
// (src/gpu/src/krylov_cuda.cu:566-677) and the solver loops of
// src/gpu/src/krylov_kernels.inc. The host here only enqueues kernels; every scalar of the
// Krylov recurrences lives on the device, and the host reads back one integer per iteration
// (the number of systems still running) one iteration behind the launches.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace cfdg {

namespace {

constexpr int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

__global__ void __launch_bounds__(k_block) k_fill(int n, double* __restrict__ x, double v) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i < n) {
        x[i] = v;
    }
}

__global__ void __launch_bounds__(k_block) k_scale_dinv(int n, const double* __restrict__ dinv, double w,
                                                        double* __restrict__ wd64, float* __restrict__ wd32) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i < n) {
        const double v = w * dinv[i];
        wd64[i] = v;
        wd32[i] = static_cast<float>(v);
    }
}

// Level-0 dot products from a finished V-cycle output v[buf] (used when no Jacobi sweep on
// level 0 could carry them: one-level hierarchies and smoothing-free cycles).
template <class T, int Dots>
__global__ void __launch_bounds__(k_block) k_level0_dots(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int buf) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const LevelDev& L = S.lv[0];
    double d1 = 0.0;
    double d2 = 0.0;
    if (row < S.n) {
        const double z = static_cast<double>(static_cast<const T*>(L.v[buf])[row]);
        if (Dots == 1) {
            d1 = static_cast<double>(static_cast<const T*>(L.b)[row]) * z;
        } else {
            d1 = S.r[row] * z;
            d2 = z * S.q[row];
        }
    }
    double v[2] = {d1, d2};
    double o[2];
    block_sums<2>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_RZ * S.nblk0 + bk_.y] = o[0];
        if (Dots == 2) {
            S.partials[SLOT_ZQ * S.nblk0 + bk_.y] = o[1];
        }
    }
}

// Number of systems whose solve is still running (one workgroup).
__global__ void __launch_bounds__(k_block) k_count_running(const SysDev* __restrict__ sys, int n_sys, int* __restrict__ out) {
    __shared__ int red[k_slices_per_block];
    int c = 0;
    for (int s = threadIdx.x; s < n_sys; s += k_block) {
        const int st = sys[s].st->status;
        c += (st == ST_RUNNING || st == ST_CONVERGED_PENDING) ? 1 : 0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off, k_wave);
    }
    if (threadIdx.x % k_wave == 0) {
        red[threadIdx.x / k_wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = red[0] + red[1] + red[2] + red[3];
    }
}

__global__ void k_gather_states(const SysDev* __restrict__ sys, int n_sys, SolveState* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sys) {
        out[s] = *sys[s].st;
    }
}

__global__ void k_init_states(const SysDev* __restrict__ sys, int n_sys, const SolveState* __restrict__ init) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_sys) {
        *sys[s].st = init[s];
    }
}

// inverse_diagonal (krylov_openmp.cpp:147-157): 1 / a_ii, 1.0 for a zero or missing diagonal.
std::vector<double> inverse_diagonal(const cfdg_csr& m) {
    std::vector<double> d(static_cast<std::size_t>(m.n_rows), 1.0);
    for (int i = 0; i < m.n_rows; ++i) {
        for (int k = m.row_ptr[i]; k < m.row_ptr[i + 1]; ++k) {
            if (m.col[k] == i && m.val[k] != 0.0) {
                d[static_cast<std::size_t>(i)] = 1.0 / m.val[k];
            }
        }
    }
    return d;
}

void check_csr(const cfdg_csr& m, int rows, int cols, const char* what) {
    require(m.n_rows == rows, what);
    require((m.n_cols == 0 ? m.n_rows : m.n_cols) == cols, what);
    require(m.nnz >= 0 && m.row_ptr != nullptr, what);
    require(m.nnz == 0 || (m.col != nullptr && m.val != nullptr), what);
    require(m.row_ptr[0] == 0 && m.row_ptr[rows] == m.nnz, what);
    for (int i = 0; i < rows; ++i) {
        require(m.row_ptr[i + 1] >= m.row_ptr[i], what);
    }
    for (int k = 0; k < m.nnz; ++k) {
        require(m.col[k] >= 0 && m.col[k] < cols, what);
    }
}

} // namespace

void require_device() {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        throw NoDevice("no HIP device available (libcfdgpu has no host fallback)");
    }
}

std::size_t DevLevel::bytes() const {
    std::size_t b = a.bytes() + p.bytes() + r.bytes() + wd64.bytes() + wd32.bytes() + b64.bytes() +
                    b32.bytes() + cinv.bytes();
    for (int i = 0; i < 3; ++i) {
        b += v64[i].bytes() + v32[i].bytes();
    }
    return b;
}

// ------------------------------------------------------------------------------------------
// System
// ------------------------------------------------------------------------------------------

System::System(const cfdg_ldu_pattern& pt) {
    require_device();
    require(pt.n >= 1 && pt.n_faces >= 0 && pt.nnz >= pt.n, "LduPattern: invalid sizes");
    require(pt.row_ptr != nullptr && pt.col != nullptr && pt.diag_pos != nullptr,
            "LduPattern: null arrays");
    require(pt.n_faces == 0 || (pt.upper_pos != nullptr && pt.lower_pos != nullptr),
            "LduPattern: null face positions");
    require(pt.row_ptr[0] == 0 && pt.row_ptr[pt.n] == pt.nnz, "LduPattern: row_ptr mismatch");
    for (int k = 0; k < pt.nnz; ++k) {
        require(pt.col[k] >= 0 && pt.col[k] < pt.n, "LduPattern: column out of range");
    }
    n_ = pt.n;
    nnz0_ = pt.nnz;
    nf_ = pt.n_faces;
    std::vector<int> csr_src(static_cast<std::size_t>(pt.nnz), -1);
    auto place = [&](int pos, int src) {
        require(pos >= 0 && pos < pt.nnz, "LduPattern: position out of range");
        csr_src[static_cast<std::size_t>(pos)] = src;
    };
    for (int i = 0; i < n_; ++i) {
        place(pt.diag_pos[i], i);
    }
    for (int f = 0; f < nf_; ++f) {
        place(pt.upper_pos[f], n_ + f);
        place(pt.lower_pos[f], n_ + nf_ + f);
    }
    const HostEll e = csr_to_ell(n_, n_, pt.row_ptr, pt.col);
    std::vector<int> src(e.src.size());
    for (std::size_t k = 0; k < src.size(); ++k) {
        src[k] = e.src[k] >= 0 ? csr_src[static_cast<std::size_t>(e.src[k])] : -1;
    }
    hipStream_t s = stream_.get();
    a0_.upload(e, nullptr, false, false, s);
    a0_.v64.resize(e.col.size());
    a0_.v32.resize(e.col.size());
    a0_src_.upload(src.data(), src.size(), s);
    ldu_.resize(static_cast<std::size_t>(n_) + 2 * static_cast<std::size_t>(nf_));
    init_vectors();
}

System::System(const cfdg_csr& a) {
    require_device();
    require(a.n_rows >= 1, "CsrView: empty matrix");
    check_csr(a, a.n_rows, a.n_rows, "CsrView: invalid square CSR matrix");
    n_ = a.n_rows;
    nnz0_ = a.nnz;
    nf_ = 0;
    const HostEll e = csr_to_ell(n_, n_, a.row_ptr, a.col);
    hipStream_t s = stream_.get();
    a0_.upload(e, a.val, true, true, s);
    init_vectors();
    const std::vector<double> d = inverse_diagonal(a);
    dinv0_.upload(d.data(), d.size(), s);
    assembled_ = true;
}

void System::init_vectors() {
    const std::size_t n = static_cast<std::size_t>(n_);
    for (DevBuf<double>* v : {&b_, &x_, &r_, &p_, &q_, &z_, &r0_, &v_, &s_, &t_, &ph_, &sh_, &dinv0_,
                              &wd0_64_}) {
        v->resize(n);
        v->zero(stream_.get());
    }
    r32_.resize(n);
    wd0_32_.resize(n);
    const int nblk0 = ceil_div(n_, k_block);
    partials_.resize(static_cast<std::size_t>(k_slots) * static_cast<std::size_t>(nblk0));
    partials_.zero(stream_.get());
    state_.resize(1);
    state_.zero(stream_.get());
    hip_check(hipStreamSynchronize(stream_.get()), "init sync");
}

void System::assemble(const double* diag, const double* upper, const double* lower) {
    require(nf_ >= 0 && a0_src_.size() > 0, "assemble: system was not created from an LDU pattern");
    require(diag != nullptr && (nf_ == 0 || (upper != nullptr && lower != nullptr)),
            "assemble: null coefficient arrays");
    hipStream_t s = stream_.get();
    const std::size_t n = static_cast<std::size_t>(n_);
    const std::size_t nf = static_cast<std::size_t>(nf_);
    double* ldu = ldu_.get();
    hip_check(hipMemcpyAsync(ldu, diag, n * sizeof(double), hipMemcpyHostToDevice, s), "H2D diag");
    if (nf > 0) {
        hip_check(hipMemcpyAsync(ldu + n, upper, nf * sizeof(double), hipMemcpyHostToDevice, s),
                  "H2D upper");
        hip_check(hipMemcpyAsync(ldu + n + nf, lower, nf * sizeof(double), hipMemcpyHostToDevice, s),
                  "H2D lower");
    }
    const long long slots = static_cast<long long>(a0_.col.size());
    k_ldu_gather<<<ceil_div(slots, k_block), k_block, 0, s>>>(slots, a0_src_.get(), ldu,
                                                               a0_.v64.get(), a0_.v32.get());
    hip_check(hipGetLastError(), "k_ldu_gather launch");
    k_diag_inverse<<<ceil_div(n_, k_block), k_block, 0, s>>>(n_, ldu, weight_, dinv0_.get(),
                                                             wd0_64_.get(), wd0_32_.get());
    hip_check(hipGetLastError(), "k_diag_inverse launch");
    hip_check(hipStreamSynchronize(s), "assemble sync");
    assembled_ = true;
}

void System::set_rhs(const double* b) {
    require(b != nullptr, "set_rhs: null b");
    b_.upload(b, static_cast<std::size_t>(n_), stream_.get());
}

void System::set_x(const double* x) {
    if (x == nullptr) { // x = 0, enqueued on the system's stream ahead of the next solve
        k_fill<<<ceil_div(n_, k_block), k_block, 0, stream_.get()>>>(n_, x_.get(), 0.0);
        hip_check(hipGetLastError(), "k_fill launch");
        return;
    }
    x_.upload(x, static_cast<std::size_t>(n_), stream_.get());
}

void System::get_x(double* x) {
    require(x != nullptr, "get_x: null x");
    hip_check(hipMemcpyAsync(x, x_.get(), static_cast<std::size_t>(n_) * sizeof(double),
                             hipMemcpyDeviceToHost, stream_.get()),
              "D2H x");
    hip_check(hipStreamSynchronize(stream_.get()), "D2H sync");
}

void System::set_amg(const cfdg_amg_view& v) {
    require(v.n_levels >= 1 && v.levels != nullptr,
            "DeviceSolver::set_amg: empty hierarchy");
    require(v.n_levels <= k_max_levels, "set_amg: too many levels");
    require(v.coarse_inverse != nullptr, "set_amg: null coarse inverse");
    require(v.smooth_sweeps >= 0 && v.smooth_sweeps <= 64, "set_amg: invalid smooth_sweeps");
    const int L = v.n_levels;
    std::vector<int> rows(static_cast<std::size_t>(L));
    rows[0] = n_;
    for (int l = 1; l < L; ++l) {
        rows[static_cast<std::size_t>(l)] = v.levels[l].a.n_rows;
        require(rows[static_cast<std::size_t>(l)] >= 1, "set_amg: empty level");
    }
    for (int l = 0; l + 1 < L; ++l) {
        const int nl = rows[static_cast<std::size_t>(l)];
        const int nn = rows[static_cast<std::size_t>(l) + 1];
        check_csr(v.levels[l].p, nl, nn, "set_amg: prolongation does not match the level sizes");
        check_csr(v.levels[l].r, nn, nl, "set_amg: restriction does not match the level sizes");
    }
    for (int l = 1; l < L; ++l) {
        const int nl = rows[static_cast<std::size_t>(l)];
        check_csr(v.levels[l].a, nl, nl, "set_amg: level operator is not square");
    }
    hipStream_t s = stream_.get();
    const double w = v.jacobi_weight;
    std::vector<DevLevel> lv(static_cast<std::size_t>(L));
    for (int l = 0; l < L; ++l) {
        DevLevel& d = lv[static_cast<std::size_t>(l)];
        d.n = rows[static_cast<std::size_t>(l)];
        d.n_next = l + 1 < L ? rows[static_cast<std::size_t>(l) + 1] : 0;
        const bool coarsest = l == L - 1;
        if (l >= 1 && !coarsest) {
            const cfdg_csr& a = v.levels[l].a;
            d.a.upload_csr(d.n, d.n, a.row_ptr, a.col, a.val, s);
            const std::vector<double> dinv = inverse_diagonal(a);
            std::vector<double> wd(dinv.size());
            std::vector<float> wd32(dinv.size());
            for (std::size_t i = 0; i < dinv.size(); ++i) {
                wd[i] = w * dinv[i];
                wd32[i] = static_cast<float>(wd[i]);
            }
            d.wd64.upload(wd.data(), wd.size(), s);
            d.wd32.upload(wd32.data(), wd32.size(), s);
        }
        if (!coarsest) {
            const cfdg_csr& p = v.levels[l].p;
            const cfdg_csr& r = v.levels[l].r;
            d.p.upload_csr(d.n, d.n_next, p.row_ptr, p.col, p.val, s);
            d.r.upload_csr(d.n_next, d.n, r.row_ptr, r.col, r.val, s);
        } else {
            d.cinv.upload(v.coarse_inverse, static_cast<std::size_t>(d.n) * static_cast<std::size_t>(d.n), s);
        }
    }
    levels_ = std::move(lv);
    sweeps_ = v.smooth_sweeps;
    weight_ = w;
    has_amg_ = true;
    ++generation_;
    // Level-0 smoother weights follow the (possibly new) weight.
    k_scale_dinv<<<ceil_div(n_, k_block), k_block, 0, s>>>(n_, dinv0_.get(), weight_, wd0_64_.get(),
                                                            wd0_32_.get());
    hip_check(hipGetLastError(), "k_scale_dinv launch");
    hip_check(hipStreamSynchronize(s), "set_amg sync");
}

void System::ensure_level_vectors(bool fp32) {
    for (std::size_t l = 0; l < levels_.size(); ++l) {
        DevLevel& d = levels_[l];
        const std::size_t n = static_cast<std::size_t>(d.n);
        for (int i = 0; i < 3; ++i) {
            if (fp32) {
                if (d.v32[i].size() != n) {
                    d.v32[i].resize(n);
                    d.v32[i].zero(stream_.get());
                }
            } else if (d.v64[i].size() != n) {
                d.v64[i].resize(n);
                d.v64[i].zero(stream_.get());
            }
        }
        if (l > 0) {
            if (fp32 && d.b32.size() != n) {
                d.b32.resize(n);
                d.b32.zero(stream_.get());
            }
            if (!fp32 && d.b64.size() != n) {
                d.b64.resize(n);
                d.b64.zero(stream_.get());
            }
        }
    }
    hip_check(hipStreamSynchronize(stream_.get()), "level vectors sync");
}

SysDev System::descriptor(bool fp32, bool use_amg) {
    require(assembled_, "solve: the matrix has not been assembled");
    SysDev d{};
    d.st = state_.get();
    d.partials = partials_.get();
    d.nblk0 = ceil_div(n_, k_block);
    d.n = n_;
    d.a64 = a0_.view64();
    d.b = b_.get();
    d.x = x_.get();
    d.r = r_.get();
    d.p = p_.get();
    d.q = q_.get();
    d.z = z_.get();
    d.r32 = r32_.get();
    d.dinv = dinv0_.get();
    d.r0 = r0_.get();
    d.v = v_.get();
    d.s = s_.get();
    d.t = t_.get();
    d.ph = ph_.get();
    d.sh = sh_.get();
    d.n_levels = 0;
    if (!use_amg) {
        return d;
    }
    require(has_amg_, "solve: use_amg requested but no hierarchy was set");
    ensure_level_vectors(fp32);
    const int L = static_cast<int>(levels_.size());
    d.n_levels = L;
    for (int l = 0; l < L; ++l) {
        DevLevel& s = levels_[static_cast<std::size_t>(l)];
        LevelDev& o = d.lv[l];
        const bool coarsest = l == L - 1;
        o.n = s.n;
        o.n_next = s.n_next;
        o.coarsest = coarsest ? 1 : 0;
        if (l == 0) {
            o.a = fp32 ? a0_.view32() : a0_.view64();
            o.b = fp32 ? static_cast<void*>(r32_.get()) : static_cast<void*>(r_.get());
            o.wd = fp32 ? static_cast<const void*>(wd0_32_.get()) : static_cast<const void*>(wd0_64_.get());
        } else {
            if (!coarsest) {
                o.a = fp32 ? s.a.view32() : s.a.view64();
                o.wd = fp32 ? static_cast<const void*>(s.wd32.get()) : static_cast<const void*>(s.wd64.get());
            }
            o.b = fp32 ? static_cast<void*>(s.b32.get()) : static_cast<void*>(s.b64.get());
        }
        if (!coarsest) {
            o.p = fp32 ? s.p.view32() : s.p.view64();
            o.r = fp32 ? s.r.view32() : s.r.view64();
        }
        for (int i = 0; i < 3; ++i) {
            o.v[i] = fp32 ? static_cast<void*>(s.v32[i].get()) : static_cast<void*>(s.v64[i].get());
        }
        o.cinv = s.cinv.get();
    }
    return d;
}

std::size_t System::footprint() const {
    std::size_t b = a0_.bytes() + a0_src_.bytes() + ldu_.bytes() + dinv0_.bytes() + wd0_64_.bytes() +
                    wd0_32_.bytes() + r32_.bytes() + partials_.bytes() + state_.bytes();
    for (const DevBuf<double>* v : {&b_, &x_, &r_, &p_, &q_, &z_, &r0_, &v_, &s_, &t_, &ph_, &sh_}) {
        b += v->bytes();
    }
    for (const DevLevel& l : levels_) {
        b += l.bytes();
    }
    return b;
}

// ------------------------------------------------------------------------------------------
// Batched driver
// ------------------------------------------------------------------------------------------

namespace {

struct Tables {
    int n_sys = 0;
    int l_max = 1;
    DevBuf<int> ids;
    std::vector<DevBuf<int2>> t;   // per level: 256-row blocks of systems having the level
    std::vector<int> tn;
    std::vector<DevBuf<int2>> c;   // per level: 4-row blocks of systems whose coarsest it is
    std::vector<int> cn;
    DevBuf<int2> t0_single;        // level-0 blocks of one-level hierarchies
    int t0_single_n = 0;
    // per level and matrix (0 operator, 1 prolongation, 2 restriction): blocks of 256 / tpr
    // rows of that matrix for the systems whose level l is not the coarsest
    std::vector<DevBuf<int2>> tm[3];
    std::vector<int> tmn[3];
    std::vector<char> a_ell; // every system's level-l operator is sliced ELL (fused presmooth)
};

void build_tables(Tables& tb, const std::vector<System*>& sys, bool use_amg, hipStream_t s) {
    tb.n_sys = static_cast<int>(sys.size());
    std::vector<int> ids(sys.size());
    std::iota(ids.begin(), ids.end(), 0);
    tb.ids.upload(ids.data(), ids.size(), s);
    tb.l_max = 1;
    if (use_amg) {
        for (System* x : sys) {
            tb.l_max = std::max(tb.l_max, x->n_levels());
        }
    }
    tb.t.clear();
    tb.c.clear();
    tb.t.resize(static_cast<std::size_t>(tb.l_max));
    tb.c.resize(static_cast<std::size_t>(tb.l_max));
    tb.tn.assign(static_cast<std::size_t>(tb.l_max), 0);
    tb.cn.assign(static_cast<std::size_t>(tb.l_max), 0);
    std::vector<int2> single;
    for (int l = 0; l < tb.l_max; ++l) {
        std::vector<int2> t;
        std::vector<int2> c;
        for (int q = 0; q < tb.n_sys; ++q) {
            System* x = sys[static_cast<std::size_t>(q)];
            const int L = use_amg ? x->n_levels() : 1;
            if (l >= L) {
                continue;
            }
            const int rows = (use_amg && l > 0) ? x->level_rows(l) : x->n();
            const int nb = ceil_div(rows, k_block);
            for (int b = 0; b < nb; ++b) {
                t.push_back(make_int2(q, b));
            }
            if (use_amg && l == L - 1) {
                const int nc = ceil_div(rows, k_slices_per_block);
                for (int b = 0; b < nc; ++b) {
                    c.push_back(make_int2(q, b));
                }
                if (l == 0) {
                    for (int b = 0; b < nb; ++b) {
                        single.push_back(make_int2(q, b));
                    }
                }
            }
        }
        tb.tn[static_cast<std::size_t>(l)] = static_cast<int>(t.size());
        tb.cn[static_cast<std::size_t>(l)] = static_cast<int>(c.size());
        tb.t[static_cast<std::size_t>(l)].upload(t.data(), t.size(), s);
        tb.c[static_cast<std::size_t>(l)].upload(c.data(), c.size(), s);
    }
    tb.t0_single_n = static_cast<int>(single.size());
    tb.t0_single.upload(single.data(), single.size(), s);
    for (int which = 0; which < 3; ++which) {
        tb.tm[which].clear();
        tb.tm[which].resize(static_cast<std::size_t>(tb.l_max));
        tb.tmn[which].assign(static_cast<std::size_t>(tb.l_max), 0);
    }
    tb.a_ell.assign(static_cast<std::size_t>(tb.l_max), 1);
    if (!use_amg) {
        return;
    }
    for (int l = 0; l + 1 < tb.l_max; ++l) {
        for (int which = 0; which < 3; ++which) {
            std::vector<int2> t;
            for (int q = 0; q < tb.n_sys; ++q) {
                System* x = sys[static_cast<std::size_t>(q)];
                if (l + 1 >= x->n_levels()) {
                    continue; // level l is this system's coarsest (or absent)
                }
                const int tpr = x->tpr(l, which);
                if (which == 0 && tpr != 1) {
                    tb.a_ell[static_cast<std::size_t>(l)] = 0;
                }
                const int rows = which == 2 ? x->level_rows(l + 1) : x->level_rows(l);
                const int nb = ceil_div(static_cast<long long>(rows) * tpr, k_block);
                for (int b = 0; b < nb; ++b) {
                    t.push_back(make_int2(q, b));
                }
            }
            tb.tmn[which][static_cast<std::size_t>(l)] = static_cast<int>(t.size());
            tb.tm[which][static_cast<std::size_t>(l)].upload(t.data(), t.size(), s);
        }
    }
}

#define CFDG_LAUNCH(kernel, grid, stream, ...)                                                    \
    do {                                                                                          \
        if ((grid) > 0) {                                                                         \
            kernel<<<(grid), k_block, 0, (stream)>>>(__VA_ARGS__);                                \
        }                                                                                         \
    } while (0)

// Kernel timing of one solve: HIP events on the solve stream bracket every launch of the
// probed kernel. The first `useful` launches are the ones that did work (later launches of
// the iteration pipeline return at once and are not counted).
thread_local Probe* g_probe = nullptr;

void probe_begin(int kind, hipStream_t s) {
    if (g_probe == nullptr || g_probe->kind != kind) {
        return;
    }
    if (g_probe->used == g_probe->events.size()) {
        hipEvent_t a = nullptr;
        hipEvent_t b = nullptr;
        hip_check(hipEventCreate(&a), "event create");
        hip_check(hipEventCreate(&b), "event create");
        g_probe->events.emplace_back(a, b);
    }
    hip_check(hipEventRecord(g_probe->events[g_probe->used].first, s), "event record");
}

void probe_end(int kind, hipStream_t s) {
    if (g_probe == nullptr || g_probe->kind != kind) {
        return;
    }
    hip_check(hipEventRecord(g_probe->events[g_probe->used].second, s), "event record");
    ++g_probe->used;
}

int final_index(int sweeps) {
    const int cur = (sweeps >= 2 && (sweeps - 2) % 2 == 1) ? 2 : 0;
    return sweeps % 2 == 0 ? cur : 1;
}

template <class T, int Dots>
void enqueue_vcycle(const SysDev* d, const Tables& tb, int sweeps, hipStream_t s) {
    const int fin = final_index(sweeps);
    const int L = tb.l_max;
    CFDG_LAUNCH(k_coarse<T>, tb.cn[0], s, d, tb.c[0].get(), 0, fin);
    std::vector<int> cur(static_cast<std::size_t>(L), 0);
    for (int l = 0; l + 1 < L; ++l) {
        const std::size_t li = static_cast<std::size_t>(l);
        const int g = tb.tn[li];                        // row-wise blocks
        const int2* blk = tb.t[li].get();
        const int ga = tb.tmn[0][li];                   // operator blocks
        const int2* ba = tb.tm[0][li].get();
        int c = 0;
        if (sweeps == 0) {
            CFDG_LAUNCH(k_zero_level<T>, g, s, d, blk, l, 0);
        } else if (sweeps == 1) {
            CFDG_LAUNCH(k_jacobi_zero<T>, g, s, d, blk, l, 0);
        } else {
            if (tb.a_ell[li] != 0) {
                CFDG_LAUNCH(k_presmooth2<T>, ga, s, d, ba, l, 0);
            } else { // long rows: form x1 = wd b once instead of gathering it per entry
                CFDG_LAUNCH(k_jacobi_zero<T>, g, s, d, blk, l, 2);
                CFDG_LAUNCH((k_jacobi<T, 0>), ga, s, d, ba, l, 2, 0);
            }
            for (int k = 2; k < sweeps; ++k) {
                const int o = c == 0 ? 2 : 0;
                CFDG_LAUNCH((k_jacobi<T, 0>), ga, s, d, ba, l, c, o);
                c = o;
            }
        }
        cur[li] = c;
        CFDG_LAUNCH(k_residual<T>, ga, s, d, ba, l, c, 1);
        CFDG_LAUNCH(k_restrict<T>, tb.tmn[2][li], s, d, tb.tm[2][li].get(), l, 1);
        CFDG_LAUNCH(k_coarse<T>, tb.cn[li + 1], s, d, tb.c[li + 1].get(), l + 1, fin);
    }
    for (int l = L - 2; l >= 0; --l) {
        const std::size_t li = static_cast<std::size_t>(l);
        const int ga = tb.tmn[0][li];
        const int2* ba = tb.tm[0][li].get();
        const int c0 = cur[li];
        CFDG_LAUNCH(k_prolong<T>, tb.tmn[1][li], s, d, tb.tm[1][li].get(), l, c0, fin);
        int c = c0;
        for (int k = 0; k < sweeps; ++k) {
            const int o = c == 1 ? c0 : 1;
            const bool last = k == sweeps - 1;
            if (l == 0 && last) {
                probe_begin(PROBE_L0_SWEEP, s);
                CFDG_LAUNCH((k_jacobi<T, Dots>), ga, s, d, ba, l, c, o);
                probe_end(PROBE_L0_SWEEP, s);
            } else {
                CFDG_LAUNCH((k_jacobi<T, 0>), ga, s, d, ba, l, c, o);
            }
            c = o;
        }
    }
    if (sweeps == 0 && L > 1) {
        CFDG_LAUNCH((k_level0_dots<T, Dots>), tb.tn[0], s, d, tb.t[0].get(), fin);
    } else {
        CFDG_LAUNCH((k_level0_dots<T, Dots>), tb.t0_single_n, s, d, tb.t0_single.get(), fin);
    }
}

template <int Op>
void fin(const SysDev* d, const Tables& tb, hipStream_t s) {
    k_finalize<Op><<<tb.n_sys, k_fin_block, 0, s>>>(d, tb.ids.get());
}

enum VecId { V_X = 0, V_R, V_P, V_Q, V_Z, V_R0, V_V, V_S, V_T, V_PH, V_SH };

struct Plan {
    cfdg_method method;
    bool amg;
    bool fp32;
    int sweeps;
};

void enqueue_init(const Plan& pl, const SysDev* d, const Tables& tb, hipStream_t s) {
    const int g0 = tb.tn[0];
    const int2* b0 = tb.t[0].get();
    if (pl.method == CFDG_METHOD_BICGSTAB) {
        CFDG_LAUNCH(k_bicg_init, g0, s, d, b0);
        fin<FIN_BICG_INIT>(d, tb, s);
        return;
    }
    if (!pl.amg) {
        CFDG_LAUNCH(k_cg_init, g0, s, d, b0);
        fin<FIN_CG_INIT>(d, tb, s);
        CFDG_LAUNCH(k_p_update<double>, g0, s, d, b0, 0, 0);
        return;
    }
    const int fz = final_index(pl.sweeps);
    if (pl.fp32) {
        CFDG_LAUNCH(k_residual0<true>, g0, s, d, b0);
        fin<FIN_AMG_INIT>(d, tb, s);
        enqueue_vcycle<float, 2>(d, tb, pl.sweeps, s);
        fin<FIN_BETA_FLEX>(d, tb, s);
        CFDG_LAUNCH(k_p_update<float>, g0, s, d, b0, 1, fz);
    } else {
        CFDG_LAUNCH(k_residual0<false>, g0, s, d, b0);
        fin<FIN_AMG_INIT>(d, tb, s);
        enqueue_vcycle<double, 1>(d, tb, pl.sweeps, s);
        fin<FIN_BETA>(d, tb, s);
        CFDG_LAUNCH(k_p_update<double>, g0, s, d, b0, 1, fz);
    }
}

void enqueue_iteration(const Plan& pl, const SysDev* d, const Tables& tb, hipStream_t s) {
    const int g0 = tb.tn[0];
    const int2* b0 = tb.t[0].get();
    if (pl.method == CFDG_METHOD_BICGSTAB) {
        fin<FIN_BICG_RHO>(d, tb, s);
        CFDG_LAUNCH(k_bicg_p, g0, s, d, b0);
        CFDG_LAUNCH(k_spmv_dot<false>, g0, s, d, b0, V_PH, V_V, V_R0, SLOT_R0V, 0, 0);
        fin<FIN_BICG_ALPHA>(d, tb, s);
        CFDG_LAUNCH(k_bicg_s, g0, s, d, b0);
        fin<FIN_BICG_S>(d, tb, s);
        CFDG_LAUNCH(k_bicg_early_x, g0, s, d, b0);
        CFDG_LAUNCH(k_spmv_dot<true>, g0, s, d, b0, V_SH, V_T, V_T, SLOT_TT, V_S, SLOT_TS);
        fin<FIN_BICG_OMEGA>(d, tb, s);
        CFDG_LAUNCH(k_bicg_xr, g0, s, d, b0);
        fin<FIN_BICG_RES>(d, tb, s);
        return;
    }
    probe_begin(PROBE_OUTER_SPMV, s);
    CFDG_LAUNCH(k_spmv_dot<false>, g0, s, d, b0, V_P, V_Q, V_P, SLOT_PQ, 0, 0);
    probe_end(PROBE_OUTER_SPMV, s);
    fin<FIN_ALPHA>(d, tb, s);
    if (!pl.amg) {
        CFDG_LAUNCH(k_update_xr<1>, g0, s, d, b0);
        fin<FIN_CG_RES_BETA>(d, tb, s);
        CFDG_LAUNCH(k_p_update<double>, g0, s, d, b0, 0, 0);
        return;
    }
    const int fz = final_index(pl.sweeps);
    if (pl.fp32) {
        CFDG_LAUNCH(k_update_xr<2>, g0, s, d, b0);
        fin<FIN_RES>(d, tb, s);
        enqueue_vcycle<float, 2>(d, tb, pl.sweeps, s);
        fin<FIN_BETA_FLEX>(d, tb, s);
        CFDG_LAUNCH(k_p_update<float>, g0, s, d, b0, 1, fz);
    } else {
        CFDG_LAUNCH(k_update_xr<0>, g0, s, d, b0);
        fin<FIN_RES>(d, tb, s);
        enqueue_vcycle<double, 1>(d, tb, pl.sweeps, s);
        fin<FIN_BETA>(d, tb, s);
        CFDG_LAUNCH(k_p_update<double>, g0, s, d, b0, 1, fz);
    }
}

} // namespace

// Everything a set of systems needs to be solved together, built once and reused: launch
// tables, device descriptors, and (unless kernels are being probed) HIP graphs of the solver's
// start-up sequence and of one iteration. Iteration k of the graph pipeline writes the number
// of still-running systems into pinned host memory slot k mod 2 from the device; the host waits
// for the previous slot while the next iteration is already queued.
class Runner {
public:
    Runner(const std::vector<System*>& systems, const BatchOptions& opt, hipStream_t s)
        : sys_(systems), s_(s), running_(2), init_host_(systems.size()), out_host_(systems.size()) {
        require(!sys_.empty(), "solve_batch: no systems");
        const bool amg = opt.use_amg && opt.method == CFDG_METHOD_CG;
        const bool fp32 = amg && opt.precision != CFDG_PREC_FP64;
        int sweeps = 2;
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            if (amg) {
                require(sys_[i]->has_amg(), "solve: use_amg requested but no hierarchy was set");
                if (i == 0) {
                    sweeps = sys_[i]->sweeps();
                }
                require(sys_[i]->sweeps() == sweeps, "solve_batch: systems differ in smooth_sweeps");
            }
            gens_.push_back(sys_[i]->generation());
        }
        pl_ = Plan{opt.method, amg, fp32, sweeps};
        build_tables(tb_, sys_, amg, s_);
        std::vector<SysDev> desc(sys_.size());
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            desc[i] = sys_[i]->descriptor(fp32, amg);
        }
        desc_.upload(desc.data(), desc.size(), s_);
        d_init_.resize(sys_.size());
        d_out_.resize(sys_.size());
        void* dev = nullptr;
        hip_check(hipHostGetDevicePointer(&dev, running_.get(), 0), "hipHostGetDevicePointer");
        running_dev_ = static_cast<int*>(dev);
        use_graphs_ = std::getenv("CFDG_NO_GRAPHS") == nullptr;
    }
    Runner(const Runner&) = delete;
    Runner& operator=(const Runner&) = delete;
    ~Runner() {
        for (hipGraphExec_t g : {init_exec_, iter_exec_[0], iter_exec_[1]}) {
            if (g != nullptr) {
                (void)hipGraphExecDestroy(g);
            }
        }
    }

    [[nodiscard]] bool matches(const std::vector<System*>& systems, const BatchOptions& opt) const {
        if (systems != sys_) {
            return false;
        }
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            if (sys_[i]->generation() != gens_[i]) {
                return false;
            }
        }
        const bool amg = opt.use_amg && opt.method == CFDG_METHOD_CG;
        return opt.method == pl_.method && amg == pl_.amg &&
               (amg && opt.precision != CFDG_PREC_FP64) == pl_.fp32;
    }

    void solve(const std::vector<cfdg_solve_control>& control, std::vector<cfdg_solve_stats>& stats,
               Probe* probe) {
        require(control.size() == sys_.size(), "solve_batch: one control per system");
        const int n_sys = static_cast<int>(sys_.size());
        int max_it = 0;
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            require(control[i].max_iterations >= 0, "solve: max_iterations must be >= 0");
            max_it = std::max(max_it, control[i].max_iterations);
            SolveState& st = init_host_[i];
            std::memset(&st, 0, sizeof(st));
            st.tol = control[i].abs_tolerance;
            st.max_it = control[i].max_iterations;
            st.status = ST_RUNNING;
        }
        const SysDev* d = desc_.get();
        struct ProbeScope { // the thread's probe for the duration of this solve
            explicit ProbeScope(Probe* p) {
                g_probe = p;
                if (p != nullptr) {
                    p->used = 0;
                }
            }
            ~ProbeScope() { g_probe = nullptr; }
        } probe_scope(probe);
        // Graphs from the second solve of a runner on: one-shot solves skip the capture cost.
        const bool graphs = use_graphs_ && probe == nullptr && solves_ > 0;
        ++solves_;
        if (graphs && init_exec_ == nullptr) {
            capture_graphs();
        }
        running_[0] = running_[1] = 1;
        hip_check(hipEventRecord(t0_.get(), s_), "event record");
        hip_check(hipMemcpyAsync(d_init_.get(), init_host_.get(), sys_.size() * sizeof(SolveState),
                                 hipMemcpyHostToDevice, s_),
                  "H2D states");
        k_init_states<<<ceil_div(n_sys, 64), 64, 0, s_>>>(d, n_sys, d_init_.get());
        if (graphs) {
            hip_check(hipGraphLaunch(init_exec_, s_), "graph launch");
        } else {
            enqueue_init(pl_, d, tb_, s_);
        }
        hip_check(hipGetLastError(), "init launch");
        // Iterations are enqueued one ahead of the convergence read-back; surplus iterations are
        // no-ops on the device because every kernel tests the system's status first.
        for (int k = 1; k <= max_it + 2; ++k) {
            const int slot = k & 1;
            if (graphs) {
                hip_check(hipGraphLaunch(iter_exec_[slot], s_), "graph launch");
            } else {
                enqueue_iteration(pl_, d, tb_, s_);
                k_count_running<<<1, k_block, 0, s_>>>(d, n_sys, running_dev_ + slot);
            }
            hip_check(hipEventRecord(ev_[slot].get(), s_), "event record");
            hip_check(hipGetLastError(), "iteration launch");
            if (k >= 2) {
                const int prev = (k - 1) & 1;
                hip_check(hipEventSynchronize(ev_[prev].get()), "event sync");
                if (running_[static_cast<std::size_t>(prev)] == 0) {
                    break;
                }
            }
        }
        CFDG_LAUNCH(k_true_residual, tb_.tn[0], s_, d, tb_.t[0].get());
        fin<FIN_TRUE>(d, tb_, s_);
        k_gather_states<<<ceil_div(n_sys, 64), 64, 0, s_>>>(d, n_sys, d_out_.get());
        hip_check(hipEventRecord(t1_.get(), s_), "event record");
        hip_check(hipMemcpyAsync(out_host_.get(), d_out_.get(), sys_.size() * sizeof(SolveState),
                                 hipMemcpyDeviceToHost, s_),
                  "D2H states");
        hip_check(hipStreamSynchronize(s_), "solve sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, t0_.get(), t1_.get()), "event elapsed");
        stats.assign(sys_.size(), cfdg_solve_stats{});
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            const SolveState& st = out_host_[i];
            cfdg_solve_stats& o = stats[i];
            o.iterations = std::min(st.it, st.max_it);
            o.converged = st.status == ST_CONVERGED ? 1 : 0;
            o.on_device = 1;
            o.exit_reason = st.status;
            o.residual = st.res;
            o.true_residual = st.true_res;
            o.rhs_norm = st.bnorm;
            o.solve_ms = ms;
        }
        if (probe != nullptr) {
            // Launches that did work: the initial V-cycle plus one per iteration that
            // preconditioned (level-0 sweep), or one per iteration (outer SpMV).
            const int its = stats[0].iterations;
            std::size_t useful = static_cast<std::size_t>(its);
            if (probe->kind == PROBE_L0_SWEEP && stats[0].exit_reason != ST_CONVERGED) {
                useful = static_cast<std::size_t>(its) + 1;
            }
            useful = std::min(useful, probe->used);
            for (std::size_t i = 0; i < useful; ++i) {
                float e = 0.0f;
                hip_check(hipEventElapsedTime(&e, probe->events[i].first, probe->events[i].second),
                          "event elapsed");
                probe->total_ms += e;
            }
            probe->launches += static_cast<long long>(useful);
        }
    }

private:
    void capture_graphs() {
        const SysDev* d = desc_.get();
        const int n_sys = static_cast<int>(sys_.size());
        auto capture = [&](auto&& enqueue) {
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "begin capture");
            enqueue();
            hip_check(hipStreamEndCapture(s_, &g), "end capture");
            hipGraphExec_t e = nullptr;
            hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "graph instantiate");
            hip_check(hipGraphDestroy(g), "graph destroy");
            return e;
        };
        init_exec_ = capture([&] { enqueue_init(pl_, d, tb_, s_); });
        for (int slot = 0; slot < 2; ++slot) {
            iter_exec_[slot] = capture([&] {
                enqueue_iteration(pl_, d, tb_, s_);
                k_count_running<<<1, k_block, 0, s_>>>(d, n_sys, running_dev_ + slot);
            });
        }
    }

    std::vector<System*> sys_;
    std::vector<unsigned> gens_;
    hipStream_t s_;
    Plan pl_{};
    Tables tb_;
    DevBuf<SysDev> desc_;
    DevBuf<SolveState> d_init_;
    DevBuf<SolveState> d_out_;
    PinnedBuf<int> running_;
    int* running_dev_ = nullptr;
    PinnedBuf<SolveState> init_host_;
    PinnedBuf<SolveState> out_host_;
    Event ev_[2];
    Event t0_;
    Event t1_;
    bool use_graphs_ = true;
    long long solves_ = 0;
    hipGraphExec_t init_exec_ = nullptr;
    hipGraphExec_t iter_exec_[2] = {nullptr, nullptr};
};

System::~System() = default;

std::unique_ptr<Runner>& System::runner_slot(cfdg_method method, bool use_amg,
                                             cfdg_precision precision) {
    int k = 0;
    if (method == CFDG_METHOD_BICGSTAB) {
        k = 3;
    } else if (!use_amg) {
        k = 2;
    } else if (precision != CFDG_PREC_FP64) {
        k = 1;
    }
    return runners_[k];
}

void solve_batch(const std::vector<System*>& systems, const BatchOptions& opt,
                 const std::vector<cfdg_solve_control>& control,
                 std::vector<cfdg_solve_stats>& stats, hipStream_t s) {
    require(!systems.empty(), "solve_batch: no systems");
    Probe* probe = (systems.size() == 1 && systems[0]->probe.kind != PROBE_OFF) ? &systems[0]->probe
                                                                             : nullptr;
    if (systems.size() == 1) { // single system: keep the runner with the system
        System& sys = *systems[0];
        std::unique_ptr<Runner>& r = sys.runner_slot(opt.method, opt.use_amg, opt.precision);
        if (!r || !r->matches(systems, opt)) {
            r.reset();
            r = std::make_unique<Runner>(systems, opt, s);
        }
        r->solve(control, stats, probe);
        return;
    }
    Runner r(systems, opt, s);
    r.solve(control, stats, nullptr);
}

} // namespace cfdg

// ------------------------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------------------------
using namespace cfdg;

struct cfdg_system {
    std::unique_ptr<System> sys;
};

namespace {

void check_control(const cfdg_solve_control* c) {
    require(c != nullptr, "null solve control");
    require(c->max_iterations >= 0, "max_iterations must be >= 0");
    require(!(c->abs_tolerance < 0.0), "abs_tolerance must be >= 0");
}

cfdg_solve_stats solve_one(System& s, cfdg_method method, bool use_amg, cfdg_precision prec,
                           const cfdg_solve_control& c) {
    require(method == CFDG_METHOD_CG || method == CFDG_METHOD_BICGSTAB, "unknown method");
    require(prec == CFDG_PREC_FP64 || prec == CFDG_PREC_MIXED_FP32 || prec == CFDG_PREC_MIXED_BF16,
            "unknown precision");
    BatchOptions opt;
    opt.method = method;
    opt.use_amg = use_amg;
    opt.precision = prec;
    std::vector<cfdg_solve_stats> st;
    solve_batch({&s}, opt, {c}, st, s.stream());
    return st[0];
}

} // namespace

extern "C" cfdg_status cfdg_system_create(const cfdg_ldu_pattern* pattern, cfdg_system** out) {
    return guarded([&] {
        require(pattern != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        auto h = std::make_unique<cfdg_system>();
        h->sys = std::make_unique<System>(*pattern);
        *out = h.release();
    });
}

extern "C" cfdg_status cfdg_system_destroy(cfdg_system* sys) {
    return guarded([&] { delete sys; });
}

extern "C" cfdg_status cfdg_system_on_device(const cfdg_system* sys, int* on_device) {
    return guarded([&] {
        require(sys != nullptr && on_device != nullptr, "null argument");
        *on_device = 1;
    });
}

extern "C" cfdg_status cfdg_system_set_amg(cfdg_system* sys, const cfdg_amg_view* amg) {
    return guarded([&] {
        require(sys != nullptr && amg != nullptr, "null argument");
        sys->sys->set_amg(*amg);
    });
}

extern "C" cfdg_status cfdg_system_assemble(cfdg_system* sys, const double* diag,
                                            const double* upper, const double* lower) {
    return guarded([&] {
        require(sys != nullptr, "null system");
        sys->sys->assemble(diag, upper, lower);
    });
}

extern "C" cfdg_status cfdg_system_set_rhs(cfdg_system* sys, const double* b) {
    return guarded([&] {
        require(sys != nullptr, "null system");
        sys->sys->set_rhs(b);
    });
}

extern "C" cfdg_status cfdg_system_set_x(cfdg_system* sys, const double* x) {
    return guarded([&] {
        require(sys != nullptr, "null system");
        sys->sys->set_x(x);
    });
}

extern "C" cfdg_status cfdg_system_get_x(cfdg_system* sys, double* x) {
    return guarded([&] {
        require(sys != nullptr, "null system");
        sys->sys->get_x(x);
    });
}

extern "C" cfdg_status cfdg_system_solve_resident(cfdg_system* sys, cfdg_method method, int use_amg,
                                                  cfdg_precision precision,
                                                  const cfdg_solve_control* control,
                                                  cfdg_solve_stats* stats) {
    return guarded([&] {
        require(sys != nullptr && stats != nullptr, "null argument");
        check_control(control);
        *stats = solve_one(*sys->sys, method, use_amg != 0, precision, *control);
    });
}

extern "C" cfdg_status cfdg_system_solve(cfdg_system* sys, const double* diag, const double* upper,
                                         const double* lower, const double* b, double* x,
                                         cfdg_method method, int use_amg,
                                         const cfdg_solve_control* control,
                                         cfdg_solve_stats* stats) {
    return guarded([&] {
        require(sys != nullptr && stats != nullptr && b != nullptr && x != nullptr, "null argument");
        check_control(control);
        System& s = *sys->sys;
        s.assemble(diag, upper, lower);
        s.set_rhs(b);
        s.set_x(x);
        // The reference runs AMG only for the CG method (device_solver.cpp:138-149).
        const bool amg = use_amg != 0 && method == CFDG_METHOD_CG && s.has_amg();
        *stats = solve_one(s, method, amg, CFDG_PREC_FP64, *control);
        s.get_x(x);
    });
}

extern "C" cfdg_status cfdg_system_probe(cfdg_system* sys, int kind) {
    return guarded([&] {
        require(sys != nullptr, "null system");
        require(kind >= PROBE_OFF && kind <= PROBE_OUTER_SPMV, "unknown probe kind");
        Probe& p = sys->sys->probe;
        p.kind = kind;
        p.total_ms = 0.0;
        p.launches = 0;
    });
}

extern "C" cfdg_status cfdg_system_probe_read(const cfdg_system* sys, int precision,
                                              double* avg_ms, long long* launches,
                                              double* bytes_per_launch) {
    return guarded([&] {
        require(sys != nullptr && avg_ms != nullptr && launches != nullptr &&
                    bytes_per_launch != nullptr,
                "null argument");
        const System& s = *sys->sys;
        const Probe& p = s.probe;
        *launches = p.launches;
        *avg_ms = p.launches > 0 ? p.total_ms / static_cast<double>(p.launches) : 0.0;
        const double n = static_cast<double>(s.n());
        const double nnz = static_cast<double>(s.nnz0());
        const bool fp32 = precision != CFDG_PREC_FP64;
        if (p.kind == PROBE_L0_SWEEP) {
            // matrix (value + column index) + x_in, b, w*dinv, x_out at the V-cycle precision;
            // the mixed sweep also reads r and q in FP64 for its two dot products.
            *bytes_per_launch = fp32 ? nnz * (4.0 + 4.0) + 4.0 * 4.0 * n + 2.0 * 8.0 * n
                                     : nnz * (8.0 + 4.0) + 4.0 * 8.0 * n;
        } else if (p.kind == PROBE_OUTER_SPMV) {
            *bytes_per_launch = nnz * (8.0 + 4.0) + 2.0 * 8.0 * n; // p in, q out
        } else {
            *bytes_per_launch = 0.0;
        }
    });
}

extern "C" cfdg_status cfdg_system_footprint(const cfdg_system* sys, size_t* bytes) {
    return guarded([&] {
        require(sys != nullptr && bytes != nullptr, "null argument");
        *bytes = sys->sys->footprint();
    });
}

extern "C" cfdg_status cfdg_solve_csr(const cfdg_csr* a, const double* b, double* x,
                                      cfdg_method method, const cfdg_solve_control* control,
                                      cfdg_solve_stats* stats) {
    return guarded([&] {
        require(a != nullptr && b != nullptr && x != nullptr && stats != nullptr, "null argument");
        check_control(control);
        System s(*a);
        s.set_rhs(b);
        s.set_x(x);
        *stats = solve_one(s, method, false, CFDG_PREC_FP64, *control);
        s.get_x(x);
    });
}

extern "C" cfdg_status cfdg_solve_amg_cg_prec(const cfdg_amg_view* amg, const double* b, double* x,
                                              cfdg_precision precision,
                                              const cfdg_solve_control* control,
                                              cfdg_solve_stats* stats) {
    return guarded([&] {
        require(amg != nullptr && b != nullptr && x != nullptr && stats != nullptr, "null argument");
        require(amg->n_levels >= 1 && amg->levels != nullptr, "solve_amg_cg: empty hierarchy");
        check_control(control);
        System s(amg->levels[0].a);
        s.set_amg(*amg);
        s.set_rhs(b);
        s.set_x(x);
        *stats = solve_one(s, CFDG_METHOD_CG, true, precision, *control);
        s.get_x(x);
    });
}

extern "C" cfdg_status cfdg_solve_amg_cg(const cfdg_amg_view* amg, const double* b, double* x,
                                         const cfdg_solve_control* control, cfdg_solve_stats* stats) {
    return cfdg_solve_amg_cg_prec(amg, b, x, CFDG_PREC_FP64, control, stats);
}

extern "C" cfdg_status cfdg_assemble_convection_diffusion(int n_cells, int n_faces, const int* owner,
                                                          const int* neighbour, const double* flux,
                                                          const double* gamma,
                                                          const double* delta_coeff,
                                                          double convection, double* upper,
                                                          double* lower, double* diag,
                                                          double* divergence) {
    return guarded([&] {
        require(n_cells >= 0 && n_faces >= 0, "negative size");
        if (n_cells == 0 && n_faces == 0) {
            return;
        }
        require_device();
        require(diag != nullptr && divergence != nullptr, "null cell arrays");
        require(n_faces == 0 || (owner && neighbour && flux && gamma && delta_coeff && upper && lower),
                "null face arrays");
        // Faces of each cell in ascending face order (counting sort keeps the order).
        std::vector<int> ptr(static_cast<std::size_t>(n_cells) + 1, 0);
        for (int f = 0; f < n_faces; ++f) {
            require(owner[f] >= 0 && owner[f] < n_cells && neighbour[f] >= 0 && neighbour[f] < n_cells,
                    "face cell index out of range");
            ++ptr[static_cast<std::size_t>(owner[f]) + 1];
            ++ptr[static_cast<std::size_t>(neighbour[f]) + 1];
        }
        for (int c = 0; c < n_cells; ++c) {
            ptr[static_cast<std::size_t>(c) + 1] += ptr[static_cast<std::size_t>(c)];
        }
        std::vector<int> faces(static_cast<std::size_t>(ptr.back()));
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int f = 0; f < n_faces; ++f) {
            faces[static_cast<std::size_t>(fill[static_cast<std::size_t>(owner[f])]++)] = f;
            faces[static_cast<std::size_t>(fill[static_cast<std::size_t>(neighbour[f])]++)] = f;
        }
        Stream st;
        hipStream_t s = st.get();
        DevBuf<int> d_ptr, d_faces, d_own;
        DevBuf<double> d_flux, d_gamma, d_delta, d_up, d_lo, d_diag, d_div;
        d_ptr.upload(ptr.data(), ptr.size(), s);
        d_faces.upload(faces.data(), faces.size(), s);
        d_diag.upload(diag, static_cast<std::size_t>(n_cells), s);
        d_div.upload(divergence, static_cast<std::size_t>(n_cells), s);
        if (n_faces > 0) {
            d_own.upload(owner, static_cast<std::size_t>(n_faces), s);
            d_flux.upload(flux, static_cast<std::size_t>(n_faces), s);
            d_gamma.upload(gamma, static_cast<std::size_t>(n_faces), s);
            d_delta.upload(delta_coeff, static_cast<std::size_t>(n_faces), s);
            d_up.resize(static_cast<std::size_t>(n_faces));
            d_lo.resize(static_cast<std::size_t>(n_faces));
            k_assemble_faces<<<ceil_div(n_faces, k_block), k_block, 0, s>>>(
                n_faces, d_flux.get(), d_gamma.get(), d_delta.get(), convection, d_up.get(), d_lo.get());
            hip_check(hipGetLastError(), "k_assemble_faces launch");
        }
        if (n_cells > 0) {
            k_assemble_cells<<<ceil_div(n_cells, k_block), k_block, 0, s>>>(
                n_cells, d_ptr.get(), d_faces.get(), d_own.get(), d_flux.get(), d_gamma.get(),
                d_delta.get(), convection, d_diag.get(), d_div.get());
            hip_check(hipGetLastError(), "k_assemble_cells launch");
        }
        auto down = [&](double* h, const DevBuf<double>& d, int n) {
            if (n > 0) {
                hip_check(hipMemcpyAsync(h, d.get(), static_cast<std::size_t>(n) * sizeof(double),
                                         hipMemcpyDeviceToHost, s),
                          "D2H");
            }
        };
        down(upper, d_up, n_faces);
        down(lower, d_lo, n_faces);
        down(diag, d_diag, n_cells);
        down(divergence, d_div, n_cells);
        hip_check(hipStreamSynchronize(s), "assemble sync");
    });
}
