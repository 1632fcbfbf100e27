// Device-resident systems, the batched solve driver and the C-ABI of the solver.
//
// Reference counterparts: DeviceSolver (src/gpu/src/device_solver.cpp), the CUDA system
// (src/gpu/src/krylov_cuda.cu:566-677) and the solver loops of
// src/gpu/src/krylov_kernels.inc. The host here only enqueues kernels; every scalar of the
// Krylov recurrences lives on the device, and the host reads back one integer per iteration
// (the number of systems still running) one iteration behind the launches.
#include "solver.hpp"
#include "fv.hpp"
#include "kernels.hpp"
#include "dense.hpp"

#include <algorithm>
#include <climits>
#include <limits>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>

namespace cfdg {

namespace {

constexpr int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

__global__ void __launch_bounds__(k_block) k_fill(int n, double* __restrict__ x, double v) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i < n) {
        x[i] = v;
    }
}

// Column formats of a sliced-ELL fill: 32-bit columns, 16-bit offsets column - row (square
// matrices), or unsigned 16-bit offsets column - cbase[slice].
enum ColFormat : int { COL_ABS = 0, COL_ROW = 1, COL_BASE = 2 };

// Sliced-ELL fill from CSR, one thread per (slice, lane): slot sp[s] + 64 k + lane holds entry
// k lanes + lane mod lanes of row s 64 / lanes + lane / lanes, or padding (value 0, a column the
// row touches anyway, or the slice's base column), exactly as csr_to_ell. src != nullptr: also
// the CSR position map (LDU patterns); va may be null (values assembled later).
template <int Fmt>
__global__ void __launch_bounds__(k_block) k_ell_fill(int n_rows, int n_cols, int n_slices, int lanes,
                                                      const int* __restrict__ sp, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, const double* __restrict__ va,
                                                      const int* __restrict__ csr_src, int* __restrict__ col,
                                                      short* __restrict__ dcol, double* __restrict__ v64,
                                                      float* __restrict__ v32, int* __restrict__ src,
                                                      const int* __restrict__ cbase = nullptr) {
    const long long t = static_cast<long long>(blockIdx.x) * k_block + threadIdx.x;
    const int slice = static_cast<int>(t / k_wave);
    if (slice >= n_slices) {
        return;
    }
    const int lane = static_cast<int>(t % k_wave);
    const int r = slice * (k_wave / lanes) + lane / lanes;
    const int sub = lane % lanes;
    const int start = r < n_rows ? rp[r] : 0;
    const int len = r < n_rows ? rp[r + 1] - start : 0;
    int pad = (n_rows == n_cols && n_rows > 0) ? n_rows - 1 : 0;
    if (r < n_rows) {
        pad = n_rows == n_cols ? r : (len > 0 ? ci[start + len - 1] : 0);
    }
    const int base = sp[slice];
    const int w = (sp[slice + 1] - base) / k_wave;
    if constexpr (Fmt == COL_BASE) {
        pad = cbase[slice];
    }
    for (int k = 0; k < w; ++k) {
        const int j = k * lanes + sub;
        const long long slot = static_cast<long long>(base) + static_cast<long long>(k) * k_wave + lane;
        const bool real = j < len;
        const int c = real ? ci[start + j] : pad;
        if constexpr (Fmt == COL_ROW) {
            dcol[slot] = static_cast<short>(c - r);
        } else if constexpr (Fmt == COL_BASE) {
            dcol[slot] = static_cast<short>(static_cast<unsigned short>(c - cbase[slice]));
        } else {
            col[slot] = c;
        }
        if (va != nullptr) {
            const double v = real ? va[start + j] : 0.0;
            v64[slot] = v;
            v32[slot] = static_cast<float>(v);
        }
        if (src != nullptr) {
            src[slot] = real ? csr_src[start + j] : -1;
        }
    }
}

// Smallest and largest column of the real entries of each slice (0 for a slice with none); the
// grid covers whole slices, so a wavefront exits as a whole.
__global__ void __launch_bounds__(k_block) k_slice_cols(int n_rows, int n_slices, int lanes, const int* __restrict__ rp,
                                                        const int* __restrict__ ci, int* __restrict__ lo,
                                                        int* __restrict__ hi) {
    const long long t = static_cast<long long>(blockIdx.x) * k_block + threadIdx.x;
    const int slice = static_cast<int>(t / k_wave);
    if (slice >= n_slices) {
        return;
    }
    const int lane = static_cast<int>(t % k_wave);
    const int r = slice * (k_wave / lanes) + lane / lanes;
    int mn = 0x7fffffff;
    int mx = -1;
    if (r < n_rows) {
        for (int j = rp[r] + lane % lanes; j < rp[r + 1]; j += lanes) {
            mn = min(mn, ci[j]);
            mx = max(mx, ci[j]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, __shfl_xor(mn, off, k_wave));
        mx = max(mx, __shfl_xor(mx, off, k_wave));
    }
    if (lane == 0) {
        lo[slice] = mx < 0 ? 0 : mn;
        hi[slice] = mx < 0 ? 0 : mx;
    }
}

// Widest per-slice stencil tried; wider unions keep the other column formats.
constexpr int k_max_stencil = 32;

// Per-slice stencil of a square matrix with one lane per row: U_s, the ascending union over the
// slice's rows of the offsets column - row. Count pass (Write false): width[s] = |U_s|, or -1 when
// a row's columns do not ascend strictly or |U_s| > k_max_stencil. Write pass (sp from those
// widths): soff[sp[s] / 64 + k] = U_s[k], and slot sp[s] + 64 k + lane holds the row's entry of
// offset U_s[k] (value, or CSR position map through csr_src), else 0 (map -1). A wavefront owns a
// slice and takes the offsets in ascending order by a wavefront minimum, so a row's entries keep
// their CSR order. The grid covers whole slices.
template <bool Write>
__global__ void __launch_bounds__(k_block) k_slice_stencil(int n_rows, int n_slices, const int* __restrict__ rp,
                                                           const int* __restrict__ ci, int* __restrict__ width,
                                                           const int* __restrict__ sp, int* __restrict__ soff,
                                                           const double* __restrict__ va,
                                                           const int* __restrict__ csr_src, double* __restrict__ v64,
                                                           float* __restrict__ v32, int* __restrict__ src) {
    const long long t = static_cast<long long>(blockIdx.x) * k_block + threadIdx.x;
    const int slice = static_cast<int>(t / k_wave);
    if (slice >= n_slices) {
        return;
    }
    const int lane = static_cast<int>(t % k_wave);
    const int r = slice * k_wave + lane;
    int j = r < n_rows ? rp[r] : 0;
    const int end = r < n_rows ? rp[r + 1] : 0;
    if constexpr (!Write) {
        bool ascending = true;
        for (int q = j + 1; q < end; ++q) {
            ascending = ascending && ci[q] > ci[q - 1];
        }
        if (!__all(ascending)) {
            if (lane == 0) {
                width[slice] = -1;
            }
            return;
        }
    }
    const int w = Write ? (sp[slice + 1] - sp[slice]) / k_wave : k_max_stencil;
    const long long base = Write ? sp[slice] : 0;
    int k = 0;
    for (;;) {
        const int mine = j < end ? ci[j] - r : INT_MAX;
        int m = mine;
        for (int off = 32; off > 0; off >>= 1) {
            m = min(m, __shfl_xor(m, off, k_wave));
        }
        if (m == INT_MAX || k >= w) { // uniform over the wavefront
            break;
        }
        const bool real = mine == m;
        if constexpr (Write) {
            if (lane == 0) {
                soff[base / k_wave + k] = m;
            }
            const long long slot = base + static_cast<long long>(k) * k_wave + lane;
            if (va != nullptr) {
                const double v = real ? va[j] : 0.0;
                v64[slot] = v;
                v32[slot] = static_cast<float>(v);
            }
            if (src != nullptr) {
                src[slot] = real ? csr_src[j] : -1;
            }
        }
        j += real ? 1 : 0;
        ++k;
    }
    if constexpr (!Write) {
        const bool more = __any(j < end);
        if (lane == 0) {
            width[slice] = more ? -1 : k;
        }
    }
}

// Slice offsets of the per-slice stencil layout of a square matrix with one lane per row (rp_d,
// ci_d: its CSR arrays on the device), or empty when a slice has none, the stencil slots exceed
// the sliced-ELL slots (ell_slots) by more than 10%, or CFDG_NO_STENCIL is set.
std::vector<int> stencil_offsets(int rows, const int* rp_d, const int* ci_d, long long ell_slots, hipStream_t s) {
    const int n_slices = (rows + k_wave - 1) / k_wave;
    if (n_slices == 0 || std::getenv("CFDG_NO_STENCIL") != nullptr) {
        return {};
    }
    DevBuf<int> width;
    width.resize(static_cast<std::size_t>(n_slices));
    const int grid = ceil_div(static_cast<long long>(n_slices) * k_wave, k_block);
    k_slice_stencil<false><<<grid, k_block, 0, s>>>(rows, n_slices, rp_d, ci_d, width.get(), nullptr, nullptr, nullptr,
                                                    nullptr, nullptr, nullptr, nullptr);
    hip_check(hipGetLastError(), "k_slice_stencil launch");
    std::vector<int> w(static_cast<std::size_t>(n_slices));
    hip_check(hipMemcpyAsync(w.data(), width.get(), w.size() * sizeof(int), hipMemcpyDeviceToHost, s),
              "D2H stencil widths");
    hip_check(hipStreamSynchronize(s), "stencil widths sync");
    std::vector<int> sp(static_cast<std::size_t>(n_slices) + 1, 0);
    long long total = 0;
    for (int k = 0; k < n_slices; ++k) {
        if (w[static_cast<std::size_t>(k)] < 0) {
            return {};
        }
        total += static_cast<long long>(w[static_cast<std::size_t>(k)]) * k_wave;
        if (total > std::numeric_limits<int>::max()) {
            return {};
        }
        sp[static_cast<std::size_t>(k) + 1] = static_cast<int>(total);
    }
    if (total == 0 || static_cast<double>(total) > 1.1 * static_cast<double>(ell_slots)) {
        return {};
    }
    return sp;
}

__device__ __forceinline__ unsigned short float_to_bf16(float f) { // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

// BF16 copy of a sliced-ELL FP64 operator: per row the scale s = max |off-diagonal| (1 if
// none), off-diagonal slots bf16(a / s), the diagonal slot 0 and the diagonal in FP32. The
// diagonal is the first slot of the row holding its own column (padding slots come after it).
// The stored diagonal absorbs the rounding of the off-diagonals, d' = d + sum_j (a_ij - s
// bf16(a_ij / s)), so that row sums (the operator applied to a constant) are kept: rounding
// errors then act like a small perturbation of the operator's differences and leave the
// near-null space of diffusion operators alone.
__global__ void __launch_bounds__(k_block) k_make_bf16(EllView m, unsigned short* __restrict__ vbf,
                                                       float* __restrict__ rscale, float* __restrict__ diag) {
    const int row = blockIdx.x * k_block + threadIdx.x;
    const int slice = row / k_wave;
    if (slice >= m.n_slices) {
        return;
    }
    const int lane = row % k_wave;
    const int base = m.sp[slice];
    const int w = (m.sp[slice + 1] - base) / k_wave;
    // Column of slot k (setup kernel, run once per assembly: a per-slot test is fine here).
    auto column = [&](int k) {
        if (m.soff != nullptr) {
            return row + m.soff[base / k_wave + k / k_wave];
        }
        return m.dcol != nullptr ? row + static_cast<int>(m.dcol[base + lane + k]) : m.col[base + lane + k];
    };
    const double* val = static_cast<const double*>(m.val) + base + lane;
    int kd = -1;
    double d = 0.0;
    double mx = 0.0;
    for (int k = 0; k < w; ++k) {
        const double v = val[k * k_wave];
        if (kd < 0 && row < m.n_rows && column(k * k_wave) == row) {
            kd = k;
            d = v;
        } else {
            mx = fmax(mx, fabs(v));
        }
    }
    const float sf = static_cast<float>(mx > 0.0 ? mx : 1.0);
    double lost = 0.0;
    for (int k = 0; k < w; ++k) {
        unsigned short h = 0;
        if (k != kd) {
            const double v = val[k * k_wave];
            h = float_to_bf16(static_cast<float>(v / static_cast<double>(sf)));
            lost += v - static_cast<double>(sf) * static_cast<double>(__uint_as_float(static_cast<unsigned>(h) << 16));
        }
        vbf[base + lane + k * k_wave] = h;
    }
    rscale[row] = sf;
    diag[row] = static_cast<float>(d + lost);
}

__global__ void __launch_bounds__(k_block) k_scale_dinv(int n, const double* __restrict__ dinv, double w64,
                                                        double w32, double* __restrict__ wd64,
                                                        float* __restrict__ wd32) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i < n) {
        wd64[i] = w64 * dinv[i];
        wd32[i] = static_cast<float>(w32 * dinv[i]);
    }
}

// Gershgorin bound of the eigenvalues of D^-1 A: max_i sum_j |a_ij| / |a_ii| (2 for the
// diffusion operators with zero row sums, which is also their largest eigenvalue on fine grids).
double gershgorin(const cfdg_csr& a) {
    double g = 0.0;
    for (int i = 0; i < a.n_rows; ++i) {
        double d = 0.0;
        double s = 0.0;
        for (int k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) {
            s += std::fabs(a.val[k]);
            if (a.col[k] == i) {
                d = std::fabs(a.val[k]);
            }
        }
        if (d > 0.0) {
            g = std::max(g, s / d);
        }
    }
    return g > 0.0 ? g : 2.0;
}

// Damping of a mixed-precision Jacobi sweep on a level whose D^-1 A spectrum lies below h: the
// centre of [h / ratio, h] (the first-degree Chebyshev smoother, which damps the upper part of
// the spectrum uniformly). A sweep diverges on the top mode only if h (1 + 1 / ratio) is below
// the largest eigenvalue.
// Damping of the mixed smoother on a level whose D^-1 A spectrum ends at h: the centre of
// [h / ratio, h]; ratio 0 selects the hierarchy's own (reference) weight w_ref instead.
double mixed_weight(double h, double w_ref) {
    const double ratio = mixed_ratio();
    return ratio == 0.0 ? w_ref : 2.0 / (h + h / ratio);
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
    int ok = 1;
    for (int i = 0; i < rows; ++i) {
        ok = std::min(ok, m.row_ptr[i + 1] >= m.row_ptr[i] ? 1 : 0);
    }
    require(ok != 0, what);
    for (int k = 0; k < m.nnz; ++k) {
        ok = std::min(ok, (m.col[k] >= 0 && m.col[k] < cols) ? 1 : 0);
    }
    require(ok != 0, what);
}

} // namespace

// Smoother of the mixed-precision V-cycle: sweeps per side (default 1, or CFDG_MIXED_SWEEPS)
// and the ratio of the damping interval (default 6, or CFDG_MIXED_RATIO; see mixed_weight),
// process-wide, settable with cfdg_set_mixed_smoother.
namespace {
std::atomic<int> g_mixed_sweeps{-1};
std::atomic<double> g_mixed_ratio{-1.0};
} // namespace

int mixed_sweeps() {
    int v = g_mixed_sweeps.load();
    if (v < 0) {
        const char* e = std::getenv("CFDG_MIXED_SWEEPS");
        v = e != nullptr ? std::max(0, std::min(64, std::atoi(e))) : 1;
        int expected = -1;
        g_mixed_sweeps.compare_exchange_strong(expected, v);
        v = g_mixed_sweeps.load();
    }
    return v;
}

double mixed_ratio() {
    double v = g_mixed_ratio.load();
    if (v < 0.0) {
        const char* e = std::getenv("CFDG_MIXED_RATIO");
        const double r = e != nullptr ? std::atof(e) : 6.0;
        v = (r > 1.0 || r == 0.0) ? r : 6.0;
        double expected = -1.0;
        g_mixed_ratio.compare_exchange_strong(expected, v);
        v = g_mixed_ratio.load();
    }
    return v;
}

void require_device() {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        throw NoDevice("no HIP device available (libcfdgpu has no host fallback)");
    }
}

std::size_t DevLevel::bytes() const {
    std::size_t b = a.bytes() + p.bytes() + r.bytes() + rt.bytes() + wd64.bytes() + wd32.bytes() + b64.bytes() +
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
    int cols_ok = 1;
    for (int k = 0; k < pt.nnz; ++k) {
        cols_ok = std::min(cols_ok, (pt.col[k] >= 0 && pt.col[k] < pt.n) ? 1 : 0);
    }
    require(cols_ok != 0, "LduPattern: column out of range");
    n_ = pt.n;
    nnz0_ = pt.nnz;
    nf_ = pt.n_faces;
    std::vector<int> csr_src(static_cast<std::size_t>(pt.nnz), -1);
    int pos_ok = 1;
    const int nnz = pt.nnz;
    for (int i = 0; i < n_; ++i) {
        const int p = pt.diag_pos[i];
        if (p >= 0 && p < nnz) {
            csr_src[static_cast<std::size_t>(p)] = i;
        } else {
            pos_ok = 0;
        }
    }
    for (int f = 0; f < nf_; ++f) {
        const int pu = pt.upper_pos[f];
        const int pl = pt.lower_pos[f];
        if (pu >= 0 && pu < nnz && pl >= 0 && pl < nnz) {
            csr_src[static_cast<std::size_t>(pu)] = n_ + f;
            csr_src[static_cast<std::size_t>(pl)] = n_ + nf_ + f;
        } else {
            pos_ok = 0;
        }
    }
    require(pos_ok != 0, "LduPattern: position out of range");
    hipStream_t s = stream_.get();
    if (std::getenv("CFDG_HOST_ELL") != nullptr) { // host conversion (same layout), for checks
        const HostEll e = csr_to_ell(n_, n_, pt.row_ptr, pt.col);
        std::vector<int> src(e.src.size());
        for (std::size_t k = 0; k < src.size(); ++k) {
            src[k] = e.src[k] >= 0 ? csr_src[static_cast<std::size_t>(e.src[k])] : -1;
        }
        a0_.upload(e, nullptr, false, false, s);
        a0_.v64.resize(e.col.size());
        a0_.v32.resize(e.col.size());
        a0_src_.upload(src.data(), src.size(), s);
        ldu_.resize(static_cast<std::size_t>(n_) + 2 * static_cast<std::size_t>(nf_));
        init_vectors();
        return;
    }
    // Sliced ELL (one lane per row) and its slot -> LDU index map, built on the device.
    const std::vector<int> offsets = ell_slice_offsets(n_, pt.row_ptr, 1);
    const std::size_t slots = static_cast<std::size_t>(offsets.back());
    a0_.n_rows = n_;
    a0_.n_cols = n_;
    a0_.n_slices = static_cast<int>(offsets.size()) - 1;
    a0_.tpr = 1;
    a0_.csr = 0;
    a0_.sp.upload(offsets.data(), offsets.size(), s);
    {
        DevBuf<int> rp_d;
        DevBuf<int> ci_d;
        DevBuf<int> src_d;
        rp_d.upload(pt.row_ptr, static_cast<std::size_t>(n_) + 1, s);
        ci_d.upload(pt.col, static_cast<std::size_t>(pt.nnz), s);
        src_d.upload(csr_src.data(), csr_src.size(), s);
        const int grid = ceil_div(static_cast<long long>(a0_.n_slices) * k_wave, k_block);
        const std::vector<int> st = stencil_offsets(n_, rp_d.get(), ci_d.get(), offsets.back(), s);
        a0_.soff.resize(0);
        a0_.cbase.resize(0);
        if (!st.empty()) {
            const std::size_t st_slots = static_cast<std::size_t>(st.back());
            a0_.sp.upload(st.data(), st.size(), s);
            a0_.soff.resize(st_slots / k_wave);
            a0_.col.resize(0);
            a0_.dcol.resize(0);
            a0_src_.resize(st_slots);
            k_slice_stencil<true><<<grid, k_block, 0, s>>>(n_, a0_.n_slices, rp_d.get(), ci_d.get(), nullptr,
                                                           a0_.sp.get(), a0_.soff.get(), nullptr, src_d.get(),
                                                           nullptr, nullptr, a0_src_.get());
            hip_check(hipGetLastError(), "k_slice_stencil launch");
            hip_check(hipStreamSynchronize(s), "pattern build sync");
            a0_.v64.resize(st_slots);
            a0_.v32.resize(st_slots);
            build_symmetric_layout(pt, csr_src);
            ldu_.resize(static_cast<std::size_t>(n_) + 2 * static_cast<std::size_t>(nf_));
            init_vectors();
            return;
        }
        const bool short_cols = short_offsets_fit(n_, n_, pt.row_ptr, pt.col);
        if (short_cols) {
            a0_.dcol.resize(slots);
            a0_.col.resize(0);
        } else {
            a0_.col.resize(slots);
            a0_.dcol.resize(0);
        }
        a0_src_.resize(slots);
        if (short_cols) {
            k_ell_fill<COL_ROW><<<grid, k_block, 0, s>>>(n_, n_, a0_.n_slices, 1, a0_.sp.get(), rp_d.get(), ci_d.get(),
                                                      nullptr, src_d.get(), nullptr, a0_.dcol.get(), nullptr,
                                                      nullptr, a0_src_.get());
        } else {
            k_ell_fill<COL_ABS><<<grid, k_block, 0, s>>>(n_, n_, a0_.n_slices, 1, a0_.sp.get(), rp_d.get(), ci_d.get(),
                                                       nullptr, src_d.get(), a0_.col.get(), nullptr, nullptr,
                                                       nullptr, a0_src_.get());
        }
        hip_check(hipGetLastError(), "k_ell_fill launch");
        hip_check(hipStreamSynchronize(s), "pattern build sync");
    }
    a0_.v64.resize(slots);
    a0_.v32.resize(slots);
    ldu_.resize(static_cast<std::size_t>(n_) + 2 * static_cast<std::size_t>(nf_));
    init_vectors();
}

void System::build_symmetric_layout(const cfdg_ldu_pattern& pt, const std::vector<int>& csr_src) {
    sym_w_ = 0;
    sym_n_ = 0;
    sym_active_ = false;
    if (std::getenv("CFDG_NO_SYM") != nullptr) {
        return;
    }
    // The stencil: the distinct column offsets of all rows (at most k_max_offsets).
    constexpr int k_max_offsets = 16;
    std::vector<int> off;
    bool small = true;
#pragma omp parallel
    {
        std::vector<int> mine;
        bool ok = true;
#pragma omp for schedule(static) nowait
        for (int i = 0; i < n_; ++i) {
            for (int k = pt.row_ptr[i]; k < pt.row_ptr[i + 1] && ok; ++k) {
                const int o = pt.col[k] - i;
                if (std::find(mine.begin(), mine.end(), o) == mine.end()) {
                    mine.push_back(o);
                    ok = static_cast<int>(mine.size()) <= k_max_offsets;
                }
            }
        }
#pragma omp critical
        {
            small = small && ok;
            for (const int o : mine) {
                if (std::find(off.begin(), off.end(), o) == off.end()) {
                    off.push_back(o);
                }
            }
        }
    }
    if (!small || static_cast<int>(off.size()) > k_max_offsets) {
        return;
    }
    std::sort(off.begin(), off.end());
    std::vector<int> up; // offsets >= 0: the stored slots
    for (const int o : off) {
        if (!std::binary_search(off.begin(), off.end(), -o)) {
            return; // not structurally symmetric
        }
        if (o >= 0) {
            up.push_back(o);
        }
    }
    const int n_off = static_cast<int>(off.size());
    const int w = static_cast<int>(up.size());
    std::vector<int> tab(static_cast<std::size_t>(2 * n_off));
    for (int k = 0; k < n_off; ++k) {
        tab[static_cast<std::size_t>(k)] = off[static_cast<std::size_t>(k)];
        const int a = off[static_cast<std::size_t>(k)] < 0 ? -off[static_cast<std::size_t>(k)] : off[static_cast<std::size_t>(k)];
        tab[static_cast<std::size_t>(n_off + k)] = static_cast<int>(std::lower_bound(up.begin(), up.end(), a) - up.begin());
    }
    const std::size_t per_slice = static_cast<std::size_t>(w) * k_wave;
    const std::size_t slots = static_cast<std::size_t>(ceil_div(n_, k_wave)) * per_slice;
    std::vector<int> src(slots, -1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_; ++i) {
        for (int k = pt.row_ptr[i]; k < pt.row_ptr[i + 1]; ++k) {
            const int o = pt.col[k] - i;
            if (o >= 0) {
                const std::size_t j = static_cast<std::size_t>(std::lower_bound(up.begin(), up.end(), o) - up.begin());
                src[static_cast<std::size_t>(i / k_wave) * per_slice + j * k_wave + static_cast<std::size_t>(i % k_wave)] =
                    csr_src[static_cast<std::size_t>(k)];
            }
        }
    }
    hipStream_t s = stream_.get();
    sym_src_.upload(src.data(), src.size(), s);
    sym_tab_.upload(tab.data(), tab.size(), s);
    sym_v64_.resize(slots);
    sym_v32_.resize(slots);
    sym_flag_.resize(1);
    hip_check(hipStreamSynchronize(s), "symmetric layout sync");
    sym_w_ = w;
    sym_n_ = n_off;
}

EllView System::a0_view(bool fp32) const {
    EllView e = fp32 ? a0_.view32() : a0_.view64();
    if (sym_active_) {
        e.sym_w = sym_w_;
        e.sym_n = sym_n_;
        e.soff = sym_tab_.get();
        e.val = fp32 ? static_cast<const void*>(sym_v32_.get()) : static_cast<const void*>(sym_v64_.get());
    }
    return e;
}

System::System(const cfdg_csr& a) {
    require_device();
    require(a.n_rows >= 1, "CsrView: empty matrix");
    check_csr(a, a.n_rows, a.n_rows, "CsrView: invalid square CSR matrix");
    n_ = a.n_rows;
    nnz0_ = a.nnz;
    nf_ = 0;
    hipStream_t s = stream_.get();
    if (std::getenv("CFDG_HOST_ELL") != nullptr) { // host conversion (same layout), for checks
        a0_.upload(csr_to_ell(n_, n_, a.row_ptr, a.col), a.val, true, true, s);
    } else {
        a0_.build_on_device(n_, n_, a.row_ptr, a.col, a.val, 1, s);
    }
    init_vectors();
    const std::vector<double> d = inverse_diagonal(a);
    dinv0_.upload(d.data(), d.size(), s);
    assembled_ = true;
}

void System::init_vectors() {
    const std::size_t n = static_cast<std::size_t>(n_);
    for (DevBuf<double>* v : {&b_, &x_, &r_, &p_, &p2_, &q_, &z_, &r0_, &v_, &s_, &t_, &ph_, &sh_, &dinv0_,
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

void System::residual_norms(double out[3]) {
    require(assembled_, "normalised residual: the matrix has not been assembled");
    const EllView a = a0_view(false);
    require(a.tpr == 1 && a.csr == 0, "normalised residual: unexpected level-0 layout");
    hipStream_t s = stream_.get();
    const int nblk = ceil_div(n_, k_block);
    if (norms_.size() < 4) {
        norms_.resize(4);
    }
    double* part = partials_.get(); // k_slots * nblk entries; no solve is queued meanwhile
    k_nr_xsum<<<nblk, k_block, 0, s>>>(n_, x_.get(), part);
    k_nr_reduce<<<1, k_fin_block, 0, s>>>(part, nblk, 1, norms_.get());
    k_nr_rows<<<nblk, k_block, 0, s>>>(a, n_, x_.get(), b_.get(), norms_.get(), part, nblk);
    k_nr_reduce<<<1, k_fin_block, 0, s>>>(part, nblk, 3, norms_.get() + 1);
    hip_check(hipGetLastError(), "normalised residual launch");
    double h[4];
    hip_check(hipMemcpyAsync(h, norms_.get(), sizeof h, hipMemcpyDeviceToHost, s), "D2H norms");
    hip_check(hipStreamSynchronize(s), "norms sync");
    out[0] = h[1];
    out[1] = h[2];
    out[2] = h[3];
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
    finish_assembly();
}

void System::assemble_device(const double* diag, const double* upper, const double* lower, const double* b) {
    require(nf_ >= 0 && a0_src_.size() > 0, "assemble: system was not created from an LDU pattern");
    hipStream_t s = stream_.get();
    const std::size_t n = static_cast<std::size_t>(n_);
    const std::size_t nf = static_cast<std::size_t>(nf_);
    double* ldu = ldu_.get();
    hip_check(hipMemcpyAsync(ldu, diag, n * sizeof(double), hipMemcpyDeviceToDevice, s), "D2D diag");
    if (nf > 0) {
        hip_check(hipMemcpyAsync(ldu + n, upper, nf * sizeof(double), hipMemcpyDeviceToDevice, s), "D2D upper");
        hip_check(hipMemcpyAsync(ldu + n + nf, lower, nf * sizeof(double), hipMemcpyDeviceToDevice, s),
                  "D2D lower");
    }
    hip_check(hipMemcpyAsync(b_.get(), b, n * sizeof(double), hipMemcpyDeviceToDevice, s), "D2D rhs");
    finish_assembly();
}

void System::finish_assembly() {
    hipStream_t s = stream_.get();
    double* ldu = ldu_.get();
    const long long slots = static_cast<long long>(a0_src_.size());
    k_ldu_gather<<<ceil_div(slots, k_block), k_block, 0, s>>>(slots, a0_src_.get(), ldu,
                                                               a0_.v64.get(), a0_.v32.get());
    hip_check(hipGetLastError(), "k_ldu_gather launch");
    if (sym_w_ > 0) { // both layouts get the values; the symmetry check picks the one kernels read
        const long long ss = static_cast<long long>(sym_src_.size());
        k_ldu_gather<<<ceil_div(ss, k_block), k_block, 0, s>>>(ss, sym_src_.get(), ldu, sym_v64_.get(),
                                                                sym_v32_.get());
        hip_check(hipGetLastError(), "k_ldu_gather launch");
        hip_check(hipMemsetAsync(sym_flag_.get(), 0, sizeof(int), s), "symmetry flag clear");
        if (nf_ > 0) {
            k_ldu_asymmetric<<<ceil_div(nf_, k_block), k_block, 0, s>>>(nf_, ldu + n_, ldu + n_ + nf_,
                                                                        sym_flag_.get());
            hip_check(hipGetLastError(), "k_ldu_asymmetric launch");
        }
    }
    k_diag_inverse<<<ceil_div(n_, k_block), k_block, 0, s>>>(n_, ldu, weight_, weight0_32_, dinv0_.get(),
                                                             wd0_64_.get(), wd0_32_.get());
    hip_check(hipGetLastError(), "k_diag_inverse launch");
    int asym = 1;
    if (sym_w_ > 0) {
        hip_check(hipMemcpyAsync(&asym, sym_flag_.get(), sizeof(int), hipMemcpyDeviceToHost, s), "D2H symmetry flag");
    }
    hip_check(hipStreamSynchronize(s), "assemble sync");
    const bool sym = sym_w_ > 0 && asym == 0;
    if (sym != sym_active_) {
        sym_active_ = sym;
        ++generation_; // cached descriptors hold the level-0 view
    }
    assembled_ = true;
    bf16_ready_ = false;
}

void System::set_rhs(const double* b) {
    require(b != nullptr, "set_rhs: null b");
    b_.upload(b, static_cast<std::size_t>(n_), stream_.get());
}

void System::set_x(const double* x) {
    if (x == nullptr) { // x = 0, enqueued on the system's stream ahead of the next solve
        k_fill<<<ceil_div(n_, k_block), k_block, 0, stream_.get()>>>(n_, x_.get(), 0.0);
        hip_check(hipGetLastError(), "k_fill launch");
        pending_ = true;
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

void System::copy_rhs_from(const double* dev, hipStream_t s) {
    hip_check(hipMemcpyAsync(b_.get(), dev, static_cast<std::size_t>(n_) * sizeof(double),
                             hipMemcpyDeviceToDevice, s),
              "D2D b");
}

void System::copy_x_to(double* dev, hipStream_t s) const {
    hip_check(hipMemcpyAsync(dev, x_.get(), static_cast<std::size_t>(n_) * sizeof(double),
                             hipMemcpyDeviceToDevice, s),
              "D2D x");
}

namespace {

// Stored slots of a matrix in the larger of its two device layouts: sliced ELL (64-row slices
// padded to their longest row) or CSR.
std::size_t layout_slots(int rows, const int* rp) {
    std::size_t ell = 0;
    for (int s0 = 0; s0 < rows; s0 += k_wave) {
        int w = 0;
        for (int i = s0; i < std::min(rows, s0 + k_wave); ++i) {
            w = std::max(w, rp[i + 1] - rp[i]);
        }
        ell += static_cast<std::size_t>(k_wave) * static_cast<std::size_t>(w);
    }
    return std::max(ell, static_cast<std::size_t>(rp[rows]));
}

} // namespace

std::size_t estimate_footprint(const cfdg_ldu_pattern& pt, const cfdg_amg_view* amg) {
    require(pt.n >= 1 && pt.row_ptr != nullptr, "estimate_footprint: invalid pattern");
    const std::size_t n = static_cast<std::size_t>(pt.n);
    const std::size_t nblk0 = static_cast<std::size_t>(ceil_div(pt.n, k_block));
    // Level 0: per slot col, dcol, FP64 / FP32 / BF16 values and the LDU source index; per row
    // 15 FP64 vectors, 3 FP32 vectors and the BF16 scales.
    std::size_t b = layout_slots(pt.n, pt.row_ptr) * (4 + 2 + 8 + 4 + 2 + 4);
    b += n * (15 * 8 + 3 * 4 + 2 * 4) + (n + 2 * static_cast<std::size_t>(pt.n_faces)) * 8;
    b += n * 5 * (4 + 8 + 4); // symmetric stencil storage (box stencils store 3 to 5 slots per row)
    b += static_cast<std::size_t>(k_slots) * nblk0 * 8 + sizeof(SolveState) + 64 * 1024;
    if (amg != nullptr && amg->levels != nullptr) {
        for (int l = 0; l < amg->n_levels; ++l) {
            const cfdg_amg_level_view& v = amg->levels[l];
            const std::size_t nl = l == 0 ? n : static_cast<std::size_t>(v.a.n_rows);
            b += nl * (5 * 8 + 5 * 4); // V-cycle vectors of both precisions
            if (l > 0 && v.a.row_ptr != nullptr) {
                b += layout_slots(v.a.n_rows, v.a.row_ptr) * (4 + 2 + 8 + 4) + nl * 8;
            }
            for (const cfdg_csr* m : {&v.p, &v.r}) {
                if (m->row_ptr != nullptr && m->n_rows > 0) {
                    b += layout_slots(m->n_rows, m->row_ptr) * (4 + 2 + 8 + 4) +
                         static_cast<std::size_t>(m->n_rows + 1) * 4;
                }
            }
            if (l == amg->n_levels - 1) {
                b += nl * nl * 8;
            }
        }
    }
    return b + b / 64; // allocation granularity
}

namespace {

// x_i = pseudo-random in [-1, 1) from the index (splitmix64), the power iteration's start.
__global__ void __launch_bounds__(k_block) k_power_start(int n, double* __restrict__ x) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i < n) {
        unsigned long long z = 0x9E3779B97F4A7C15ULL * static_cast<unsigned long long>(i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        x[i] = static_cast<double>(z >> 11) * 0x1.0p-52 - 1.0;
    }
}

// y = c D^-1 A x (D^-1 = wd / w) and the block partials of y.y; 256 / tpr rows per block.
__global__ void __launch_bounds__(k_block) k_power_step(EllView m, const double* __restrict__ wd, double c,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        double* __restrict__ partials) {
    const int tpr = m.tpr;
    const int row = static_cast<int>(blockIdx.x) * (k_block / tpr) + static_cast<int>(threadIdx.x) / tpr;
    const int sub = static_cast<int>(threadIdx.x) & (tpr - 1);
    const double ax = mat_row<double>(m, row, sub, x);
    double yy = 0.0;
    if (sub == 0 && row < m.n_rows) {
        const double v = c * wd[row] * ax;
        y[row] = v;
        yy = v * v;
    }
    double in[1] = {yy};
    double out[1];
    block_sums<1>(in, out);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = out[0];
    }
}

// Largest eigenvalue of D^-1 A for an uploaded square operator (wd = w D^-1): 20 power
// iterations from the pseudo-random start; the norms are block partials summed on the host in
// block order, so the estimate is reproducible.
double device_lambda(const DevEll& a, const DevBuf<double>& wd, double w, hipStream_t s) {
    const int n = a.n_rows;
    if (n <= 0) {
        return 2.0;
    }
    const int blocks = ceil_div(static_cast<long long>(n) * a.tpr, k_block);
    DevBuf<double> x;
    DevBuf<double> y;
    DevBuf<double> part;
    x.resize(static_cast<std::size_t>(n));
    y.resize(static_cast<std::size_t>(n));
    part.resize(static_cast<std::size_t>(blocks));
    std::vector<double> host(static_cast<std::size_t>(blocks));
    k_power_start<<<ceil_div(n, k_block), k_block, 0, s>>>(n, x.get());
    hip_check(hipGetLastError(), "k_power_start launch");
    double c = 1.0 / w;  // first step: the start vector unnormalised
    double lambda = 0.0;
    for (int it = 0; it < 21; ++it) {
        k_power_step<<<blocks, k_block, 0, s>>>(a.view64(), wd.get(), c, x.get(), y.get(), part.get());
        hip_check(hipGetLastError(), "k_power_step launch");
        hip_check(hipMemcpyAsync(host.data(), part.get(), host.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                  "D2H power partials");
        hip_check(hipStreamSynchronize(s), "power sync");
        double yy = 0.0;
        for (double v : host) {
            yy += v;
        }
        const double norm = std::sqrt(yy);
        if (!(norm > 0.0) || !std::isfinite(norm)) {
            return 2.0;
        }
        if (it > 0) {
            lambda = norm; // the input had unit norm
        }
        c = 1.0 / (w * norm); // next input: y / |y|
        std::swap(x, y);
    }
    return lambda > 0.0 ? lambda : 2.0;
}

} // namespace

void DevEll::build_on_device(int rows, int cols, const int* rp, const int* ci, const double* va, int lanes,
                             hipStream_t s) {
    const std::vector<int> offsets = ell_slice_offsets(rows, rp, lanes);
    n_rows = rows;
    n_cols = cols;
    n_slices = static_cast<int>(offsets.size()) - 1;
    tpr = lanes;
    csr = 0;
    const std::size_t slots = static_cast<std::size_t>(offsets.back());
    const std::size_t nnz = static_cast<std::size_t>(rp[rows]);
    sp.upload(offsets.data(), offsets.size(), s);
    DevBuf<int> rp_d;
    DevBuf<int> ci_d;
    DevBuf<double> va_d;
    rp_d.upload(rp, static_cast<std::size_t>(rows) + 1, s);
    ci_d.upload(ci, nnz, s);
    va_d.upload(va, nnz, s);
    const int grid = ceil_div(static_cast<long long>(n_slices) * k_wave, k_block);
    // Square matrices: 16-bit offsets from the row when every column lies within 32767 of it.
    // Others (transfer operators): unsigned 16-bit offsets from the slice's smallest column when
    // every slice's columns span at most 65535.
    int fmt = COL_ABS;
    cbase.resize(0);
    soff.resize(0);
    if (rows == cols && lanes == 1 && grid > 0) {
        const std::vector<int> st = stencil_offsets(rows, rp_d.get(), ci_d.get(), offsets.back(), s);
        if (!st.empty()) {
            const std::size_t st_slots = static_cast<std::size_t>(st.back());
            sp.upload(st.data(), st.size(), s);
            soff.resize(st_slots / k_wave);
            col.resize(0);
            dcol.resize(0);
            v64.resize(st_slots);
            v32.resize(st_slots);
            k_slice_stencil<true><<<grid, k_block, 0, s>>>(rows, n_slices, rp_d.get(), ci_d.get(), nullptr, sp.get(),
                                                           soff.get(), va_d.get(), nullptr, v64.get(), v32.get(),
                                                           nullptr);
            hip_check(hipGetLastError(), "k_slice_stencil launch");
            hip_check(hipStreamSynchronize(s), "ELL build sync");
            return;
        }
    }
    if (rows == cols) {
        fmt = short_offsets_fit(rows, cols, rp, ci) ? COL_ROW : COL_ABS;
    } else if (grid > 0 && std::getenv("CFDG_NO_SHORT_COLS") == nullptr) {
        DevBuf<int> hi;
        cbase.resize(static_cast<std::size_t>(n_slices));
        hi.resize(static_cast<std::size_t>(n_slices));
        k_slice_cols<<<grid, k_block, 0, s>>>(rows, n_slices, lanes, rp_d.get(), ci_d.get(), cbase.get(), hi.get());
        hip_check(hipGetLastError(), "k_slice_cols launch");
        std::vector<int> lo_h(static_cast<std::size_t>(n_slices));
        std::vector<int> hi_h(static_cast<std::size_t>(n_slices));
        hip_check(hipMemcpyAsync(lo_h.data(), cbase.get(), lo_h.size() * sizeof(int), hipMemcpyDeviceToHost, s),
                  "D2H slice columns");
        hip_check(hipMemcpyAsync(hi_h.data(), hi.get(), hi_h.size() * sizeof(int), hipMemcpyDeviceToHost, s),
                  "D2H slice columns");
        hip_check(hipStreamSynchronize(s), "slice columns sync");
        bool fits = true;
        for (std::size_t k = 0; k < lo_h.size() && fits; ++k) {
            fits = hi_h[k] - lo_h[k] <= 65535;
        }
        fmt = fits ? COL_BASE : COL_ABS;
        if (!fits) {
            cbase.resize(0);
        }
    }
    if (fmt != COL_ABS) {
        dcol.resize(slots);
        col.resize(0);
    } else {
        col.resize(slots);
        dcol.resize(0);
    }
    v64.resize(slots);
    v32.resize(slots);
    if (grid > 0) {
        if (fmt == COL_ROW) {
            k_ell_fill<COL_ROW><<<grid, k_block, 0, s>>>(rows, cols, n_slices, lanes, sp.get(), rp_d.get(),
                                                         ci_d.get(), va_d.get(), nullptr, nullptr, dcol.get(),
                                                         v64.get(), v32.get(), nullptr);
        } else if (fmt == COL_BASE) {
            k_ell_fill<COL_BASE><<<grid, k_block, 0, s>>>(rows, cols, n_slices, lanes, sp.get(), rp_d.get(),
                                                          ci_d.get(), va_d.get(), nullptr, nullptr, dcol.get(),
                                                          v64.get(), v32.get(), nullptr, cbase.get());
        } else {
            k_ell_fill<COL_ABS><<<grid, k_block, 0, s>>>(rows, cols, n_slices, lanes, sp.get(), rp_d.get(),
                                                         ci_d.get(), va_d.get(), nullptr, col.get(), nullptr,
                                                         v64.get(), v32.get(), nullptr);
        }
        hip_check(hipGetLastError(), "k_ell_fill launch");
    }
    hip_check(hipStreamSynchronize(s), "ELL build sync"); // the staging buffers go out of scope
}

void System::set_amg(const cfdg_amg_view& v, bool fuse_wanted) {
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
    setup_phase("set_amg: checks", 0);
    // With one sweep per side in the FP32 cycle (the smoother applying to this hierarchy), the
    // down leg of levels 1 .. L-2 with at most fuse_rows rows is the fused restriction
    // R (I - A diag(wd32)): on those levels a launch costs more than the bytes, while on larger
    // ones the product moves about as many bytes as the two passes and takes the longest to form.
    // CFDG_FUSED_RESTRICTION: 0 off, N > 1 the row limit (default 65536).
    long long fuse_rows = 65536;
    if (const char* fenv = std::getenv("CFDG_FUSED_RESTRICTION")) {
        fuse_rows = std::atoll(fenv);
    }
    const bool fuse = fuse_wanted && mixed_sweeps() == 1 && fuse_rows > 0;
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
            for (std::size_t i = 0; i < dinv.size(); ++i) {
                wd[i] = w * dinv[i];
            }
            d.wd64.upload(wd.data(), wd.size(), s);
            // Upper end of the spectrum of D^-1 A: the power-iteration estimate (4-8% low on the
            // box hierarchies) with a 15% margin.
            setup_phase("set_amg: operator upload", l);
            const double w32 = mixed_ratio() == 0.0 ? w : mixed_weight(1.15 * device_lambda(d.a, d.wd64, w, s), w);
            setup_phase("set_amg: spectrum estimate", l);
            std::vector<float> wd32(dinv.size());
            for (std::size_t i = 0; i < dinv.size(); ++i) {
                wd32[i] = static_cast<float>(w32 * dinv[i]);
            }
            d.wd32.upload(wd32.data(), wd32.size(), s);
            if (fuse && d.n <= fuse_rows) {
                const Csr rt = fused_restriction(v.levels[l].r, a, wd32.data());
                setup_phase("set_amg: fused restriction product", l);
                d.rt.upload_csr(d.n_next, d.n, rt.rp.data(), rt.ci.data(), rt.va.data(), s);
                setup_phase("set_amg: fused restriction upload", l);
            }
        }
        if (!coarsest) {
            const cfdg_csr& p = v.levels[l].p;
            const cfdg_csr& r = v.levels[l].r;
            d.p.upload_csr(d.n, d.n_next, p.row_ptr, p.col, p.val, s);
            d.r.upload_csr(d.n_next, d.n, r.row_ptr, r.col, r.val, s);
            setup_phase("set_amg: transfer upload", l);
        } else {
            d.cinv.upload(v.coarse_inverse, static_cast<std::size_t>(d.n) * static_cast<std::size_t>(d.n), s);
        }
    }
    // Dense tail of the FP32 cycle: a level below the finest that is not the coarsest is solved
    // exactly with its dense inverse (computed on the matrix cores) instead of being smoothed and
    // restricted further: the levels under it are launch-bound, and one dense product replaces
    // their launches. The first level with at most kTailOrder rows takes it, or one with at most
    // kDenseWideMaxOrder rows whose inverse holds at most a tenth as many entries as the finest
    // operator: in a batch the launches are shared by all members while every member reads its
    // own inverse, so the larger tails pay off only against a large finest level (CFDG_DENSE_TAIL=0:
    // off).
    int tail = 0;
    const char* tenv = std::getenv("CFDG_DENSE_TAIL");
    if (!(tenv != nullptr && std::string(tenv) == "0")) {
        const long long nnz0 = v.levels[0].a.row_ptr != nullptr && v.levels[0].a.nnz > 0
                                   ? static_cast<long long>(v.levels[0].a.nnz)
                                   : 7LL * static_cast<long long>(n_);
        for (int l = 1; l + 1 < L; ++l) {
            const int nl = rows[static_cast<std::size_t>(l)];
            const long long cells = static_cast<long long>(nl) * nl;
            if (nl > kDenseWideMaxOrder || (nl > kTailOrder && 10 * cells > nnz0)) {
                continue;
            }
            const cfdg_csr& a = v.levels[l].a;
            std::vector<double> dense(static_cast<std::size_t>(nl) * static_cast<std::size_t>(nl), 0.0);
            for (int i = 0; i < nl; ++i) {
                for (int k = a.row_ptr[i]; k < a.row_ptr[i + 1]; ++k) {
                    dense[static_cast<std::size_t>(i) * static_cast<std::size_t>(nl) + static_cast<std::size_t>(a.col[k])] +=
                        a.val[k];
                }
            }
            std::vector<double> inv(dense.size());
            const std::vector<int> ok = dense_inverse_batch({nl}, {dense.data()}, {inv.data()}, s);
            if (ok[0] != 0) {
                lv[static_cast<std::size_t>(l)].cinv_tail.upload(inv.data(), inv.size(), s);
                tail = l;
            }
            break;
        }
        setup_phase("set_amg: dense tail", tail);
    }
    // Level 0's operator is the assembled one (it may differ from the hierarchy's level 0, which
    // only serves the bound); diffusion operators have lambda(D^-1 A) <= 2.
    const cfdg_csr& a0 = v.levels[0].a;
    const double lambda0 = (L > 1 && a0.row_ptr != nullptr && a0.n_rows == n_) ? gershgorin(a0) : 2.0;
    levels_ = std::move(lv);
    tail_ = tail;
    sweeps_ = v.smooth_sweeps;
    weight_ = w;
    weight0_32_ = mixed_weight(lambda0, w);
    mixed_ratio_ = mixed_ratio();
    has_amg_ = true;
    ++generation_;
    // Level-0 smoother weights follow the (possibly new) weights.
    k_scale_dinv<<<ceil_div(n_, k_block), k_block, 0, s>>>(n_, dinv0_.get(), weight_, weight0_32_, wd0_64_.get(),
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

void System::make_bf16() {
    const std::size_t slots = a0_.v64.size();
    const std::size_t rows = static_cast<std::size_t>(a0_.n_slices) * k_wave;
    a0_.vbf.resize(slots);
    a0_.rscale.resize(rows);
    a0_.diag32.resize(rows);
    k_make_bf16<<<ceil_div(static_cast<long long>(rows), k_block), k_block, 0, stream_.get()>>>(
        a0_.view64(), a0_.vbf.get(), a0_.rscale.get(), a0_.diag32.get());
    hip_check(hipGetLastError(), "k_make_bf16 launch");
    hip_check(hipStreamSynchronize(stream_.get()), "bf16 sync");
    bf16_ready_ = true;
}

SysDev System::descriptor(bool fp32, bool use_amg, bool bf16) {
    require(assembled_, "solve: the matrix has not been assembled");
    if (bf16 && !bf16_ready_) {
        make_bf16();
    }
    SysDev d{};
    d.st = state_.get();
    d.partials = partials_.get();
    d.nblk0 = ceil_div(n_, k_block);
    d.n = n_;
    d.a64 = a0_view(false);
    d.b = b_.get();
    d.x = x_.get();
    d.r = r_.get();
    d.p = p_.get();
    d.p2 = p2_.get();
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
    const int L = n_levels(fp32);
    d.n_levels = L;
    for (int l = 0; l < L; ++l) {
        DevLevel& s = levels_[static_cast<std::size_t>(l)];
        LevelDev& o = d.lv[l];
        const bool coarsest = l == L - 1;
        o.n = s.n;
        o.n_next = s.n_next;
        o.coarsest = coarsest ? 1 : 0;
        if (l == 0) {
            o.a = bf16 ? a0_.viewbf() : a0_view(fp32);
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
            if (fp32 && s.rt.n_rows > 0) {
                o.rt = s.rt.view32();
            }
        }
        for (int i = 0; i < 3; ++i) {
            o.v[i] = fp32 ? static_cast<void*>(s.v32[i].get()) : static_cast<void*>(s.v64[i].get());
        }
        o.cinv = (coarsest && l != static_cast<int>(levels_.size()) - 1) ? s.cinv_tail.get() : s.cinv.get();
    }
    return d;
}

std::size_t System::footprint() const {
    std::size_t b = a0_.bytes() + a0_src_.bytes() + sym_src_.bytes() + sym_tab_.bytes() + sym_v64_.bytes() +
                    sym_v32_.bytes() + sym_flag_.bytes() + ldu_.bytes() + dinv0_.bytes() + wd0_64_.bytes() +
                    wd0_32_.bytes() + r32_.bytes() + partials_.bytes() + state_.bytes();
    for (const DevBuf<double>* v : {&b_, &x_, &r_, &p_, &p2_, &q_, &z_, &r0_, &v_, &s_, &t_, &ph_, &sh_}) {
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

// A launch table in device memory: (system, block) pairs inside the owning Tables' arena.
struct TableRef {
    const int2* p = nullptr;
    [[nodiscard]] const int2* get() const { return p; }
};

struct IdsRef {
    const int* p = nullptr;
    [[nodiscard]] const int* get() const { return p; }
};

struct Tables {
    int n_sys = 0;
    int l_max = 1;
    IdsRef ids;                    // descriptor index of each system (finalize workgroups)
    std::vector<TableRef> t;       // per level: 256-row blocks of systems having the level
    std::vector<int> tn;
    std::vector<TableRef> c;       // per level: 4-row blocks of systems whose coarsest it is
    std::vector<int> cn;
    TableRef t0_single;            // level-0 blocks of one-level hierarchies
    int t0_single_n = 0;
    // per level and matrix (0 operator, 1 prolongation, 2 restriction, 3 fused restriction):
    // blocks of 256 / tpr rows of that matrix for the systems whose level l is not the coarsest
    // (3: those of them that have the fused restriction at level l)
    std::vector<TableRef> tm[4];
    std::vector<int> tmn[4];
    std::vector<char> a_ell; // every system's level-l operator is sliced ELL (fused presmooth)
    std::vector<char> x1;    // the level's residual forms x1 = wd b per gather (short rows or small)
    std::vector<char> fused; // every system has the fused restriction at level l
    DevBuf<int2> arena;      // every table above, back to back
    PinnedBuf<int2> stage;   // host image of the arena for asynchronous uploads
};

// Launch tables of a set of systems. Entries name a system by its index into the descriptor
// array: its position in `sys`, or slots[position] when slots are given. A synchronous build
// returns with the tables in device memory; an asynchronous one queues the upload on `s`
// (the tables' previous contents must no longer be in use by queued work).
void build_tables(Tables& tb, const std::vector<System*>& sys, bool use_amg, bool fp32, hipStream_t s,
                  const std::vector<int>* slots = nullptr, bool async = false) {
    tb.n_sys = static_cast<int>(sys.size());
    std::vector<int> ids(sys.size());
    std::iota(ids.begin(), ids.end(), 0);
    if (slots != nullptr) {
        require(slots->size() == sys.size(), "build_tables: one slot per system");
        ids = *slots;
    }
    tb.l_max = 1;
    if (use_amg) {
        for (System* x : sys) {
            tb.l_max = std::max(tb.l_max, x->n_levels(fp32));
        }
    }
    const std::size_t L = static_cast<std::size_t>(tb.l_max);
    // Host image: ids (two per int2), then the tables; offsets recorded per table.
    std::vector<int2> img((ids.size() + 1) / 2, make_int2(0, 0));
    std::memcpy(img.data(), ids.data(), ids.size() * sizeof(int));
    std::vector<std::size_t> off_t(L), off_c(L), off_m[4];
    tb.tn.assign(L, 0);
    tb.cn.assign(L, 0);
    std::vector<int2> single;
    for (int l = 0; l < tb.l_max; ++l) {
        const std::size_t lu = static_cast<std::size_t>(l);
        off_t[lu] = img.size();
        std::vector<int2> c;
        for (int q = 0; q < tb.n_sys; ++q) {
            System* x = sys[static_cast<std::size_t>(q)];
            const int id = ids[static_cast<std::size_t>(q)];
            const int nl = use_amg ? x->n_levels(fp32) : 1;
            if (l >= nl) {
                continue;
            }
            const int rows = (use_amg && l > 0) ? x->level_rows(l) : x->n();
            const int nb = ceil_div(rows, k_block);
            for (int b = 0; b < nb; ++b) {
                img.push_back(make_int2(id, b));
            }
            if (use_amg && l == nl - 1) {
                const int nc = ceil_div(rows, k_slices_per_block);
                for (int b = 0; b < nc; ++b) {
                    c.push_back(make_int2(id, b));
                }
                if (l == 0) {
                    for (int b = 0; b < nb; ++b) {
                        single.push_back(make_int2(id, b));
                    }
                }
            }
        }
        tb.tn[lu] = static_cast<int>(img.size() - off_t[lu]);
        off_c[lu] = img.size();
        tb.cn[lu] = static_cast<int>(c.size());
        img.insert(img.end(), c.begin(), c.end());
    }
    const std::size_t off_single = img.size();
    tb.t0_single_n = static_cast<int>(single.size());
    img.insert(img.end(), single.begin(), single.end());
    tb.a_ell.assign(L, 1);
    tb.x1.assign(L, 1);
    tb.fused.assign(L, 0);
    for (int which = 0; which < 4; ++which) {
        tb.tmn[which].assign(L, 0);
        off_m[which].assign(L, img.size());
    }
    for (int l = 0; use_amg && l + 1 < tb.l_max; ++l) {
        const std::size_t lu = static_cast<std::size_t>(l);
        tb.fused[lu] = l >= 1 ? 1 : 0;
        for (int q = 0; q < tb.n_sys; ++q) {
            System* x = sys[static_cast<std::size_t>(q)];
            if (l + 1 < x->n_levels(fp32) && !x->has_fused_restriction(l)) {
                tb.fused[lu] = 0;
            }
        }
        for (int which = 0; which < 4; ++which) {
            off_m[which][lu] = img.size();
            if (which == 3 && tb.fused[lu] == 0) {
                continue;
            }
            for (int q = 0; q < tb.n_sys; ++q) {
                System* x = sys[static_cast<std::size_t>(q)];
                if (l + 1 >= x->n_levels(fp32)) {
                    continue; // level l is this system's coarsest (or absent)
                }
                const int tpr = x->tpr(l, which);
                // The fused first two sweeps gather wd and b per entry: worth it only for the
                // short rows of the finest operator.
                if (which == 0 && (tpr != 1 || l > 0)) {
                    tb.a_ell[lu] = 0;
                }
                // Forming x1 per gather doubles the gathers of a row; for many long rows a
                // stored x1 (one more row-wise launch) is cheaper.
                if (which == 0 && tpr != 1 && static_cast<long long>(x->level_rows(l)) * tpr >= (1LL << 19)) {
                    tb.x1[lu] = 0;
                }
                const int rows = which >= 2 ? x->level_rows(l + 1) : x->level_rows(l);
                const int nb = ceil_div(static_cast<long long>(rows) * tpr, k_block);
                for (int b = 0; b < nb; ++b) {
                    img.push_back(make_int2(ids[static_cast<std::size_t>(q)], b));
                }
            }
            tb.tmn[which][lu] = static_cast<int>(img.size() - off_m[which][lu]);
        }
    }
    if (tb.arena.size() < img.size()) {
        if (async) {
            hip_check(hipStreamSynchronize(s), "tables: drain before growing");
        }
        tb.arena.resize(std::max(img.size(), 2 * tb.arena.size()));
    }
    if (async) {
        if (tb.stage.size() < img.size()) {
            tb.stage = PinnedBuf<int2>(std::max(img.size(), 2 * tb.stage.size()));
        }
        std::memcpy(tb.stage.get(), img.data(), img.size() * sizeof(int2));
        hip_check(hipMemcpyAsync(tb.arena.get(), tb.stage.get(), img.size() * sizeof(int2),
                                 hipMemcpyHostToDevice, s),
                  "tables H2D");
    } else {
        hip_check(hipMemcpyAsync(tb.arena.get(), img.data(), img.size() * sizeof(int2),
                                 hipMemcpyHostToDevice, s),
                  "tables H2D");
        hip_check(hipStreamSynchronize(s), "tables H2D sync");
    }
    const int2* base = tb.arena.get();
    tb.ids.p = reinterpret_cast<const int*>(base);
    tb.t.assign(L, TableRef{});
    tb.c.assign(L, TableRef{});
    for (std::size_t l = 0; l < L; ++l) {
        tb.t[l].p = base + off_t[l];
        tb.c[l].p = base + off_c[l];
    }
    tb.t0_single.p = base + off_single;
    for (int which = 0; which < 4; ++which) {
        tb.tm[which].assign(L, TableRef{});
        for (std::size_t l = 0; l < L; ++l) {
            tb.tm[which][l].p = base + off_m[which][l];
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

#define CFDG_LAUNCH_PAIR(kernel, grid, stream, ...)                                               \
    do {                                                                                          \
        if ((grid) > 0) {                                                                         \
            kernel<<<(grid), k_block_pair, 0, (stream)>>>(__VA_ARGS__);                           \
        }                                                                                         \
    } while (0)

// One of two kernels of the same arithmetic: the pair variant (k_block_pair threads, two slices
// per wavefront) or the one-slice kernel.
#define CFDG_LAUNCH_EITHER(pair, kpair, kone, grid, stream, ...)                                  \
    do {                                                                                          \
        if (pair) {                                                                               \
            CFDG_LAUNCH_PAIR(kpair, grid, stream, __VA_ARGS__);                                   \
        } else {                                                                                  \
            CFDG_LAUNCH(kone, grid, stream, __VA_ARGS__);                                         \
        }                                                                                         \
    } while (0)

// Level-0 FP32 post-sweep with two slices per wavefront (bit-identical; CFDG_NO_PAIR_SWEEPS,
// read whenever a V-cycle is queued, keeps one slice per wavefront, for comparisons).
bool pair_sweeps() { return std::getenv("CFDG_NO_PAIR_SWEEPS") == nullptr; }

int final_index(int sweeps) {
    const int cur = (sweeps >= 2 && (sweeps - 2) % 2 == 1) ? 2 : 0;
    return sweeps % 2 == 0 ? cur : 1;
}

// pair: the level-0 prolongation and last post-sweep run two slices per wavefront (FP32 values
// only). Measured slower elsewhere: on BF16 rows, on the x1-forming residual (two gathers per
// entry), and on every coarser level or multi-lane operator (the halved workgroups cost more
// than the extra loads in flight gain there).
// Factors of the sweeps' damping on every level: the degree-k Chebyshev smoother over the
// damping interval [h / ratio, h] is k damped-Jacobi sweeps whose weights are the reciprocals of
// the polynomial's roots, tau_j = c + d cos(pi (2j - 1) / (2k)) with c = (h + h / ratio) / 2 and
// d = (h - h / ratio) / 2. Relative to the level's stored weight 1 / c (degree 1), sweep j takes
// c / tau_j = 1 / (1 + q cos(pi (2j - 1) / (2k))), q = (ratio - 1) / (ratio + 1), independent of
// h. Pre-smoothing runs j = 1..k, post-smoothing k..1 (a symmetric cycle). ratio 0 (the
// reference weight) and k = 1 give factors of exactly 1, which leave the arithmetic unchanged.
struct SweepFactors {
    double pre[64];
    [[nodiscard]] double post(int j, int k) const { return pre[k - 1 - j]; }
};

// Iterations a runner keeps queued ahead of its convergence read-back (CFDG_QUEUE_DEPTH, 1 to
// k_max_queue_depth; default 1).
constexpr int k_max_queue_depth = 7;

int queue_depth() {
    static const int d = [] {
        const char* e = std::getenv("CFDG_QUEUE_DEPTH");
        const int v = e != nullptr ? std::atoi(e) : 1;
        return std::min(std::max(v, 1), k_max_queue_depth);
    }();
    return d;
}

SweepFactors sweep_factors(int k, double ratio) {
    SweepFactors f{};
    const double q = ratio > 1.0 ? (ratio - 1.0) / (ratio + 1.0) : 0.0;
    for (int j = 0; j < 64; ++j) {
        f.pre[j] = 1.0;
    }
    if (q > 0.0 && k > 1) {
        const double pi = 3.14159265358979323846;
        for (int j = 0; j < k && j < 64; ++j) {
            f.pre[j] = 1.0 / (1.0 + q * std::cos(pi * (2.0 * j + 1.0) / (2.0 * k)));
        }
    }
    return f;
}

template <class T, int Dots>
void enqueue_vcycle(const SysDev* d, const Tables& tb, int sweeps, hipStream_t s, bool pair, const SweepFactors& sf) {
    const int fin = final_index(sweeps);
    const T f1 = static_cast<T>(sf.pre[0]);
    const int L = tb.l_max;
    CFDG_LAUNCH(k_coarse<T>, tb.cn[0], s, d, tb.c[0].get(), 0, fin);
    std::vector<int> cur(static_cast<std::size_t>(L), 0);
    // Where the first sweep from zero of level l is stored as v[x1_slot(l)] = wd b (-1: not
    // stored, it is formed per gathered entry); for l >= 1 the restriction that produces b
    // stores it.
    // The fused restriction (one sweep per side, FP32 cycle) replaces a level's residual and
    // restriction passes; nothing reads the level's stored first sweep then.
    auto fused = [&](int l) { return std::is_same_v<T, float> && sweeps == 1 && tb.fused[static_cast<std::size_t>(l)] != 0; };
    auto x1_slot = [&](int l) {
        const std::size_t li = static_cast<std::size_t>(l);
        if (sweeps == 1) {
            return tb.x1[li] != 0 || fused(l) ? -1 : 0;
        }
        if (sweeps >= 2) {
            return tb.a_ell[li] != 0 ? -1 : 2;
        }
        return -1;
    };
    for (int l = 0; l + 1 < L; ++l) {
        const std::size_t li = static_cast<std::size_t>(l);
        const int g = tb.tn[li];                        // row-wise blocks
        const int2* blk = tb.t[li].get();
        const int ga = tb.tmn[0][li];                   // operator blocks
        const int2* ba = tb.tm[0][li].get();
        const int x1_next = l + 2 < L ? x1_slot(l + 1) : -1;
        int c = 0;
        if (sweeps == 0) {
            CFDG_LAUNCH(k_zero_level<T>, g, s, d, blk, l, 0);
        } else if (sweeps == 1) {
            // x1 = wd b is formed inside the residual (or stored, for many long rows) and
            // inside the prolongation; bitwise the same as a stored sweep
            cur[li] = 0;
            if (fused(l)) {
                CFDG_LAUNCH(k_restrict_fused<T>, tb.tmn[3][li], s, d, tb.tm[3][li].get(), l, x1_next, f1);
                CFDG_LAUNCH(k_coarse<T>, tb.cn[li + 1], s, d, tb.c[li + 1].get(), l + 1, fin);
                continue;
            }
            if (tb.x1[li] != 0) {
                CFDG_LAUNCH(k_residual_x1<T>, ga, s, d, ba, l, 1, f1);
            } else {
                if (l == 0) {
                    CFDG_LAUNCH(k_jacobi_zero<T>, g, s, d, blk, l, 0, f1);
                }
                CFDG_LAUNCH(k_residual<T>, ga, s, d, ba, l, 0, 1);
            }
            CFDG_LAUNCH(k_restrict<T>, tb.tmn[2][li], s, d, tb.tm[2][li].get(), l, 1, x1_next, f1);
            CFDG_LAUNCH(k_coarse<T>, tb.cn[li + 1], s, d, tb.c[li + 1].get(), l + 1, fin);
            continue;
        } else {
            const T f2 = static_cast<T>(sf.pre[1]);
            if (tb.a_ell[li] != 0) {
                CFDG_LAUNCH(k_presmooth2<T>, ga, s, d, ba, l, 0, f1, f2);
            } else { // long rows: form x1 = wd b once instead of gathering it per entry
                if (l == 0) {
                    CFDG_LAUNCH(k_jacobi_zero<T>, g, s, d, blk, l, 2, f1);
                }
                CFDG_LAUNCH((k_jacobi<T, 0>), ga, s, d, ba, l, 2, 0, f2);
            }
            for (int k = 2; k < sweeps; ++k) {
                const int o = c == 0 ? 2 : 0;
                CFDG_LAUNCH((k_jacobi<T, 0>), ga, s, d, ba, l, c, o, static_cast<T>(sf.pre[k]));
                c = o;
            }
        }
        cur[li] = c;
        CFDG_LAUNCH(k_residual<T>, ga, s, d, ba, l, c, 1);
        CFDG_LAUNCH(k_restrict<T>, tb.tmn[2][li], s, d, tb.tm[2][li].get(), l, 1, x1_next, f1);
        CFDG_LAUNCH(k_coarse<T>, tb.cn[li + 1], s, d, tb.c[li + 1].get(), l + 1, fin);
    }
    for (int l = L - 2; l >= 0; --l) {
        const std::size_t li = static_cast<std::size_t>(l);
        const int ga = tb.tmn[0][li];
        const int2* ba = tb.tm[0][li].get();
        const int c0 = cur[li];
        if (sweeps == 1) {
            CFDG_LAUNCH_EITHER(pair && l == 0, (k_prolong_pair<T, true>), (k_prolong<T, true>), tb.tmn[1][li], s, d,
                               tb.tm[1][li].get(), l, c0, fin, f1);
        } else {
            CFDG_LAUNCH_EITHER(pair && l == 0, (k_prolong_pair<T, false>), (k_prolong<T, false>), tb.tmn[1][li], s,
                               d, tb.tm[1][li].get(), l, c0, fin, T(1));
        }
        int c = c0;
        for (int k = 0; k < sweeps; ++k) {
            const int o = c == 1 ? c0 : 1;
            const bool last = k == sweeps - 1;
            const T fk = static_cast<T>(sf.post(k, sweeps));
            if (l == 0 && last) {
                probe_begin(PROBE_L0_SWEEP, s);
                if (pair) {
                    CFDG_LAUNCH_PAIR((k_jacobi_pair<T, Dots>), ga, s, d, ba, l, c, o, fk);
                } else {
                    CFDG_LAUNCH((k_jacobi<T, Dots>), ga, s, d, ba, l, c, o, fk);
                }
                probe_end(PROBE_L0_SWEEP, s);
            } else {
                CFDG_LAUNCH((k_jacobi<T, 0>), ga, s, d, ba, l, c, o, fk);
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
void fin(const SysDev* d, const Tables& tb, hipStream_t s, int* status_out = nullptr) {
    k_finalize<Op><<<tb.n_sys, k_fin_block, 0, s>>>(d, tb.ids.get(), status_out);
}

enum VecId { V_X = 0, V_R, V_P, V_Q, V_Z, V_R0, V_V, V_S, V_T, V_PH, V_SH };

struct Plan {
    cfdg_method method;
    bool amg;
    bool fp32;
    int sweeps;
    bool bf16 = false;
    SweepFactors sf = sweep_factors(1, 0.0); // all 1: plain damped Jacobi
};

void enqueue_init(const Plan& pl, const SysDev* d, const Tables& tb, hipStream_t s) {
    const int g0 = tb.tn[0];
    const int2* b0 = tb.t[0].get();
    if (pl.method == CFDG_METHOD_BICGSTAB) {
        CFDG_LAUNCH(k_bicg_init<false>, g0, s, d, b0);
        fin<FIN_BICG_INIT>(d, tb, s);
        return;
    }
    // CG start: r, z and beta = 0; p = z + 0 p_old (p_old = 0) is formed by the first
    // iteration's product (k_spmv_pz).
    if (!pl.amg) {
        CFDG_LAUNCH(k_cg_init<false>, g0, s, d, b0);
        fin<FIN_CG_INIT>(d, tb, s);
        return;
    }
    if (pl.fp32) {
        CFDG_LAUNCH(k_residual0<true>, g0, s, d, b0);
        fin<FIN_AMG_INIT>(d, tb, s);
        enqueue_vcycle<float, 2>(d, tb, pl.sweeps, s, !pl.bf16 && pair_sweeps(), pl.sf);
        fin<FIN_BETA_FLEX>(d, tb, s);
    } else {
        CFDG_LAUNCH(k_residual0<false>, g0, s, d, b0);
        fin<FIN_AMG_INIT>(d, tb, s);
        enqueue_vcycle<double, 1>(d, tb, pl.sweeps, s, false, pl.sf);
        fin<FIN_BETA>(d, tb, s);
    }
}

// One iteration; parity = iteration number mod 2 selects the search-direction buffers.
// One iteration for every listed system; its last step stores each system's status at
// status_out[id].
void enqueue_iteration(const Plan& pl, const SysDev* d, const Tables& tb, hipStream_t s, int parity,
                       int* status_out) {
    const int g0 = tb.tn[0];
    const int2* b0 = tb.t[0].get();
    if (pl.method == CFDG_METHOD_BICGSTAB) {
        CFDG_LAUNCH(k_bicg_p, g0, s, d, b0); // rho and beta: ended the previous finalize (bicg_rho)
        CFDG_LAUNCH((k_spmv_dot<false, false>), g0, s, d, b0, V_PH, V_V, V_R0, SLOT_R0V, 0, 0);
        fin<FIN_BICG_ALPHA>(d, tb, s);
        CFDG_LAUNCH(k_bicg_s, g0, s, d, b0);
        fin<FIN_BICG_S>(d, tb, s);
        CFDG_LAUNCH(k_bicg_early_x, g0, s, d, b0);
        CFDG_LAUNCH((k_spmv_dot<true, false>), g0, s, d, b0, V_SH, V_T, V_T, SLOT_TT, V_S, SLOT_TS);
        fin<FIN_BICG_OMEGA>(d, tb, s);
        CFDG_LAUNCH(k_bicg_xr, g0, s, d, b0);
        fin<FIN_BICG_RES>(d, tb, s, status_out);
        return;
    }
    const int fz = final_index(pl.sweeps);
    probe_begin(PROBE_OUTER_SPMV, s);
    if (!pl.amg) {
        CFDG_LAUNCH((k_spmv_pz<double, false>), g0, s, d, b0, parity, 0, 0);
    } else if (pl.fp32) {
        CFDG_LAUNCH((k_spmv_pz<float, false>), g0, s, d, b0, parity, 1, fz);
    } else {
        CFDG_LAUNCH((k_spmv_pz<double, false>), g0, s, d, b0, parity, 1, fz);
    }
    probe_end(PROBE_OUTER_SPMV, s);
    fin<FIN_ALPHA>(d, tb, s);
    if (!pl.amg) {
        CFDG_LAUNCH(k_update_xr<1>, g0, s, d, b0, parity);
        fin<FIN_CG_RES_BETA>(d, tb, s, status_out);
        return;
    }
    if (pl.fp32) {
        CFDG_LAUNCH(k_update_xr<2>, g0, s, d, b0, parity);
        fin<FIN_RES>(d, tb, s);
        enqueue_vcycle<float, 2>(d, tb, pl.sweeps, s, !pl.bf16 && pair_sweeps(), pl.sf);
        fin<FIN_BETA_FLEX>(d, tb, s, status_out);
    } else {
        CFDG_LAUNCH(k_update_xr<0>, g0, s, d, b0, parity);
        fin<FIN_RES>(d, tb, s);
        enqueue_vcycle<double, 1>(d, tb, pl.sweeps, s, false, pl.sf);
        fin<FIN_BETA>(d, tb, s, status_out);
    }
}

} // namespace

// Everything a set of systems needs to be solved together, built once and reused: launch
// tables, device descriptors, and (unless kernels are being probed) HIP graphs of the solver's
// start-up sequence and of one iteration. Iteration k of the graph pipeline writes every
// system's status into pinned host memory slot k mod 2 from the device; the host waits for the
// previous slot while the next iteration is already queued.
class Runner {
public:
    Runner(const std::vector<System*>& systems, const BatchOptions& opt, hipStream_t s)
        : sys_(systems), s_(s), running_(2 * systems.size()), init_host_(systems.size()),
          out_host_(systems.size()) {
        require(!sys_.empty(), "solve_batch: no systems");
        const bool amg = opt.use_amg && opt.method == CFDG_METHOD_CG;
        const bool fp32 = amg && opt.precision != CFDG_PREC_FP64;
        int sweeps = 2;
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            if (amg) {
                require(sys_[i]->has_amg(), "solve: use_amg requested but no hierarchy was set");
                if (i == 0) {
                    sweeps = sys_[i]->sweeps(fp32);
                }
                require(sys_[i]->sweeps(fp32) == sweeps, "solve_batch: systems differ in smooth_sweeps");
            }
            gens_.push_back(sys_[i]->generation());
            uids_.push_back(sys_[i]->uid());
        }
        pl_ = Plan{opt.method, amg, fp32, sweeps};
        pl_.bf16 = amg && opt.precision == CFDG_PREC_MIXED_BF16;
        if (fp32) {
            const double ratio = sys_[0]->smoother_ratio();
            for (const System* x : sys_) {
                require(x->smoother_ratio() == ratio, "solve_batch: systems differ in the mixed smoother");
            }
            pl_.sf = sweep_factors(sweeps, ratio);
        }
        precision_ = amg ? opt.precision : CFDG_PREC_FP64;
        const bool bf16 = amg && opt.precision == CFDG_PREC_MIXED_BF16;
        build_tables(tb_, sys_, amg, pl_.fp32, s_);
        std::vector<SysDev> desc(sys_.size());
        for (std::size_t i = 0; i < sys_.size(); ++i) {
            desc[i] = sys_[i]->descriptor(fp32, amg, bf16);
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
            if (sys_[i]->generation() != gens_[i] || sys_[i]->uid() != uids_[i]) {
                return false;
            }
        }
        const bool amg = opt.use_amg && opt.method == CFDG_METHOD_CG;
        if (amg && sys_[0]->sweeps(pl_.fp32) != pl_.sweeps) {
            return false; // the mixed smoother's sweeps were changed
        }
        return opt.method == pl_.method && amg == pl_.amg &&
               (amg ? opt.precision : CFDG_PREC_FP64) == precision_;
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
        for (std::size_t i = 0; i < running_.size(); ++i) {
            running_[i] = ST_RUNNING;
        }
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
        // Iterations are enqueued `depth` ahead of the convergence read-back (CFDG_QUEUE_DEPTH,
        // 1 to 7; default 1); surplus iterations are no-ops on the device because every kernel
        // tests the system's status first. Iteration k reports into status slot k & 1; with
        // depth 2 or more the read after iteration j's event may see a later iteration's report
        // in the same slot, which is only more recent (a status only ever leaves RUNNING).
        const int depth = queue_depth();
        for (int k = 1; k <= max_it + 1 + depth; ++k) {
            const int slot = k & 1;
            if (graphs) {
                hip_check(hipGraphLaunch(iter_exec_[slot], s_), "graph launch");
            } else {
                enqueue_iteration(pl_, d, tb_, s_, slot, running_dev_ + slot * n_sys);
            }
            hip_check(hipEventRecord(ev_[k % k_events].get(), s_), "event record");
            hip_check(hipGetLastError(), "iteration launch");
            if (k > depth) {
                const int j = k - depth;
                spin_wait(ev_[j % k_events].get(), "event sync");
                if (!any_running(j & 1)) {
                    break;
                }
            }
        }
        CFDG_LAUNCH(k_true_residual<false>, tb_.tn[0], s_, d, tb_.t[0].get());
        fin<FIN_TRUE>(d, tb_, s_);
        k_gather_states<<<ceil_div(n_sys, 64), 64, 0, s_>>>(d, n_sys, d_out_.get());
        hip_check(hipEventRecord(t1_.get(), s_), "event record");
        hip_check(hipMemcpyAsync(out_host_.get(), d_out_.get(), sys_.size() * sizeof(SolveState),
                                 hipMemcpyDeviceToHost, s_),
                  "D2H states");
        hip_check(hipEventRecord(done_.get(), s_), "event record");
        spin_wait(done_.get(), "solve sync");
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
    // Statuses that iteration slot `slot` reported (host-mapped): is any system still going?
    [[nodiscard]] bool any_running(int slot) const {
        const std::size_t n = sys_.size();
        for (std::size_t i = 0; i < n; ++i) {
            const int st = running_[static_cast<std::size_t>(slot) * n + i];
            if (st == ST_RUNNING || st == ST_CONVERGED_PENDING) {
                return true;
            }
        }
        return false;
    }

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
                enqueue_iteration(pl_, d, tb_, s_, slot, running_dev_ + slot * n_sys);
            });
        }
    }

public:
    // x = 0 for every system, one launch on the runner's stream (ordered before its next solve).
    void zero_x() {
        CFDG_LAUNCH(k_zero_x, tb_.tn[0], s_, desc_.get(), tb_.t[0].get());
    }

private:
    std::vector<System*> sys_;
    std::vector<unsigned> gens_;
    std::vector<unsigned long long> uids_;
    hipStream_t s_;
    Plan pl_{};
    Tables tb_;
    DevBuf<SysDev> desc_;
    DevBuf<SolveState> d_init_;
    DevBuf<SolveState> d_out_;
    PinnedBuf<int> running_;     // per iteration slot (2) and system: status after the iteration
    int* running_dev_ = nullptr;
    PinnedBuf<SolveState> init_host_;
    PinnedBuf<SolveState> out_host_;
    static constexpr int k_events = k_max_queue_depth + 1; // one per queued iteration and the one waited on
    Event ev_[k_events];
    Event done_;
    Event t0_;
    Event t1_;
    bool use_graphs_ = true;
    long long solves_ = 0;
    cfdg_precision precision_ = CFDG_PREC_FP64;
    hipGraphExec_t init_exec_ = nullptr;
    hipGraphExec_t iter_exec_[2] = {nullptr, nullptr};
};

System::~System() {
    // Buffers go back to the device pool, where they are reused at once.
    (void)hipStreamSynchronize(stream_.get());
}

unsigned long long System::next_uid() {
    static std::atomic<unsigned long long> counter{0};
    return ++counter;
}

namespace {

enum SlotPhase : int { SLOT_FREE = 0, SLOT_RUNNING = 1, SLOT_FINISHING = 2 };

} // namespace

struct SlotMachinery {
    int k = 0;
    Plan pl{};
    double ratio = 0.0; // mixed smoother interval ratio of the plan's sweep factors
    cfdg_precision precision = CFDG_PREC_FP64;
    hipStream_t s = nullptr;
    DevBuf<SysDev> desc;
    PinnedBuf<SysDev> desc_host;
    PinnedBuf<SolveState> init_host;
    PinnedBuf<SolveState> out_host;
    PinnedBuf<int> status;      // two regions of k entries, one per iteration parity
    int* status_dev = nullptr;
    std::vector<System*> sys;   // per slot
    std::vector<int> phase;
    std::vector<Tables> single; // per slot: tables of the slot's system alone
    std::vector<unsigned long long> single_of; // uid of the system the single tables describe
    std::vector<unsigned> single_gen;
    std::vector<long long> epoch; // per slot: number of the solve occupying it
    std::vector<std::unique_ptr<Event>> done_ev;
    std::vector<std::chrono::steady_clock::time_point> t_start;
    std::vector<int> running;   // slots covered by the next iteration
    Tables merged[2];
    int cur = 0;
    bool dirty = false;
    long long g = 1;            // number of the next iteration
    std::vector<std::pair<int, long long>> covered[2]; // (slot, epoch) of each covered solve
    long long solves_started = 0;
    std::vector<long long> inflight; // iterations queued whose flags have not been read
    Event ev[2];
};

SlotRunner::SlotRunner(int slots, const BatchOptions& opt, int sweeps, hipStream_t s)
    : m_(std::make_unique<SlotMachinery>()) {
    require(slots >= 1 && slots <= 4096, "slot runner: slots must be in [1, 4096]");
    require(sweeps >= 0 && sweeps <= 64, "slot runner: invalid smooth_sweeps");
    SlotMachinery& m = *m_;
    m.k = slots;
    const bool amg = opt.use_amg && opt.method == CFDG_METHOD_CG;
    const bool fp32 = amg && opt.precision != CFDG_PREC_FP64;
    m.pl = Plan{opt.method, amg, fp32, fp32 ? mixed_sweeps() : sweeps};
    m.pl.bf16 = amg && opt.precision == CFDG_PREC_MIXED_BF16;
    if (fp32) {
        m.pl.sf = sweep_factors(m.pl.sweeps, mixed_ratio());
        m.ratio = mixed_ratio();
    }
    m.precision = amg ? opt.precision : CFDG_PREC_FP64;
    m.s = s;
    const std::size_t k = static_cast<std::size_t>(slots);
    m.desc.resize(k);
    m.desc_host = PinnedBuf<SysDev>(k);
    m.init_host = PinnedBuf<SolveState>(k);
    m.out_host = PinnedBuf<SolveState>(k);
    m.status = PinnedBuf<int>(2 * k);
    void* dev = nullptr;
    hip_check(hipHostGetDevicePointer(&dev, m.status.get(), 0), "hipHostGetDevicePointer");
    m.status_dev = static_cast<int*>(dev);
    m.sys.assign(k, nullptr);
    m.phase.assign(k, SLOT_FREE);
    m.single = std::vector<Tables>(k);
    m.single_of.assign(k, 0);
    m.single_gen.assign(k, 0);
    m.epoch.assign(k, -1);
    m.t_start.resize(k);
    for (std::size_t i = 0; i < k; ++i) {
        m.done_ev.push_back(std::make_unique<Event>());
    }
}

SlotRunner::~SlotRunner() {
    if (m_ && m_->s != nullptr) {
        (void)hipStreamSynchronize(m_->s); // queued work reads the tables and pinned buffers
    }
}

int SlotRunner::slots() const { return m_->k; }

int SlotRunner::free_slots() const {
    return static_cast<int>(std::count(m_->phase.begin(), m_->phase.end(), SLOT_FREE));
}

bool SlotRunner::busy() const {
    return std::any_of(m_->phase.begin(), m_->phase.end(), [](int p) { return p != SLOT_FREE; });
}

int SlotRunner::start(System* x, const cfdg_solve_control& c) {
    SlotMachinery& m = *m_;
    require(x != nullptr, "slot runner: null system");
    require(c.max_iterations >= 0, "solve: max_iterations must be >= 0");
    const auto it = std::find(m.phase.begin(), m.phase.end(), SLOT_FREE);
    require(it != m.phase.end(), "slot runner: no free slot");
    if (m.pl.amg) {
        require(x->has_amg(), "solve: use_amg requested but no hierarchy was set");
        require(x->sweeps(m.pl.fp32) == m.pl.sweeps, "slot runner: system's smooth_sweeps differ from the runner's");
        require(!m.pl.fp32 || x->smoother_ratio() == m.ratio,
                "slot runner: system's mixed smoother differs from the runner's");
    }
    const int j = static_cast<int>(it - m.phase.begin());
    const std::size_t ju = static_cast<std::size_t>(j);
    const bool bf16 = m.pl.amg && m.precision == CFDG_PREC_MIXED_BF16;
    if (bf16) {
        x->ensure_bf16();
    }
    SysDev d = x->descriptor(m.pl.fp32, m.pl.amg, bf16);
    hip_check(hipStreamSynchronize(x->stream()), "slot runner: system stream sync");
    // The solve's first iteration is global iteration g; it must see parity 1.
    d.flip = static_cast<int>((m.g & 1) ^ 1);
    m.desc_host[ju] = d;
    hip_check(hipMemcpyAsync(m.desc.get() + j, m.desc_host.get() + j, sizeof(SysDev),
                             hipMemcpyHostToDevice, m.s),
              "slot descriptor H2D");
    if (m.single_of[ju] != x->uid() || m.single_gen[ju] != x->generation()) {
        std::vector<System*> one{x};
        std::vector<int> ids{j};
        build_tables(m.single[ju], one, m.pl.amg, m.pl.fp32, m.s, &ids, true);
        m.single_of[ju] = x->uid();
        m.single_gen[ju] = x->generation();
    }
    SolveState& st = m.init_host[ju];
    std::memset(&st, 0, sizeof(st));
    st.tol = c.abs_tolerance;
    st.max_it = c.max_iterations;
    st.status = ST_RUNNING;
    hip_check(hipMemcpyAsync(x->state(), &st, sizeof(SolveState), hipMemcpyHostToDevice, m.s),
              "slot state H2D");
    enqueue_init(m.pl, m.desc.get(), m.single[ju], m.s);
    hip_check(hipGetLastError(), "slot init launch");
    m.sys[ju] = x;
    m.epoch[ju] = m.solves_started++;
    m.phase[ju] = SLOT_RUNNING;
    m.t_start[ju] = std::chrono::steady_clock::now();
    m.running.push_back(j);
    m.dirty = true;
    return j;
}

void SlotRunner::step(std::vector<Done>& done) {
    SlotMachinery& m = *m_;
    const SysDev* d = m.desc.get();
    // 1. Queue the next iteration over the running solves.
    if (!m.running.empty()) {
        if (m.dirty) {
            std::vector<System*> xs;
            for (int j : m.running) {
                xs.push_back(m.sys[static_cast<std::size_t>(j)]);
            }
            m.cur ^= 1; // the other table set's last users finished before the previous step
            build_tables(m.merged[m.cur], xs, m.pl.amg, m.pl.fp32, m.s, &m.running, true);
            m.dirty = false;
        }
        const Tables& tb = m.merged[m.cur];
        const int parity = static_cast<int>(m.g & 1);
        enqueue_iteration(m.pl, d, tb, m.s, parity, m.status_dev + static_cast<std::size_t>(parity) * m.k);
        hip_check(hipEventRecord(m.ev[parity].get(), m.s), "event record");
        hip_check(hipGetLastError(), "slot iteration launch");
        m.covered[parity].clear();
        for (int j : m.running) {
            m.covered[parity].emplace_back(j, m.epoch[static_cast<std::size_t>(j)]);
        }
        m.inflight.push_back(m.g);
        ++m.g;
    }
    // 2. Read the flags of the oldest queued iteration once the newest one is queued behind
    // it (or at once when nothing else is queued); retire the solves that ended there.
    if (m.inflight.size() >= 2 || (m.running.empty() && !m.inflight.empty())) {
        const long long h = m.inflight.front();
        m.inflight.erase(m.inflight.begin());
        const int hp = static_cast<int>(h & 1);
        spin_wait(m.ev[hp].get(), "event sync");
        const int* flags = m.status.get() + static_cast<std::size_t>(hp) * m.k;
        for (const auto& [j, ep] : m.covered[hp]) {
            const std::size_t ju = static_cast<std::size_t>(j);
            const int f = flags[ju];
            // Flags of a slot's earlier solve say nothing about the one that occupies it now.
            if (m.phase[ju] != SLOT_RUNNING || m.epoch[ju] != ep || f == ST_RUNNING ||
                f == ST_CONVERGED_PENDING) {
                continue;
            }
            const Tables& one = m.single[ju];
            CFDG_LAUNCH(k_true_residual<false>, one.tn[0], m.s, d, one.t[0].get());
            fin<FIN_TRUE>(d, one, m.s);
            hip_check(hipMemcpyAsync(m.out_host.get() + j, m.sys[ju]->state(), sizeof(SolveState),
                                     hipMemcpyDeviceToHost, m.s),
                      "slot state D2H");
            hip_check(hipEventRecord(m.done_ev[ju]->get(), m.s), "event record");
            hip_check(hipGetLastError(), "slot finish launch");
            m.phase[ju] = SLOT_FINISHING;
            m.running.erase(std::find(m.running.begin(), m.running.end(), j));
            m.dirty = true;
        }
    }
    // 3. Report finished solves whose state has reached the host (wait when nothing else is
    // left to drive the device).
    const bool idle = m.running.empty() && m.inflight.empty();
    for (int j = 0; j < m.k; ++j) {
        const std::size_t ju = static_cast<std::size_t>(j);
        if (m.phase[ju] != SLOT_FINISHING) {
            continue;
        }
        if (idle) {
            spin_wait(m.done_ev[ju]->get(), "event sync");
        } else {
            const hipError_t q = hipEventQuery(m.done_ev[ju]->get());
            if (q == hipErrorNotReady) {
                continue;
            }
            hip_check(q, "event query");
        }
        const SolveState& st = m.out_host[ju];
        Done r{};
        r.slot = j;
        r.stats.iterations = std::min(st.it, st.max_it);
        r.stats.converged = st.status == ST_CONVERGED ? 1 : 0;
        r.stats.on_device = 1;
        r.stats.exit_reason = st.status;
        r.stats.residual = st.res;
        r.stats.true_residual = st.true_res;
        r.stats.rhs_norm = st.bnorm;
        r.stats.solve_ms = std::chrono::duration<double, std::milli>(
                               std::chrono::steady_clock::now() - m.t_start[ju])
                               .count();
        done.push_back(r);
        m.phase[ju] = SLOT_FREE;
        m.sys[ju] = nullptr;
    }
}

std::unique_ptr<Runner>& System::runner_slot(cfdg_method method, bool use_amg,
                                             cfdg_precision precision) {
    int k = 0;
    if (method == CFDG_METHOD_BICGSTAB) {
        k = 3;
    } else if (!use_amg) {
        k = 2;
    } else if (precision == CFDG_PREC_MIXED_BF16) {
        k = 4;
    } else if (precision != CFDG_PREC_FP64) {
        k = 1;
    }
    return runners_[k];
}

void solve_batch(const std::vector<System*>& systems, const BatchOptions& opt,
                 const std::vector<cfdg_solve_control>& control,
                 std::vector<cfdg_solve_stats>& stats, hipStream_t s) {
    require(!systems.empty(), "solve_batch: no systems");
    if (opt.use_amg && opt.method == CFDG_METHOD_CG && opt.precision == CFDG_PREC_MIXED_BF16) {
        for (System* x : systems) {
            x->ensure_bf16();
        }
    }
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

cfdg_solve_stats cfdg::solve_system(System& s, cfdg_method method, bool use_amg, cfdg_precision precision,
                                    const cfdg_solve_control& control) {
    check_control(&control);
    return solve_one(s, method, use_amg, precision, control);
}

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

extern "C" cfdg_status cfdg_system_assemble_pressure(cfdg_system* sys, const cfdg_fv* fv) {
    return guarded([&] {
        require(sys != nullptr && fv != nullptr, "assemble_pressure: null handle");
        const FvPressureArrays a = fv_pressure_arrays(fv);
        require(a.diag != nullptr, "assemble_pressure: no pressure equation assembled on this DeviceFv");
        require(a.n_cells == sys->sys->n() && a.n_internal_faces == sys->sys->n_faces(),
                "assemble_pressure: the system's pattern is not the DeviceFv mesh's");
        sys->sys->assemble_device(a.diag, a.upper, a.lower, a.source);
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
        // With use_amg and a hierarchy the reference runs AMG-preconditioned CG whatever the
        // method (device_solver.cpp:138-147, krylov_cuda.cu:659-667); otherwise the method's
        // Jacobi-preconditioned Krylov solver.
        const bool amg = use_amg != 0 && s.has_amg();
        *stats = solve_one(s, amg ? CFDG_METHOD_CG : method, amg, CFDG_PREC_FP64, *control);
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
        const bool bf16 = precision == CFDG_PREC_MIXED_BF16;
        // Column indices count 4 bytes per stored entry whatever their stored width (16-bit
        // offsets here), values their stored width; vectors once per launch.
        const double idx = 4.0;
        if (p.kind == PROBE_L0_SWEEP) {
            // matrix (value + column index) + x_in, b, w*dinv, x_out at the V-cycle precision;
            // the mixed sweep also reads r and q in FP64 for its two dot products, and the
            // BF16 operator its FP32 row scales and diagonal.
            if (bf16) {
                *bytes_per_launch = nnz * (2.0 + idx) + 2.0 * 4.0 * n + 4.0 * 4.0 * n + 2.0 * 8.0 * n;
            } else if (fp32) {
                *bytes_per_launch = nnz * (4.0 + idx) + 4.0 * 4.0 * n + 2.0 * 8.0 * n;
            } else {
                *bytes_per_launch = nnz * (8.0 + idx) + 4.0 * 8.0 * n;
            }
        } else if (p.kind == PROBE_OUTER_SPMV) {
            // z (V-cycle precision) and p_old in, p_new and q out
            *bytes_per_launch = nnz * (8.0 + idx) + (fp32 ? 4.0 : 8.0) * n + 3.0 * 8.0 * n;
        } else {
            *bytes_per_launch = 0.0;
        }
    });
}

namespace {

// y = A x for one matrix in any layout (layout timing).
template <class T>
__global__ void __launch_bounds__(k_block) k_spmv_view(EllView m, const T* __restrict__ x, T* __restrict__ y) {
    const int tpr = m.tpr;
    const int row = static_cast<int>(blockIdx.x) * (k_block / tpr) + static_cast<int>(threadIdx.x) / tpr;
    const int sub = static_cast<int>(threadIdx.x) & (tpr - 1);
    const T v = mat_row<T>(m, row, sub, x);
    if (sub == 0 && row < m.n_rows) {
        y[row] = v;
    }
}

} // namespace

extern "C" cfdg_status cfdg_set_mixed_smoother(int sweeps, double ratio) {
    return guarded([&] {
        require(sweeps >= 0 && sweeps <= 64, "mixed smoother: sweeps must be in [0, 64]");
        require(ratio > 1.0 || ratio == 0.0, "mixed smoother: ratio must exceed 1 (or be 0: reference weight)");
        g_mixed_sweeps.store(sweeps);
        g_mixed_ratio.store(ratio);
    });
}

extern "C" cfdg_status cfdg_layout_of(int n_rows, const int* row_ptr, int* lanes, int* csr) {
    return guarded([&] {
        require(n_rows >= 0 && row_ptr != nullptr && lanes != nullptr && csr != nullptr, "null argument");
        const Layout l = choose_layout(n_rows, row_ptr);
        *lanes = l.lanes;
        *csr = l.csr;
    });
}

extern "C" cfdg_status cfdg_set_layout_rules(const cfdg_layout_rule* rules, int count) {
    return guarded([&] {
        require(count >= 0 && (count == 0 || rules != nullptr), "layout rules: null table");
        set_layout_rules(rules, count);
    });
}

extern "C" cfdg_status cfdg_spmv_time(const cfdg_csr* a, int lanes, int csr, int fp32, int reps, double* avg_ms) {
    return guarded([&] {
        require(a != nullptr && avg_ms != nullptr && reps >= 1, "null argument");
        require(lanes >= 1 && lanes <= k_wave && (lanes & (lanes - 1)) == 0, "lanes must be a power of two <= 64");
        require(csr == 0 || csr == 1, "csr must be 0 or 1");
        require_device();
        const int cols = a->n_cols == 0 ? a->n_rows : a->n_cols;
        check_csr(*a, a->n_rows, cols, "spmv_time: invalid CSR matrix");
        Stream stream;
        hipStream_t s = stream.get();
        DevEll m;
        const Layout lay{lanes, csr};
        m.upload_csr(a->n_rows, cols, a->row_ptr, a->col, a->val, s, &lay);
        const std::vector<double> ones(static_cast<std::size_t>(cols), 1.0);
        const std::vector<float> ones32(static_cast<std::size_t>(cols), 1.0f);
        DevBuf<double> x64;
        DevBuf<double> y64;
        DevBuf<float> x32;
        DevBuf<float> y32;
        x64.upload(ones.data(), ones.size(), s);
        x32.upload(ones32.data(), ones32.size(), s);
        y64.resize(static_cast<std::size_t>(a->n_rows));
        y32.resize(static_cast<std::size_t>(a->n_rows));
        const int grid = ceil_div(static_cast<long long>(a->n_rows) * lanes, k_block);
        auto launch = [&] {
            if (grid == 0) {
                return;
            }
            if (fp32 != 0) {
                k_spmv_view<float><<<grid, k_block, 0, s>>>(m.view32(), x32.get(), y32.get());
            } else {
                k_spmv_view<double><<<grid, k_block, 0, s>>>(m.view64(), x64.get(), y64.get());
            }
        };
        for (int i = 0; i < 3; ++i) {
            launch();
        }
        hip_check(hipGetLastError(), "spmv launch");
        Event e0;
        Event e1;
        hip_check(hipEventRecord(e0.get(), s), "event record");
        for (int i = 0; i < reps; ++i) {
            launch();
        }
        hip_check(hipEventRecord(e1.get(), s), "event record");
        hip_check(hipEventSynchronize(e1.get()), "event sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0.get(), e1.get()), "event elapsed");
        *avg_ms = static_cast<double>(ms) / reps;
    });
}

struct cfdg_batch {
    std::vector<System*> systems;
    BatchOptions opt;
    Stream stream;
    std::unique_ptr<Runner> runner;
};

extern "C" cfdg_status cfdg_batch_create(cfdg_system* const* systems, int count, cfdg_method method,
                                         int use_amg, cfdg_precision precision, cfdg_batch** out) {
    return guarded([&] {
        require(out != nullptr && systems != nullptr && count >= 1, "batch: need at least one system");
        require(method == CFDG_METHOD_CG || method == CFDG_METHOD_BICGSTAB, "unknown method");
        *out = nullptr;
        auto b = std::make_unique<cfdg_batch>();
        for (int i = 0; i < count; ++i) {
            require(systems[i] != nullptr, "batch: null system");
            b->systems.push_back(systems[i]->sys.get());
        }
        b->opt.method = method;
        b->opt.use_amg = use_amg != 0 && method == CFDG_METHOD_CG;
        b->opt.precision = precision;
        *out = b.release();
    });
}

extern "C" cfdg_status cfdg_batch_destroy(cfdg_batch* batch) {
    return guarded([&] { delete batch; });
}

// Fills queued on the members' streams are completed (the batch runs on its own stream) and the
// runner is (re)built when the members or their hierarchies changed.
static void batch_prepare(cfdg_batch& batch) {
    for (System* s : batch.systems) {
        if (batch.opt.use_amg && batch.opt.precision == CFDG_PREC_MIXED_BF16) {
            s->ensure_bf16();
        }
        s->finish_pending();
    }
    if (!batch.runner || !batch.runner->matches(batch.systems, batch.opt)) {
        batch.runner.reset();
        batch.runner = std::make_unique<Runner>(batch.systems, batch.opt, batch.stream.get());
    }
}

extern "C" cfdg_status cfdg_batch_zero_x(cfdg_batch* batch) {
    return guarded([&] {
        require(batch != nullptr, "null argument");
        batch_prepare(*batch);
        batch->runner->zero_x();
        hip_check(hipGetLastError(), "batch zero x");
    });
}

extern "C" cfdg_status cfdg_batch_solve(cfdg_batch* batch, const cfdg_solve_control* controls,
                                        cfdg_solve_stats* stats) {
    return guarded([&] {
        require(batch != nullptr && controls != nullptr && stats != nullptr, "null argument");
        const std::size_t n = batch->systems.size();
        std::vector<cfdg_solve_control> c(controls, controls + n);
        for (const cfdg_solve_control& x : c) {
            check_control(&x);
        }
        batch_prepare(*batch);
        std::vector<cfdg_solve_stats> st;
        batch->runner->solve(c, st, nullptr);
        std::copy(st.begin(), st.end(), stats);
    });
}

extern "C" cfdg_status cfdg_system_normalised_residual(cfdg_system* sys, double* normalised, double* res,
                                                       double* norm) {
    return guarded([&] {
        require(sys != nullptr, "null argument");
        double o[3];
        sys->sys->residual_norms(o);
        if (normalised != nullptr) {
            *normalised = o[0] / (o[1] + 1e-20); // k_small (ldu_matrix.cpp:22, 65-68)
        }
        if (res != nullptr) {
            *res = o[0];
        }
        if (norm != nullptr) {
            *norm = o[1];
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

// ------------------------------------------------------------------------------------------
// Allocation-free CSR solve (no counterpart in the reference). The caller's device arrays are
// read in place; work vectors, block partials, the solve state, the descriptor and the launch
// table live in `scratch`, a static device array of fixed size, so the call allocates nothing.
// ------------------------------------------------------------------------------------------
namespace cfdg {
namespace {

constexpr std::size_t round_up(std::size_t a, std::size_t b) { return (a + b - 1) / b * b; }

constexpr int k_na_max_rows = CFDG_NOALLOC_MAX_ROWS;
constexpr int k_na_vectors = 12; // r p p2 q z dinv r0 v s t ph sh
constexpr std::size_t k_na_max_blocks = static_cast<std::size_t>(k_na_max_rows) / k_block;
static_assert(k_na_max_rows % k_block == 0, "whole blocks");
// Byte offsets of the scratch regions.
constexpr std::size_t k_na_state = 0;   // SolveState
constexpr std::size_t k_na_flag = 256;  // int: 1 when the CSR check failed
constexpr std::size_t k_na_ids = 512;   // int: system id 0 (finalize launches)
constexpr std::size_t k_na_desc = 1024; // SysDev
constexpr std::size_t k_na_blocks = round_up(k_na_desc + sizeof(SysDev), 256);
constexpr std::size_t k_na_partials = round_up(k_na_blocks + k_na_max_blocks * sizeof(int2), 256);
constexpr std::size_t k_na_vecs =
    round_up(k_na_partials + static_cast<std::size_t>(k_slots) * k_na_max_blocks * sizeof(double), 256);
constexpr std::size_t k_na_bytes =
    k_na_vecs + static_cast<std::size_t>(k_na_vectors) * static_cast<std::size_t>(k_na_max_rows) * sizeof(double);
static_assert(k_na_bytes <= CFDG_NOALLOC_SCRATCH_BYTES, "scratch size stated in cfdgpu.h");
static_assert(sizeof(SolveState) <= k_na_flag - k_na_state, "state region");

// Checks a caller's CSR matrix in place: row offsets from 0 to nnz and non-decreasing, columns
// in [0, n). Any violation stores 1 at *bad (every writer stores the same value). Thread i
// checks row offset i (i <= n) and column i (i < nnz).
__global__ void __launch_bounds__(k_block) k_na_check(int n, int nnz, const int* __restrict__ rp,
                                                      const int* __restrict__ col, int* __restrict__ bad) {
    const long long i = static_cast<long long>(blockIdx.x) * k_block + threadIdx.x;
    bool ok = true;
    if (i <= n) {
        const int a = rp[i];
        ok = a >= 0 && a <= nnz && (i != 0 || a == 0) && (i != n || a == nnz) && (i == n || rp[i + 1] >= a);
    }
    if (i < nnz) {
        const int c = col[i];
        ok = ok && c >= 0 && c < n;
    }
    if (!ok) {
        *bad = 1;
    }
}

// Launch table (system 0, block b), system id 0, and the inverse diagonal of the checked matrix
// (1 / a_ii, 1.0 for a zero or missing diagonal, the row's last nonzero diagonal entry winning,
// as the host conversion of the one-shot solve; krylov_openmp.cpp:147-157).
__global__ void __launch_bounds__(k_block) k_na_setup(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                                      const double* __restrict__ val, int nblk, int2* __restrict__ blocks,
                                                      int* __restrict__ ids, double* __restrict__ dinv) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i < nblk) {
        blocks[i] = make_int2(0, i);
    }
    if (i == 0) {
        ids[0] = 0;
    }
    if (i < n) {
        double d = 1.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (col[k] == i && val[k] != 0.0) {
                d = 1.0 / val[k];
            }
        }
        dinv[i] = d;
    }
}

constexpr int k_na_max_devices = 64;
std::mutex g_na_mutex;                          // one allocation-free solve at a time per process
void* g_na_scratch[k_na_max_devices] = {};      // device address of `scratch` per device

} // namespace
} // namespace cfdg

// The scratch area of cfdg_solve_csr_noalloc: one copy per device, held until the process ends.
// Host code never passes this name to a kernel (on the host it is only a stand-in); it uses the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[CFDG_NOALLOC_SCRATCH_BYTES];

namespace cfdg {
namespace {

// Device address of `scratch` on device `dev` (current), looked up once per device.
unsigned char* na_scratch(int dev) {
    require(dev >= 0 && dev < k_na_max_devices, "solve_csr_noalloc: device index out of range");
    if (g_na_scratch[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        require(p != nullptr, "solve_csr_noalloc: no device address for the scratch area");
        g_na_scratch[dev] = p;
    }
    return static_cast<unsigned char*>(g_na_scratch[dev]);
}

template <int Op>
void na_fin(const SysDev* d, const int* ids) {
    k_finalize<Op><<<1, k_fin_block, 0, nullptr>>>(d, ids, nullptr);
}

int na_status(const SolveState* st) {
    int status = 0;
    hip_check(hipMemcpy(&status, &st->status, sizeof(int), hipMemcpyDeviceToHost), "D2H status");
    return status;
}

} // namespace
} // namespace cfdg

extern "C" cfdg_status cfdg_solve_csr_noalloc(const cfdg_csr* a, const double* b, double* x,
                                              cfdg_method method, const cfdg_solve_control* control,
                                              cfdg_solve_stats* stats) {
    return guarded([&] {
        require(a != nullptr && b != nullptr && x != nullptr && stats != nullptr, "null argument");
        check_control(control);
        require(method == CFDG_METHOD_CG || method == CFDG_METHOD_BICGSTAB, "unknown method");
        require(a->n_rows >= 1 && (a->n_cols == 0 || a->n_cols == a->n_rows),
                "CsrView: invalid square CSR matrix");
        require(a->nnz >= 0 && a->row_ptr != nullptr && (a->nnz == 0 || (a->col != nullptr && a->val != nullptr)),
                "CsrView: invalid square CSR matrix");
        if (a->n_rows > k_na_max_rows) {
            throw CapacityError("solve_csr_noalloc: " + std::to_string(a->n_rows) + " rows exceed the maximum of " +
                                std::to_string(k_na_max_rows) + " (fixed scratch area)");
        }
        require_device();
        const auto t_start = std::chrono::steady_clock::now();
        std::lock_guard<std::mutex> lock(g_na_mutex);
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        unsigned char* base = na_scratch(dev);
        const int n = a->n_rows;
        const int nnz = a->nnz;
        const int nblk = ceil_div(n, k_block);
        const std::size_t stride = round_up(static_cast<std::size_t>(n), k_wave) * sizeof(double);
        auto vec = [&](int i) { return reinterpret_cast<double*>(base + k_na_vecs + static_cast<std::size_t>(i) * stride); };
        auto* st = reinterpret_cast<SolveState*>(base + k_na_state);
        auto* flag = reinterpret_cast<int*>(base + k_na_flag);
        auto* ids = reinterpret_cast<int*>(base + k_na_ids);
        auto* desc = reinterpret_cast<SysDev*>(base + k_na_desc);
        auto* blocks = reinterpret_cast<int2*>(base + k_na_blocks);
        auto* partials = reinterpret_cast<double*>(base + k_na_partials);

        // The matrix is checked on the device before any kernel gathers through it.
        const int zero = 0;
        hip_check(hipMemcpy(flag, &zero, sizeof(int), hipMemcpyHostToDevice), "H2D flag");
        const long long checks = std::max<long long>(static_cast<long long>(n) + 1, nnz);
        k_na_check<<<ceil_div(checks, k_block), k_block, 0, nullptr>>>(n, nnz, a->row_ptr, a->col, flag);
        hip_check(hipGetLastError(), "k_na_check launch");
        int bad = 0;
        hip_check(hipMemcpy(&bad, flag, sizeof(int), hipMemcpyDeviceToHost), "D2H flag");
        require(bad == 0, "CsrView: invalid square CSR matrix (device check)");

        SysDev d{};
        d.st = st;
        d.partials = partials;
        d.nblk0 = nblk;
        d.n = n;
        d.n_levels = 0;
        d.flip = 0;
        d.a64.n_rows = n;
        d.a64.tpr = 1;
        d.a64.csr = 1;
        d.a64.sp = a->row_ptr;
        d.a64.col = a->col;
        d.a64.val = a->val;
        d.b = b;
        d.x = x;
        d.r = vec(0);
        d.p = vec(1);
        d.p2 = vec(2);
        d.q = vec(3);
        d.z = vec(4);
        double* dinv = vec(5);
        d.dinv = dinv;
        d.r0 = vec(6);
        d.v = vec(7);
        d.s = vec(8);
        d.t = vec(9);
        d.ph = vec(10);
        d.sh = vec(11);
        SolveState s0{};
        s0.tol = control->abs_tolerance;
        s0.max_it = control->max_iterations;
        s0.status = ST_RUNNING;
        hip_check(hipMemcpy(desc, &d, sizeof(SysDev), hipMemcpyHostToDevice), "H2D descriptor");
        hip_check(hipMemcpy(st, &s0, sizeof(SolveState), hipMemcpyHostToDevice), "H2D state");
        k_na_setup<<<nblk, k_block, 0, nullptr>>>(n, a->row_ptr, a->col, a->val, nblk, blocks, ids, dinv);
        hip_check(hipGetLastError(), "k_na_setup launch");

        // The solver's own kernels over the CSR arrays, one iteration at a time; the host reads
        // the status after each iteration (surplus iterations would be no-ops on the device).
        const int max_it = control->max_iterations;
        if (method == CFDG_METHOD_CG) {
            k_cg_init<true><<<nblk, k_block, 0, nullptr>>>(desc, blocks);
            na_fin<FIN_CG_INIT>(desc, ids);
        } else {
            k_bicg_init<true><<<nblk, k_block, 0, nullptr>>>(desc, blocks);
            na_fin<FIN_BICG_INIT>(desc, ids);
        }
        hip_check(hipGetLastError(), "init launch");
        for (int k = 1; k <= max_it + 2; ++k) {
            const int parity = k & 1;
            if (method == CFDG_METHOD_CG) {
                k_spmv_pz<double, true><<<nblk, k_block, 0, nullptr>>>(desc, blocks, parity, 0, 0);
                na_fin<FIN_ALPHA>(desc, ids);
                k_update_xr<1><<<nblk, k_block, 0, nullptr>>>(desc, blocks, parity);
                na_fin<FIN_CG_RES_BETA>(desc, ids);
            } else {
                k_bicg_p<<<nblk, k_block, 0, nullptr>>>(desc, blocks);
                k_spmv_dot<false, true><<<nblk, k_block, 0, nullptr>>>(desc, blocks, V_PH, V_V, V_R0, SLOT_R0V, 0, 0);
                na_fin<FIN_BICG_ALPHA>(desc, ids);
                k_bicg_s<<<nblk, k_block, 0, nullptr>>>(desc, blocks);
                na_fin<FIN_BICG_S>(desc, ids);
                k_bicg_early_x<<<nblk, k_block, 0, nullptr>>>(desc, blocks);
                k_spmv_dot<true, true><<<nblk, k_block, 0, nullptr>>>(desc, blocks, V_SH, V_T, V_T, SLOT_TT, V_S, SLOT_TS);
                na_fin<FIN_BICG_OMEGA>(desc, ids);
                k_bicg_xr<<<nblk, k_block, 0, nullptr>>>(desc, blocks);
                na_fin<FIN_BICG_RES>(desc, ids);
            }
            hip_check(hipGetLastError(), "iteration launch");
            const int status = na_status(st);
            if (status != ST_RUNNING && status != ST_CONVERGED_PENDING) {
                break;
            }
        }
        k_true_residual<true><<<nblk, k_block, 0, nullptr>>>(desc, blocks);
        na_fin<FIN_TRUE>(desc, ids);
        hip_check(hipGetLastError(), "true residual launch");
        SolveState out{};
        hip_check(hipMemcpy(&out, st, sizeof(SolveState), hipMemcpyDeviceToHost), "D2H state");
        *stats = cfdg_solve_stats{};
        stats->iterations = std::min(out.it, out.max_it);
        stats->converged = out.status == ST_CONVERGED ? 1 : 0;
        stats->on_device = 1;
        stats->exit_reason = out.status;
        stats->residual = out.res;
        stats->true_residual = out.true_res;
        stats->rhs_norm = out.bnorm;
        stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
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
        s.set_amg(*amg, false); // one solve: the fused products would not pay off
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
