// Batched dense inversion of the AMG coarsest-level operators on the matrix cores (gfx950
// v_mfma_f64_16x16x4f64). The reference computes the coarsest inverse on the host, n_c LDL^T
// solves (src/fv/src/amg.cpp:185-189, 227-240); here every matrix of a batch is inverted by one
// workgroup with in-place block Gauss-Jordan elimination in 16 x 16 tiles: the 16 x 16 pivot
// block is inverted in LDS, the pivot row and column panels are multiplied by it and the
// trailing matrix takes the rank-16 update, each 16 x 16 x 16 product as four FP64 MFMAs. The
// elimination has no pivoting, as LDL^T has none: its pivots are LDL^T's (in exact arithmetic),
// and a zero or non-finite pivot marks the matrix for the host path (which then falls back to
// the pseudo-inverse, amg.cpp:227-240).
#include "dense.hpp"
#include "hip_util.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace cfdg {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kT = 16;         // tile edge
constexpr int kThreads = 256;  // four wavefronts; thread t holds pivot-block entry (t / 16, t % 16)
constexpr int kWaves = kThreads / 64;

struct InvJob {
    int n;            // order
    int N;            // order padded to a multiple of 16 (identity in the padding)
    const double* a;  // n x n row-major input
    double* w;        // N x N row-major workspace
    double* inv;      // n x n row-major result
};

// Operands of v_mfma_f64_16x16x4f64 for lane l = 16 q + p (p < 16, q < 4): the A operand is
// element (p, q) of a 16 x 4 block, the B operand element (q, p) of a 4 x 16 block, and the four
// accumulators are elements (4 r + q, p), r < 4, of the 16 x 16 result (unlike the FP32
// 16 x 16 products, whose accumulators hold rows 4 q + r).
__device__ __forceinline__ f64x4 mma(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__global__ void __launch_bounds__(kThreads) k_dense_inverse(const InvJob* __restrict__ jobs, int* __restrict__ status) {
    const InvJob J = jobs[blockIdx.x];
    const int n = J.n;
    const int N = J.N;
    const int nb = N / kT;
    double* __restrict__ W = J.w;
    const int t = static_cast<int>(threadIdx.x);
    const int lane = t % 64;
    const int wave = t / 64;
    const int p = lane % 16;
    const int q = lane / 16;
    __shared__ double dinv[kT * kT];
    __shared__ double rowp[kDenseMaxOrder * kT]; // block row k after the pivot-row update
    __shared__ double colp[kDenseMaxOrder * kT]; // block column k before its update
    __shared__ int bad;
    if (t == 0) {
        bad = 0;
    }
    for (long long e = t; e < static_cast<long long>(N) * N; e += kThreads) {
        const int i = static_cast<int>(e / N);
        const int j = static_cast<int>(e % N);
        W[e] = (i < n && j < n) ? J.a[static_cast<long long>(i) * n + j] : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    const int di = t / kT;
    const int dj = t % kT;
    for (int k = 0; k < nb; ++k) {
        const int k0 = k * kT;
        // Pivot block to LDS and the old block column k to LDS.
        dinv[t] = W[static_cast<long long>(k0 + di) * N + k0 + dj];
        for (int e = t; e < N * kT; e += kThreads) {
            const int i = e / kT; // row of the matrix
            colp[e] = W[static_cast<long long>(i) * N + k0 + e % kT];
        }
        __syncthreads();
        // Gauss-Jordan inversion of the pivot block, one entry per thread.
        for (int s = 0; s < kT; ++s) {
            const double piv = dinv[s * kT + s];
            const double m_is = dinv[di * kT + s];
            const double m_sj = dinv[s * kT + dj];
            const double m_ij = dinv[t];
            __syncthreads();
            if (t == 0 && (piv == 0.0 || !isfinite(piv))) {
                bad = 1;
            }
            double v;
            if (di == s) {
                v = dj == s ? 1.0 / piv : m_sj / piv;
            } else {
                v = dj == s ? -m_is / piv : m_ij - m_is * (m_sj / piv);
            }
            dinv[t] = v;
            __syncthreads();
        }
        // Pivot row: W[k][j] = P^-1 W[k][j] (j != k), kept in LDS for the trailing update.
        for (int j = wave; j < nb; j += kWaves) {
            if (j == k) {
                continue;
            }
            const int j0 = j * kT;
            f64x4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < kT / 4; ++kk) {
                const double a = dinv[p * kT + 4 * kk + q];
                const double b = W[static_cast<long long>(k0 + 4 * kk + q) * N + j0 + p];
                c = mma(a, b, c);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 4 * r + q;
                rowp[j0 * kT + row * kT + p] = c[r];
                W[static_cast<long long>(k0 + row) * N + j0 + p] = c[r];
            }
        }
        __syncthreads();
        // Trailing update W[i][j] -= W_old[i][k] W_new[k][j] (i, j != k), and the pivot column
        // W[i][k] = -W_old[i][k] P^-1 (i != k), the tiles shared out over the wavefronts.
        const int tiles = nb * nb;
        for (int tile = wave; tile < tiles; tile += kWaves) {
            const int i = tile / nb;
            const int j = tile % nb;
            if (i == k) {
                continue;
            }
            const int i0 = i * kT;
            const int j0 = j * kT;
            f64x4 c;
            if (j == k) {
                c = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < kT / 4; ++kk) {
                    const double a = -colp[(i0 + p) * kT + 4 * kk + q];
                    const double b = dinv[(4 * kk + q) * kT + p];
                    c = mma(a, b, c);
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    c[r] = W[static_cast<long long>(i0 + 4 * r + q) * N + j0 + p];
                }
#pragma unroll
                for (int kk = 0; kk < kT / 4; ++kk) {
                    const double a = -colp[(i0 + p) * kT + 4 * kk + q];
                    const double b = rowp[j0 * kT + (4 * kk + q) * kT + p];
                    c = mma(a, b, c);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                W[static_cast<long long>(i0 + 4 * r + q) * N + j0 + p] = c[r];
            }
        }
        W[static_cast<long long>(k0 + di) * N + k0 + dj] = dinv[t];
        __syncthreads();
    }
    for (long long e = t; e < static_cast<long long>(n) * n; e += kThreads) {
        const int i = static_cast<int>(e / n);
        const int j = static_cast<int>(e % n);
        J.inv[e] = W[static_cast<long long>(i) * N + j];
    }
    if (t == 0) {
        status[blockIdx.x] = bad;
    }
}

// Orders above kDenseMaxOrder: the same block Gauss-Jordan spread over the device, two launches
// per 16-row block step k. k_gj_row (one wavefront per block column j): inverts the pivot block
// in LDS (every workgroup the same way, so every copy is identical), forms the new pivot row tile
// W[k][j] = P^-1 W[k][j] and saves the old pivot column tile W[j][k] to colp; workgroup k keeps
// P^-1 in pinv (W[k][k] itself must stay the pivot block while the others read it).
// k_gj_update (one wavefront per tile (i, j), i != k): W[i][j] -= colp[i] W[k][j],
// W[i][k] = -colp[i] P^-1, and one wavefront stores W[k][k] = P^-1. Within a launch only the
// wavefront of a tile writes it, and what a launch reads of other tiles is not written by it.
__device__ __forceinline__ void gj_pivot_inverse(double* __restrict__ P, int lane, int* __restrict__ bad) {
    // P: 16 x 16 in LDS, lane holds entries lane + 64 e, e < 4 (row (lane + 64 e) / 16).
    for (int s = 0; s < kT; ++s) {
        const double piv = P[s * kT + s];
        double v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int t = lane + 64 * e;
            const int di = t / kT;
            const int dj = t % kT;
            const double m_is = P[di * kT + s];
            const double m_sj = P[s * kT + dj];
            const double m_ij = P[t];
            if (di == s) {
                v[e] = dj == s ? 1.0 / piv : m_sj / piv;
            } else {
                v[e] = dj == s ? -m_is / piv : m_ij - m_is * (m_sj / piv);
            }
        }
        __syncthreads();
        if (lane == 0 && (piv == 0.0 || !isfinite(piv))) {
            *bad = 1;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            P[lane + 64 * e] = v[e];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(64) k_gj_row(double* __restrict__ W, int N, int k, double* __restrict__ colp,
                                               double* __restrict__ pinv, int* __restrict__ bad) {
    const int j = static_cast<int>(blockIdx.x);
    const int lane = static_cast<int>(threadIdx.x);
    const int p = lane % 16;
    const int q = lane / 16;
    const int k0 = k * kT;
    const int j0 = j * kT;
    __shared__ double P[kT * kT];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int t = lane + 64 * e;
        P[t] = W[static_cast<long long>(k0 + t / kT) * N + k0 + t % kT];
    }
    __syncthreads();
    gj_pivot_inverse(P, lane, bad);
    if (j == k) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            pinv[lane + 64 * e] = P[lane + 64 * e];
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) { // the old pivot column tile (j, k), before the update launch
        const int t = lane + 64 * e;
        colp[static_cast<long long>(j0 + t / kT) * kT + t % kT] = W[static_cast<long long>(j0 + t / kT) * N + k0 + t % kT];
    }
    f64x4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < kT / 4; ++kk) {
        const double a = P[p * kT + 4 * kk + q];
        const double b = W[static_cast<long long>(k0 + 4 * kk + q) * N + j0 + p];
        c = mma(a, b, c);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        W[static_cast<long long>(k0 + 4 * r + q) * N + j0 + p] = c[r];
    }
}

__global__ void __launch_bounds__(kThreads) k_gj_update(double* __restrict__ W, int N, int k,
                                                        const double* __restrict__ colp,
                                                        const double* __restrict__ pinv) {
    const int nb = N / kT;
    const int lane = static_cast<int>(threadIdx.x) % 64;
    const int wave = static_cast<int>(threadIdx.x) / 64;
    const long long tile = static_cast<long long>(blockIdx.x) * kWaves + wave;
    const int k0 = k * kT;
    if (tile == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int t = lane + 64 * e;
            W[static_cast<long long>(k0 + t / kT) * N + k0 + t % kT] = pinv[t];
        }
    }
    if (tile >= static_cast<long long>(nb - 1) * nb) {
        return;
    }
    const int ii = static_cast<int>(tile / nb);
    const int i = ii < k ? ii : ii + 1; // every block row but the pivot row
    const int j = static_cast<int>(tile % nb);
    const int p = lane % 16;
    const int q = lane / 16;
    const int i0 = i * kT;
    const int j0 = j * kT;
    f64x4 c;
    if (j == k) {
        c = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < kT / 4; ++kk) {
            const double a = -colp[static_cast<long long>(i0 + p) * kT + 4 * kk + q];
            const double b = pinv[(4 * kk + q) * kT + p];
            c = mma(a, b, c);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c[r] = W[static_cast<long long>(i0 + 4 * r + q) * N + j0 + p];
        }
#pragma unroll
        for (int kk = 0; kk < kT / 4; ++kk) {
            const double a = -colp[static_cast<long long>(i0 + p) * kT + 4 * kk + q];
            const double b = W[static_cast<long long>(k0 + 4 * kk + q) * N + j0 + p]; // new pivot row
            c = mma(a, b, c);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        W[static_cast<long long>(i0 + 4 * r + q) * N + j0 + p] = c[r];
    }
}

__global__ void k_gj_load(const double* __restrict__ a, int n, int N, double* __restrict__ W) {
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<long long>(N) * N) {
        return;
    }
    const int i = static_cast<int>(e / N);
    const int j = static_cast<int>(e % N);
    W[e] = (i < n && j < n) ? a[static_cast<long long>(i) * n + j] : (i == j ? 1.0 : 0.0);
}

__global__ void k_gj_store(const double* __restrict__ W, int n, int N, double* __restrict__ inv) {
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<long long>(n) * n) {
        return;
    }
    inv[e] = W[static_cast<long long>(e / n) * N + e % n];
}

// One matrix of order kDenseMaxOrder < n <= kDenseWideMaxOrder; true when the device inverted it.
bool wide_inverse(int n, const double* host_a, double* host_inv, hipStream_t s) {
    const int N = (n + kT - 1) / kT * kT;
    const int nb = N / kT;
    const std::size_t nn = static_cast<std::size_t>(n) * static_cast<std::size_t>(n);
    DevBuf<double> a(nn);
    DevBuf<double> W(static_cast<std::size_t>(N) * static_cast<std::size_t>(N));
    DevBuf<double> colp(static_cast<std::size_t>(N) * kT);
    DevBuf<double> pinv(static_cast<std::size_t>(kT) * kT);
    DevBuf<int> bad(1);
    a.upload(host_a, nn, s);
    bad.zero(s);
    const long long cells = static_cast<long long>(N) * N;
    k_gj_load<<<static_cast<unsigned>((cells + 255) / 256), 256, 0, s>>>(a.get(), n, N, W.get());
    const long long tiles = static_cast<long long>(nb - 1) * nb;
    for (int k = 0; k < nb; ++k) {
        k_gj_row<<<static_cast<unsigned>(nb), 64, 0, s>>>(W.get(), N, k, colp.get(), pinv.get(), bad.get());
        const long long blocks = tiles > 0 ? (tiles + kWaves - 1) / kWaves : 1; // at least the W[k][k] store
        k_gj_update<<<static_cast<unsigned>(blocks), kThreads, 0, s>>>(W.get(), N, k, colp.get(), pinv.get());
    }
    hip_check(hipGetLastError(), "Gauss-Jordan launches");
    k_gj_store<<<static_cast<unsigned>((static_cast<long long>(nn) + 255) / 256), 256, 0, s>>>(W.get(), n, N, a.get());
    hip_check(hipGetLastError(), "Gauss-Jordan store");
    int flag = 1;
    hip_check(hipMemcpyAsync(&flag, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s), "D2H flag");
    d2h_staged(host_inv, a.get(), nn * sizeof(double), s);
    hip_check(hipStreamSynchronize(s), "Gauss-Jordan sync");
    return flag == 0;
}

} // namespace

std::vector<int> dense_inverse_batch(const std::vector<int>& order, const std::vector<const double*>& mats,
                                     const std::vector<double*>& out, hipStream_t s) {
    const std::size_t count = order.size();
    require(mats.size() == count && out.size() == count, "dense inverse: one matrix and output per order");
    std::vector<int> done(count, 0);
    std::vector<std::size_t> dev_ids;
    std::size_t a_total = 0;
    std::size_t w_total = 0;
    for (std::size_t i = 0; i < count; ++i) {
        const int n = order[i];
        require(n >= 0, "dense inverse: negative order");
        if (n > kDenseMaxOrder && n <= kDenseWideMaxOrder) {
            done[i] = wide_inverse(n, mats[i], out[i], s) ? 1 : 0;
            continue;
        }
        if (n >= 1 && n <= kDenseMaxOrder) {
            dev_ids.push_back(i);
            const std::size_t nn = static_cast<std::size_t>(n) * static_cast<std::size_t>(n);
            const std::size_t pad = static_cast<std::size_t>((n + kT - 1) / kT * kT);
            a_total += nn;
            w_total += pad * pad;
        }
    }
    if (dev_ids.empty()) {
        return done;
    }
    DevBuf<double> a(a_total);
    DevBuf<double> inv(a_total);
    DevBuf<double> w(w_total);
    DevBuf<InvJob> jobs(dev_ids.size());
    DevBuf<int> status(dev_ids.size());
    std::vector<InvJob> host_jobs(dev_ids.size());
    std::size_t ao = 0;
    std::size_t wo = 0;
    for (std::size_t k = 0; k < dev_ids.size(); ++k) {
        const std::size_t i = dev_ids[k];
        const int n = order[i];
        const int pad = (n + kT - 1) / kT * kT;
        const std::size_t nn = static_cast<std::size_t>(n) * static_cast<std::size_t>(n);
        hip_check(hipMemcpyAsync(a.get() + ao, mats[i], nn * sizeof(double), hipMemcpyHostToDevice, s), "H2D matrix");
        host_jobs[k] = InvJob{n, pad, a.get() + ao, w.get() + wo, inv.get() + ao};
        ao += nn;
        wo += static_cast<std::size_t>(pad) * static_cast<std::size_t>(pad);
    }
    jobs.upload(host_jobs.data(), host_jobs.size(), s);
    k_dense_inverse<<<static_cast<unsigned>(dev_ids.size()), kThreads, 0, s>>>(jobs.get(), status.get());
    hip_check(hipGetLastError(), "k_dense_inverse launch");
    std::vector<int> st(dev_ids.size(), 1);
    hip_check(hipMemcpyAsync(st.data(), status.get(), st.size() * sizeof(int), hipMemcpyDeviceToHost, s), "D2H status");
    ao = 0;
    for (std::size_t k = 0; k < dev_ids.size(); ++k) {
        const std::size_t i = dev_ids[k];
        const std::size_t nn = static_cast<std::size_t>(order[i]) * static_cast<std::size_t>(order[i]);
        hip_check(hipMemcpyAsync(out[i], inv.get() + ao, nn * sizeof(double), hipMemcpyDeviceToHost, s), "D2H inverse");
        ao += nn;
    }
    hip_check(hipStreamSynchronize(s), "dense inverse sync");
    for (std::size_t k = 0; k < dev_ids.size(); ++k) {
        done[dev_ids[k]] = st[k] == 0 ? 1 : 0;
    }
    return done;
}

} // namespace cfdg
