// Sparse products of the smoothed-aggregation setup on the device: the smoothed prolongation
// P = P_tent - omega D^-1 A P_tent (pruned), R = P^T and the Galerkin operator R (A P) (pruned)
// of fv::AmgPreconditioner::setup (src/fv/src/amg.cpp:150-180), with the aggregation itself on
// the host (amg_setup.cpp), since the reference's greedy aggregation is sequential by definition.
//
// Every product C = A B is formed the way the host setup forms it (and Eigen's conservative
// product before it): entry (i, j) is the sum of a_ik b_kj over k in the storage order of row i
// of A, starting from the first product, explicit zeros kept. On the device the products of a
// row are written in that order with the key (i, j), a stable radix sort by key brings equal keys
// together without reordering them, and one thread per key adds its run in order; small levels
// whose rows have long product lists (stalled coarsening) accumulate each row densely instead,
// exactly as the host does. The result is bit-identical to the host products for any matrix, so
// the pruned patterns are identical too.
// The operator of the current level stays on the device from one level to the next.
#include "amg_device.hpp"
#include "hip_util.hpp"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <vector>

namespace cfdg {
namespace {

constexpr int k_blk = 256;

int blocks_for(long long n) { return static_cast<int>((n + k_blk - 1) / k_blk); }

int bits_for(long long n) { // bits of the largest value in [0, n)
    int b = 1;
    while ((1LL << b) < n) {
        ++b;
    }
    return b;
}

using Key = unsigned long long;

__device__ __forceinline__ Key make_key(int row, int col) {
    return (static_cast<Key>(static_cast<unsigned>(row)) << 32) | static_cast<unsigned>(col);
}

// Number of products of each row of A B.
__global__ void k_count_products(int n, const int* __restrict__ arp, const int* __restrict__ aci,
                                 const int* __restrict__ brp, long long* __restrict__ cnt) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n) {
        return;
    }
    long long c = 0;
    for (int ka = arp[i]; ka < arp[i + 1]; ++ka) {
        const int k = aci[ka];
        c += brp[k + 1] - brp[k];
    }
    cnt[i] = c;
}

// The products of each row in the order of the host setup: b_kj * y, y = lscale_i * a_ik (or a_ik).
__global__ void k_emit_products(int n, const int* __restrict__ arp, const int* __restrict__ aci,
                                const double* __restrict__ ava, const int* __restrict__ brp,
                                const int* __restrict__ bci, const double* __restrict__ bva,
                                const double* __restrict__ lscale, const long long* __restrict__ off,
                                Key* __restrict__ key, double* __restrict__ val) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n) {
        return;
    }
    long long q = off[i];
    const double s = lscale != nullptr ? lscale[i] : 1.0;
    for (int ka = arp[i]; ka < arp[i + 1]; ++ka) {
        const int k = aci[ka];
        const double y = lscale != nullptr ? s * ava[ka] : ava[ka];
        for (int kb = brp[k]; kb < brp[k + 1]; ++kb) {
            key[q] = make_key(i, bci[kb]);
            val[q] = bva[kb] * y;
            ++q;
        }
    }
}

__global__ void k_heads(long long n, const Key* __restrict__ key, int* __restrict__ head) {
    const long long q = static_cast<long long>(blockIdx.x) * k_blk + threadIdx.x;
    if (q < n) {
        head[q] = (q == 0 || key[q] != key[q - 1]) ? 1 : 0;
    }
}

// run[q] = number of heads up to q (inclusive): the run of item q is run[q] - 1.
__global__ void k_run_starts(long long n, const int* __restrict__ head, const int* __restrict__ run,
                             long long* __restrict__ start) {
    const long long q = static_cast<long long>(blockIdx.x) * k_blk + threadIdx.x;
    if (q < n && head[q] != 0) {
        start[run[q] - 1] = q;
    }
}

// One thread per run of equal keys: its values added in order, starting from the first.
__global__ void k_run_sums(int runs, long long n, const long long* __restrict__ start, const Key* __restrict__ key,
                           const double* __restrict__ val, Key* __restrict__ okey, double* __restrict__ oval) {
    const int r = blockIdx.x * k_blk + threadIdx.x;
    if (r >= runs) {
        return;
    }
    const long long a = start[r];
    const long long b = r + 1 < runs ? start[r + 1] : n;
    double s = val[a];
    for (long long q = a + 1; q < b; ++q) {
        s += val[q];
    }
    okey[r] = key[a];
    oval[r] = s;
}

// Row offsets of a key-sorted entry list: rp[i] = first entry with row >= i.
__global__ void k_row_ptr(int n_rows, int n, const Key* __restrict__ key, int* __restrict__ rp) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i > n_rows) {
        return;
    }
    int lo = 0;
    int hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (static_cast<int>(key[mid] >> 32) < i) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    rp[i] = lo;
}

__global__ void k_split_keys(int n, const Key* __restrict__ key, int* __restrict__ col) {
    const int q = blockIdx.x * k_blk + threadIdx.x;
    if (q < n) {
        col[q] = static_cast<int>(key[q] & 0xffffffffULL);
    }
}

// prune(0.0): keep[q] = !(|v| <= 0).
__global__ void k_keep(int n, const double* __restrict__ val, int* __restrict__ keep) {
    const int q = blockIdx.x * k_blk + threadIdx.x;
    if (q < n) {
        keep[q] = !(fabs(val[q]) <= 0.0) ? 1 : 0;
    }
}

__global__ void k_compact(int n, const int* __restrict__ keep, const int* __restrict__ pos, const Key* __restrict__ key,
                          const double* __restrict__ val, Key* __restrict__ okey, double* __restrict__ oval) {
    const int q = blockIdx.x * k_blk + threadIdx.x;
    if (q < n && keep[q] != 0) {
        okey[pos[q] - 1] = key[q];
        oval[pos[q] - 1] = val[q];
    }
}

// Entries of P = P_tent - omega S over the union pattern of each row, pruned (the host setup's
// merge, amg_setup.cpp): count (out == nullptr) or write.
__device__ int merge_row(int i, int ja, const int* srp, const int* sci, const double* sva, double omega, Key* okey,
                         double* oval) {
    bool a_done = false;
    int m = 0;
    for (int k = srp[i]; k < srp[i + 1]; ++k) {
        const int j = sci[k];
        const double ws = omega * sva[k];
        if (!a_done && ja < j) {
            if (okey != nullptr) {
                okey[m] = make_key(i, ja);
                oval[m] = 1.0 - 0.0;
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
        if (!(fabs(v) <= 0.0)) {
            if (okey != nullptr) {
                okey[m] = make_key(i, j);
                oval[m] = v;
            }
            ++m;
        }
    }
    if (!a_done) {
        if (okey != nullptr) {
            okey[m] = make_key(i, ja);
            oval[m] = 1.0;
        }
        ++m;
    }
    return m;
}

__global__ void k_prolong_count(int n, const int* __restrict__ agg, const int* __restrict__ srp, const int* __restrict__ sci,
                                const double* __restrict__ sva, double omega, long long* __restrict__ cnt) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i < n) {
        cnt[i] = merge_row(i, agg[i], srp, sci, sva, omega, nullptr, nullptr);
    }
}

__global__ void k_prolong_write(int n, const int* __restrict__ agg, const int* __restrict__ srp, const int* __restrict__ sci,
                                const double* __restrict__ sva, double omega, const long long* __restrict__ off,
                                Key* __restrict__ key, double* __restrict__ val) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i < n) {
        merge_row(i, agg[i], srp, sci, sva, omega, key + off[i], val + off[i]);
    }
}

// Transposed keys (column, row) of every entry, in storage order.
__global__ void k_transpose_keys(int n_rows, const int* __restrict__ rp, const int* __restrict__ ci,
                                 Key* __restrict__ key) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n_rows) {
        return;
    }
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        key[k] = make_key(ci[k], i);
    }
}

__global__ void k_fill_iota(int n, int* __restrict__ a) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i <= n) {
        a[i] = i;
    }
}

__global__ void k_fill_one(int n, double* __restrict__ a) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i < n) {
        a[i] = 1.0;
    }
}

// Dense-row products for small levels with long product lists: one wavefront per row of A B,
// the row accumulated in LDS (a value and a first-touch mark per column; lane l owns the columns
// j = l mod 64). The row's products are taken 64 at a time in the host's order and applied one
// after the other, each by the lane that owns its column, so every column's sum has the host's
// operand order (first product, then +=). The wavefront then writes the row's values and marks
// to its scratch row (coalesced) and its count of columns. k_wave_dense = 64 threads per block.
constexpr int k_wave_dense = 64;
constexpr int k_dense_max_cols = 5400; // 12 bytes of LDS per column, within 64 KiB

__global__ void __launch_bounds__(k_wave_dense) k_dense_rows(int n, int n_cols, const int* __restrict__ arp,
                                                             const int* __restrict__ aci,
                                                             const double* __restrict__ ava,
                                                             const int* __restrict__ brp,
                                                             const int* __restrict__ bci,
                                                             const double* __restrict__ bva,
                                                             const double* __restrict__ lscale,
                                                             double* __restrict__ acc_all, int* __restrict__ mark_all,
                                                             long long* __restrict__ cnt) {
    extern __shared__ double lds[];
    double* acc = lds;
    int* mark = reinterpret_cast<int*>(lds + n_cols);
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= n) {
        return;
    }
    for (int j = lane; j < n_cols; j += k_wave_dense) {
        mark[j] = 0;
    }
    __syncthreads();
    const double s = lscale != nullptr ? lscale[i] : 1.0;
    for (int ka = arp[i]; ka < arp[i + 1]; ++ka) {
        const int k = aci[ka];
        const double y = lscale != nullptr ? s * ava[ka] : ava[ka];
        for (int kb0 = brp[k]; kb0 < brp[k + 1]; kb0 += k_wave_dense) {
            const int m = min(k_wave_dense, brp[k + 1] - kb0);
            int jl = 0;
            double pl = 0.0;
            if (lane < m) {
                jl = bci[kb0 + lane];
                pl = bva[kb0 + lane] * y;
            }
            for (int t = 0; t < m; ++t) { // in order; the owner lane of column j applies it
                const int j = __shfl(jl, t, k_wave_dense);
                const double prod = __shfl(pl, t, k_wave_dense);
                if ((j & (k_wave_dense - 1)) == lane) {
                    if (mark[j] == 0) {
                        mark[j] = 1;
                        acc[j] = prod;
                    } else {
                        acc[j] += prod;
                    }
                }
            }
        }
    }
    __syncthreads();
    double* acc_row = acc_all + static_cast<long long>(i) * n_cols;
    int* mark_row = mark_all + static_cast<long long>(i) * n_cols;
    long long c = 0;
    for (int j = lane; j < n_cols; j += k_wave_dense) {
        mark_row[j] = mark[j];
        acc_row[j] = mark[j] != 0 ? acc[j] : 0.0;
        c += mark[j];
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off, k_wave_dense);
    }
    if (lane == 0) {
        cnt[i] = c;
    }
}

// The marked columns of each scratch row, ascending, at the row's offset: one wavefront per
// row, 64 columns per step, positions by ballot.
__global__ void __launch_bounds__(k_wave_dense) k_dense_write(int n, int n_cols, const double* __restrict__ acc_all,
                                                              const int* __restrict__ mark_all,
                                                              const long long* __restrict__ off,
                                                              Key* __restrict__ key, double* __restrict__ val) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= n) {
        return;
    }
    const double* acc = acc_all + static_cast<long long>(i) * n_cols;
    const int* mark = mark_all + static_cast<long long>(i) * n_cols;
    long long q = off[i];
    const unsigned long long below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    for (int j0 = 0; j0 < n_cols; j0 += k_wave_dense) {
        const int j = j0 + lane;
        const bool on = j < n_cols && mark[j] != 0;
        const unsigned long long bits = __ballot(on);
        if (on) {
            const long long pos = q + __popcll(bits & below);
            key[pos] = make_key(i, j);
            val[pos] = acc[j];
        }
        q += __popcll(bits);
    }
}

// ---- Aggregation (amg.cpp:12-78), passes 1 and 2 ----

// Strength flags per entry, |a_ij| > theta sqrt(|d_i d_j|) with i != j (amg.cpp:16-18), the
// count of strong entries per row and, for the transposed pattern, keys (j, i) of them.
__global__ void k_strength(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ va, const double* __restrict__ d, double theta,
                           unsigned char* __restrict__ strong, long long* __restrict__ cnt) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double di = d[i];
    long long c = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        const bool st = i != j && fabs(va[k]) > theta * sqrt(fabs(di * d[j]));
        strong[k] = st ? 1 : 0;
        c += st ? 1 : 0;
    }
    cnt[i] = c;
}

__global__ void k_strong_keys(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                              const unsigned char* __restrict__ strong, const long long* __restrict__ off,
                              Key* __restrict__ key) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n) {
        return;
    }
    long long q = off[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        if (strong[k] != 0) {
            key[q++] = make_key(ci[k], i);
        }
    }
}

// Roots of pass 1 in rounds (amg.cpp:22-46). Node i is a root iff it has a strong neighbour and
// no root s < i claimed i or one of its strong neighbours j before i's turn, i.e. no root s < i in
// T(i), S(i) or T(S(i)) (S: strong neighbours, T: the nodes whose strong set holds the node). A
// node decides as soon as such a root is known (not a root) or every s < i there is decided (a
// root). Statuses: 0 undecided, 1 root, 2 not a root; a round reads `in` and writes `out`, so each
// round is deterministic. The smallest undecided node always decides, so the rounds end, and the
// roots are those of the sequential loop.
__global__ void k_root_round(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const unsigned char* __restrict__ strong, const int* __restrict__ trp,
                             const int* __restrict__ tci, const int* __restrict__ in, int* __restrict__ out) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n) {
        return;
    }
    const int st = in[i];
    if (st != 0) {
        out[i] = st;
        return;
    }
    bool known = true;
    bool claimed = false;
    auto look = [&](int v) {
        if (v < i) {
            const int x = in[v];
            claimed = claimed || x == 1;
            known = known && x != 0;
        }
    };
    for (int t = trp[i]; t < trp[i + 1] && !claimed; ++t) {
        look(tci[t]);
    }
    for (int k = rp[i]; k < rp[i + 1] && !claimed; ++k) {
        if (strong[k] == 0) {
            continue;
        }
        const int j = ci[k];
        look(j);
        for (int t = trp[j]; t < trp[j + 1] && !claimed; ++t) {
            look(tci[t]);
        }
    }
    out[i] = claimed ? 2 : (known ? 1 : 0);
}

// Per-block count of undecided nodes.
__global__ void k_count_undecided(int n, const int* __restrict__ st, int* __restrict__ part) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    const int c = __syncthreads_count(i < n && st[i] == 0);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = c;
    }
}

__global__ void k_root_flags(int n, const int* __restrict__ st, int* __restrict__ flag) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i < n) {
        flag[i] = st[i] == 1 ? 1 : 0;
    }
}

// Pass-1 aggregates: a root's index among the roots, given to it and its strong neighbours (the
// strong sets of roots are disjoint); -1 elsewhere (agg1 filled with -1 before).
__global__ void k_claim(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                        const unsigned char* __restrict__ strong, const int* __restrict__ st,
                        const int* __restrict__ id, int* __restrict__ agg1) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n || st[i] != 1) {
        return;
    }
    const int a = id[i];
    agg1[i] = a;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        if (strong[k] != 0) {
            agg1[ci[k]] = a;
        }
    }
}

// Pass 2 (amg.cpp:48-63): a node left free takes the pass-1 aggregate of its strongest strong
// neighbour (first maximum of |a_ij| in row order), from the pass-1 snapshot.
__global__ void k_pass2(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                        const double* __restrict__ va, const unsigned char* __restrict__ strong,
                        const int* __restrict__ agg1, int* __restrict__ agg) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i >= n) {
        return;
    }
    if (agg1[i] >= 0) {
        agg[i] = agg1[i];
        return;
    }
    double best = 0.0;
    int choice = -1;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        if (strong[k] != 0 && agg1[j] >= 0 && fabs(va[k]) > best) {
            best = fabs(va[k]);
            choice = agg1[j];
        }
    }
    agg[i] = choice;
}

// Undecided (0), or not a root (2) for a row without a strong entry.
__global__ void k_init_status(int n, const long long* __restrict__ cnt, int* __restrict__ st) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i < n) {
        st[i] = cnt[i] > 0 ? 0 : 2;
    }
}

__global__ void k_fill_int(int n, int* __restrict__ x, int v) {
    const int i = blockIdx.x * k_blk + threadIdx.x;
    if (i < n) {
        x[i] = v;
    }
}

template <class T>
void exclusive_sum(const T* in, T* out, long long n, hipStream_t s) {
    std::size_t bytes = 0;
    hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, static_cast<int>(n), s), "scan size");
    DevBuf<char> tmp(bytes > 0 ? bytes : 1);
    hip_check(hipcub::DeviceScan::ExclusiveSum(tmp.get(), bytes, in, out, static_cast<int>(n), s), "scan");
    hip_check(hipStreamSynchronize(s), "scan sync"); // the scratch goes back to the pool
}

template <class T>
void inclusive_sum(const T* in, T* out, long long n, hipStream_t s) {
    std::size_t bytes = 0;
    hip_check(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, static_cast<int>(n), s), "scan size");
    DevBuf<char> tmp(bytes > 0 ? bytes : 1);
    hip_check(hipcub::DeviceScan::InclusiveSum(tmp.get(), bytes, in, out, static_cast<int>(n), s), "scan");
    hip_check(hipStreamSynchronize(s), "scan sync"); // the scratch goes back to the pool
}

// Stable sort of (key, value) pairs by the low end_bit bits of the key.
void sort_pairs(DevBuf<Key>& key, DevBuf<double>& val, long long n, int end_bit, hipStream_t s) {
    if (n <= 1) {
        return;
    }
    DevBuf<Key> k2(static_cast<std::size_t>(n));
    DevBuf<double> v2(static_cast<std::size_t>(n));
    std::size_t bytes = 0;
    hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.get(), k2.get(), val.get(), v2.get(),
                                                 static_cast<int>(n), 0, end_bit, s),
              "radix sort size");
    DevBuf<char> tmp(bytes > 0 ? bytes : 1);
    hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.get(), bytes, key.get(), k2.get(), val.get(), v2.get(),
                                                 static_cast<int>(n), 0, end_bit, s),
              "radix sort");
    hip_check(hipStreamSynchronize(s), "radix sort sync"); // the inputs and scratch go back to the pool
    key = std::move(k2);
    val = std::move(v2);
}

} // namespace

// Device CSR matrix (sorted columns).
struct DeviceSetup::DCsr {
    int n_rows = 0;
    int n_cols = 0;
    int nnz = 0;
    DevBuf<int> rp;
    DevBuf<int> ci;
    DevBuf<double> va;
};

namespace {

using DCsr = DeviceSetup::DCsr;

void upload(const Csr& a, DCsr& d, hipStream_t s) {
    d.n_rows = a.n_rows;
    d.n_cols = a.n_cols;
    d.nnz = a.nnz();
    d.rp.upload(a.rp.data(), a.rp.size(), s);
    d.ci.upload(a.ci.empty() ? nullptr : a.ci.data(), a.ci.size(), s);
    d.va.upload(a.va.empty() ? nullptr : a.va.data(), a.va.size(), s);
}

void download(const DCsr& d, Csr& a, hipStream_t s) {
    a.n_rows = d.n_rows;
    a.n_cols = d.n_cols;
    a.rp.resize(static_cast<std::size_t>(d.n_rows) + 1);
    a.ci.resize(static_cast<std::size_t>(d.nnz));
    a.va.resize(static_cast<std::size_t>(d.nnz));
    d2h_staged(a.rp.data(), d.rp.get(), a.rp.size() * sizeof(int), s);
    if (d.nnz > 0) {
        d2h_staged(a.ci.data(), d.ci.get(), a.ci.size() * sizeof(int), s);
        d2h_staged(a.va.data(), d.va.get(), a.va.size() * sizeof(double), s);
    }
}

long long last_of(const long long* dev_off, const long long* dev_cnt, int n, hipStream_t s) {
    long long a = 0;
    long long b = 0;
    hip_check(hipMemcpyAsync(&a, dev_off + (n - 1), sizeof(long long), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(&b, dev_cnt + (n - 1), sizeof(long long), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "D2H sync");
    return a + b;
}

// Key-sorted (row, column) entries -> CSR; with prune, entries with |v| <= 0 are dropped first.
void entries_to_csr(DevBuf<Key>& key, DevBuf<double>& val, int n, int n_rows, int n_cols, bool prune, DCsr& c,
                    hipStream_t s) {
    if (prune && n > 0) {
        DevBuf<int> keep(static_cast<std::size_t>(n));
        DevBuf<int> pos(static_cast<std::size_t>(n));
        k_keep<<<blocks_for(n), k_blk, 0, s>>>(n, val.get(), keep.get());
        inclusive_sum(keep.get(), pos.get(), n, s);
        int kept = 0;
        hip_check(hipMemcpyAsync(&kept, pos.get() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "D2H sync");
        DevBuf<Key> k2(static_cast<std::size_t>(kept > 0 ? kept : 1));
        DevBuf<double> v2(static_cast<std::size_t>(kept > 0 ? kept : 1));
        k_compact<<<blocks_for(n), k_blk, 0, s>>>(n, keep.get(), pos.get(), key.get(), val.get(), k2.get(), v2.get());
        hip_check(hipGetLastError(), "compact launch");
        hip_check(hipStreamSynchronize(s), "compact sync"); // the inputs go back to the pool
        key = std::move(k2);
        val = std::move(v2);
        n = kept;
    }
    c.n_rows = n_rows;
    c.n_cols = n_cols;
    c.nnz = n;
    c.rp.resize(static_cast<std::size_t>(n_rows) + 1);
    k_row_ptr<<<blocks_for(static_cast<long long>(n_rows) + 1), k_blk, 0, s>>>(n_rows, n, key.get(), c.rp.get());
    c.ci.resize(static_cast<std::size_t>(n > 0 ? n : 1));
    if (n > 0) {
        k_split_keys<<<blocks_for(n), k_blk, 0, s>>>(n, key.get(), c.ci.get());
    }
    c.va = std::move(val);
    hip_check(hipGetLastError(), "csr launch");
    hip_check(hipStreamSynchronize(s), "csr sync");
}

// C = A B (optionally diag(lscale) A B), summed in the host setup's order; explicit zeros kept
// unless prune.
void spgemm(const DCsr& a, const DCsr& b, const double* lscale, bool prune, DCsr& c, hipStream_t s) {
    const int n = a.n_rows;
    DevBuf<long long> cnt(static_cast<std::size_t>(n));
    DevBuf<long long> off(static_cast<std::size_t>(n));
    k_count_products<<<blocks_for(n), k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), b.rp.get(), cnt.get());
    hip_check(hipGetLastError(), "count launch");
    exclusive_sum(cnt.get(), off.get(), n, s);
    const long long total = last_of(off.get(), cnt.get(), n, s);
    // Small level with long product lists: dense accumulation per row instead of sorting them.
    const long long cells = static_cast<long long>(n) * b.n_cols;
    if (n > 0 && cells <= (1LL << 25) && b.n_cols <= k_dense_max_cols && total >= 32LL * n) {
        DevBuf<double> acc(static_cast<std::size_t>(cells > 0 ? cells : 1));
        DevBuf<int> mark(static_cast<std::size_t>(cells > 0 ? cells : 1));
        const std::size_t lds = static_cast<std::size_t>(b.n_cols) * (sizeof(double) + sizeof(int));
        k_dense_rows<<<n, k_wave_dense, lds, s>>>(n, b.n_cols, a.rp.get(), a.ci.get(), a.va.get(), b.rp.get(),
                                                  b.ci.get(), b.va.get(), lscale, acc.get(), mark.get(), cnt.get());
        hip_check(hipGetLastError(), "dense rows launch");
        exclusive_sum(cnt.get(), off.get(), n, s);
        const long long kept = last_of(off.get(), cnt.get(), n, s);
        DevBuf<Key> dkey(static_cast<std::size_t>(kept > 0 ? kept : 1));
        DevBuf<double> dval(static_cast<std::size_t>(kept > 0 ? kept : 1));
        k_dense_write<<<n, k_wave_dense, 0, s>>>(n, b.n_cols, acc.get(), mark.get(), off.get(), dkey.get(),
                                                 dval.get());
        hip_check(hipGetLastError(), "dense write launch");
        hip_check(hipStreamSynchronize(s), "dense rows sync"); // the scratch goes back to the pool
        entries_to_csr(dkey, dval, static_cast<int>(kept), n, b.n_cols, prune, c, s);
        return;
    }
    require(total < static_cast<long long>(INT_MAX), "device AMG setup: too many products for one level");
    DevBuf<Key> key(static_cast<std::size_t>(total > 0 ? total : 1));
    DevBuf<double> val(static_cast<std::size_t>(total > 0 ? total : 1));
    if (total > 0) {
        k_emit_products<<<blocks_for(n), k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), a.va.get(), b.rp.get(), b.ci.get(),
                                                        b.va.get(), lscale, off.get(), key.get(), val.get());
        hip_check(hipGetLastError(), "emit launch");
        sort_pairs(key, val, total, 32 + bits_for(n), s);
    }
    // runs of equal keys, each added in order
    int runs = 0;
    DevBuf<Key> rkey(static_cast<std::size_t>(total > 0 ? total : 1));
    DevBuf<double> rval(static_cast<std::size_t>(total > 0 ? total : 1));
    if (total > 0) {
        DevBuf<int> head(static_cast<std::size_t>(total));
        DevBuf<int> run(static_cast<std::size_t>(total));
        k_heads<<<blocks_for(total), k_blk, 0, s>>>(total, key.get(), head.get());
        inclusive_sum(head.get(), run.get(), total, s);
        hip_check(hipMemcpyAsync(&runs, run.get() + (total - 1), sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "D2H sync");
        DevBuf<long long> start(static_cast<std::size_t>(runs));
        k_run_starts<<<blocks_for(total), k_blk, 0, s>>>(total, head.get(), run.get(), start.get());
        k_run_sums<<<blocks_for(runs), k_blk, 0, s>>>(runs, total, start.get(), key.get(), val.get(), rkey.get(),
                                                       rval.get());
        hip_check(hipGetLastError(), "run launch");
        hip_check(hipStreamSynchronize(s), "run sync"); // the scratch goes out of scope
    }
    entries_to_csr(rkey, rval, runs, n, b.n_cols, prune, c, s);
}

} // namespace

DeviceSetup::DeviceSetup(const Csr& a0) : cur_(std::make_unique<DCsr>()) {
    require_device();
    upload(a0, *cur_, stream_.get());
}

DeviceSetup::~DeviceSetup() { (void)hipStreamSynchronize(stream_.get()); }

std::vector<int> DeviceSetup::aggregate12(const std::vector<double>& d, double theta, int& n_roots) {
    hipStream_t s = stream_.get();
    const DCsr& a = *cur_;
    const int n = a.n_rows;
    require(static_cast<int>(d.size()) == n, "device aggregation: diagonal does not match the level");
    std::vector<int> agg(static_cast<std::size_t>(n), -1);
    n_roots = 0;
    if (n == 0) {
        return agg;
    }
    DevBuf<double> dd;
    dd.upload(d.data(), d.size(), s);
    DevBuf<unsigned char> strong(static_cast<std::size_t>(a.nnz > 0 ? a.nnz : 1));
    DevBuf<long long> cnt(static_cast<std::size_t>(n));
    DevBuf<long long> off(static_cast<std::size_t>(n));
    k_strength<<<blocks_for(n), k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), a.va.get(), dd.get(), theta, strong.get(),
                                               cnt.get());
    hip_check(hipGetLastError(), "strength launch");
    exclusive_sum(cnt.get(), off.get(), n, s);
    const long long m = last_of(off.get(), cnt.get(), n, s);
    require(m < static_cast<long long>(INT_MAX), "device aggregation: too many strong entries");
    // T, the transposed strength pattern: rows j listing the i with j in S(i), ascending.
    DCsr t;
    {
        DevBuf<Key> key(static_cast<std::size_t>(m > 0 ? m : 1));
        DevBuf<double> val(static_cast<std::size_t>(m > 0 ? m : 1));
        k_strong_keys<<<blocks_for(n), k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), strong.get(), off.get(), key.get());
        hip_check(hipGetLastError(), "strong keys launch");
        sort_pairs(key, val, m, 32 + bits_for(n), s);
        entries_to_csr(key, val, static_cast<int>(m), n, n, false, t, s);
    }
    // Pass 1 in rounds; the nodes without a strong neighbour are decided (not roots) at once.
    DevBuf<int> st0(static_cast<std::size_t>(n));
    DevBuf<int> st1(static_cast<std::size_t>(n));
    const int nb = blocks_for(n);
    k_init_status<<<nb, k_blk, 0, s>>>(n, cnt.get(), st0.get());
    hip_check(hipGetLastError(), "status launch");
    DevBuf<int> part(static_cast<std::size_t>(nb));
    std::vector<int> part_h(static_cast<std::size_t>(nb));
    long long rounds = 0;
    constexpr int k_rounds_per_check = 16;
    for (;;) {
        for (int r = 0; r < k_rounds_per_check; ++r) {
            k_root_round<<<nb, k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), strong.get(), t.rp.get(), t.ci.get(), st0.get(),
                                              st1.get());
            std::swap(st0, st1);
        }
        hip_check(hipGetLastError(), "root round launch");
        rounds += k_rounds_per_check;
        k_count_undecided<<<nb, k_blk, 0, s>>>(n, st0.get(), part.get());
        hip_check(hipMemcpyAsync(part_h.data(), part.get(), part_h.size() * sizeof(int), hipMemcpyDeviceToHost, s),
                  "D2H");
        hip_check(hipStreamSynchronize(s), "root rounds sync");
        long long undecided = 0;
        for (int v : part_h) {
            undecided += v;
        }
        if (undecided == 0) {
            break;
        }
        // every round decides at least the smallest undecided node
        require(rounds <= static_cast<long long>(n) + k_rounds_per_check, "device aggregation: rounds did not end");
    }
    DevBuf<int> flag(static_cast<std::size_t>(n));
    DevBuf<int> id(static_cast<std::size_t>(n));
    k_root_flags<<<nb, k_blk, 0, s>>>(n, st0.get(), flag.get());
    hip_check(hipGetLastError(), "root flags launch");
    exclusive_sum(flag.get(), id.get(), n, s);
    int last_id = 0;
    int last_flag = 0;
    hip_check(hipMemcpyAsync(&last_id, id.get() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipMemcpyAsync(&last_flag, flag.get() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "roots sync");
    n_roots = last_id + last_flag;
    DevBuf<int> agg1(static_cast<std::size_t>(n));
    DevBuf<int> agg2(static_cast<std::size_t>(n));
    k_fill_int<<<nb, k_blk, 0, s>>>(n, agg1.get(), -1);
    k_claim<<<nb, k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), strong.get(), st0.get(), id.get(), agg1.get());
    k_pass2<<<nb, k_blk, 0, s>>>(n, a.rp.get(), a.ci.get(), a.va.get(), strong.get(), agg1.get(), agg2.get());
    hip_check(hipGetLastError(), "aggregation launch");
    hip_check(hipMemcpyAsync(agg.data(), agg2.get(), agg.size() * sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "aggregation sync"); // the buffers go back to the pool
    return agg;
}

void DeviceSetup::level(const std::vector<double>& inv_diag, const std::vector<int>& agg, int n_agg, double omega,
                        Csr& p_out, Csr& r_out, Csr& coarse_out) {
    hipStream_t s = stream_.get();
    const DCsr& a = *cur_;
    const int n = a.n_rows;
    require(static_cast<int>(agg.size()) == n && static_cast<int>(inv_diag.size()) == n,
            "device AMG setup: aggregates / diagonal do not match the level");
    // P_tent: row i has the single entry (i, agg_i) = 1 (amg.cpp:152-158).
    DCsr pt;
    pt.n_rows = n;
    pt.n_cols = n_agg;
    pt.nnz = n;
    pt.rp.resize(static_cast<std::size_t>(n) + 1);
    pt.ci.upload(agg.data(), agg.size(), s);
    pt.va.resize(static_cast<std::size_t>(n));
    k_fill_iota<<<blocks_for(static_cast<long long>(n) + 1), k_blk, 0, s>>>(n, pt.rp.get());
    k_fill_one<<<blocks_for(n), k_blk, 0, s>>>(n, pt.va.get());
    hip_check(hipGetLastError(), "tentative launch");
    DevBuf<double> dinv;
    dinv.upload(inv_diag.data(), inv_diag.size(), s);
    // S = D^-1 A P_tent, then P = P_tent - omega S, pruned (amg.cpp:169-172)
    DCsr sm;
    setup_phase("device: tentative + uploads", -1);
    spgemm(a, pt, dinv.get(), false, sm, s);
    setup_phase("device: D^-1 A P_tent", -1);
    DCsr p;
    {
        DevBuf<long long> cnt(static_cast<std::size_t>(n));
        DevBuf<long long> off(static_cast<std::size_t>(n));
        k_prolong_count<<<blocks_for(n), k_blk, 0, s>>>(n, pt.ci.get(), sm.rp.get(), sm.ci.get(), sm.va.get(), omega,
                                                         cnt.get());
        hip_check(hipGetLastError(), "prolongation count launch");
        exclusive_sum(cnt.get(), off.get(), n, s);
        const long long total = last_of(off.get(), cnt.get(), n, s);
        DevBuf<Key> key(static_cast<std::size_t>(total > 0 ? total : 1));
        DevBuf<double> val(static_cast<std::size_t>(total > 0 ? total : 1));
        k_prolong_write<<<blocks_for(n), k_blk, 0, s>>>(n, pt.ci.get(), sm.rp.get(), sm.ci.get(), sm.va.get(), omega,
                                                         off.get(), key.get(), val.get());
        hip_check(hipGetLastError(), "prolongation launch");
        entries_to_csr(key, val, static_cast<int>(total), n, n_agg, false, p, s);
    }
    setup_phase("device: P", -1);
    // R = P^T, rows listing the fine rows in ascending order (amg.cpp:175)
    DCsr r;
    {
        DevBuf<Key> key(static_cast<std::size_t>(p.nnz > 0 ? p.nnz : 1));
        k_transpose_keys<<<blocks_for(n), k_blk, 0, s>>>(n, p.rp.get(), p.ci.get(), key.get());
        hip_check(hipGetLastError(), "transpose launch");
        DevBuf<double> val(static_cast<std::size_t>(p.nnz > 0 ? p.nnz : 1));
        if (p.nnz > 0) {
            hip_check(hipMemcpyAsync(val.get(), p.va.get(), static_cast<std::size_t>(p.nnz) * sizeof(double),
                                     hipMemcpyDeviceToDevice, s),
                      "D2D");
        }
        sort_pairs(key, val, p.nnz, 32 + bits_for(n_agg), s);
        entries_to_csr(key, val, p.nnz, n_agg, n, false, r, s);
    }
    setup_phase("device: R = P^T", -1);
    // A_c = R (A P), pruned (amg.cpp:176-178)
    DCsr ap;
    spgemm(a, p, nullptr, false, ap, s);
    setup_phase("device: A P", -1);
    auto coarse = std::make_unique<DCsr>();
    spgemm(r, ap, nullptr, true, *coarse, s);
    setup_phase("device: R (A P)", -1);
    download(p, p_out, s);
    download(r, r_out, s);
    download(*coarse, coarse_out, s);
    setup_phase("device: downloads", -1);
    cur_ = std::move(coarse);
}

} // namespace cfdg
