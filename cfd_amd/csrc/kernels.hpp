// Device kernels of the AMG-PCG / Krylov backend for gfx950 (CDNA4, wave64).
//
// Layout. Every matrix is sliced ELL with a slice height of 64 rows, one wavefront per slice:
// entry k of row (64 s + lane) sits at sp[s] + 64 k + lane, so each load instruction of a wave
// reads 64 consecutive column indices / values. Rows keep the CSR (ascending column) order of
// the reference and padding entries (value 0) come last, so an FP64 row sum equals the
// reference's serial sum (src/gpu/src/krylov_kernels.inc:12-25) bit for bit. A workgroup is
// four wavefronts = four slices = 256 rows.
//
// Work mapping. A launch covers the rows of one level for a set of systems: block b processes
// the 256-row block blocks[b].y of system blocks[b].x. A single solve is a set of one system.
//
// Reductions. Dot products never use atomics: each block writes one partial sum (wave
// butterfly, then the four wave sums in order), and a finalize kernel with one workgroup per
// system adds the partials in a fixed order. Results are reproducible run to run and do not
// depend on which other systems share the launch.
//
// Control. Each system carries a SolveState; kernels of a system whose status is non-zero
// return at once, so the host may enqueue iterations ahead of the convergence test.
#pragma once

#include "layout.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cfdg {

// ------------------------------------------------------------------------------------------
// Arithmetic helpers. The library is compiled with -ffp-contract=off; FP32 kernels use
// explicit fused multiply-adds, FP64 kernels keep the reference's separate multiply and add.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ double madd(double a, double b, double c) { return c + a * b; }
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, k_wave);
    }
    return v;
}

// Sums one value per thread over the block; thread 0 receives the result.
template <int N>
__device__ __forceinline__ void block_sums(double (&v)[N], double (&out)[N]) {
    __shared__ double red[N][k_slices_per_block];
    const int wave = threadIdx.x / k_wave;
    const int lane = threadIdx.x % k_wave;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const double w = wave_sum(v[q]);
        if (lane == 0) {
            red[q][wave] = w;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            double s = red[q][0];
#pragma unroll
            for (int w = 1; w < k_slices_per_block; ++w) {
                s += red[q][w];
            }
            out[q] = s;
        }
    }
}

__device__ __forceinline__ float bf16_to_float(unsigned short h) {
    return __uint_as_float(static_cast<unsigned>(h) << 16);
}

// Device buffers seen in the global address space: loads through such pointers compile to
// global_load with a scalar base instead of flat loads (which also count against the LDS
// counter and need a full 64-bit address per lane). Descriptor pointers are generic.
template <class T>
using gmem_t = __attribute__((address_space(1))) T;

template <class T>
__device__ __forceinline__ gmem_t<T>* gm(T* p) {
    return (gmem_t<T>*)p;
}

// Column index of slot k: 32-bit columns (ColAbs), row + 16-bit offset (ColOff, square
// matrices), the slice's base column + unsigned 16-bit offset (ColBase, transfer operators) or
// row + the slice's stencil offset of round k (ColStencil, square matrices, one lane per row).
// The format is chosen once per matrix (a uniform branch), never per slot: a per-lane pointer
// test would turn every column load into a divergent branch with its own wait.
struct ColAbs {
    const gmem_t<const int>* __restrict__ col;
    __device__ __forceinline__ int operator()(int k) const { return col[k]; }
};

struct ColOff {
    const gmem_t<const short>* __restrict__ dcol;
    int row;
    __device__ __forceinline__ int operator()(int k) const { return row + static_cast<int>(dcol[k]); }
};

struct ColBase {
    const gmem_t<const unsigned short>* __restrict__ dcol;
    int base;
    __device__ __forceinline__ int operator()(int k) const { return base + static_cast<int>(dcol[k]); }
};

// Stencil column: the slice's offset table (uniform over the wavefront) added to the row,
// clamped to [0, last] for the value-0 slots of rows that lack an offset.
struct ColStencil {
    const gmem_t<const int>* __restrict__ off;
    int row;
    int last;
    __device__ __forceinline__ int operator()(int k) const { return min(max(row + off[k / k_wave], 0), last); }
};

template <class T, class V>
__device__ __forceinline__ T value_at(const gmem_t<const V>* __restrict__ val, int i) {
    if constexpr (sizeof(V) == 2) {
        return bf16_to_float(val[i]);
    } else {
        return val[i];
    }
}

#ifndef CFDG_ELL_GROUP
#define CFDG_ELL_GROUP 4
#endif
constexpr int k_ell_group = CFDG_ELL_GROUP;

// Storage-order row sum of sliced-ELL slots with values val (type V) converted to T. Loads are
// issued k_ell_group entries at a time, the tail (w mod k_ell_group, uniform over the slice's
// wavefront) as one more group with its unused loads repeating a valid slot; only real entries
// are summed.
template <class T, class V, class C, class F>
__device__ __forceinline__ T ell_sum(const C& ci, const gmem_t<const V>* __restrict__ val, int w, F load) {
    constexpr int G = k_ell_group;
    T s = T(0);
    int k = 0;
    for (; k + G <= w; k += G) {
        int c[G];
        T a[G];
        T x[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            c[g] = ci((k + g) * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            a[g] = value_at<T, V>(val, (k + g) * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            x[g] = load(c[g]);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            s = madd(a[g], x[g], s);
        }
    }
    const int rem = w - k;
    if (rem > 0) {
        int c[G - 1];
        T a[G - 1];
        T x[G - 1];
#pragma unroll
        for (int g = 0; g < G - 1; ++g) {
            c[g] = ci((g < rem ? k + g : k) * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G - 1; ++g) {
            a[g] = value_at<T, V>(val, (g < rem ? k + g : k) * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G - 1; ++g) {
            x[g] = load(c[g]);
        }
        s = madd(a[0], x[0], s);
#pragma unroll
        for (int g = 1; g < G - 1; ++g) {
            if (g < rem) {
                s = madd(a[g], x[g], s);
            }
        }
    }
    return s;
}

// Sum over the CSR entries k, k + step, ... < end of one lane, in that order, four loads at a
// time (the lane's share of a row read by step lanes).
template <class T, class F>
__device__ __forceinline__ T csr_lane_sum(const gmem_t<const int>* __restrict__ col,
                                          const gmem_t<const T>* __restrict__ val, int k, int end, int step,
                                          F load) {
    T s = T(0);
    for (; k + 3 * step < end; k += 4 * step) {
        const int c0 = col[k];
        const int c1 = col[k + step];
        const int c2 = col[k + 2 * step];
        const int c3 = col[k + 3 * step];
        const T a0 = val[k];
        const T a1 = val[k + step];
        const T a2 = val[k + 2 * step];
        const T a3 = val[k + 3 * step];
        const T x0 = load(c0);
        const T x1 = load(c1);
        const T x2 = load(c2);
        const T x3 = load(c3);
        s = madd(a0, x0, s);
        s = madd(a1, x1, s);
        s = madd(a2, x2, s);
        s = madd(a3, x3, s);
    }
    if (k < end) {
        const bool has1 = k + step < end;
        const bool has2 = k + 2 * step < end;
        const int k1 = has1 ? k + step : k;
        const int k2 = has2 ? k + 2 * step : k;
        const int c0 = col[k];
        const int c1 = col[k1];
        const int c2 = col[k2];
        const T a0 = val[k];
        const T a1 = val[k1];
        const T a2 = val[k2];
        const T x0 = load(c0);
        const T x1 = load(c1);
        const T x2 = load(c2);
        s = madd(a0, x0, s);
        if (has1) {
            s = madd(a1, x1, s);
        }
        if (has2) {
            s = madd(a2, x2, s);
        }
    }
    return s;
}

#ifndef CFDG_SYM_GROUP
#define CFDG_SYM_GROUP 8
#endif
constexpr int k_sym_group = CFDG_SYM_GROUP; // entries whose loads are issued together

// One entry of a row in symmetric stencil storage: offset o, stored slot j of |o|. The value of
// o >= 0 is in the row's own slot, of o < 0 in row + o's slot of -o (zero where row + o < 0); the
// column is row + o clamped to the matrix, as in the per-slice stencil layout (whose missing
// offsets hold zeros).
template <class T, class F>
__device__ __forceinline__ void sym_entry(const gmem_t<const T>* __restrict__ val, int stride, int last, int row, int o,
                                          int j, F load, T& a, T& x) {
    const int r = o < 0 ? row + o : row;
    const int rr = max(r, 0);
    const T v = val[(rr >> 6) * stride + j * k_wave + (rr & (k_wave - 1))];
    a = r < 0 ? T(0) : v;
    x = load(min(max(row + o, 0), last));
}

// Row sums (rows row0, and row0 + 64 when Rows == 2) of a stencil of N offsets known at compile
// time: the table is read with constant indices (one batch of scalar loads, one wait) and every
// entry's loads are issued before the sums, which run in offset order.
template <int N, int Rows, class T, class F>
__device__ __forceinline__ void sym_rows_fixed(const EllView& m, int row0, F load, T (&y)[2]) {
    const auto tab = gm(m.soff); // offsets [0, N), stored slots [N, 2 N)
    const auto val = gm(static_cast<const T*>(m.val));
    const int stride = m.sym_w * k_wave;
    const int last = m.n_rows - 1;
    int o[N];
    int j[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        o[k] = tab[k];
        j[k] = tab[N + k];
    }
    T a[Rows][N];
    T x[Rows][N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int h = 0; h < Rows; ++h) {
            sym_entry<T>(val, stride, last, row0 + h * k_wave, o[k], j[k], load, a[h][k], x[h][k]);
        }
    }
#pragma unroll
    for (int h = 0; h < Rows; ++h) {
        T s = T(0);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            s = madd(a[h][k], x[h][k], s);
        }
        y[h] = s;
    }
}

// Any stencil size: the table read per entry, loads issued k_sym_group entries at a time.
template <int Rows, class T, class F>
__device__ __forceinline__ void sym_rows_any(const EllView& m, int row0, F load, T (&y)[2]) {
    const int n = m.sym_n;
    const auto tab = gm(m.soff);
    const auto val = gm(static_cast<const T*>(m.val));
    const int stride = m.sym_w * k_wave;
    const int last = m.n_rows - 1;
    T s[Rows];
#pragma unroll
    for (int h = 0; h < Rows; ++h) {
        s[h] = T(0);
    }
    for (int k = 0; k < n; k += k_sym_group) {
        T a[Rows][k_sym_group];
        T x[Rows][k_sym_group];
#pragma unroll
        for (int g = 0; g < k_sym_group; ++g) {
            const int kk = min(k + g, n - 1); // past the stencil: repeats its last entry, not added
            const int o = tab[kk];
            const int j = tab[n + kk];
#pragma unroll
            for (int h = 0; h < Rows; ++h) {
                sym_entry<T>(val, stride, last, row0 + h * k_wave, o, j, load, a[h][g], x[h][g]);
            }
        }
#pragma unroll
        for (int g = 0; g < k_sym_group; ++g) {
            if (k + g < n) {
#pragma unroll
                for (int h = 0; h < Rows; ++h) {
                    s[h] = madd(a[h][g], x[h][g], s[h]);
                }
            }
        }
    }
#pragma unroll
    for (int h = 0; h < Rows; ++h) {
        y[h] = s[h];
    }
}

template <int Rows, class T, class F>
__device__ __forceinline__ void sym_rows(const EllView& m, int row0, F load, T (&y)[2]) {
    switch (m.sym_n) { // uniform: the stencils of 3-D, 2-D and 1-D box meshes unrolled
    case 7:
        sym_rows_fixed<7, Rows, T>(m, row0, load, y);
        return;
    case 5:
        sym_rows_fixed<5, Rows, T>(m, row0, load, y);
        return;
    case 3:
        sym_rows_fixed<3, Rows, T>(m, row0, load, y);
        return;
    default:
        sym_rows_any<Rows, T>(m, row0, load, y);
    }
}

// Row sum of a row in symmetric stencil storage (EllView::sym_w > 0), entries in offset order.
template <class T, class F>
__device__ __forceinline__ T ell_row_sym(const EllView& m, int slice, int lane, F load) {
    T y[2];
    sym_rows<1, T>(m, slice * k_wave + lane, load, y);
    return y[0];
}

// ell_row_sym for the rows of slices s0 and s0 + 1 (lane's row in each), their loads issued
// together: the same sums as two ell_row_sym calls.
template <class T, class F>
__device__ __forceinline__ void ell_row_sym_pair(const EllView& m, int s0, int lane, F load, T& ya, T& yb) {
    T y[2];
    sym_rows<2, T>(m, s0 * k_wave + lane, load, y);
    ya = y[0];
    yb = y[1];
}

// Row sum of an ELL row with the x entry of column c produced by load(c), in storage order.
// BF16 rows (fmt 1, FP32 kernels only) give diag_i x_i + rscale_i sum_k bf16_k x_ck.
template <class T, class F>
__device__ __forceinline__ T ell_row_f(const EllView& m, int slice, int lane, F load) {
    if (m.sym_w > 0) {
        return ell_row_sym<T>(m, slice, lane, load);
    }
    const int base = m.sp[slice];
    const int w = (m.sp[slice + 1] - base) / k_wave;
    const int lg = __builtin_ctz(m.tpr); // lanes per row = 2^lg
    const int row = slice * (k_wave >> lg) + (lane >> lg);
    auto sum = [&](const auto& ci) -> T {
        if constexpr (sizeof(T) == 4) {
            if (m.fmt == 1) {
                const auto val = gm(static_cast<const unsigned short*>(m.val)) + base + lane;
                const T s = ell_sum<T, unsigned short>(ci, val, w, load);
                const T dx = row < m.n_rows ? gm(m.diag)[row] * load(row) : T(0);
                return madd(gm(m.rscale)[row], s, dx);
            }
        }
        return ell_sum<T, T>(ci, gm(static_cast<const T*>(m.val)) + base + lane, w, load);
    };
    if (m.soff != nullptr) {
        const int tab = __builtin_amdgcn_readfirstlane(base / k_wave); // uniform: one slice per wavefront
        return sum(ColStencil{gm(m.soff) + tab, row, m.n_rows - 1});
    }
    if (m.cbase != nullptr) {
        return sum(ColBase{gm(reinterpret_cast<const unsigned short*>(m.dcol)) + base + lane, gm(m.cbase)[slice]});
    }
    if (m.dcol != nullptr) {
        return sum(ColOff{gm(m.dcol) + base + lane, row});
    }
    return sum(ColAbs{gm(m.col) + base + lane});
}

// Row sum of an ELL row: sum_k val_k * x[col_k] in storage order (any column / value format).
template <class T>
__device__ __forceinline__ T ell_row(const EllView& m, int slice, int lane, const T* __restrict__ x) {
    const auto gx = gm(x);
    return ell_row_f<T>(m, slice, lane, [&](int c) { return gx[c]; });
}

// Row sum with the first Jacobi iterate x1_j = wd_j * b_j formed on the fly (the reference's
// first sweep multiplies A by x = 0, so x1 = (w dinv) b exactly).
template <class T>
__device__ __forceinline__ T ell_row_x1(const EllView& m, int slice, int lane, const gmem_t<const T>* __restrict__ wd,
                                        const gmem_t<const T>* __restrict__ b, T f) {
    return ell_row_f<T>(m, slice, lane, [&](int c) { return (f * wd[c]) * b[c]; });
}

// Row sums of two slices with equal widths and one column format, their loads issued together
// (k_ell_group entries of each slice, then the sums in storage order per row): twice the loads in
// flight per wavefront, the same arithmetic as two ell_sum calls.
template <class T, class V, class C, class F>
__device__ __forceinline__ void ell_sum2(const C& ca, const C& cb, const gmem_t<const V>* __restrict__ va,
                                         const gmem_t<const V>* __restrict__ vb, int w, F load, T& sa, T& sb) {
    constexpr int G = k_ell_group;
    sa = T(0);
    sb = T(0);
    int k = 0;
    for (; k + G <= w; k += G) {
        int c0[G];
        int c1[G];
        T a0[G];
        T a1[G];
        T x0[G];
        T x1[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            c0[g] = ca((k + g) * k_wave);
            c1[g] = cb((k + g) * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            a0[g] = value_at<T, V>(va, (k + g) * k_wave);
            a1[g] = value_at<T, V>(vb, (k + g) * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            x0[g] = load(c0[g]);
            x1[g] = load(c1[g]);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            sa = madd(a0[g], x0[g], sa);
            sb = madd(a1[g], x1[g], sb);
        }
    }
    const int rem = w - k;
    if (rem > 0) {
        int c0[G - 1];
        int c1[G - 1];
        T a0[G - 1];
        T a1[G - 1];
        T x0[G - 1];
        T x1[G - 1];
#pragma unroll
        for (int g = 0; g < G - 1; ++g) {
            const int kk = g < rem ? k + g : k;
            c0[g] = ca(kk * k_wave);
            c1[g] = cb(kk * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G - 1; ++g) {
            const int kk = g < rem ? k + g : k;
            a0[g] = value_at<T, V>(va, kk * k_wave);
            a1[g] = value_at<T, V>(vb, kk * k_wave);
        }
#pragma unroll
        for (int g = 0; g < G - 1; ++g) {
            x0[g] = load(c0[g]);
            x1[g] = load(c1[g]);
        }
        sa = madd(a0[0], x0[0], sa);
        sb = madd(a1[0], x1[0], sb);
#pragma unroll
        for (int g = 1; g < G - 1; ++g) {
            if (g < rem) {
                sa = madd(a0[g], x0[g], sa);
                sb = madd(a1[g], x1[g], sb);
            }
        }
    }
}

// The lane's share of its rows in slices s0 and s0 + 1 of a sliced-ELL matrix (a slice past the
// end gives 0; with tpr > 1 the caller adds the row groups): equal to ell_row_f on each slice,
// bit for bit. Slices of different widths are summed one after the other. Stencil columns and
// BF16 rows occur with one lane per row only.
template <class T, class F>
__device__ __forceinline__ void ell_row_pair(const EllView& m, int s0, int lane, F load, T& ya, T& yb) {
    ya = T(0);
    yb = T(0);
    if (s0 >= m.n_slices) {
        return;
    }
    if (s0 + 1 >= m.n_slices) {
        ya = ell_row_f<T>(m, s0, lane, load);
        return;
    }
    if (m.sym_w > 0) {
        ell_row_sym_pair<T>(m, s0, lane, load, ya, yb);
        return;
    }
    const int b0 = m.sp[s0];
    const int b1 = m.sp[s0 + 1];
    const int w = (b1 - b0) / k_wave;
    if (w != (m.sp[s0 + 2] - b1) / k_wave) {
        ya = ell_row_f<T>(m, s0, lane, load);
        yb = ell_row_f<T>(m, s0 + 1, lane, load);
        return;
    }
    const int lg = __builtin_ctz(m.tpr);
    const int r0 = s0 * (k_wave >> lg) + (lane >> lg);
    const int r1 = r0 + (k_wave >> lg);
    auto pair = [&](const auto& c0, const auto& c1) {
        if constexpr (sizeof(T) == 4) {
            if (m.fmt == 1) {
                const auto val = gm(static_cast<const unsigned short*>(m.val));
                T sa;
                T sb;
                ell_sum2<T, unsigned short>(c0, c1, val + b0 + lane, val + b1 + lane, w, load, sa, sb);
                const T dxa = r0 < m.n_rows ? gm(m.diag)[r0] * load(r0) : T(0);
                const T dxb = r1 < m.n_rows ? gm(m.diag)[r1] * load(r1) : T(0);
                ya = madd(gm(m.rscale)[r0], sa, dxa);
                yb = madd(gm(m.rscale)[r1], sb, dxb);
                return;
            }
        }
        const auto val = gm(static_cast<const T*>(m.val));
        ell_sum2<T, T>(c0, c1, val + b0 + lane, val + b1 + lane, w, load, ya, yb);
    };
    if (m.soff != nullptr) {
        const int t0 = __builtin_amdgcn_readfirstlane(b0 / k_wave);
        const int t1 = __builtin_amdgcn_readfirstlane(b1 / k_wave);
        pair(ColStencil{gm(m.soff) + t0, r0, m.n_rows - 1}, ColStencil{gm(m.soff) + t1, r1, m.n_rows - 1});
    } else if (m.cbase != nullptr) {
        const auto dc = gm(reinterpret_cast<const unsigned short*>(m.dcol));
        pair(ColBase{dc + b0 + lane, gm(m.cbase)[s0]}, ColBase{dc + b1 + lane, gm(m.cbase)[s0 + 1]});
    } else if (m.dcol != nullptr) {
        pair(ColOff{gm(m.dcol) + b0 + lane, r0}, ColOff{gm(m.dcol) + b1 + lane, r1});
    } else {
        pair(ColAbs{gm(m.col) + b0 + lane}, ColAbs{gm(m.col) + b1 + lane});
    }
}

// Sum over groups of tpr consecutive lanes (tpr a power of two <= 64); every lane of the
// wavefront must take part.
template <class T>
__device__ __forceinline__ T group_sum(T v, int tpr) {
    for (int off = tpr >> 1; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, k_wave);
    }
    return v;
}

// Row product of a matrix in either layout for the row / sub-lane of the calling thread; the
// result is complete in every lane of the row's group. Rows past the end give 0.
template <class T>
__device__ __forceinline__ T mat_row(const EllView& m, int row, int sub, const T* __restrict__ x) {
    if (m.csr == 0) {
        const int lg = __builtin_ctz(m.tpr);
        const int slice = row >> (6 - lg); // 64 / tpr rows per slice
        const T s = slice < m.n_slices ? ell_row<T>(m, slice, static_cast<int>(threadIdx.x) & (k_wave - 1), x)
                                       : T(0);
        return m.tpr > 1 ? group_sum(s, m.tpr) : s;
    }
    T s = T(0);
    if (row < m.n_rows) {
        const auto gx = gm(x);
        s = csr_lane_sum<T>(gm(m.col), gm(static_cast<const T*>(m.val)), m.sp[row] + sub, m.sp[row + 1],
                            m.tpr, [&](int c) { return gx[c]; });
    }
    return group_sum(s, m.tpr);
}

template <class T>
__device__ __forceinline__ T mat_row_x1(const EllView& m, int row, int sub, const gmem_t<const T>* __restrict__ wd,
                                        const gmem_t<const T>* __restrict__ b, T f) {
    if (m.csr == 0) {
        const int lg = __builtin_ctz(m.tpr);
        const int slice = row >> (6 - lg);
        const T s = slice < m.n_slices
                        ? ell_row_x1<T>(m, slice, static_cast<int>(threadIdx.x) & (k_wave - 1), wd, b, f)
                        : T(0);
        return m.tpr > 1 ? group_sum(s, m.tpr) : s;
    }
    T s = T(0);
    if (row < m.n_rows) {
        s = csr_lane_sum<T>(gm(m.col), gm(static_cast<const T*>(m.val)), m.sp[row] + sub, m.sp[row + 1],
                            m.tpr, [&](int c) { return (f * wd[c]) * b[c]; });
    }
    return group_sum(s, m.tpr);
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i runs on XCD i mod 8) and
// every XCD has its own L2. This bijection hands each XCD a contiguous range of logical blocks,
// so rows that share neighbours (and their x entries) meet in the same L2.
constexpr int k_xcds = 8;

__device__ __forceinline__ int xcd_block() {
    const int g = static_cast<int>(gridDim.x);
    const int b = static_cast<int>(blockIdx.x);
    const int q = g / k_xcds;
    const int r = g % k_xcds;
    const int xcd = b % k_xcds;
    const int k = b / k_xcds;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// The launch table's (system, block) entry of this workgroup, read as one 8-byte load: both
// halves arrive with one wait (the block half is otherwise fetched again after the status test).
__device__ __forceinline__ int2 block_entry(const int2* __restrict__ blocks) {
    const auto e = gm(reinterpret_cast<const unsigned long long*>(blocks))[xcd_block()];
    return make_int2(static_cast<int>(static_cast<unsigned>(e)), static_cast<int>(static_cast<unsigned>(e >> 32)));
}

// Outer vectors by id: 0 x, 1 r, 2 p, 3 q, 4 z, 5 r0, 6 v, 7 s, 8 t, 9 ph, 10 sh.
__device__ __forceinline__ double* outer_vec(const SysDev& S, int id) {
    switch (id) {
    case 0: return S.x;
    case 1: return S.r;
    case 2: return S.p;
    case 3: return S.q;
    case 4: return S.z;
    case 5: return S.r0;
    case 6: return S.v;
    case 7: return S.s;
    case 8: return S.t;
    case 9: return S.ph;
    default: return S.sh;
    }
}

// Level-0 FP64 row product of the calling lane's row. Csr = false: the library's sliced-ELL
// layout (every lane of a live slice takes part). Csr = true: a caller's CSR arrays read in
// place (a64.sp = row offsets, a64.col, a64.val; one lane per row, entries in CSR order), which
// gives the same sums: the sliced-ELL padding and stencil gaps add exact zeros after or between
// the same products in the same order. Rows past the end give 0.
template <bool Csr>
__device__ __forceinline__ bool a0_live(const EllView& m, int slice) {
    return Csr || slice < m.n_slices;
}

template <bool Csr, class F>
__device__ __forceinline__ double a0_row_f(const EllView& m, int slice, int lane, int row, F load) {
    if constexpr (Csr) {
        if (row >= m.n_rows) {
            return 0.0;
        }
        const auto sp = gm(m.sp);
        return csr_lane_sum<double>(gm(m.col), gm(static_cast<const double*>(m.val)), sp[row], sp[row + 1], 1, load);
    } else {
        return ell_row_f<double>(m, slice, lane, load);
    }
}

template <bool Csr>
__device__ __forceinline__ double a0_row(const EllView& m, int slice, int lane, int row, const double* __restrict__ x) {
    const auto gx = gm(x);
    return a0_row_f<Csr>(m, slice, lane, row, [&](int c) { return gx[c]; });
}

#define CFDG_BLOCK_PROLOGUE(sysarr, blocks)                                                     \
    const int2 bk_ = block_entry(blocks);                                                         \
    const SysDev& S = (sysarr)[bk_.x];                                                            \
    const int wave_ = threadIdx.x / k_wave;                                                       \
    const int lane = threadIdx.x % k_wave;                                                        \
    const int slice = bk_.y * k_slices_per_block + wave_;                                         \
    const int row = slice * k_wave + lane;                                                        \
    (void)row;

// ------------------------------------------------------------------------------------------
// Outer Krylov kernels (level 0, FP64)
// ------------------------------------------------------------------------------------------

// r = b - A x, partials r.r and b.b; p = q = 0. For mixed precision also r32 = (float) r.
template <bool Mixed>
__global__ void __launch_bounds__(k_block) k_residual0(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    double rr = 0.0;
    double bb = 0.0;
    if (slice < S.a64.n_slices) {
        const double ax = ell_row<double>(S.a64, slice, lane, S.x);
        if (row < S.n) {
            const double bi = gm(S.b)[row];
            const double ri = bi - ax;
            gm(S.r)[row] = ri;
            gm(S.p)[row] = 0.0;
            gm(S.q)[row] = 0.0;
            if (Mixed) {
                gm(S.r32)[row] = static_cast<float>(ri);
            }
            rr = ri * ri;
            bb = bi * bi;
        }
    }
    double v[2] = {rr, bb};
    double o[2];
    block_sums<2>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_RR * S.nblk0 + bk_.y] = o[0];
        S.partials[SLOT_BB * S.nblk0 + bk_.y] = o[1];
    }
}

// Jacobi CG start: r = b - A x, z = dinv r, p = 0; partials r.r, r.z, b.b.
template <bool Csr>
__global__ void __launch_bounds__(k_block) k_cg_init(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    double rr = 0.0;
    double rz = 0.0;
    double bb = 0.0;
    if (a0_live<Csr>(S.a64, slice)) {
        const double ax = a0_row<Csr>(S.a64, slice, lane, row, S.x);
        if (row < S.n) {
            const double bi = gm(S.b)[row];
            const double ri = bi - ax;
            const double zi = ri * gm(S.dinv)[row];
            gm(S.r)[row] = ri;
            gm(S.z)[row] = zi;
            gm(S.p)[row] = 0.0;
            rr = ri * ri;
            rz = ri * zi;
            bb = bi * bi;
        }
    }
    double v[3] = {rr, rz, bb};
    double o[3];
    block_sums<3>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_RR * S.nblk0 + bk_.y] = o[0];
        S.partials[SLOT_RZ * S.nblk0 + bk_.y] = o[1];
        S.partials[SLOT_BB * S.nblk0 + bk_.y] = o[2];
    }
}

// y = A x (level-0 FP64 operator) with the partial sum of w.y into `slot`; with Two, also the
// partial of y.w2 into slot2. Used for q = A p (p.q), v = A ph (r0.v), t = A sh (t.t, t.s).
template <bool Two, bool Csr>
__global__ void __launch_bounds__(k_block) k_spmv_dot(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int which_x,
                           int which_y, int which_w, int slot, int which_w2, int slot2) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const double* x = outer_vec(S, which_x);
    double* y = outer_vec(S, which_y);
    const double* w = outer_vec(S, which_w);
    double d1 = 0.0;
    double d2 = 0.0;
    if (a0_live<Csr>(S.a64, slice)) {
        const double yi = a0_row<Csr>(S.a64, slice, lane, row, x);
        if (row < S.n) {
            y[row] = yi;
            d1 = w[row] * yi;
            if (Two) {
                d2 = yi * outer_vec(S, which_w2)[row];
            }
        }
    }
    if (Two) {
        double v[2] = {d1, d2};
        double o[2];
        block_sums<2>(v, o);
        if (threadIdx.x == 0) {
            S.partials[slot * S.nblk0 + bk_.y] = o[0];
            S.partials[slot2 * S.nblk0 + bk_.y] = o[1];
        }
    } else {
        double v[1] = {d1};
        double o[1];
        block_sums<1>(v, o);
        if (threadIdx.x == 0) {
            S.partials[slot * S.nblk0 + bk_.y] = o[0];
        }
    }
}

// x += alpha p; r -= alpha q; partial r.r. Jacobi CG also forms z = dinv r and r.z; the
// mixed solver writes r32 = (float) r for the FP32 V-cycle.
// Search directions alternate between two buffers: iteration k reads p_old = parity ? p : p2
// and writes p_new = parity ? p2 : p (parity = k mod 2, iterations counted from 1).
__device__ __forceinline__ double* p_old_of(const SysDev& S, int parity) { return (parity ^ S.flip) ? S.p : S.p2; }
__device__ __forceinline__ double* p_new_of(const SysDev& S, int parity) { return (parity ^ S.flip) ? S.p2 : S.p; }

// q = A p_new with p_new = z + beta p_old formed per gathered column (the reference's separate
// p update, krylov_kernels.inc:293-296, folded into the next product; identical arithmetic),
// p_new written for the own rows, and the partial p_new.q. z is the level-0 V-cycle output
// v[z_buf] (which = 1, FP64 or FP32) or the Jacobi-preconditioned residual S.z (which = 0).
template <class TZ, bool Csr>
__global__ void __launch_bounds__(k_block) k_spmv_pz(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int parity,
                          int which, int z_buf) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    const EllView A0 = S.a64; // requested with the status (one wait)
    if ((S.st->status != ST_RUNNING) | (A0.tpr <= 0)) {
        return;
    }
    const auto z = gm(which == 0 ? reinterpret_cast<const TZ*>(S.z) : static_cast<const TZ*>(S.lv[0].v[z_buf]));
    const auto po = gm(static_cast<const double*>(p_old_of(S, parity)));
    const auto pn = gm(p_new_of(S, parity));
    const double beta = S.st->beta;
    double d = 0.0;
    if (a0_live<Csr>(A0, slice)) {
        const double qi = a0_row_f<Csr>(A0, slice, lane, row, [&](int c) {
            return static_cast<double>(z[c]) + beta * po[c];
        });
        if (row < S.n) {
            const double pi = static_cast<double>(z[row]) + beta * po[row];
            pn[row] = pi;
            gm(S.q)[row] = qi;
            d = pi * qi;
        }
    }
    double v[1] = {d};
    double o[1];
    block_sums<1>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_PQ * S.nblk0 + bk_.y] = o[0];
    }
}

// x = 0 for every system of a launch table (one launch for a whole batch).
__global__ void __launch_bounds__(k_block) k_zero_x(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (row < S.n) {
        gm(S.x)[row] = 0.0;
    }
}

template <int Mode> // 0 AMG FP64, 1 Jacobi CG, 2 mixed AMG
__global__ void __launch_bounds__(k_block) k_update_xr(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int parity) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const double alpha = S.st->alpha;
    double rr = 0.0;
    double rz = 0.0;
    if (row < S.n) {
        gm(S.x)[row] = gm(S.x)[row] + alpha * p_new_of(S, parity)[row];
        const double ri = gm(S.r)[row] - alpha * gm(S.q)[row];
        gm(S.r)[row] = ri;
        rr = ri * ri;
        if (Mode == 1) {
            const double zi = ri * gm(S.dinv)[row];
            gm(S.z)[row] = zi;
            rz = ri * zi;
        }
        if (Mode == 2) {
            gm(S.r32)[row] = static_cast<float>(ri);
        }
    }
    if (Mode == 1) {
        double v[2] = {rr, rz};
        double o[2];
        block_sums<2>(v, o);
        if (threadIdx.x == 0) {
            S.partials[SLOT_RR * S.nblk0 + bk_.y] = o[0];
            S.partials[SLOT_RZ * S.nblk0 + bk_.y] = o[1];
        }
    } else {
        double v[1] = {rr};
        double o[1];
        block_sums<1>(v, o);
        if (threadIdx.x == 0) {
            S.partials[SLOT_RR * S.nblk0 + bk_.y] = o[0];
        }
    }
}

// p = z + beta p. which = 0: z is the Jacobi-preconditioned residual S.z (FP64); which = 1:
// z is the level-0 V-cycle output v[z_buf] (FP64, or FP32 in the mixed solver).
template <class TZ>
__global__ void __launch_bounds__(k_block) k_p_update(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int which,
                           int z_buf) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const TZ* z = which == 0 ? reinterpret_cast<const TZ*>(S.z)
                             : static_cast<const TZ*>(S.lv[0].v[z_buf]);
    const double beta = S.st->beta;
    if (row < S.n) {
        gm(S.p)[row] = static_cast<double>(z[row]) + beta * gm(S.p)[row];
    }
}

// True residual |b - A x|^2 partials (always runs, after the solve).
template <bool Csr>
__global__ void __launch_bounds__(k_block) k_true_residual(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    double t = 0.0;
    if (a0_live<Csr>(S.a64, slice)) {
        const double ax = a0_row<Csr>(S.a64, slice, lane, row, S.x);
        if (row < S.n) {
            const double ri = gm(S.b)[row] - ax;
            t = ri * ri;
        }
    }
    double v[1] = {t};
    double o[1];
    block_sums<1>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_TRUE * S.nblk0 + bk_.y] = o[0];
    }
}

// ------------------------------------------------------------------------------------------
// BiCGStab kernels (krylov_kernels.inc:96-169)
// ------------------------------------------------------------------------------------------

// r = b - A x; r0 = r; p = v = 0; partials r.r, b.b and r0.r.
template <bool Csr>
__global__ void __launch_bounds__(k_block) k_bicg_init(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    double rr = 0.0;
    double bb = 0.0;
    if (a0_live<Csr>(S.a64, slice)) {
        const double ax = a0_row<Csr>(S.a64, slice, lane, row, S.x);
        if (row < S.n) {
            const double bi = gm(S.b)[row];
            const double ri = bi - ax;
            gm(S.r)[row] = ri;
            S.r0[row] = ri;
            gm(S.p)[row] = 0.0;
            gm(S.v)[row] = 0.0;
            rr = ri * ri;
            bb = bi * bi;
        }
    }
    double v[2] = {rr, bb};
    double o[2];
    block_sums<2>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_RR * S.nblk0 + bk_.y] = o[0];
        S.partials[SLOT_BB * S.nblk0 + bk_.y] = o[1];
        S.partials[SLOT_R0R * S.nblk0 + bk_.y] = o[0];
    }
}

// p = r + beta (p - omega v); ph = dinv p.
__global__ void __launch_bounds__(k_block) k_bicg_p(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const double beta = S.st->beta;
    const double omega = S.st->omega;
    if (row < S.n) {
        const double pi = gm(S.r)[row] + beta * (gm(S.p)[row] - omega * gm(S.v)[row]);
        gm(S.p)[row] = pi;
        gm(S.ph)[row] = pi * gm(S.dinv)[row];
    }
}

// s = r - alpha v; sh = dinv s; partial s.s.
__global__ void __launch_bounds__(k_block) k_bicg_s(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const double alpha = S.st->alpha;
    double ss = 0.0;
    if (row < S.n) {
        const double si = gm(S.r)[row] - alpha * gm(S.v)[row];
        S.s[row] = si;
        gm(S.sh)[row] = si * gm(S.dinv)[row];
        ss = si * si;
    }
    double v[1] = {ss};
    double o[1];
    block_sums<1>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_SS * S.nblk0 + bk_.y] = o[0];
    }
}

// Early exit of BiCGStab: x += alpha ph (status ST_CONVERGED_PENDING only).
__global__ void __launch_bounds__(k_block) k_bicg_early_x(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_CONVERGED_PENDING) {
        return;
    }
    if (row < S.n) {
        gm(S.x)[row] = gm(S.x)[row] + S.st->alpha * gm(S.ph)[row];
    }
}

// x += alpha ph + omega sh; r = s - omega t; partials r.r and r0.r.
__global__ void __launch_bounds__(k_block) k_bicg_xr(const SysDev* __restrict__ sys, const int2* __restrict__ blocks) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const double alpha = S.st->alpha;
    const double omega = S.st->omega;
    double rr = 0.0;
    double r0r = 0.0;
    if (row < S.n) {
        gm(S.x)[row] = gm(S.x)[row] + (alpha * gm(S.ph)[row] + omega * gm(S.sh)[row]);
        const double ri = S.s[row] - omega * S.t[row];
        gm(S.r)[row] = ri;
        rr = ri * ri;
        r0r = S.r0[row] * ri;
    }
    double v[2] = {rr, r0r};
    double o[2];
    block_sums<2>(v, o);
    if (threadIdx.x == 0) {
        S.partials[SLOT_RR * S.nblk0 + bk_.y] = o[0];
        S.partials[SLOT_R0R * S.nblk0 + bk_.y] = o[1];
    }
}

// ------------------------------------------------------------------------------------------
// V-cycle kernels (krylov_kernels.inc:182-238), element type T (double or float)
// ------------------------------------------------------------------------------------------

// Row mapping of a V-cycle kernel over matrix M: 256 / tpr rows per workgroup, tpr lanes per
// row. Early returns here are uniform over the workgroup.
// The system's status and the level's descriptor fields are requested before either is tested
// (one wait instead of a chain of dependent loads: the small levels' kernels are latency-bound).
#define CFDG_LEVEL_PROLOGUE(sysarr, blocks, level, MAT)                                            \
    const int2 bk_ = block_entry(blocks);                                                         \
    const SysDev& S = (sysarr)[bk_.x];                                                            \
    const LevelDev& L = S.lv[level];                                                              \
    const EllView M = L.MAT;                                                                      \
    const int tpr = M.tpr;                                                                        \
    if ((S.st->status != ST_RUNNING) | (L.coarsest != 0) | (tpr <= 0)) {                          \
        return;                                                                                   \
    }                                                                                             \
    const int row = bk_.y * (k_block / tpr) + static_cast<int>(threadIdx.x) / tpr;                \
    const int sub = static_cast<int>(threadIdx.x) & (tpr - 1);                                    \
    (void)sub;

// First two Jacobi sweeps from x = 0: v[out] = x1 + (f2 wd) (b - A x1), x1 = (f1 wd) b
// (sliced-ELL operators; the gathers of wd and b stay cache-resident for short rows). f1, f2:
// the sweeps' factors on the level's damping (Chebyshev weights; 1 for plain damped Jacobi,
// which leaves the arithmetic unchanged).
template <class T>
__global__ void __launch_bounds__(k_block) k_presmooth2(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int out,
                                                        T f1, T f2) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, a)
    const auto b = gm(static_cast<const T*>(L.b));
    const auto wd = gm(static_cast<const T*>(L.wd));
    const T ax = mat_row_x1<T>(M, row, sub, wd, b, f1);
    if (sub == 0 && row < L.n) {
        const T x1 = (f1 * wd[row]) * b[row];
        gm(static_cast<T*>(L.v[out]))[row] = x1 + (f2 * wd[row]) * (b[row] - ax);
    }
}

// ------------------------------------------------------------------------------------------
// V-cycle operations on one row (or one lane of a row's group), used by the level kernels below.
// ------------------------------------------------------------------------------------------

// v[out] = (f wd) b.
template <class T>
__device__ __forceinline__ void op_jacobi_zero(const LevelDev& L, int row, int out, T f) {
    gm(static_cast<T*>(L.v[out]))[row] = (f * gm(static_cast<const T*>(L.wd))[row]) * gm(static_cast<const T*>(L.b))[row];
}

// v[out] = v[in] + (f wd) (b - A v[in]); returns the new entry (0 for rows past the end /
// sub > 0).
template <class T>
__device__ __forceinline__ T op_jacobi(const LevelDev& L, const EllView& M, int row, int sub, int in, int out, T f) {
    const T* x = static_cast<const T*>(L.v[in]);
    const auto b = gm(static_cast<const T*>(L.b));
    const auto wd = gm(static_cast<const T*>(L.wd));
    const T ax = mat_row<T>(M, row, sub, x);
    T z = T(0);
    if (sub == 0 && row < L.n) {
        z = x[row] + (f * wd[row]) * (b[row] - ax);
        gm(static_cast<T*>(L.v[out]))[row] = z;
    }
    return z;
}

// v[out] = b - A v[in].
template <class T>
__device__ __forceinline__ void op_residual(const LevelDev& L, const EllView& M, int row, int sub, int in, int out) {
    const T ax = mat_row<T>(M, row, sub, static_cast<const T*>(L.v[in]));
    if (sub == 0 && row < L.n) {
        gm(static_cast<T*>(L.v[out]))[row] = gm(static_cast<const T*>(L.b))[row] - ax;
    }
}

// v[out] = b - A x1 with x1 = (f wd) b formed per gathered column.
template <class T>
__device__ __forceinline__ void op_residual_x1(const LevelDev& L, const EllView& M, int row, int sub, int out, T f) {
    const auto b = gm(static_cast<const T*>(L.b));
    const auto wd = gm(static_cast<const T*>(L.wd));
    const T ax = mat_row_x1<T>(M, row, sub, wd, b, f);
    if (sub == 0 && row < L.n) {
        gm(static_cast<T*>(L.v[out]))[row] = b[row] - ax;
    }
}

// b_{l+1} = M x for a row of level l + 1 (M = R_l on a level vector, or the fused restriction
// on b_l); with x1_out >= 0 also v_{l+1}[x1_out] = (f wd) b unless level l + 1 is the coarsest.
template <class T>
__device__ __forceinline__ void op_restrict(const LevelDev& L, const LevelDev& N, const EllView& M, int row, int sub,
                                            const T* x, int x1_out, T f) {
    const T y = mat_row<T>(M, row, sub, x);
    if (sub == 0 && row < L.n_next) {
        gm(static_cast<T*>(N.b))[row] = y;
        if (x1_out >= 0 && !N.coarsest) {
            gm(static_cast<T*>(N.v[x1_out]))[row] = (f * gm(static_cast<const T*>(N.wd))[row]) * y;
        }
    }
}

// v_l[io] += P_l v_{l+1}[coarse_buf]; X1: v_l[io] = (f wd) b + P_l v_{l+1}[coarse_buf] (the
// level's single pre-sweep from zero was not stored).
template <class T, bool X1>
__device__ __forceinline__ void op_prolong(const LevelDev& L, const LevelDev& N, const EllView& M, int row, int sub,
                                           int io, int coarse_buf, T f) {
    const T t = mat_row<T>(M, row, sub, static_cast<const T*>(N.v[coarse_buf]));
    if (sub == 0 && row < L.n) {
        const auto x = gm(static_cast<T*>(L.v[io]));
        if constexpr (X1) {
            x[row] = (f * gm(static_cast<const T*>(L.wd))[row]) * gm(static_cast<const T*>(L.b))[row] + t;
        } else {
            x[row] = x[row] + t;
        }
    }
}

// Coarsest level, row i: v[out]_i = sum_j Cinv_ij b_j in FP64 by one wavefront (lane-strided
// partial sums, then the wavefront butterfly).
template <class T>
__device__ __forceinline__ void op_coarse_row(const LevelDev& L, int i, int lane, int out) {
    const int n = L.n;
    const double* m = L.cinv + static_cast<long long>(i) * n;
    const auto b = gm(static_cast<const T*>(L.b));
    double s = 0.0;
    for (int j = lane; j < n; j += k_wave) {
        s += m[j] * static_cast<double>(b[j]);
    }
    s = wave_sum(s);
    if (lane == 0) {
        gm(static_cast<T*>(L.v[out]))[i] = static_cast<T>(s);
    }
}

// One Jacobi sweep from x = 0: v[out] = (f wd) b (row-wise, 256 rows per workgroup).
template <class T>
__global__ void __launch_bounds__(k_block) k_jacobi_zero(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int out,
                                                         T f) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    const LevelDev& L = S.lv[level]; // fields requested with the status (one wait)
    if ((S.st->status != ST_RUNNING) | (L.coarsest != 0) | (row >= L.n)) {
        return;
    }
    op_jacobi_zero<T>(L, row, out, f);
}

// v[out] = 0 (smoothing-free V-cycles; row-wise).
template <class T>
__global__ void __launch_bounds__(k_block) k_zero_level(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int out) {
    CFDG_BLOCK_PROLOGUE(sys, blocks)
    if (S.st->status != ST_RUNNING) {
        return;
    }
    const LevelDev& L = S.lv[level];
    if (L.coarsest || row >= L.n) {
        return;
    }
    gm(static_cast<T*>(L.v[out]))[row] = T(0);
}

// Jacobi sweep v[out] = v[in] + (f wd) (b - A v[in]). Dots (level 0 only, sliced ELL): 0 none;
// 1 partial r.z with r = b (FP64 level 0, where b aliases r); 2 partials r.z and z.q with r, q
// the FP64 outer vectors (mixed level 0).
template <class T, int Dots>
__global__ void __launch_bounds__(k_block) k_jacobi(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int in, int out,
                                                    T f) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, a)
    const T z = op_jacobi<T>(L, M, row, sub, in, out, f);
    double d1 = 0.0;
    double d2 = 0.0;
    if (Dots != 0 && sub == 0 && row < L.n) {
        if (Dots == 1) {
            d1 = static_cast<double>(gm(static_cast<const T*>(L.b))[row]) * static_cast<double>(z);
        }
        if (Dots == 2) {
            const double zd = static_cast<double>(z);
            d1 = gm(S.r)[row] * zd;
            d2 = zd * gm(S.q)[row];
        }
    }
    if (Dots == 1) {
        double v[1] = {d1};
        double o[1];
        block_sums<1>(v, o);
        if (threadIdx.x == 0) {
            S.partials[SLOT_RZ * S.nblk0 + bk_.y] = o[0];
        }
    }
    if (Dots == 2) {
        double v[2] = {d1, d2};
        double o[2];
        block_sums<2>(v, o);
        if (threadIdx.x == 0) {
            S.partials[SLOT_RZ * S.nblk0 + bk_.y] = o[0];
            S.partials[SLOT_ZQ * S.nblk0 + bk_.y] = o[1];
        }
    }
}

// Level-0 FP32 post-sweep with two slices per wavefront (the finest operator is sliced ELL with
// one lane per row): a workgroup of k_block_pair threads covers the same 256 rows as a k_block
// one, wave w serving slices 4 y + 2 w and 4 y + 2 w + 1, so every table entry keeps its meaning.
// Each row is computed as in k_jacobi and the block partials add the four slices' wavefront sums
// in slice order: the results are bit-identical to k_jacobi's.
constexpr int k_block_pair = k_block / 2;

#define CFDG_PAIR_PROLOGUE(sysarr, blocks, level)                                                 \
    const int2 bk_ = block_entry(blocks);                                                         \
    const SysDev& S = (sysarr)[bk_.x];                                                            \
    const LevelDev& L = S.lv[level];                                                              \
    const EllView M = L.a;                                                                        \
    if ((S.st->status != ST_RUNNING) | (L.coarsest != 0) | (M.tpr <= 0)) {                        \
        return;                                                                                   \
    }                                                                                             \
    const int lane = static_cast<int>(threadIdx.x) % k_wave;                                      \
    const int wave_ = static_cast<int>(threadIdx.x) / k_wave;                                     \
    const int s0 = bk_.y * k_slices_per_block + 2 * wave_;                                        \
    const int row0 = s0 * k_wave + lane;                                                          \
    const int row1 = row0 + k_wave;

// Jacobi sweep v[out] = v[in] + (f wd) (b - A v[in]) of level 0 with the partial dots of k_jacobi.
template <class T, int Dots>
__global__ void __launch_bounds__(k_block_pair) k_jacobi_pair(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int in,
                                                              int out, T f) {
    CFDG_PAIR_PROLOGUE(sys, blocks, level)
    const T* x = static_cast<const T*>(L.v[in]);
    const auto gx = gm(x);
    const auto b = gm(static_cast<const T*>(L.b));
    const auto wd = gm(static_cast<const T*>(L.wd));
    T ya;
    T yb;
    ell_row_pair<T>(M, s0, lane, [&](int c) { return gx[c]; }, ya, yb);
    const auto o = gm(static_cast<T*>(L.v[out]));
    double d1[2] = {0.0, 0.0};
    double d2[2] = {0.0, 0.0};
    const int rows[2] = {row0, row1};
    const T ys[2] = {ya, yb};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int row = rows[h];
        if (row < L.n) {
            const T z = gx[row] + (f * wd[row]) * (b[row] - ys[h]);
            o[row] = z;
            if (Dots == 1) {
                d1[h] = static_cast<double>(b[row]) * static_cast<double>(z);
            }
            if (Dots == 2) {
                const double zd = static_cast<double>(z);
                d1[h] = gm(S.r)[row] * zd;
                d2[h] = zd * gm(S.q)[row];
            }
        }
    }
    if (Dots != 0) {
        __shared__ double red[2][k_slices_per_block];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const double w1 = wave_sum(d1[h]);
            const double w2 = Dots == 2 ? wave_sum(d2[h]) : 0.0;
            if (lane == 0) {
                red[0][2 * wave_ + h] = w1;
                red[1][2 * wave_ + h] = w2;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double t1 = red[0][0];
            double t2 = red[1][0];
#pragma unroll
            for (int w = 1; w < k_slices_per_block; ++w) {
                t1 += red[0][w];
                t2 += red[1][w];
            }
            S.partials[SLOT_RZ * S.nblk0 + bk_.y] = t1;
            if (Dots == 2) {
                S.partials[SLOT_ZQ * S.nblk0 + bk_.y] = t2;
            }
        }
    }
}

// Rows of one table entry (the 256 / tpr rows a k_block workgroup covers) for a k_block_pair
// workgroup: sliced ELL reads slices 4 y + 2 w and 4 y + 2 w + 1 together in wave w; CSR rows run
// as the two halves of the k_block lanes. epi(row, sub, value) gets every lane's row product
// (complete in every lane of the row's group); rows past the end give 0. Bit-identical to
// mat_row on the same rows.
template <class T, class F, class E>
__device__ __forceinline__ void pair_rows(const EllView& m, int y, F load, E epi) {
    const int lane = static_cast<int>(threadIdx.x) % k_wave;
    const int tpr = m.tpr;
    if (m.csr == 0) {
        const int lg = __builtin_ctz(tpr);
        const int s0 = y * k_slices_per_block + 2 * (static_cast<int>(threadIdx.x) / k_wave);
        T ya;
        T yb;
        ell_row_pair<T>(m, s0, lane, load, ya, yb);
        if (tpr > 1) {
            ya = group_sum(ya, tpr);
            yb = group_sum(yb, tpr);
        }
        const int r0 = s0 * (k_wave >> lg) + (lane >> lg);
        epi(r0, lane & (tpr - 1), ya);
        epi(r0 + (k_wave >> lg), lane & (tpr - 1), yb);
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int vt = static_cast<int>(threadIdx.x) + h * k_block_pair;
        const int row = y * (k_block / tpr) + vt / tpr;
        T v = T(0);
        if (row < m.n_rows) {
            v = csr_lane_sum<T>(gm(m.col), gm(static_cast<const T*>(m.val)), m.sp[row] + (vt & (tpr - 1)),
                                m.sp[row + 1], tpr, load);
        }
        epi(row, vt & (tpr - 1), group_sum(v, tpr));
    }
}

#define CFDG_PAIR_LEVEL_PROLOGUE(sysarr, blocks, level)                                           \
    const int2 bk_ = block_entry(blocks);                                                         \
    const SysDev& S = (sysarr)[bk_.x];                                                            \
    const LevelDev& L = S.lv[level];                                                              \
    if ((S.st->status != ST_RUNNING) | (L.coarsest != 0)) {                                       \
        return;                                                                                   \
    }

// k_prolong with pair_rows: v_l[io] += P_l v_{l+1}[coarse_buf]; X1: v_l[io] = (f wd) b +
// P_l v_{l+1}.
template <class T, bool X1>
__global__ void __launch_bounds__(k_block_pair) k_prolong_pair(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int io,
                                                               int coarse_buf, T f) {
    CFDG_PAIR_LEVEL_PROLOGUE(sys, blocks, level)
    const LevelDev& N = S.lv[level + 1];
    const auto e = gm(static_cast<const T*>(N.v[coarse_buf]));
    const auto x = gm(static_cast<T*>(L.v[io]));
    pair_rows<T>(L.p, bk_.y, [&](int c) { return e[c]; }, [&](int row, int sub, T t) {
        if (sub == 0 && row < L.n) {
            if constexpr (X1) {
                x[row] = (f * gm(static_cast<const T*>(L.wd))[row]) * gm(static_cast<const T*>(L.b))[row] + t;
            } else {
                x[row] = x[row] + t;
            }
        }
    });
}

// v[out] = b - A v[in].
template <class T>
__global__ void __launch_bounds__(k_block) k_residual(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int in, int out) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, a)
    op_residual<T>(L, M, row, sub, in, out);
}

// b_{l+1} = R_l v_l[in]; rows of level l + 1. With x1_out >= 0 the next level's first
// Jacobi sweep from zero is stored too, v_{l+1}[x1_out] = (f wd) b_{l+1} (what k_jacobi_zero
// would compute from the stored b), unless level l + 1 is the coarsest.
template <class T>
__global__ void __launch_bounds__(k_block) k_restrict(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int in,
                                                      int x1_out, T f) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, r)
    op_restrict<T>(L, S.lv[level + 1], M, row, sub, static_cast<const T*>(L.v[in]), x1_out, f);
}

// Down leg of a level with one Jacobi sweep from zero as one product: b_{l+1} = R (I - A W) b_l,
// W = diag(wd) (x1 = W b, r = b - A x1, b_{l+1} = R r in exact arithmetic); rows of level l + 1;
// x1_out as k_restrict.
template <class T>
__global__ void __launch_bounds__(k_block) k_restrict_fused(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level,
                                                            int x1_out, T f) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, rt)
    op_restrict<T>(L, S.lv[level + 1], M, row, sub, static_cast<const T*>(L.b), x1_out, f);
}

// v_l[io] += P_l v_{l+1}[coarse_buf]. X1: the level's single pre-sweep from zero was not
// stored; its result (f wd) b is formed here (v[io] = (f wd) b + P v_{l+1}, the same sums).
template <class T, bool X1 = false>
__global__ void __launch_bounds__(k_block) k_prolong(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int io,
                          int coarse_buf, T f) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, p)
    op_prolong<T, X1>(L, S.lv[level + 1], M, row, sub, io, coarse_buf, f);
}

// v[out] = b - A x1 with x1 = (f wd) b, the single Jacobi sweep from zero, formed per
// gathered column instead of stored (one pre-sweep per side).
template <class T>
__global__ void __launch_bounds__(k_block) k_residual_x1(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int out,
                                                         T f) {
    CFDG_LEVEL_PROLOGUE(sys, blocks, level, a)
    op_residual_x1<T>(L, M, row, sub, out, f);
}

// Coarsest level: v[out] = C^-1 b with the dense FP64 inverse, one wavefront per row; blocks
// index groups of four rows.
template <class T>
__global__ void __launch_bounds__(k_block) k_coarse(const SysDev* __restrict__ sys, const int2* __restrict__ blocks, int level, int out) {
    const int2 bk = block_entry(blocks);
    const SysDev& S = sys[bk.x];
    const LevelDev& L = S.lv[level]; // fields requested with the status (one wait)
    const int wave = threadIdx.x / k_wave;
    const int lane = threadIdx.x % k_wave;
    const int i = bk.y * k_slices_per_block + wave;
    if ((S.st->status != ST_RUNNING) | (L.coarsest == 0) | (i >= L.n)) {
        return;
    }
    op_coarse_row<T>(L, i, lane, out);
}

// ------------------------------------------------------------------------------------------
// Finalize kernels: one workgroup per system, partials added in a fixed order.
// ------------------------------------------------------------------------------------------
constexpr int k_fin_block = 1024;

__device__ __forceinline__ double reduce_slot(const SysDev& S, int slot) {
    __shared__ double red[k_fin_block / k_wave];
    const double* p = S.partials + static_cast<long long>(slot) * S.nblk0;
    double s = 0.0;
    for (int i = threadIdx.x; i < S.nblk0; i += k_fin_block) {
        s += p[i];
    }
    s = wave_sum(s);
    if (threadIdx.x % k_wave == 0) {
        red[threadIdx.x / k_wave] = s;
    }
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int w = 1; w < k_fin_block / k_wave; ++w) {
        t += red[w];
    }
    __syncthreads();
    return t;
}

enum FinOp : int {
    FIN_AMG_INIT = 0,   // res0, |b|; converged before iterating?
    FIN_ALPHA = 1,      // pq; breakdown if 0; alpha = rz / pq
    FIN_RES = 2,        // res; converged?
    FIN_BETA = 3,       // rz_new; beta = rz_new / rz (0 first); it++ / limit
    FIN_BETA_FLEX = 4,  // rz_new, zq; beta = -zq / pq (0 first); it++ / limit
    FIN_CG_INIT = 5,    // Jacobi CG start: res0, |b|, rz; beta = 0; it = 1
    FIN_CG_RES_BETA = 6,// Jacobi CG: res check then beta
    FIN_TRUE = 7,       // true residual
    FIN_BICG_INIT = 8,
    FIN_BICG_ALPHA = 10,
    FIN_BICG_S = 11,
    FIN_BICG_OMEGA = 12,
    FIN_BICG_RES = 13
};

// Start of a BiCGStab iteration (krylov_kernels.inc:96-169): rho_new = r0.r (kept in rz), the
// iteration limit, the breakdown guards rho_new == 0 and omega == 0, and beta. Run by thread 0 of
// the finalize that ends the previous step (the start-up or the previous iteration's residual),
// so an iteration needs no launch of its own for it.
__device__ __forceinline__ void bicg_rho(SolveState* st) {
    const double rho_new = st->rz;
    if (st->it > st->max_it) {
        st->status = ST_MAXIT;
    } else if (rho_new == 0.0 || st->omega == 0.0) {
        st->status = ST_BREAKDOWN;
    } else {
        st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
        st->rho = rho_new;
    }
}

template <int Op>
__device__ __forceinline__ void finalize_step(const SysDev& S) {
    SolveState* st = S.st;
    const int status = st->status;
    if (Op == FIN_TRUE) {
        const double t = reduce_slot(S, SLOT_TRUE);
        if (threadIdx.x == 0) {
            st->true_res = sqrt(t);
        }
        return;
    }
    if (status != ST_RUNNING && !(Op == FIN_BICG_OMEGA && status == ST_CONVERGED_PENDING)) {
        return;
    }
    if (Op == FIN_AMG_INIT || Op == FIN_CG_INIT || Op == FIN_BICG_INIT) {
        const double rr = reduce_slot(S, SLOT_RR);
        const double bb = reduce_slot(S, SLOT_BB);
        double rz = 0.0;
        if (Op == FIN_CG_INIT) {
            rz = reduce_slot(S, SLOT_RZ);
        }
        double r0r = 0.0;
        if (Op == FIN_BICG_INIT) {
            r0r = reduce_slot(S, SLOT_R0R);
        }
        if (threadIdx.x == 0) {
            const double res = sqrt(rr);
            st->res = res;
            st->bnorm = sqrt(bb);
            st->it = 0;
            if (res <= st->tol) {
                st->status = ST_CONVERGED;
            } else if (Op == FIN_CG_INIT) {
                st->rz = rz;
                st->beta = 0.0;
                st->it = 1;
                if (st->it > st->max_it) {
                    st->status = ST_MAXIT;
                }
            } else if (Op == FIN_BICG_INIT) {
                st->rho = 1.0;
                st->alpha = 1.0;
                st->omega = 1.0;
                st->rz = r0r; // rho_new of the first iteration
                st->it = 1;
                bicg_rho(st);
            }
        }
        return;
    }
    if (Op == FIN_ALPHA) {
        const double pq = reduce_slot(S, SLOT_PQ);
        if (threadIdx.x == 0) {
            st->pq = pq;
            if (pq == 0.0) {
                st->status = ST_BREAKDOWN;
            } else {
                st->alpha = st->rz / pq;
            }
        }
        return;
    }
    if (Op == FIN_RES) {
        const double rr = reduce_slot(S, SLOT_RR);
        if (threadIdx.x == 0) {
            const double res = sqrt(rr);
            st->res = res;
            if (res <= st->tol) {
                st->status = ST_CONVERGED;
            }
        }
        return;
    }
    if (Op == FIN_BETA || Op == FIN_BETA_FLEX) {
        const double rz_new = reduce_slot(S, SLOT_RZ);
        double zq = 0.0;
        if (Op == FIN_BETA_FLEX) {
            zq = reduce_slot(S, SLOT_ZQ);
        }
        if (threadIdx.x == 0) {
            if (st->it == 0) {
                st->beta = 0.0;
            } else if (Op == FIN_BETA) {
                st->beta = rz_new / st->rz;
            } else {
                st->beta = -zq / st->pq;
            }
            st->rz = rz_new;
            st->it += 1;
            if (st->it > st->max_it) {
                st->status = ST_MAXIT;
            }
        }
        return;
    }
    if (Op == FIN_CG_RES_BETA) {
        const double rr = reduce_slot(S, SLOT_RR);
        const double rz_new = reduce_slot(S, SLOT_RZ);
        if (threadIdx.x == 0) {
            const double res = sqrt(rr);
            st->res = res;
            if (res <= st->tol) {
                st->status = ST_CONVERGED;
            } else {
                st->beta = rz_new / st->rz;
                st->rz = rz_new;
                st->it += 1;
                if (st->it > st->max_it) {
                    st->status = ST_MAXIT;
                }
            }
        }
        return;
    }
    if (Op == FIN_BICG_ALPHA) {
        const double r0v = reduce_slot(S, SLOT_R0V);
        if (threadIdx.x == 0) {
            if (r0v == 0.0) {
                st->status = ST_BREAKDOWN;
            } else {
                st->alpha = st->rho / r0v;
            }
        }
        return;
    }
    if (Op == FIN_BICG_S) {
        const double ss = reduce_slot(S, SLOT_SS);
        if (threadIdx.x == 0) {
            const double s_norm = sqrt(ss);
            if (s_norm <= st->tol) {
                st->res = s_norm;
                st->status = ST_CONVERGED_PENDING;
            }
        }
        return;
    }
    if (Op == FIN_BICG_OMEGA) {
        const double tt = reduce_slot(S, SLOT_TT);
        const double ts = reduce_slot(S, SLOT_TS);
        if (threadIdx.x == 0) {
            if (status == ST_CONVERGED_PENDING) {
                st->status = ST_CONVERGED;
            } else {
                st->omega = tt > 0.0 ? ts / tt : 0.0;
            }
        }
        return;
    }
    if (Op == FIN_BICG_RES) {
        const double rr = reduce_slot(S, SLOT_RR);
        const double r0r = reduce_slot(S, SLOT_R0R);
        if (threadIdx.x == 0) {
            const double res = sqrt(rr);
            st->res = res;
            if (res <= st->tol) {
                st->status = ST_CONVERGED;
            } else {
                st->rz = r0r;
                st->it += 1;
                bicg_rho(st);
            }
        }
        return;
    }
}

// One workgroup per system (ids: descriptor indices). With status_out, the system's status
// after the step is also stored at status_out[id]: the last step of an iteration reports every
// system's state to host-mapped memory this way, with no launch of its own.
template <int Op>
__global__ void __launch_bounds__(k_fin_block) k_finalize(const SysDev* __restrict__ sys, const int* __restrict__ ids,
                                                         int* __restrict__ status_out) {
    const int id = ids[blockIdx.x];
    finalize_step<Op>(sys[id]);
    if (status_out != nullptr) {
        __syncthreads();
        if (threadIdx.x == 0) {
            status_out[id] = sys[id].st->status;
        }
    }
}

// ------------------------------------------------------------------------------------------
// Normalised residual (ldu_matrix.cpp:65-103): res = sum_c |b_c - (Ax)_c| and
// norm = sum_c |(Ax)_c - (A xref)_c| + |b_c - (A xref)_c|, with A xref = rowsum(A) * mean(x).
// Three passes with fixed-order sums: block sums of x, their total, then per row the two terms
// and (b - Ax)^2 (the level-0 FP64 operator, one lane per row, rows of 256-row blocks).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(k_block) k_nr_xsum(int n, const double* __restrict__ x, double* __restrict__ part) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    double v[1] = {i < n ? x[i] : 0.0};
    double o[1];
    block_sums<1>(v, o);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = o[0];
    }
}

// sums[0] = sum of x (k_nr_reduce); part receives three arrays of nblk block sums.
__global__ void __launch_bounds__(k_block) k_nr_rows(EllView a, int n, const double* __restrict__ x,
                                                     const double* __restrict__ b, const double* __restrict__ sums,
                                                     double* __restrict__ part, int nblk) {
    const int wave = threadIdx.x / k_wave;
    const int lane = threadIdx.x % k_wave;
    const int slice = blockIdx.x * k_slices_per_block + wave;
    const int row = slice * k_wave + lane;
    const double mean = sums[0] / static_cast<double>(n);
    double t[3] = {0.0, 0.0, 0.0};
    if (slice < a.n_slices) {
        const double ax = ell_row<double>(a, slice, lane, x);
        const double rs = ell_row_f<double>(a, slice, lane, [](int) { return 1.0; });
        if (row < n) {
            const double bi = b[row];
            const double ax_ref = rs * mean;
            const double r = bi - ax;
            t[0] = fabs(r);
            t[1] = fabs(ax - ax_ref) + fabs(bi - ax_ref);
            t[2] = r * r;
        }
    }
    double o[3];
    block_sums<3>(t, o);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = o[0];
        part[nblk + blockIdx.x] = o[1];
        part[2 * nblk + blockIdx.x] = o[2];
    }
}

// out[q] = sum over i < nblk of part[q nblk + i], q < count, in a fixed order (one workgroup).
__global__ void __launch_bounds__(k_fin_block) k_nr_reduce(const double* __restrict__ part, int nblk, int count,
                                                           double* __restrict__ out) {
    __shared__ double red[k_fin_block / k_wave];
    for (int q = 0; q < count; ++q) {
        const double* p = part + static_cast<long long>(q) * nblk;
        double s = 0.0;
        for (int i = threadIdx.x; i < nblk; i += k_fin_block) {
            s += p[i];
        }
        s = wave_sum(s);
        if (threadIdx.x % k_wave == 0) {
            red[threadIdx.x / k_wave] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = red[0];
            for (int w = 1; w < k_fin_block / k_wave; ++w) {
                t += red[w];
            }
            out[q] = t;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Assembly kernels
// ------------------------------------------------------------------------------------------

// Level-0 values from LDU coefficients: val[k] = ldu[src[k]] (0 for padding), with
// ldu = [diag | upper | lower]; dinv = 1 / diag (1.0 for a zero diagonal,
// krylov_cuda.cu:557-562); wd = w dinv in FP64 and FP32; FP32 copy of the values.
__global__ void __launch_bounds__(k_block) k_ldu_gather(long long n_slots, const int* __restrict__ src, const double* __restrict__ ldu,
                             double* __restrict__ val64, float* __restrict__ val32) {
    const long long k = static_cast<long long>(blockIdx.x) * k_block + threadIdx.x;
    if (k >= n_slots) {
        return;
    }
    const int s = src[k];
    const double v = s >= 0 ? ldu[s] : 0.0;
    val64[k] = v;
    if (val32 != nullptr) {
        val32[k] = static_cast<float>(v);
    }
}

// flag = 1 if some face's upper and lower coefficients differ in any bit (the operator is not
// exactly symmetric); flag is cleared by the caller.
__global__ void __launch_bounds__(k_block) k_ldu_asymmetric(int nf, const double* __restrict__ upper,
                                                            const double* __restrict__ lower, int* __restrict__ flag) {
    const int f = blockIdx.x * k_block + threadIdx.x;
    if (f < nf && __double_as_longlong(upper[f]) != __double_as_longlong(lower[f])) {
        *flag = 1;
    }
}

// dinv = 1 / a_ii (1.0 for a zero diagonal), wd64 = w64 dinv (FP64 smoother), wd32 = w32 dinv
// (mixed smoother).
__global__ void __launch_bounds__(k_block) k_diag_inverse(int n, const double* __restrict__ diag, double w64, double w32,
                                                          double* __restrict__ dinv, double* __restrict__ wd64,
                                                          float* __restrict__ wd32) {
    const int i = blockIdx.x * k_block + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double d = diag[i];
    const double di = d != 0.0 ? 1.0 / d : 1.0;
    dinv[i] = di;
    wd64[i] = w64 * di;
    if (wd32 != nullptr) {
        wd32[i] = static_cast<float>(w32 * di);
    }
}

// Face coefficients of the convection-diffusion operator (krylov_openmp.cpp:289-295).
__global__ void __launch_bounds__(k_block) k_assemble_faces(int nf, const double* __restrict__ flux, const double* __restrict__ gamma,
                                 const double* __restrict__ delta, double convection,
                                 double* __restrict__ upper, double* __restrict__ lower) {
    const int f = blockIdx.x * k_block + threadIdx.x;
    if (f >= nf) {
        return;
    }
    const double fl = flux[f];
    const double g = gamma[f] * delta[f];
    const double pos = fl > 0.0 ? convection * fl : 0.0;
    const double neg = fl < 0.0 ? convection * fl : 0.0;
    upper[f] = neg - g;
    lower[f] = -pos - g;
}

// Per-cell accumulation of diag / divergence over the cell's faces in ascending face order,
// the order of the reference's serial face loop (krylov_openmp.cpp:296-299).
__global__ void __launch_bounds__(k_block) k_assemble_cells(int nc, const int* __restrict__ cf_ptr, const int* __restrict__ cf_face,
                                 const int* __restrict__ owner, const double* __restrict__ flux,
                                 const double* __restrict__ gamma, const double* __restrict__ delta,
                                 double convection, double* __restrict__ diag,
                                 double* __restrict__ divergence) {
    const int c = blockIdx.x * k_block + threadIdx.x;
    if (c >= nc) {
        return;
    }
    double d = diag[c];
    double dv = divergence[c];
    for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
        const int f = cf_face[k];
        const double fl = flux[f];
        const double g = gamma[f] * delta[f];
        const double pos = fl > 0.0 ? convection * fl : 0.0;
        const double neg = fl < 0.0 ? convection * fl : 0.0;
        if (owner[f] == c) {
            d += pos + g;
            dv += fl;
        } else {
            d += -neg + g;
            dv -= fl;
        }
    }
    diag[c] = d;
    divergence[c] = dv;
}

} // namespace cfdg
