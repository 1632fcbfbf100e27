// Finite-volume cell kernels on the device: cell gradients (Green-Gauss, weighted least
// squares, skewness-corrected Gauss), the Barth-Jespersen / Venkatakrishnan limiter and the
// explicit internal-face corrections of the scalar and momentum equations (the operations of
// gpu::DeviceFv, src/gpu/include/cfd/gpu/fv_device.hpp:34-71; per-cell math of
// src/gpu/src/fv_cell.inc:25-260, sequencing of src/gpu/src/fv_openmp.cpp:130-322).
//
// One thread per cell runs over the cell's faces in ascending face order, with the reference's
// operation order and no contraction of multiply-adds, so every result equals the serial CPU face
// loops bit for bit; nothing is atomic. Quantities of a face that both its cells need (the
// explicit correction fluxes, the pressure equation's face flux and coefficient) are computed
// once per face by a face pass on face-major arrays and summed per cell in face order. The face
// data a cell needs is laid
// out once, at creation, in per-cell slots, slot-major over the 64 cells of a wavefront and one
// array per component (the sliced layout of the solver's matrices): every kernel streams it
// coalesced instead of gathering 3-vectors of faces shared with other cells through the face list
// (which re-read 2.6-5x the useful bytes). The three components of a vector gradient are formed in
// one pass over the slots. Workgroups map to XCDs in contiguous cell ranges, so the neighbour
// values a cell gathers sit in the same L2.
#include "common.hpp"
#include "fv.hpp"
#include "hip_util.hpp"

#include <cfdgpu.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

namespace cfdg {
namespace {

constexpr int k_fv_block = 256;
constexpr int k_fv_xcds = 8;

template <class T>
using gptr = const __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> g(const T* p) {
    return (gptr<T>)p;
}

// Workgroup b runs on XCD b mod 8; hand each XCD a contiguous range of cell blocks.
__device__ __forceinline__ int fv_cell() {
    const int n = static_cast<int>(gridDim.x);
    const int b = static_cast<int>(blockIdx.x);
    const int q = n / k_fv_xcds;
    const int r = n % k_fv_xcds;
    const int x = b % k_fv_xcds;
    const int k = b / k_fv_xcds;
    const int blk = x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
    return blk * k_fv_block + static_cast<int>(threadIdx.x);
}

// The mesh as per-cell face slots. Slot k of cell 64 s + lane (slice s = one wavefront) sits at
// base[s] + 64 k + lane, so a wavefront's loads of one slot are 64 consecutive entries; a slice
// has as many slots as its cell with the most faces, and each cell's slots keep its ascending
// face order (padding, face -1, comes last). A face's data is stored in the slots of both its
// cells, component by component: every kernel streams it once, coalesced, instead of gathering
// 3-vectors of faces that other cells share through the face list.
struct FvSlots {
    int nc;
    int nif;
    const int* base;           // n_slices + 1
    const int* face;           // face id (per-call face arrays: flux, gamma, nu), -1 padding
    const int* other;          // internal face: the other cell; boundary face: -1 - boundary index
    const unsigned char* owns; // 1 when the slot's cell owns the face
    const double* w;           // interpolation weight (internal faces)
    const double* sf0;
    const double* sf1;
    const double* sf2;
    const double* ld0;         // least-squares vector d / |d|^2
    const double* ld1;
    const double* ld2;
    const double* cf0;         // face centre
    const double* cf1;
    const double* cf2;
    const double* kv0;         // non-orthogonal correction vector
    const double* kv1;
    const double* kv2;
    const double* cc;          // 3 per cell
    const double* vol;
    const double* lsq;         // 9 per cell, column-major
};

// Slot range of the calling cell's slice: first slot index of this lane and the slice's slot count.
__device__ __forceinline__ void slot_range(const FvSlots& m, int c, int& first, int& count) {
    const int s = c >> 6;
    const int b0 = g(m.base)[s];
    first = b0 + (c & 63);
    count = (g(m.base)[s + 1] - b0) >> 6;
}

// One slot's operands for a gradient: the face (-1 for padding), the slot's vector (least-squares
// d / |d|^2 or the area vector), the weight and the other side's NC values (the neighbour's, or the
// boundary face's). Every load stays in bounds for padding slots too: their `other` is cell 0 and
// their vectors are zero, and boundary indices were validated on the host.
template <bool Lsq, int NC>
struct GradSlot {
    int f;
    bool mine;
    bool internal;
    double v0, v1, v2, w;
    double u[NC];
};

template <bool Lsq, int NC>
__device__ __forceinline__ GradSlot<Lsq, NC> load_grad_slot(const FvSlots& m, int t, const double* phi,
                                                            const double* bv, int stride) {
    GradSlot<Lsq, NC> sl;
    sl.f = g(m.face)[t];
    const int o = g(m.other)[t];
    sl.mine = g(m.owns)[t] != 0;
    sl.internal = o >= 0;
    sl.v0 = Lsq ? g(m.ld0)[t] : g(m.sf0)[t];
    sl.v1 = Lsq ? g(m.ld1)[t] : g(m.sf1)[t];
    sl.v2 = Lsq ? g(m.ld2)[t] : g(m.sf2)[t];
    sl.w = Lsq ? 0.0 : g(m.w)[t];
    const long long at = o >= 0 ? static_cast<long long>(o) * stride : static_cast<long long>(-1 - o) * stride;
    const double* src = o >= 0 ? phi : bv;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        sl.u[i] = g(src)[at + i];
    }
    return sl;
}

// Adds one face's term in the reference's arithmetic: dphi = phi_N - phi_P (the boundary value
// minus the cell's for boundary faces) times d, or +-S_f phi_f with phi_f = w phi_P + (1 - w) phi_N.
// pc holds the calling cell's values, which equal phi[c * stride + i] bit for bit.
template <bool Lsq, int NC>
__device__ __forceinline__ void add_grad_slot(const GradSlot<Lsq, NC>& sl, const double* pc, double* g0, double* g1,
                                              double* g2) {
    if (sl.f < 0) {
        return; // padding
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (Lsq) {
            const double dv = !sl.internal ? sl.u[i] - pc[i] : (sl.mine ? sl.u[i] - pc[i] : pc[i] - sl.u[i]);
            g0[i] += sl.v0 * dv;
            g1[i] += sl.v1 * dv;
            g2[i] += sl.v2 * dv;
        } else if (sl.internal) {
            const double pf = sl.mine ? sl.w * pc[i] + (1.0 - sl.w) * sl.u[i] : sl.w * sl.u[i] + (1.0 - sl.w) * pc[i];
            if (sl.mine) {
                g0[i] += sl.v0 * pf;
                g1[i] += sl.v1 * pf;
                g2[i] += sl.v2 * pf;
            } else {
                g0[i] -= sl.v0 * pf;
                g1[i] -= sl.v1 * pf;
                g2[i] -= sl.v2 * pf;
            }
        } else {
            g0[i] += sl.v0 * sl.u[i];
            g1[i] += sl.v1 * sl.u[i];
            g2[i] += sl.v2 * sl.u[i];
        }
    }
}

// Gradient of NC components of a field with `stride` values per cell and boundary face; derivative
// j of component i goes to out[c * ostride + (3 i + j) * qstride] (qstride 1: interleaved per cell;
// ostride 1 and qstride n_cells: component-major). Least squares: A (sum_f d_f dphi_f), A the cell's
// column-major inverse; Green-Gauss: sum_f S_f phi_f / V with phi_f = w phi_P + (1 - w) phi_N
// (fv_cell.inc:25-73). Slots are loaded two at a time ahead of their terms, which are added in
// slot (ascending face) order; padding slots add nothing.
template <bool Lsq, int NC>
__device__ __forceinline__ void cell_gradient(const FvSlots& m, int c, const double* phi, const double* bv, int stride,
                                              double* out, int ostride, int qstride) {
    const auto fp = g(phi);
    double g0[NC];
    double g1[NC];
    double g2[NC];
    double pc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        g0[i] = 0.0;
        g1[i] = 0.0;
        g2[i] = 0.0;
        pc[i] = fp[c * stride + i];
    }
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    int k = 0;
    for (; k + 2 <= count; k += 2) {
        const int t = first + (k << 6);
        const auto a = load_grad_slot<Lsq, NC>(m, t, phi, bv, stride);
        const auto b = load_grad_slot<Lsq, NC>(m, t + 64, phi, bv, stride);
        add_grad_slot<Lsq, NC>(a, pc, g0, g1, g2);
        add_grad_slot<Lsq, NC>(b, pc, g0, g1, g2);
    }
    if (k < count) {
        add_grad_slot<Lsq, NC>(load_grad_slot<Lsq, NC>(m, first + (k << 6), phi, bv, stride), pc, g0, g1, g2);
    }
    double* out_c = out + static_cast<long long>(c) * ostride;
    if (Lsq) {
        const auto a = g(m.lsq) + 9LL * c;
        const double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6], a7 = a[7],
                     a8 = a[8];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            out_c[static_cast<long long>(3 * i) * qstride] = a0 * g0[i] + a3 * g1[i] + a6 * g2[i];
            out_c[static_cast<long long>(3 * i + 1) * qstride] = a1 * g0[i] + a4 * g1[i] + a7 * g2[i];
            out_c[static_cast<long long>(3 * i + 2) * qstride] = a2 * g0[i] + a5 * g1[i] + a8 * g2[i];
        }
    } else {
        const double v = g(m.vol)[c];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            out_c[static_cast<long long>(3 * i) * qstride] = g0[i] / v;
            out_c[static_cast<long long>(3 * i + 1) * qstride] = g1[i] / v;
            out_c[static_cast<long long>(3 * i + 2) * qstride] = g2[i] / v;
        }
    }
}


// CompMajor: the output is component-major with qstride between components; otherwise each cell's
// 3 NC derivatives are contiguous (the component stride is the constant 1).
template <bool Lsq, int NC, bool CompMajor = false>
__global__ void __launch_bounds__(k_fv_block) k_fv_gradient(FvSlots m, const double* phi, const double* bv, int stride,
                                                            double* out, int ostride, int qstride) {
    const int c = fv_cell();
    if (c < m.nc) {
        cell_gradient<Lsq, NC>(m, c, phi, bv, stride, out, ostride, CompMajor ? qstride : 1);
    }
}


// Limiter of the three components of u (fv_cell.inc:123-164): bounds from the cell and its face
// neighbours (boundary values for boundary faces), then per face and component the ratio of the
// admissible change to the extrapolated one, Barth-Jespersen min(1, bound / d) or the
// Venkatakrishnan function with eps^2 per component; |d| < 1e-300 is skipped.
__global__ void __launch_bounds__(k_fv_block) k_fv_limiter(FvSlots m, const double* u, const double* ub,
                                                           const double* grad, int venkat, double e0, double e1,
                                                           double e2, double* lim) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    const auto fu = g(u);
    const double eps2[3] = {e0, e1, e2};
    double lo[3];
    double hi[3];
    double uc[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        uc[i] = fu[3 * c + i];
        lo[i] = uc[i];
        hi[i] = uc[i];
    }
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        if (g(m.face)[t] < 0) {
            break;
        }
        const int o = g(m.other)[t];
        const auto v = o >= 0 ? fu + 3 * o : g(ub) + 3 * (-1 - o);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double x = v[i];
            lo[i] = x < lo[i] ? x : lo[i];
            hi[i] = x > hi[i] ? x : hi[i];
        }
    }
    double l[3] = {1.0, 1.0, 1.0};
    const auto gc = g(grad) + 9LL * c;
    double gm[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        gm[j] = gc[j];
    }
    const double c0 = g(m.cc)[3 * c];
    const double c1 = g(m.cc)[3 * c + 1];
    const double c2 = g(m.cc)[3 * c + 2];
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        if (g(m.face)[t] < 0) {
            break;
        }
        const double r0 = g(m.cf0)[t] - c0;
        const double r1 = g(m.cf1)[t] - c1;
        const double r2 = g(m.cf2)[t] - c2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double d = gm[3 * i] * r0 + gm[3 * i + 1] * r1 + gm[3 * i + 2] * r2;
            if (d < 1e-300 && d > -1e-300) {
                continue;
            }
            const double bound = d > 0.0 ? hi[i] - uc[i] : lo[i] - uc[i];
            double q;
            if (venkat != 0) {
                q = (bound * bound + eps2[i] + 2.0 * d * bound) / (bound * bound + 2.0 * d * d + d * bound + eps2[i]);
            } else {
                q = bound / d > 0.0 ? bound / d : 0.0;
            }
            l[i] = q < l[i] ? q : l[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        lim[3 * c + i] = l[i];
    }
}


// Face-major mesh data for the face passes: owner / neighbour (checked at creation) and the
// weight, area vector, face centre and correction vector copied from the owners' slots.
struct FvFaces {
    int nif;
    const int* own;
    const int* nei;
    const double* w;
    const double* sf; // 3 per face
    const double* cf; // 3 per face
    const double* kv; // 3 per face
};

// Explicit flux of an internal face for a scalar equation (fv_cell.inc:167-195), one thread per
// face: e_f = -F_f g_up . (x_f - C_up) [linear upwind] + gamma_f k_f . (w g_P + (1 - w) g_N); the
// cells add it with k_fv_face_sum (+ for the owner, - for the neighbour).
__global__ void __launch_bounds__(k_fv_block) k_fv_scalar_face(FvFaces fa, const double* __restrict__ cc,
                                                               const double* __restrict__ grad,
                                                               const double* __restrict__ flux,
                                                               const double* __restrict__ gamma, int linear_upwind,
                                                               double* __restrict__ e_out) {
    const int f = fv_cell();
    if (f >= fa.nif) {
        return;
    }
    const auto gg = g(grad);
    const auto cg = g(cc);
    const int p = g(fa.own)[f];
    const int n = g(fa.nei)[f];
    const double fl = g(flux)[f];
    double e = 0.0;
    if (linear_upwind != 0) {
        const int up = fl >= 0.0 ? p : n;
        const double r0 = g(fa.cf)[3LL * f] - cg[3 * up];
        const double r1 = g(fa.cf)[3LL * f + 1] - cg[3 * up + 1];
        const double r2 = g(fa.cf)[3LL * f + 2] - cg[3 * up + 2];
        e -= fl * (gg[3 * up] * r0 + gg[3 * up + 1] * r1 + gg[3 * up + 2] * r2);
    }
    const double wf = g(fa.w)[f];
    const double gf0 = wf * gg[3 * p] + (1.0 - wf) * gg[3 * n];
    const double gf1 = wf * gg[3 * p + 1] + (1.0 - wf) * gg[3 * n + 1];
    const double gf2 = wf * gg[3 * p + 2] + (1.0 - wf) * gg[3 * n + 2];
    e += g(gamma)[f] * (g(fa.kv)[3LL * f] * gf0 + g(fa.kv)[3LL * f + 1] * gf1 + g(fa.kv)[3LL * f + 2] * gf2);
    e_out[f] = e;
}

// Explicit flux of an internal face for the momentum equation (fv_cell.inc:199-260), one thread
// per face, per component: -F_f lim_up (g_up^T (x_f - C_up)) [linear upwind, blended with the
// linear face value when central > 0, limited when lim is given] + nu_f (g_f^T k_f)
// [+ nu_f g_f S_f], g_f the interpolated 3x3 gradient; component i goes to e_out[i * nif + f] and
// the cells add it with k_fv_face_sum.
__global__ void __launch_bounds__(k_fv_block) k_fv_momentum_face(FvFaces fa, const double* __restrict__ cc,
                                                                 const double* __restrict__ grad,
                                                                 const double* __restrict__ lim, int limited,
                                                                 const double* __restrict__ flux,
                                                                 const double* __restrict__ nu_face,
                                                                 int linear_upwind, double central,
                                                                 const double* __restrict__ u, int transpose_term,
                                                                 double* __restrict__ e_out) {
    const int f = fv_cell();
    if (f >= fa.nif) {
        return;
    }
    const auto gg = g(grad);
    const auto fu = g(u);
    const auto cg = g(cc);
    const int p = g(fa.own)[f];
    const int n = g(fa.nei)[f];
    const double fl = g(flux)[f];
    const double wf = g(fa.w)[f];
    double e[3] = {0.0, 0.0, 0.0};
    if (linear_upwind != 0) {
        const int up = fl >= 0.0 ? p : n;
        const double r0 = g(fa.cf)[3LL * f] - cg[3 * up];
        const double r1 = g(fa.cf)[3LL * f + 1] - cg[3 * up + 1];
        const double r2 = g(fa.cf)[3LL * f + 2] - cg[3 * up + 2];
        const auto gu = gg + 9LL * up;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double d = gu[3 * i] * r0 + gu[3 * i + 1] * r1 + gu[3 * i + 2] * r2;
            if (central > 0.0) {
                const double uf = wf * fu[3 * p + i] + (1.0 - wf) * fu[3 * n + i];
                d = (1.0 - central) * d + central * (uf - fu[3 * up + i]);
            }
            if (limited != 0) {
                d = g(lim)[3 * up + i] * d;
            }
            e[i] -= fl * d;
        }
    }
    double gf[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        gf[j] = wf * gg[9LL * p + j] + (1.0 - wf) * gg[9LL * n + j];
    }
    const double nu = g(nu_face)[f];
    const double k0 = g(fa.kv)[3LL * f];
    const double kk1 = g(fa.kv)[3LL * f + 1];
    const double k2 = g(fa.kv)[3LL * f + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        e[i] += nu * (gf[3 * i] * k0 + gf[3 * i + 1] * kk1 + gf[3 * i + 2] * k2);
    }
    if (transpose_term != 0) {
        const double a0 = g(fa.sf)[3LL * f];
        const double a1 = g(fa.sf)[3LL * f + 1];
        const double a2 = g(fa.sf)[3LL * f + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            e[i] += nu * (gf[i] * a0 + gf[i + 3] * a1 + gf[i + 6] * a2);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        e_out[f + static_cast<long long>(i) * fa.nif] = e[i];
    }
}

// source_c += sum over the cell's internal faces in face order of +e_f (owner) or -e_f
// (neighbour), NC components, in the order of the reference's face loops; each face's flux comes
// from its face pass, component i of face f at e_f[i * nif + f].
template <int NC>
__global__ void __launch_bounds__(k_fv_block) k_fv_face_sum(FvSlots m, const double* __restrict__ e_f,
                                                            double* __restrict__ source) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    double s[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        s[i] = 0.0;
    }
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        const int o = g(m.other)[t];
        const int f = g(m.face)[t];
        if (f < 0 || o < 0) {
            break; // boundary faces follow the internal ones
        }
        const bool mine = g(m.owns)[t] != 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double e = g(e_f)[f + static_cast<long long>(i) * m.nif];
            s[i] = mine ? s[i] + e : s[i] - e;
        }
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        source[NC * c + i] += s[i];
    }
}

// Face pass of the skewness-corrected Gauss gradient (fv_cell.inc:77-118): the corrected face
// value of each internal face, NC components, phi_f = w phi_P + (1 - w) phi_N + g_f . (x_f - x_f'),
// g_f the interpolated least-squares gradient lg and x_f' = w C_P + (1 - w) C_N, as each cell
// of the face would form it. Component i of face f goes to pf_out[i * nif + f]; derivative q of
// cell c's gradient is lg[c * gcs + q * gqs].
template <int NC>
__global__ void __launch_bounds__(k_fv_block) k_fv_face_value(FvFaces fa, const double* __restrict__ cc,
                                                              const double* __restrict__ phi, int stride,
                                                              const double* __restrict__ lg, int gcs,
                                                              int gqs, double* __restrict__ pf_out) {
    const int f = fv_cell();
    if (f >= fa.nif) {
        return;
    }
    const auto fp = g(phi);
    const auto gl = g(lg);
    const auto cg = g(cc);
    const int p = g(fa.own)[f];
    const int n = g(fa.nei)[f];
    const double wf = g(fa.w)[f];
    double r[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double xip = wf * cg[3 * p + j] + (1.0 - wf) * cg[3 * n + j];
        r[j] = g(fa.cf)[3LL * f + j] - xip;
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        double pf = wf * fp[p * stride + i] + (1.0 - wf) * fp[n * stride + i];
        double dot = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const long long q = static_cast<long long>(3 * i + j) * gqs;
            const double gf = wf * gl[static_cast<long long>(p) * gcs + q] + (1.0 - wf) * gl[static_cast<long long>(n) * gcs + q];
            dot = j == 0 ? gf * r[0] : dot + gf * r[j];
        }
        pf_out[f + static_cast<long long>(i) * fa.nif] = pf + dot;
    }
}

// Cell pass of the corrected Gauss gradient: sum over the cell's faces in face order of
// +-S_f phi_f (internal faces, values from k_fv_face_value) and S_f phi_b (boundary faces),
// divided by the volume.
// One slot's operands for k_fv_gauss_from_faces: the area vector and the face value (internal
// faces; component i of face f at pf_f[i * nif + f]) or the boundary value; padding loads no
// value.
template <int NC>
struct GaussSlot {
    int f;
    bool internal;
    bool mine;
    double s0, s1, s2;
    double u[NC];
};

template <int NC>
__device__ __forceinline__ GaussSlot<NC> load_gauss_slot(const FvSlots& m, int t, const double* pf_f, const double* bv,
                                                         int stride) {
    GaussSlot<NC> sl;
    sl.f = g(m.face)[t];
    const int o = g(m.other)[t];
    sl.internal = o >= 0;
    sl.mine = g(m.owns)[t] != 0;
    sl.s0 = g(m.sf0)[t];
    sl.s1 = g(m.sf1)[t];
    sl.s2 = g(m.sf2)[t];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        sl.u[i] = sl.f < 0 ? 0.0
                  : o >= 0 ? g(pf_f)[sl.f + static_cast<long long>(i) * m.nif]
                           : g(bv)[static_cast<long long>(-1 - o) * stride + i];
    }
    return sl;
}

template <int NC>
__device__ __forceinline__ void add_gauss_slot(const GaussSlot<NC>& sl, double* g0, double* g1, double* g2) {
    if (sl.f < 0) {
        return; // padding
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (sl.internal && !sl.mine) {
            g0[i] -= sl.s0 * sl.u[i];
            g1[i] -= sl.s1 * sl.u[i];
            g2[i] -= sl.s2 * sl.u[i];
        } else {
            g0[i] += sl.s0 * sl.u[i];
            g1[i] += sl.s1 * sl.u[i];
            g2[i] += sl.s2 * sl.u[i];
        }
    }
}

template <int NC>
__global__ void __launch_bounds__(k_fv_block) k_fv_gauss_from_faces(FvSlots m, const double* __restrict__ pf_f,
                                                                    const double* __restrict__ bv, int stride,
                                                                    double* __restrict__ out, int ostride) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    double g0[NC];
    double g1[NC];
    double g2[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        g0[i] = 0.0;
        g1[i] = 0.0;
        g2[i] = 0.0;
    }
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    // Faces in ascending order, two slots loaded ahead of their terms.
    int k = 0;
    for (; k + 2 <= count; k += 2) {
        const int t = first + (k << 6);
        const auto a = load_gauss_slot<NC>(m, t, pf_f, bv, stride);
        const auto b = load_gauss_slot<NC>(m, t + 64, pf_f, bv, stride);
        add_gauss_slot<NC>(a, g0, g1, g2);
        add_gauss_slot<NC>(b, g0, g1, g2);
    }
    if (k < count) {
        add_gauss_slot<NC>(load_gauss_slot<NC>(m, first + (k << 6), pf_f, bv, stride), g0, g1, g2);
    }
    double* out_c = out + static_cast<long long>(c) * ostride;
    const double v = g(m.vol)[c];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        out_c[3 * i] = g0[i] / v;
        out_c[3 * i + 1] = g1[i] / v;
        out_c[3 * i + 2] = g2[i] / v;
    }
}

// Eigen's Vec3 dot (fixed size 3: the first two terms as one SSE2 packet, then the third).
__device__ __forceinline__ double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
    return (a0 * b0 + a1 * b1) + a2 * b2;
}

// H(U)/A and rAU of the pressure correction (flow_solver.cpp:688-706). Per component i the
// momentum matrix's off-diagonal product accumulates over the cell's internal faces in face
// order, upper_f u_i(N) where the cell owns f and lower_f u_i(P) where it is the neighbour
// (LduMatrix::off_diagonal_product, ldu_matrix.cpp:39-49); hbya_i = (source_i - off_i) / diag_i
// and rAU = V / mean(diag) with Eigen's (x + y) + z.
__global__ void __launch_bounds__(k_fv_block) k_fv_hbya(FvSlots m, const double* __restrict__ upper,
                                                        const double* __restrict__ lower,
                                                        const double* __restrict__ diag3,
                                                        const double* __restrict__ source3,
                                                        const double* __restrict__ u, double* __restrict__ hbya,
                                                        double* __restrict__ rau) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    const auto fu = g(u);
    double off0 = 0.0;
    double off1 = 0.0;
    double off2 = 0.0;
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        const int f = g(m.face)[t];
        const int o = g(m.other)[t];
        if (f < 0 || o < 0) {
            break; // boundary faces and padding follow the internal faces
        }
        const double a = g(m.owns)[t] != 0 ? g(upper)[f] : g(lower)[f];
        off0 += a * fu[3 * o];
        off1 += a * fu[3 * o + 1];
        off2 += a * fu[3 * o + 2];
    }
    const double d0 = g(diag3)[3 * c];
    const double d1 = g(diag3)[3 * c + 1];
    const double d2 = g(diag3)[3 * c + 2];
    hbya[3 * c] = (g(source3)[3 * c] - off0) / d0;
    hbya[3 * c + 1] = (g(source3)[3 * c + 1] - off1) / d1;
    hbya[3 * c + 2] = (g(source3)[3 * c + 2] - off2) / d2;
    rau[c] = g(m.vol)[c] / (((d0 + d1) + d2) / 3.0);
}

struct PressureIn {
    const double* hbya;
    const double* rau;
    const double* dc;
    const int* kind;          // per boundary face
    const double* fixed_u;    // 3 per boundary face
    const double* outlet_p;   // per boundary face
    const double* u_prev;     // 3 per cell (majumdar only)
    const double* phi_prev;   // per face (majumdar only)
    double beta;              // 1 - alpha_u
    int majumdar;
    int pin;
};

struct PressureOut {
    double* diag;
    double* upper;
    double* lower;
    double* source;
    double* phi;
    double* rau_f;
};

// Face-major copies of the weight, area vector, face centre and correction vector of every
// face, taken from the owner's slots (each listed face has exactly one): the face passes read
// them coalesced.
__global__ void __launch_bounds__(k_fv_block) k_fv_faces_from_slots(FvSlots m, double* __restrict__ w,
                                                                    double* __restrict__ sf, double* __restrict__ cf,
                                                                    double* __restrict__ kv) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        const int f = g(m.face)[t];
        if (f < 0) {
            break;
        }
        if (g(m.owns)[t] == 0) {
            continue;
        }
        if (f < m.nif) {
            w[f] = g(m.w)[t];
        }
        sf[3LL * f] = g(m.sf0)[t];
        sf[3LL * f + 1] = g(m.sf1)[t];
        sf[3LL * f + 2] = g(m.sf2)[t];
        cf[3LL * f] = g(m.cf0)[t];
        cf[3LL * f + 1] = g(m.cf1)[t];
        cf[3LL * f + 2] = g(m.cf2)[t];
        kv[3LL * f] = g(m.kv0)[t];
        kv[3LL * f + 1] = g(m.kv1)[t];
        kv[3LL * f + 2] = g(m.kv2)[t];
    }
}

struct PressureFaces {
    int nif;
    int nf;
    const int* own;
    const int* nei;
    const double* w;
    const double* sf; // 3 per face
};

// Face pass of the pressure equation (flow_solver.cpp:708-753 and the face loop of
// assemble_convection_diffusion, krylov_openmp.cpp:289-295, convection 0), one thread per face:
// phiHbyA = (w HbyA_P + (1 - w) HbyA_N) . S and rAU_f (internal), the boundary kinds' fluxes
// with rAU_f = rAU_P, the Majumdar term beta (phi_prev - u_prev,f . S), the Laplacian
// coefficient g = rAU_f dc (outlets only on the boundary; 0 elsewhere) and upper / lower.
__global__ void __launch_bounds__(k_fv_block) k_fv_face_flux(PressureFaces fa, PressureIn in, PressureOut out,
                                                             double* __restrict__ gface) {
    const int f = fv_cell();
    if (f >= fa.nf) {
        return;
    }
    const auto hb = g(in.hbya);
    const auto ra = g(in.rau);
    const int p = g(fa.own)[f];
    const double s0 = g(fa.sf)[3LL * f];
    const double s1 = g(fa.sf)[3LL * f + 1];
    const double s2 = g(fa.sf)[3LL * f + 2];
    double phi = 0.0;
    double rf = 0.0;
    double gc = 0.0;
    if (f < fa.nif) {
        const int n = g(fa.nei)[f];
        const double w = g(fa.w)[f];
        const double v0 = w * hb[3 * p] + (1.0 - w) * hb[3 * n];
        const double v1 = w * hb[3 * p + 1] + (1.0 - w) * hb[3 * n + 1];
        const double v2 = w * hb[3 * p + 2] + (1.0 - w) * hb[3 * n + 2];
        phi = dot3(v0, v1, v2, s0, s1, s2);
        rf = w * ra[p] + (1.0 - w) * ra[n];
        if (in.majumdar != 0) {
            const auto up = g(in.u_prev);
            const double uf0 = w * up[3 * p] + (1.0 - w) * up[3 * n];
            const double uf1 = w * up[3 * p + 1] + (1.0 - w) * up[3 * n + 1];
            const double uf2 = w * up[3 * p + 2] + (1.0 - w) * up[3 * n + 2];
            phi += in.beta * (g(in.phi_prev)[f] - dot3(uf0, uf1, uf2, s0, s1, s2));
        }
        gc = rf * g(in.dc)[f];
        const double pos = phi > 0.0 ? 0.0 * phi : 0.0;
        const double neg = phi < 0.0 ? 0.0 * phi : 0.0;
        out.upper[f] = neg - gc;
        out.lower[f] = -pos - gc;
    } else {
        const int b = f - fa.nif;
        const int kind = g(in.kind)[b];
        if (kind == CFDG_BC_VELOCITY_INLET || kind == CFDG_BC_WALL) {
            const auto fu = g(in.fixed_u);
            phi = dot3(fu[3 * b], fu[3 * b + 1], fu[3 * b + 2], s0, s1, s2);
        } else if (kind == CFDG_BC_PRESSURE_OUTLET) {
            phi = dot3(hb[3 * p], hb[3 * p + 1], hb[3 * p + 2], s0, s1, s2);
            if (in.majumdar != 0) {
                const auto up = g(in.u_prev);
                phi += in.beta * (g(in.phi_prev)[f] - dot3(up[3 * p], up[3 * p + 1], up[3 * p + 2], s0, s1, s2));
            }
        }
        rf = ra[p];
        if (kind == CFDG_BC_PRESSURE_OUTLET) {
            gc = rf * g(in.dc)[f];
        }
    }
    out.phi[f] = phi;
    out.rau_f[f] = rf;
    gface[f] = gc;
}

// Cell pass of the pressure equation (assemble_convection_diffusion's accumulation, then
// flow_solver.cpp:760-796): over the cell's faces in face order, the diagonal sum of the
// faces' coefficients (pos + g as owner, -neg + g as neighbour, convection 0) and the divergence
// of phiHbyA, then source = -divergence before the first boundary face, the boundary terms
// (-phi; outlets: diagonal + g, source + g p_out) and set_reference(0, 0) for cell 0. The faces'
// values come from the face pass, so every cell adds the same numbers in the order of the
// serial loops.
__global__ void __launch_bounds__(k_fv_block) k_fv_pressure_cells(FvSlots m, const double* __restrict__ phi_f,
                                                                  const double* __restrict__ gface, PressureIn in,
                                                                  PressureOut out) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    double diag = 0.0;
    double div = 0.0;
    double src = 0.0;
    bool boundary = false;
    int first = 0;
    int count = 0;
    slot_range(m, c, first, count);
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        const int f = g(m.face)[t];
        if (f < 0) {
            break;
        }
        const int o = g(m.other)[t];
        const double phi = g(phi_f)[f];
        const double gc = g(gface)[f];
        if (o >= 0) {
            if (g(m.owns)[t] != 0) {
                const double pos = phi > 0.0 ? 0.0 * phi : 0.0;
                diag += pos + gc;
                div += phi;
            } else {
                const double neg = phi < 0.0 ? 0.0 * phi : 0.0;
                diag += -neg + gc;
                div -= phi;
            }
            continue;
        }
        if (!boundary) { // source -= divergence, before the boundary terms
            src = 0.0 - div;
            boundary = true;
        }
        const int b = -1 - o;
        src -= phi;
        if (g(in.kind)[b] == CFDG_BC_PRESSURE_OUTLET) {
            diag += gc;
            src += gc * g(in.outlet_p)[b];
        }
    }
    if (!boundary) {
        src = 0.0 - div;
    }
    if (in.pin != 0 && c == 0) {
        src += diag * 0.0;
        diag += diag;
    }
    out.diag[c] = diag;
    out.source[c] = src;
}

// Fills the slots of every cell of every slice (cells past the last one get padding) from the
// face-indexed mesh arrays; values are copied, never recomputed.
struct FvFaceArrays {
    int nc;
    int nif;
    const int* start;
    const int* faces;
    const int* own;
    const int* nei;
    const double* w;
    const double* sf;
    const double* lsq_d;
    const double* cf;
    const double* kv;
};

struct FvSlotOut {
    int* face;
    int* other;
    unsigned char* owns;
    double* w;
    double* v[12]; // sf0..2, ld0..2, cf0..2, kv0..2
};

__global__ void __launch_bounds__(k_fv_block) k_fv_fill_slots(FvFaceArrays a, const int* base, int n_cells_padded,
                                                              FvSlotOut o) {
    const int c = static_cast<int>(blockIdx.x) * k_fv_block + static_cast<int>(threadIdx.x);
    if (c >= n_cells_padded) {
        return;
    }
    const int s = c >> 6;
    const int first = base[s] + (c & 63);
    const int count = (base[s + 1] - base[s]) >> 6;
    const int k0 = c < a.nc ? a.start[c] : 0;
    const int deg = c < a.nc ? a.start[c + 1] - k0 : 0;
    for (int k = 0; k < count; ++k) {
        const int t = first + (k << 6);
        if (k >= deg) {
            o.face[t] = -1;
            o.other[t] = 0;
            o.owns[t] = 0;
            o.w[t] = 0.0;
            for (int j = 0; j < 12; ++j) {
                o.v[j][t] = 0.0;
            }
            continue;
        }
        const int f = a.faces[k0 + k];
        o.face[t] = f;
        if (f < a.nif) {
            const int p = a.own[f];
            o.other[t] = p == c ? a.nei[f] : p;
            o.owns[t] = p == c ? 1 : 0;
            o.w[t] = a.w[f];
        } else {
            o.other[t] = -1 - (f - a.nif);
            o.owns[t] = 1;
            o.w[t] = 0.0;
        }
        for (int j = 0; j < 3; ++j) {
            o.v[j][t] = a.sf[3 * f + j];
            o.v[3 + j][t] = a.lsq_d[3 * f + j];
            o.v[6 + j][t] = a.cf[3 * f + j];
            o.v[9 + j][t] = a.kv[3 * f + j];
        }
    }
}

template <class T>
void need(const T* p, std::size_t n, const char* what) {
    if (n > 0 && p == nullptr) {
        throw InvalidArgument(std::string("DeviceFv: missing mesh array ") + what);
    }
}

std::size_t at_least_one(long long n) { return static_cast<std::size_t>(n > 0 ? n : 1); }

} // namespace

// Mesh on the device plus the field buffers of the reference's backend (fv_openmp.cpp:70-76).
class FvDevice {
public:
    explicit FvDevice(const cfdg_fv_mesh& v) {
        require_device();
        // DeviceFv's checks (fv_device.cpp:24-48); indices are checked too, since a kernel must
        // never gather out of bounds.
        if (v.n_cells <= 0 || v.n_internal_faces < 0 || v.n_faces < v.n_internal_faces ||
            v.cell_offsets == nullptr) {
            throw InvalidArgument("DeviceFv: invalid mesh sizes");
        }
        nc_ = v.n_cells;
        nif_ = v.n_internal_faces;
        nf_ = v.n_faces;
        const std::size_t c = static_cast<std::size_t>(nc_);
        const std::size_t f = static_cast<std::size_t>(nf_);
        const std::size_t fi = static_cast<std::size_t>(nif_);
        need(v.owner, f, "owner");
        need(v.neighbour, fi, "neighbour");
        const int n_list = v.cell_offsets[nc_];
        need(v.cell_faces, static_cast<std::size_t>(n_list > 0 ? n_list : 0), "cell_faces");
        need(v.weight, fi, "weight");
        need(v.sf, 3 * f, "sf");
        need(v.face_centre, 3 * f, "face_centre");
        need(v.cell_centre, 3 * c, "cell_centre");
        need(v.volume, c, "volume");
        need(v.correction, 3 * f, "correction");
        need(v.lsq_d, 3 * f, "lsq_d");
        need(v.lsq_inverse, 9 * c, "lsq_inverse");
        require(v.cell_offsets[0] == 0, "DeviceFv: cell_offsets must start at 0");
        for (int i = 0; i < nc_; ++i) {
            require(v.cell_offsets[i + 1] >= v.cell_offsets[i], "DeviceFv: cell_offsets must not decrease");
        }
        for (int k = 0; k < n_list; ++k) {
            require(v.cell_faces[k] >= 0 && v.cell_faces[k] < nf_, "DeviceFv: cell face out of range");
        }
        for (int k = 0; k < nf_; ++k) {
            require(v.owner[k] >= 0 && v.owner[k] < nc_, "DeviceFv: owner out of range");
        }
        for (int k = 0; k < nif_; ++k) {
            require(v.neighbour[k] >= 0 && v.neighbour[k] < nc_, "DeviceFv: neighbour out of range");
        }
        hipStream_t s = stream_.get();
        // Slice s (cells 64 s .. 64 s + 63) gets 64 x (most faces of its cells) slots.
        n_slices_ = (nc_ + 63) / 64;
        std::vector<int> base(static_cast<std::size_t>(n_slices_) + 1, 0);
        long long total = 0;
        for (int sl = 0; sl < n_slices_; ++sl) {
            int width = 0;
            for (int cell = 64 * sl; cell < std::min(nc_, 64 * sl + 64); ++cell) {
                width = std::max(width, v.cell_offsets[cell + 1] - v.cell_offsets[cell]);
            }
            total += 64LL * width;
            require(total <= 0x7fffffffLL, "DeviceFv: mesh too large for 32-bit face slots");
            base[static_cast<std::size_t>(sl) + 1] = static_cast<int>(total);
        }
        const std::size_t slots = at_least_one(total);
        base_.upload(base.data(), base.size(), s);
        sface_.resize(slots);
        sother_.resize(slots);
        sowns_.resize(slots);
        sw_.resize(slots);
        for (DevBuf<double>& a : sv_) {
            a.resize(slots);
        }
        {
            // face-indexed copies, needed only to fill the slots
            DevBuf<int> start;
            DevBuf<int> faces;
            DevBuf<double> w;
            DevBuf<double> sf;
            DevBuf<double> lsq_d;
            DevBuf<double> cf;
            DevBuf<double> kv;
            const int zero_i = 0;
            const double zero_d = 0.0;
            start.upload(v.cell_offsets, c + 1, s);
            faces.upload(n_list > 0 ? v.cell_faces : &zero_i, at_least_one(n_list), s);
            fown_.upload(v.owner, f, s); // kept: checked indices for the pressure face pass
            fnei_.upload(nif_ > 0 ? v.neighbour : &zero_i, at_least_one(nif_), s);
            w.upload(nif_ > 0 ? v.weight : &zero_d, at_least_one(nif_), s);
            sf.upload(v.sf, 3 * f, s);
            lsq_d.upload(v.lsq_d, 3 * f, s);
            cf.upload(v.face_centre, 3 * f, s);
            kv.upload(v.correction, 3 * f, s);
            const FvFaceArrays a{nc_, nif_, start.get(), faces.get(), fown_.get(), fnei_.get(), w.get(),
                                 sf.get(), lsq_d.get(), cf.get(), kv.get()};
            FvSlotOut o{sface_.get(), sother_.get(), sowns_.get(), sw_.get(), {}};
            for (int j = 0; j < 12; ++j) {
                o.v[j] = sv_[static_cast<std::size_t>(j)].get();
            }
            const int padded = 64 * n_slices_;
            k_fv_fill_slots<<<(padded + k_fv_block - 1) / k_fv_block, k_fv_block, 0, s>>>(a, base_.get(), padded, o);
            hip_check(hipGetLastError(), "k_fv_fill_slots launch");
            hip_check(hipStreamSynchronize(s), "DeviceFv slots sync"); // the copies go out of scope
        }
        cc_.upload(v.cell_centre, 3 * c, s);
        vol_.upload(v.volume, c, s);
        lsq_.upload(v.lsq_inverse, 9 * c, s);
        const std::size_t nb = at_least_one(nf_ - nif_);
        phi_.resize(c);
        bv_.resize(nb);
        sgrad_.resize(3 * c);
        vphi_.resize(3 * c);
        vbv_.resize(3 * nb);
        vgrad_.resize(9 * c);
        lim_.resize(3 * c);
        flux_.resize(at_least_one(nif_));
        coef_.resize(at_least_one(nif_));
        src_.resize(3 * c);
        tmp_.resize(9 * c);
    }
    ~FvDevice() { (void)hipStreamSynchronize(stream_.get()); }
    FvDevice(const FvDevice&) = delete;
    FvDevice& operator=(const FvDevice&) = delete;

    void gradient(const double* phi, const double* bv, int scheme, double* grad) {
        check_scheme(scheme);
        require(phi != nullptr && grad != nullptr && (nb() == 0 || bv != nullptr), "DeviceFv::gradient: null array");
        begin();
        put(phi_, phi, nc_);
        put(bv_, bv, nb());
        gradient_loop<1>(phi_.get(), bv_.get(), scheme, sgrad_.get(), 3);
        get(sgrad_, grad, 3LL * nc_);
    }

    void vector_gradient(const double* u, const double* ub, int scheme, double* grad) {
        check_scheme(scheme);
        require(u != nullptr && grad != nullptr && (nb() == 0 || ub != nullptr),
                "DeviceFv::vector_gradient: null array");
        begin();
        put(vphi_, u, 3LL * nc_);
        put(vbv_, ub, 3LL * nb());
        gradient_loop<3>(vphi_.get(), vbv_.get(), scheme, vgrad_.get(), 9);
        get(vgrad_, grad, 9LL * nc_);
        have_vector_ = true;
    }

    void scalar_corrections(const double* phi, const double* bv, int scheme, const double* flux, const double* gamma,
                            bool linear_upwind, double* source) {
        check_scheme(scheme);
        require(phi != nullptr && source != nullptr && (nb() == 0 || bv != nullptr) &&
                    (nif_ == 0 || (flux != nullptr && gamma != nullptr)),
                "DeviceFv::scalar_corrections: null array");
        begin();
        put(phi_, phi, nc_);
        put(bv_, bv, nb());
        gradient_loop<1>(phi_.get(), bv_.get(), scheme, sgrad_.get(), 3);
        put(flux_, flux, nif_);
        put(coef_, gamma, nif_);
        put(src_, source, nc_);
        ensure_faces();
        if (eface_.size() == 0) {
            eface_.resize(at_least_one(3LL * nif_));
        }
        if (nif_ > 0) { // no internal face: nothing to add
            launch([&](hipStream_t s) {
                k_fv_scalar_face<<<face_blocks(nif_), k_fv_block, 0, s>>>(faces(), cc_.get(), sgrad_.get(),
                                                                          flux_.get(), coef_.get(),
                                                                          linear_upwind ? 1 : 0, eface_.get());
            });
        }
        launch([&](hipStream_t s) { k_fv_face_sum<1><<<blocks(), k_fv_block, 0, s>>>(mesh(), eface_.get(), src_.get()); });
        get(src_, source, nc_);
    }

    void momentum_corrections(const double* flux, const double* nu_face, bool linear_upwind, bool limited, bool venkat,
                              const double* eps2, double central, bool transpose_term, double* source) {
        if (!have_vector_) {
            throw InvalidArgument("DeviceFv::momentum_corrections before vector_gradient");
        }
        require(source != nullptr && (nif_ == 0 || (flux != nullptr && nu_face != nullptr)) &&
                    (!venkat || eps2 != nullptr),
                "DeviceFv::momentum_corrections: null array");
        const double e0 = venkat ? eps2[0] : 0.0;
        const double e1 = venkat ? eps2[1] : 0.0;
        const double e2 = venkat ? eps2[2] : 0.0;
        begin();
        put(flux_, flux, nif_);
        put(coef_, nu_face, nif_);
        put(src_, source, 3LL * nc_);
        const bool use_lim = limited && linear_upwind;
        if (use_lim) {
            launch([&](hipStream_t s) {
                k_fv_limiter<<<blocks(), k_fv_block, 0, s>>>(mesh(), vphi_.get(), vbv_.get(), vgrad_.get(),
                                                             venkat ? 1 : 0, e0, e1, e2, lim_.get());
            });
        }
        ensure_faces();
        if (eface_.size() == 0) {
            eface_.resize(at_least_one(3LL * nif_));
        }
        if (nif_ > 0) { // no internal face: nothing to add
            launch([&](hipStream_t s) {
                k_fv_momentum_face<<<face_blocks(nif_), k_fv_block, 0, s>>>(
                    faces(), cc_.get(), vgrad_.get(), lim_.get(), use_lim ? 1 : 0, flux_.get(), coef_.get(),
                    linear_upwind ? 1 : 0, central, vphi_.get(), transpose_term ? 1 : 0, eface_.get());
            });
        }
        launch([&](hipStream_t s) { k_fv_face_sum<3><<<blocks(), k_fv_block, 0, s>>>(mesh(), eface_.get(), src_.get()); });
        get(src_, source, 3LL * nc_);
    }

    // Pressure equation (flow_solver.cpp:688-796) from host inputs; the results stay on the
    // device (pressure()) and are copied to the non-null outputs.
    void pressure_equation(const cfdg_pressure_inputs& in, cfdg_pressure_outputs* out) {
        const long long c = nc_;
        const long long fi = nif_;
        const long long f = nf_;
        const long long b = nb();
        require((fi == 0 || (in.upper != nullptr && in.lower != nullptr)) && in.diag != nullptr &&
                    in.source != nullptr && in.u != nullptr && in.delta_coeffs != nullptr &&
                    (b == 0 || (in.boundary_kind != nullptr && in.fixed_velocity != nullptr &&
                                in.outlet_pressure != nullptr)),
                "pressure_equation: null array");
        for (long long k = 0; k < b; ++k) {
            require(in.boundary_kind[k] >= CFDG_BC_VELOCITY_INLET && in.boundary_kind[k] <= CFDG_BC_EMPTY,
                    "pressure_equation: boundary kind out of range");
        }
        const bool majumdar = in.u_prev != nullptr && in.phi_prev != nullptr && in.alpha_u < 1.0;
        if (pupper_.size() == 0) {
            const std::size_t c1 = at_least_one(c);
            const std::size_t fi1 = at_least_one(fi);
            const std::size_t f1 = at_least_one(f);
            const std::size_t b1 = at_least_one(b);
            mupper_.resize(fi1);
            mlower_.resize(fi1);
            mdiag_.resize(3 * c1);
            msrc_.resize(3 * c1);
            mu_.resize(3 * c1);
            dc_.resize(f1);
            kind_.resize(b1);
            fixed_u_.resize(3 * b1);
            outlet_p_.resize(b1);
            hbya_.resize(3 * c1);
            rau_.resize(c1);
            phih_.resize(f1);
            rauf_.resize(f1);
            pdiag_.resize(c1);
            pupper_.resize(fi1);
            plower_.resize(fi1);
            psrc_.resize(c1);
        }
        if (majumdar && uprev_.size() == 0) {
            uprev_.resize(at_least_one(3 * c));
            phiprev_.resize(at_least_one(f));
        }
        begin();
        put(mupper_, in.upper, fi);
        put(mlower_, in.lower, fi);
        put(mdiag_, in.diag, 3 * c);
        put(msrc_, in.source, 3 * c);
        put(mu_, in.u, 3 * c);
        put(dc_, in.delta_coeffs, f);
        if (b > 0) {
            hip_check(hipMemcpyAsync(kind_.get(), in.boundary_kind, static_cast<std::size_t>(b) * sizeof(int),
                                     hipMemcpyHostToDevice, stream_.get()),
                      "DeviceFv H2D");
        }
        put(fixed_u_, in.fixed_velocity, 3 * b);
        put(outlet_p_, in.outlet_pressure, b);
        if (majumdar) {
            put(uprev_, in.u_prev, 3 * c);
            put(phiprev_, in.phi_prev, f);
        }
        const FvSlots m = mesh();
        launch([&](hipStream_t s) {
            k_fv_hbya<<<blocks(), k_fv_block, 0, s>>>(m, mupper_.get(), mlower_.get(), mdiag_.get(), msrc_.get(),
                                                      mu_.get(), hbya_.get(), rau_.get());
        });
        const PressureIn pin{hbya_.get(),     rau_.get(),
                             dc_.get(),       kind_.get(),
                             fixed_u_.get(),  outlet_p_.get(),
                             majumdar ? uprev_.get() : nullptr,
                             majumdar ? phiprev_.get() : nullptr,
                             1.0 - in.alpha_u, majumdar ? 1 : 0,
                             in.pin_reference != 0 ? 1 : 0};
        const PressureOut pout{pdiag_.get(), pupper_.get(), plower_.get(), psrc_.get(), phih_.get(), rauf_.get()};
        ensure_faces();
        if (gface_.size() == 0) {
            gface_.resize(at_least_one(f));
        }
        const PressureFaces fa{nif_, nf_, fown_.get(), fnei_.get(), fw_.get(), fsf_.get()};
        if (f > 0) {
            launch([&](hipStream_t s) {
                k_fv_face_flux<<<face_blocks(f), k_fv_block, 0, s>>>(fa, pin, pout, gface_.get());
            });
        }
        launch([&](hipStream_t s) {
            k_fv_pressure_cells<<<blocks(), k_fv_block, 0, s>>>(m, phih_.get(), gface_.get(), pin, pout);
        });
        have_pressure_ = true;
        const auto back = [&](const DevBuf<double>& buf, double* dst, long long n) {
            if (dst != nullptr && n > 0) {
                hip_check(hipMemcpyAsync(dst, buf.get(), static_cast<std::size_t>(n) * sizeof(double),
                                         hipMemcpyDeviceToHost, stream_.get()),
                          "DeviceFv D2H");
            }
        };
        if (out != nullptr) {
            back(pdiag_, out->diag, c);
            back(pupper_, out->upper, fi);
            back(plower_, out->lower, fi);
            back(psrc_, out->source, c);
            back(phih_, out->phi_hbya, f);
            back(rauf_, out->rau_f, f);
            back(hbya_, out->hbya, 3 * c);
            back(rau_, out->rau, c);
        }
        hip_check(hipStreamSynchronize(stream_.get()), "DeviceFv sync");
    }

    // The last pressure equation on the device (all null before the first one).
    [[nodiscard]] FvPressureArrays pressure() const {
        if (!have_pressure_) {
            return FvPressureArrays{nc_, nif_, nullptr, nullptr, nullptr, nullptr};
        }
        return FvPressureArrays{nc_, nif_, pdiag_.get(), pupper_.get(), plower_.get(), psrc_.get()};
    }

    void set_timing(bool on) { timing_ = on; }
    [[nodiscard]] double last_kernel_ms() const { return last_ms_; }

private:
    [[nodiscard]] int nb() const { return nf_ - nif_; }
    [[nodiscard]] int face_blocks(long long n) const { return static_cast<int>((n + k_fv_block - 1) / k_fv_block); }

    // Face-major weights, area vectors, face centres and correction vectors, once (zero for a
    // face no cell lists).
    void ensure_faces() {
        if (fw_.size() != 0) {
            return;
        }
        hipStream_t s = stream_.get();
        fw_.resize(at_least_one(nif_));
        fsf_.resize(at_least_one(3LL * nf_));
        fcf_.resize(at_least_one(3LL * nf_));
        fkv_.resize(at_least_one(3LL * nf_));
        for (DevBuf<double>* b : {&fw_, &fsf_, &fcf_, &fkv_}) {
            b->zero(s);
        }
        k_fv_faces_from_slots<<<blocks(), k_fv_block, 0, s>>>(mesh(), fw_.get(), fsf_.get(), fcf_.get(), fkv_.get());
        hip_check(hipGetLastError(), "k_fv_faces_from_slots launch");
    }

    [[nodiscard]] FvFaces faces() const {
        return FvFaces{nif_, fown_.get(), fnei_.get(), fw_.get(), fsf_.get(), fcf_.get(), fkv_.get()};
    }
    [[nodiscard]] int blocks() const { return (nc_ + k_fv_block - 1) / k_fv_block; }

    static void check_scheme(int scheme) {
        require(scheme >= 0 && scheme <= 2, "DeviceFv: gradient scheme must be 0, 1 or 2");
    }

    [[nodiscard]] FvSlots mesh() const {
        const auto v = [&](int j) { return static_cast<const double*>(sv_[static_cast<std::size_t>(j)].get()); };
        return FvSlots{nc_,   nif_, base_.get(), sface_.get(), sother_.get(), sowns_.get(), sw_.get(),
                       v(0),  v(1), v(2),        v(3),         v(4),          v(5),         v(6),
                       v(7),  v(8), v(9),        v(10),        v(11),         cc_.get(),    vol_.get(),
                       lsq_.get()};
    }

    // Gradient of NC interleaved components (fv_openmp.cpp:257-311): least squares, Green-Gauss,
    // or least squares into tmp (component-major, so that the face pass gathers each derivative
    // of the two cells coalesced) followed by the corrected Gauss gradient.
    template <int NC>
    void gradient_loop(const double* phi, const double* bv, int scheme, double* out, int ostride) {
        const FvSlots m = mesh();
        if (scheme == 2) {
            double* tmp = tmp_.get();
            launch([&](hipStream_t s) {
                k_fv_gradient<true, NC, true><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, tmp, 1, nc_);
            });
            ensure_faces();
            if (eface_.size() < static_cast<std::size_t>(at_least_one(3LL * nif_))) {
                eface_.resize(at_least_one(3LL * nif_));
            }
            if (nif_ > 0) {
                launch([&](hipStream_t s) {
                    k_fv_face_value<NC><<<face_blocks(nif_), k_fv_block, 0, s>>>(faces(), cc_.get(), phi, NC, tmp, 1,
                                                                                 nc_, eface_.get());
                });
            }
            launch([&](hipStream_t s) {
                k_fv_gauss_from_faces<NC><<<blocks(), k_fv_block, 0, s>>>(m, eface_.get(), bv, NC, out, ostride);
            });
            return;
        }
        if (scheme == 1) {
            launch([&](hipStream_t s) {
                k_fv_gradient<true, NC><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, out, ostride, 1);
            });
        } else {
            launch([&](hipStream_t s) {
                k_fv_gradient<false, NC><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, out, ostride, 1);
            });
        }
    }

    void begin() { last_ms_ = 0.0; }

    // Launches on the object's stream; with timing on, HIP events bracket the launch and its
    // duration is added to last_ms_.
    template <class F>
    void launch(F&& f) {
        hipStream_t s = stream_.get();
        if (timing_) {
            hip_check(hipEventRecord(e0_.get(), s), "event record");
        }
        f(s);
        hip_check(hipGetLastError(), "DeviceFv kernel launch");
        if (timing_) {
            hip_check(hipEventRecord(e1_.get(), s), "event record");
            hip_check(hipEventSynchronize(e1_.get()), "event sync");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, e0_.get(), e1_.get()), "event elapsed");
            last_ms_ += static_cast<double>(ms);
        }
    }

    void put(DevBuf<double>& buf, const double* data, long long n) {
        if (n > 0) {
            hip_check(hipMemcpyAsync(buf.get(), data, static_cast<std::size_t>(n) * sizeof(double),
                                     hipMemcpyHostToDevice, stream_.get()),
                      "DeviceFv H2D");
        }
    }

    void get(const DevBuf<double>& buf, double* data, long long n) {
        if (n > 0) {
            hip_check(hipMemcpyAsync(data, buf.get(), static_cast<std::size_t>(n) * sizeof(double),
                                     hipMemcpyDeviceToHost, stream_.get()),
                      "DeviceFv D2H");
        }
        hip_check(hipStreamSynchronize(stream_.get()), "DeviceFv sync");
    }

    int nc_ = 0;
    int nif_ = 0;
    int nf_ = 0;
    int n_slices_ = 0;
    Stream stream_;
    DevBuf<int> base_, sface_, sother_;
    DevBuf<unsigned char> sowns_;
    DevBuf<double> sw_;
    DevBuf<double> sv_[12]; // sf, lsq_d, face centre, correction: three components each
    DevBuf<double> cc_, vol_, lsq_;
    DevBuf<double> phi_, bv_, sgrad_, vphi_, vbv_, vgrad_, lim_, flux_, coef_, src_, tmp_;
    // pressure equation: momentum inputs, boundary data, intermediate and output arrays
    DevBuf<double> mupper_, mlower_, mdiag_, msrc_, mu_, dc_, fixed_u_, outlet_p_, uprev_, phiprev_;
    DevBuf<int> kind_;
    DevBuf<double> hbya_, rau_, phih_, rauf_, pdiag_, pupper_, plower_, psrc_;
    DevBuf<int> fown_, fnei_;         // face-major owner / neighbour (pressure face pass)
    DevBuf<double> fw_, fsf_, fcf_, fkv_; // face-major weight, area vector, face centre, correction
    DevBuf<double> gface_, eface_;        // per face: Laplacian coefficient; explicit correction flux
    bool have_pressure_ = false;
    Event e0_;
    Event e1_;
    bool have_vector_ = false;
    bool timing_ = false;
    double last_ms_ = 0.0;
};

} // namespace cfdg

// ------------------------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------------------------
using cfdg::FvDevice;

struct cfdg_fv {
    std::unique_ptr<FvDevice> d;
};

extern "C" cfdg_status cfdg_fv_create(const cfdg_fv_mesh* mesh, cfdg_fv** out) {
    return cfdg::guarded([&] {
        cfdg::require(mesh != nullptr && out != nullptr, "DeviceFv: null argument");
        auto h = std::make_unique<cfdg_fv>();
        h->d = std::make_unique<FvDevice>(*mesh);
        *out = h.release();
    });
}

extern "C" cfdg_status cfdg_fv_destroy(cfdg_fv* fv) {
    return cfdg::guarded([&] { delete fv; });
}

extern "C" cfdg_status cfdg_fv_gradient(cfdg_fv* fv, const double* phi, const double* bv, int scheme, double* grad) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr, "DeviceFv: null handle");
        fv->d->gradient(phi, bv, scheme, grad);
    });
}

extern "C" cfdg_status cfdg_fv_vector_gradient(cfdg_fv* fv, const double* u, const double* ub, int scheme,
                                               double* grad) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr, "DeviceFv: null handle");
        fv->d->vector_gradient(u, ub, scheme, grad);
    });
}

extern "C" cfdg_status cfdg_fv_scalar_corrections(cfdg_fv* fv, const double* phi, const double* bv, int scheme,
                                                  const double* flux, const double* gamma, int linear_upwind,
                                                  double* source) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr, "DeviceFv: null handle");
        fv->d->scalar_corrections(phi, bv, scheme, flux, gamma, linear_upwind != 0, source);
    });
}

extern "C" cfdg_status cfdg_fv_momentum_corrections(cfdg_fv* fv, const double* flux, const double* nu_face,
                                                    int linear_upwind, int limited, int venkat, const double* eps2,
                                                    double central, int transpose_term, double* source) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr, "DeviceFv: null handle");
        fv->d->momentum_corrections(flux, nu_face, linear_upwind != 0, limited != 0, venkat != 0, eps2, central,
                                    transpose_term != 0, source);
    });
}

extern "C" cfdg_status cfdg_fv_pressure_equation(cfdg_fv* fv, const cfdg_pressure_inputs* in,
                                                 cfdg_pressure_outputs* out) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr && in != nullptr, "DeviceFv: null argument");
        fv->d->pressure_equation(*in, out);
    });
}

namespace cfdg {
FvPressureArrays fv_pressure_arrays(const cfdg_fv* fv) { return fv->d->pressure(); }
} // namespace cfdg

extern "C" cfdg_status cfdg_fv_on_device(const cfdg_fv* fv, int* on_device) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr && on_device != nullptr, "DeviceFv: null argument");
        *on_device = 1;
    });
}

extern "C" cfdg_status cfdg_fv_timing(cfdg_fv* fv, int enable) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr, "DeviceFv: null handle");
        fv->d->set_timing(enable != 0);
    });
}

extern "C" cfdg_status cfdg_fv_last_kernel_ms(const cfdg_fv* fv, double* ms) {
    return cfdg::guarded([&] {
        cfdg::require(fv != nullptr && ms != nullptr, "DeviceFv: null argument");
        *ms = fv->d->last_kernel_ms();
    });
}
