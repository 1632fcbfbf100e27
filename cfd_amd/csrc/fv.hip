// Finite-volume cell kernels on the device: cell gradients (Green-Gauss, weighted least
// squares, skewness-corrected Gauss), the Barth-Jespersen / Venkatakrishnan limiter and the
// explicit internal-face corrections of the scalar and momentum equations (the operations of
// gpu::DeviceFv, src/gpu/include/cfd/gpu/fv_device.hpp:34-71; per-cell math of
// src/gpu/src/fv_cell.inc:25-260, sequencing of src/gpu/src/fv_openmp.cpp:130-322).
//
// One thread per cell gathers over the cell's faces in ascending face order (the mesh's CSR
// face list), with the reference's operation order and no contraction of multiply-adds, so every
// result equals the serial CPU face loops bit for bit; nothing is scattered, nothing is atomic.
// The three components of a vector gradient are formed in one pass over the faces (the face
// geometry is read once instead of three times). Workgroups map to XCDs in contiguous row
// ranges, so the neighbours a cell gathers from sit in the same L2.
#include "common.hpp"
#include "hip_util.hpp"

#include <cfdgpu.h>

#include <hip/hip_runtime.h>

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

// Device copies of the mesh arrays.
struct FvMesh {
    int nc;
    int nif;
    int nf;
    const int* own;
    const int* nei;
    const int* start;
    const int* faces;
    const double* w;
    const double* sf;
    const double* cf;
    const double* cc;
    const double* vol;
    const double* kv;
    const double* lsq_d;
    const double* lsq;
};

// Gradient of NC components (comp0 .. comp0 + NC - 1) of a field with `stride` values per cell
// and boundary face; component i goes to out[c * ostride + ooff + 3 i + j]. Least squares:
// A (sum_f d_f dphi_f) with A the cell's column-major inverse; Green-Gauss: sum_f S_f phi_f / V
// with phi_f = w phi_P + (1 - w) phi_N (fv_cell.inc:25-73).
template <bool Lsq, int NC>
__device__ __forceinline__ void cell_gradient(const FvMesh& m, int c, const double* phi, const double* bv,
                                              int stride, int comp0, double* out, int ostride, int ooff) {
    const auto own = g(m.own);
    const auto nei = g(m.nei);
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
        pc[i] = fp[c * stride + comp0 + i];
    }
    const int k1 = g(m.start)[c + 1];
    for (int k = g(m.start)[c]; k < k1; ++k) {
        const int f = g(m.faces)[k];
        if (Lsq) {
            const double d0 = g(m.lsq_d)[3 * f];
            const double d1 = g(m.lsq_d)[3 * f + 1];
            const double d2 = g(m.lsq_d)[3 * f + 2];
            if (f < m.nif) {
                const int p = own[f];
                const int n = nei[f];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double dv = fp[n * stride + comp0 + i] - fp[p * stride + comp0 + i];
                    g0[i] += d0 * dv;
                    g1[i] += d1 * dv;
                    g2[i] += d2 * dv;
                }
            } else {
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double dv = g(bv)[(f - m.nif) * stride + comp0 + i] - pc[i];
                    g0[i] += d0 * dv;
                    g1[i] += d1 * dv;
                    g2[i] += d2 * dv;
                }
            }
        } else {
            const double s0 = g(m.sf)[3 * f];
            const double s1 = g(m.sf)[3 * f + 1];
            const double s2 = g(m.sf)[3 * f + 2];
            if (f < m.nif) {
                const int p = own[f];
                const int n = nei[f];
                const double wf = g(m.w)[f];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double pf = wf * fp[p * stride + comp0 + i] + (1.0 - wf) * fp[n * stride + comp0 + i];
                    if (p == c) {
                        g0[i] += s0 * pf;
                        g1[i] += s1 * pf;
                        g2[i] += s2 * pf;
                    } else {
                        g0[i] -= s0 * pf;
                        g1[i] -= s1 * pf;
                        g2[i] -= s2 * pf;
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double b = g(bv)[(f - m.nif) * stride + comp0 + i];
                    g0[i] += s0 * b;
                    g1[i] += s1 * b;
                    g2[i] += s2 * b;
                }
            }
        }
    }
    double* o = out + static_cast<long long>(c) * ostride + ooff;
    if (Lsq) {
        const auto a = g(m.lsq) + 9LL * c;
        const double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6], a7 = a[7],
                     a8 = a[8];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            o[3 * i] = a0 * g0[i] + a3 * g1[i] + a6 * g2[i];
            o[3 * i + 1] = a1 * g0[i] + a4 * g1[i] + a7 * g2[i];
            o[3 * i + 2] = a2 * g0[i] + a5 * g1[i] + a8 * g2[i];
        }
    } else {
        const double v = g(m.vol)[c];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            o[3 * i] = g0[i] / v;
            o[3 * i + 1] = g1[i] / v;
            o[3 * i + 2] = g2[i] / v;
        }
    }
}

// Skewness-corrected Green-Gauss gradient with the least-squares gradients lg (3 per component
// and cell, component i of cell c at lg[c * 3 NC + 3 i]): face values w phi_P + (1 - w) phi_N +
// (w g_P + (1 - w) g_N) . (x_f - x_ip), x_ip = w C_P + (1 - w) C_N (fv_cell.inc:77-118).
template <int NC>
__device__ __forceinline__ void cell_gradient_corrected(const FvMesh& m, int c, const double* phi, const double* bv,
                                                        int stride, int comp0, const double* lg, double* out,
                                                        int ostride, int ooff) {
    const auto own = g(m.own);
    const auto nei = g(m.nei);
    const auto fp = g(phi);
    const auto gl = g(lg);
    const auto cc = g(m.cc);
    double g0[NC];
    double g1[NC];
    double g2[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        g0[i] = 0.0;
        g1[i] = 0.0;
        g2[i] = 0.0;
    }
    const int k1 = g(m.start)[c + 1];
    for (int k = g(m.start)[c]; k < k1; ++k) {
        const int f = g(m.faces)[k];
        const double s0 = g(m.sf)[3 * f];
        const double s1 = g(m.sf)[3 * f + 1];
        const double s2 = g(m.sf)[3 * f + 2];
        if (f < m.nif) {
            const int p = own[f];
            const int n = nei[f];
            const double wf = g(m.w)[f];
            double r[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double xip = wf * cc[3 * p + j] + (1.0 - wf) * cc[3 * n + j];
                r[j] = g(m.cf)[3 * f + j] - xip;
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                double pf = wf * fp[p * stride + comp0 + i] + (1.0 - wf) * fp[n * stride + comp0 + i];
                double dot = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double gf = wf * gl[3 * NC * p + 3 * i + j] + (1.0 - wf) * gl[3 * NC * n + 3 * i + j];
                    dot = j == 0 ? gf * r[0] : dot + gf * r[j];
                }
                pf += dot;
                if (p == c) {
                    g0[i] += s0 * pf;
                    g1[i] += s1 * pf;
                    g2[i] += s2 * pf;
                } else {
                    g0[i] -= s0 * pf;
                    g1[i] -= s1 * pf;
                    g2[i] -= s2 * pf;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double b = g(bv)[(f - m.nif) * stride + comp0 + i];
                g0[i] += s0 * b;
                g1[i] += s1 * b;
                g2[i] += s2 * b;
            }
        }
    }
    double* o = out + static_cast<long long>(c) * ostride + ooff;
    const double v = g(m.vol)[c];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        o[3 * i] = g0[i] / v;
        o[3 * i + 1] = g1[i] / v;
        o[3 * i + 2] = g2[i] / v;
    }
}

template <bool Lsq, int NC>
__global__ void __launch_bounds__(k_fv_block) k_fv_gradient(FvMesh m, const double* phi, const double* bv, int stride,
                                                            double* out, int ostride) {
    const int c = fv_cell();
    if (c < m.nc) {
        cell_gradient<Lsq, NC>(m, c, phi, bv, stride, 0, out, ostride, 0);
    }
}

template <int NC>
__global__ void __launch_bounds__(k_fv_block) k_fv_gradient_corrected(FvMesh m, const double* phi, const double* bv,
                                                                      int stride, const double* lg, double* out,
                                                                      int ostride) {
    const int c = fv_cell();
    if (c < m.nc) {
        cell_gradient_corrected<NC>(m, c, phi, bv, stride, 0, lg, out, ostride, 0);
    }
}

// Limiter of the three components of u (fv_cell.inc:123-164): bounds from the cell and its
// face neighbours (boundary values for boundary faces), then per face and component the ratio
// of the admissible change to the extrapolated one, Barth-Jespersen min(1, bound / d) or the
// Venkatakrishnan function with eps^2 per component; |d| < 1e-300 is skipped.
__global__ void __launch_bounds__(k_fv_block) k_fv_limiter(FvMesh m, const double* u, const double* ub,
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
    const int k0 = g(m.start)[c];
    const int k1 = g(m.start)[c + 1];
    for (int k = k0; k < k1; ++k) {
        const int f = g(m.faces)[k];
        gptr<double> v;
        if (f < m.nif) {
            const int p = g(m.own)[f];
            v = fu + 3 * (p == c ? g(m.nei)[f] : p);
        } else {
            v = g(ub) + 3 * (f - m.nif);
        }
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
    for (int k = k0; k < k1; ++k) {
        const int f = g(m.faces)[k];
        const double r0 = g(m.cf)[3 * f] - c0;
        const double r1 = g(m.cf)[3 * f + 1] - c1;
        const double r2 = g(m.cf)[3 * f + 2] - c2;
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

// Explicit internal-face fluxes of a scalar equation into source (fv_cell.inc:167-195):
// e_f = -F_f g_up . (x_f - C_up) [linear upwind] + gamma_f k_f . (w g_P + (1 - w) g_N),
// added for the owner and subtracted for the neighbour.
__global__ void __launch_bounds__(k_fv_block) k_fv_scalar_corrections(FvMesh m, const double* grad, const double* flux,
                                                                      const double* gamma, int linear_upwind,
                                                                      double* source) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    const auto gg = g(grad);
    const auto cc = g(m.cc);
    double s = 0.0;
    const int k1 = g(m.start)[c + 1];
    for (int k = g(m.start)[c]; k < k1; ++k) {
        const int f = g(m.faces)[k];
        if (f >= m.nif) {
            break; // boundary faces follow the internal ones
        }
        const int p = g(m.own)[f];
        const int n = g(m.nei)[f];
        const double fl = g(flux)[f];
        double e = 0.0;
        if (linear_upwind != 0) {
            const int up = fl >= 0.0 ? p : n;
            const double r0 = g(m.cf)[3 * f] - cc[3 * up];
            const double r1 = g(m.cf)[3 * f + 1] - cc[3 * up + 1];
            const double r2 = g(m.cf)[3 * f + 2] - cc[3 * up + 2];
            e -= fl * (gg[3 * up] * r0 + gg[3 * up + 1] * r1 + gg[3 * up + 2] * r2);
        }
        const double wf = g(m.w)[f];
        const double gf0 = wf * gg[3 * p] + (1.0 - wf) * gg[3 * n];
        const double gf1 = wf * gg[3 * p + 1] + (1.0 - wf) * gg[3 * n + 1];
        const double gf2 = wf * gg[3 * p + 2] + (1.0 - wf) * gg[3 * n + 2];
        e += g(gamma)[f] * (g(m.kv)[3 * f] * gf0 + g(m.kv)[3 * f + 1] * gf1 + g(m.kv)[3 * f + 2] * gf2);
        s = p == c ? s + e : s - e;
    }
    source[c] += s;
}

// Explicit internal-face fluxes of the momentum equation (fv_cell.inc:199-260): per component
// -F_f lim_up (g_up^T (x_f - C_up)) [linear upwind, blended with the linear face value when
// central > 0, limited when lim is given] + nu_f (g_f^T k_f) [+ nu_f g_f S_f], g_f the
// interpolated 3x3 gradient.
__global__ void __launch_bounds__(k_fv_block) k_fv_momentum_corrections(FvMesh m, const double* grad, const double* lim,
                                                                        int limited, const double* flux,
                                                                        const double* nu_face, int linear_upwind,
                                                                        double central, const double* u,
                                                                        int transpose_term, double* source) {
    const int c = fv_cell();
    if (c >= m.nc) {
        return;
    }
    const auto gg = g(grad);
    const auto fu = g(u);
    const auto cc = g(m.cc);
    double s[3] = {0.0, 0.0, 0.0};
    const int k1 = g(m.start)[c + 1];
    for (int k = g(m.start)[c]; k < k1; ++k) {
        const int f = g(m.faces)[k];
        if (f >= m.nif) {
            break;
        }
        const int p = g(m.own)[f];
        const int n = g(m.nei)[f];
        const double fl = g(flux)[f];
        const double wf = g(m.w)[f];
        double e[3] = {0.0, 0.0, 0.0};
        if (linear_upwind != 0) {
            const int up = fl >= 0.0 ? p : n;
            const double r0 = g(m.cf)[3 * f] - cc[3 * up];
            const double r1 = g(m.cf)[3 * f + 1] - cc[3 * up + 1];
            const double r2 = g(m.cf)[3 * f + 2] - cc[3 * up + 2];
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
        const double k0 = g(m.kv)[3 * f];
        const double kk1 = g(m.kv)[3 * f + 1];
        const double k2 = g(m.kv)[3 * f + 2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            e[i] += nu * (gf[3 * i] * k0 + gf[3 * i + 1] * kk1 + gf[3 * i + 2] * k2);
        }
        if (transpose_term != 0) {
            const double a0 = g(m.sf)[3 * f];
            const double a1 = g(m.sf)[3 * f + 1];
            const double a2 = g(m.sf)[3 * f + 2];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                e[i] += nu * (gf[i] * a0 + gf[i + 3] * a1 + gf[i + 6] * a2);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            s[i] = p == c ? s[i] + e[i] : s[i] - e[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        source[3 * c + i] += s[i];
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
        own_.upload(v.owner, f, s);
        const int zero_i = 0;
        const double zero_d = 0.0;
        nei_.upload(nif_ > 0 ? v.neighbour : &zero_i, at_least_one(nif_), s);
        w_.upload(nif_ > 0 ? v.weight : &zero_d, at_least_one(nif_), s);
        start_.upload(v.cell_offsets, c + 1, s);
        faces_.upload(n_list > 0 ? v.cell_faces : &zero_i, at_least_one(n_list), s);
        sf_.upload(v.sf, 3 * f, s);
        cf_.upload(v.face_centre, 3 * f, s);
        cc_.upload(v.cell_centre, 3 * c, s);
        vol_.upload(v.volume, c, s);
        kv_.upload(v.correction, 3 * f, s);
        lsq_d_.upload(v.lsq_d, 3 * f, s);
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
        launch([&](hipStream_t s) {
            k_fv_scalar_corrections<<<blocks(), k_fv_block, 0, s>>>(mesh(), sgrad_.get(), flux_.get(), coef_.get(),
                                                                    linear_upwind ? 1 : 0, src_.get());
        });
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
        launch([&](hipStream_t s) {
            k_fv_momentum_corrections<<<blocks(), k_fv_block, 0, s>>>(
                mesh(), vgrad_.get(), lim_.get(), use_lim ? 1 : 0, flux_.get(), coef_.get(), linear_upwind ? 1 : 0,
                central, vphi_.get(), transpose_term ? 1 : 0, src_.get());
        });
        get(src_, source, 3LL * nc_);
    }

    void set_timing(bool on) { timing_ = on; }
    [[nodiscard]] double last_kernel_ms() const { return last_ms_; }

private:
    [[nodiscard]] int nb() const { return nf_ - nif_; }
    [[nodiscard]] int blocks() const { return (nc_ + k_fv_block - 1) / k_fv_block; }

    static void check_scheme(int scheme) {
        require(scheme >= 0 && scheme <= 2, "DeviceFv: gradient scheme must be 0, 1 or 2");
    }

    [[nodiscard]] FvMesh mesh() const {
        return FvMesh{nc_,         nif_,         nf_,       own_.get(), nei_.get(), start_.get(),
                      faces_.get(), w_.get(),    sf_.get(), cf_.get(),  cc_.get(),  vol_.get(),
                      kv_.get(),    lsq_d_.get(), lsq_.get()};
    }

    // Gradient of NC interleaved components (fv_openmp.cpp:257-311): least squares, Green-Gauss,
    // or least squares into tmp followed by the corrected Gauss gradient.
    template <int NC>
    void gradient_loop(const double* phi, const double* bv, int scheme, double* out, int ostride) {
        const FvMesh m = mesh();
        if (scheme == 2) {
            double* tmp = tmp_.get();
            launch([&](hipStream_t s) {
                k_fv_gradient<true, NC><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, tmp, 3 * NC);
            });
            launch([&](hipStream_t s) {
                k_fv_gradient_corrected<NC><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, tmp, out, ostride);
            });
            return;
        }
        if (scheme == 1) {
            launch([&](hipStream_t s) {
                k_fv_gradient<true, NC><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, out, ostride);
            });
        } else {
            launch([&](hipStream_t s) {
                k_fv_gradient<false, NC><<<blocks(), k_fv_block, 0, s>>>(m, phi, bv, NC, out, ostride);
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
    Stream stream_;
    DevBuf<int> own_, nei_, start_, faces_;
    DevBuf<double> w_, sf_, cf_, cc_, vol_, kv_, lsq_d_, lsq_;
    DevBuf<double> phi_, bv_, sgrad_, vphi_, vbv_, vgrad_, lim_, flux_, coef_, src_, tmp_;
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
