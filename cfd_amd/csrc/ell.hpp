// Host-side conversion of CSR matrices to the device layouts of the kernels.
#pragma once

#include "hip_util.hpp"
#include "layout.hpp"

#include <vector>

namespace cfdg {

struct HostEll {
    int n_rows = 0;
    int n_cols = 0;
    int n_slices = 0;
    int lanes = 1;        // lanes per row
    std::vector<int> sp;  // n_slices + 1
    std::vector<int> col; // sp[n_slices]
    std::vector<int> src; // CSR position per slot, -1 for padding
};

// Slices of 64 / lanes rows; lane l of slice s serves row s 64 / lanes + l / lanes and slot
// (64 k + l) holds that row's entry k lanes + l mod lanes in CSR order, or padding (value 0)
// pointing at a column that the row touches anyway: the row itself for square matrices, else
// the row's last column (rows past the end: the last row, or column 0).
HostEll csr_to_ell(int n_rows, int n_cols, const int* rp, const int* ci, int lanes = 1);

// Stored slots of the sliced-ELL layout with the given lanes per row (no conversion).
long long ell_slots(int n_rows, const int* rp, int lanes);

// Slice offsets (n_slices + 1) of the sliced-ELL layout with the given lanes per row.
std::vector<int> ell_slice_offsets(int n_rows, const int* rp, int lanes);

// True when every entry of a square CSR matrix lies within 32767 of its row (16-bit offsets).
bool short_offsets_fit(int n_rows, int n_cols, const int* rp, const int* ci);

template <class T>
std::vector<T> ell_values(const HostEll& e, const double* va) {
    std::vector<T> v(e.col.size());
    const long long n = static_cast<long long>(v.size());
    for (long long k = 0; k < n; ++k) {
        const int s = e.src[static_cast<std::size_t>(k)];
        v[static_cast<std::size_t>(k)] = s >= 0 ? static_cast<T>(va[s]) : T(0);
    }
    return v;
}

// Lanes per row for a matrix: 1 (sliced ELL) for short rows, else a power of two near a
// quarter of the mean row length, at most 64.
int pick_tpr(int n_rows, long long nnz);

// Layout of a matrix on the device: lanes per row and sliced ELL (csr = 0) or CSR (csr = 1).
struct Layout {
    int lanes;
    int csr;
};

// The layout for a matrix with these rows and entries: the first rule of the tuned table that
// matches (set_layout_rules), else the built-in rule (pick_tpr, then sliced ELL unless its
// padding exceeds 1.5 times the entries).
Layout choose_layout(int n_rows, const int* rp);
void set_layout_rules(const cfdg_layout_rule* rules, int count);

// Device copy of a sparse matrix (sliced ELL or CSR, see EllView) with FP64 and / or FP32
// values.
struct DevEll {
    int n_rows = 0;
    int n_cols = 0;
    int n_slices = 0;
    int tpr = 1;
    int csr = 0;
    DevBuf<int> sp;
    DevBuf<int> col;
    DevBuf<short> dcol;          // 16-bit column offsets (replaces col when present)
    DevBuf<int> cbase;           // per slice, non-square matrices: dcol holds col - cbase (unsigned)
    DevBuf<int> soff;            // per-slice stencil offsets (replaces col / dcol when present)
    DevBuf<double> v64;
    DevBuf<float> v32;
    DevBuf<unsigned short> vbf;  // BF16 values scaled per row (diagonal slot 0)
    DevBuf<float> rscale;        // per padded row
    DevBuf<float> diag32;        // per padded row

    void upload(const HostEll& e, const double* va, bool fp64, bool fp32, hipStream_t s);
    // Sliced ELL with the given lanes built on the device from the CSR arrays (the same layout
    // as upload(csr_to_ell(...)), FP64 and FP32 values).
    void build_on_device(int n_rows, int n_cols, const int* rp, const int* ci, const double* va,
                         int lanes, hipStream_t s);
    // Chooses the layout from the row lengths (choose_layout), or uses the one given.
    void upload_csr(int n_rows, int n_cols, const int* rp, const int* ci, const double* va,
                    hipStream_t s, const Layout* forced = nullptr);
    [[nodiscard]] EllView view(const void* v, int fmt = 0) const {
        EllView e{};
        e.n_rows = n_rows;
        e.tpr = tpr;
        e.n_slices = n_slices;
        e.fmt = fmt;
        e.csr = csr;
        e.sp = sp.get();
        e.col = col.size() > 0 ? col.get() : nullptr;
        e.dcol = dcol.size() > 0 ? dcol.get() : nullptr;
        e.cbase = cbase.size() > 0 ? cbase.get() : nullptr;
        e.soff = soff.size() > 0 ? soff.get() : nullptr;
        e.val = v;
        e.rscale = rscale.get();
        e.diag = diag32.get();
        return e;
    }
    [[nodiscard]] EllView view64() const { return view(v64.get()); }
    [[nodiscard]] EllView view32() const { return view(v32.get()); }
    [[nodiscard]] EllView viewbf() const { return view(vbf.get(), 1); }
    [[nodiscard]] std::size_t bytes() const {
        return sp.bytes() + col.bytes() + dcol.bytes() + cbase.bytes() + soff.bytes() + v64.bytes() + v32.bytes() +
               vbf.bytes() + rscale.bytes() + diag32.bytes();
    }
};

} // namespace cfdg
