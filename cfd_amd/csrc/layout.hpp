// Device data layout shared by the host driver and the kernels: sliced-ELL views, per-level and
// per-system descriptors, and the per-system solve state.
#pragma once

#include <hip/hip_runtime.h>

namespace cfdg {

constexpr int k_wave = 64;
constexpr int k_block = 256;
constexpr int k_slices_per_block = k_block / k_wave;
constexpr int k_max_levels = 25;

enum Slot : int {
    SLOT_PQ = 0,
    SLOT_RR = 1,
    SLOT_RZ = 2,
    SLOT_ZQ = 3,
    SLOT_BB = 4,
    SLOT_TRUE = 5,
    SLOT_R0R = 6,
    SLOT_R0V = 7,
    SLOT_SS = 8,
    SLOT_TT = 9,
    SLOT_TS = 10,
    k_slots = 11
};

enum Status : int {
    ST_RUNNING = 0,
    ST_CONVERGED = 1,
    ST_BREAKDOWN = 2,
    ST_MAXIT = 3,
    ST_CONVERGED_PENDING = 4 // BiCGStab early exit: x += alpha ph still to apply
};

struct SolveState {
    double tol;
    double rz;
    double pq;
    double alpha;
    double beta;
    double res;
    double bnorm;
    double true_res;
    double rho;
    double omega;
    double zq;
    int it;
    int max_it;
    int status;
    int pad;
};

// A sparse matrix as the kernels see it; tpr (a power of two) lanes read each row and a
// workgroup covers 256 / tpr rows.
// Sliced ELL (csr == 0): a slice is one wavefront, 64 / tpr rows of tpr lanes each (ptr = slice
// offsets); lane l of slice s serves row s 64 / tpr + l / tpr, and its round k (slot
// ptr[s] + 64 k + l) holds the row's entry k tpr + l mod tpr in CSR order, or padding (value 0,
// a column the row touches anyway). With tpr == 1 a lane owns a row and sums it in CSR order.
// CSR (csr == 1, ptr = row offsets): lane j of a row's group sums its entries j, j + tpr, ...
// Both give the same sums for the same tpr. dcol (if not null) replaces col by 16-bit offsets
// column - row. soff (if not null; square, tpr == 1) replaces both by a per-slice stencil: round
// k of every row of slice s reads column row + soff[sp[s] / 64 + k] (clamped to the matrix), the
// slice's distinct offsets in ascending order, so a row's entries keep their CSR order and the
// offsets a row lacks hold value 0. fmt 1 stores the values as BF16 scaled per row (a_ij = rscale_i * bf16_ij) with
// the diagonal slot zero and the diagonal in FP32 in diag (rscale and diag padded to whole
// slices; tpr == 1 only).
// Symmetric stencil storage (sym_w > 0; square, exactly symmetric, tpr == 1): one stencil of
// sym_n offsets for all rows, soff = [offsets ascending | stored slot of each], and only the
// entries with offset >= 0 stored: sym_w slots per row, slice s holding rows 64 s .. 64 s + 63
// at val[s sym_w 64 + j 64 + lane] for the j-th offset >= 0. An entry with offset o < 0 is read
// from row + o's stored slot of -o (a_ij = a_ji); a row's entries are summed in offset order,
// the order (and result) of the per-slice stencil layout.
struct EllView {
    int n_rows;
    int tpr;
    int n_slices;    // ELL only
    int fmt;         // 0 values of the kernel's type, 1 BF16 + row scale + FP32 diagonal
    int csr;         // 1: CSR rows, 0: sliced ELL
    int sym_w;       // symmetric stencil storage: stored slots per row (0: not used)
    int sym_n;       // symmetric stencil storage: offsets of the stencil
    const int* sp;   // ELL slice offsets or CSR row offsets
    const int* col;
    const short* dcol;  // 16-bit columns: col = row + dcol, or col = cbase[slice] + (unsigned) dcol
    const int* cbase;   // per slice: first column of the 16-bit window (non-square matrices)
    const int* soff;    // per-slice stencil offsets (square matrices, one lane per row)
    const void* val; // double, float or BF16 bits
    const float* rscale;
    const float* diag;
};

struct LevelDev {
    EllView a;        // n x n
    EllView p;        // n x n_next
    EllView r;        // n_next x n
    EllView rt;       // n_next x n: R (I - A diag(wd)), the down leg from x = 0 with one sweep
    int n;
    int n_next;
    int coarsest;     // 1 on the coarsest level
    int pad;
    void* v[3];       // work vectors (iterate / residual / ping-pong), element type T
    void* b;          // right-hand side of the level, element type T
    const void* wd;   // jacobi_weight * dinv, element type T
    const double* cinv; // dense row-major inverse (coarsest level only)
};

struct SysDev {
    SolveState* st;
    double* partials; // k_slots * nblk0
    int nblk0;        // 256-row blocks of level 0
    int n;
    int n_levels;
    int flip;               // XOR on the launch's iteration parity: 0 unless the system joined
                            // a running set at an even iteration
    EllView a64;            // level-0 operator in FP64
    const double* b;        // right-hand side
    double* x;              // iterate
    double* r;
    double* p;
    double* p2;             // second search-direction buffer (alternates with p)
    double* q;
    double* z;              // Jacobi-preconditioned residual / BiCGStab scratch
    float* r32;             // FP32 copy of r (mixed V-cycle input)
    const double* dinv;     // 1 / a_ii (1.0 where a_ii == 0)
    double* r0;             // BiCGStab
    double* v;
    double* s;
    double* t;
    double* ph;
    double* sh;
    LevelDev lv[k_max_levels];
};

} // namespace cfdg
