// Device-resident linear systems and the batched solve driver.
#pragma once

#include "ell.hpp"
#include "hip_util.hpp"
#include "layout.hpp"

#include <memory>
#include <vector>

namespace cfdg {

// One AMG level on the device. Level 0 uses the system's assembled operator instead of `a`.
struct DevLevel {
    int n = 0;
    int n_next = 0;
    DevEll a;
    DevEll p;
    DevEll r;
    DevEll rt; // fused restriction R (I - A diag(wd32)) (levels >= 1 of the FP32 cycle), if formed
    DevBuf<double> wd64;
    DevBuf<float> wd32;
    DevBuf<double> v64[3];
    DevBuf<double> b64;
    DevBuf<float> v32[3];
    DevBuf<float> b32;
    DevBuf<double> cinv; // coarsest level only
    DevBuf<double> cinv_tail; // dense inverse of this level's operator where the FP32 cycle ends
    [[nodiscard]] std::size_t bytes() const;
};

enum ProbeKind : int {
    PROBE_OFF = 0,
    PROBE_L0_SWEEP = 1,   // last level-0 Jacobi sweep of the V-cycle (fused with the dots)
    PROBE_OUTER_SPMV = 2  // outer FP64 q = A p fused with p.q
};

// HIP-event timing of one kernel class over the solves of a system.
struct Probe {
    int kind = PROBE_OFF;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::size_t used = 0;     // launches recorded in the current solve
    double total_ms = 0.0;    // accumulated over solves (useful launches only)
    long long launches = 0;
    ~Probe() {
        for (auto& e : events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
};

class Runner;

// Sweeps per side of the mixed-precision V-cycle and the ratio of its damping interval
// (process-wide settings).
int mixed_sweeps();
double mixed_ratio();

// A linear system resident in HBM: level-0 operator (FP64 + FP32 values), right-hand side,
// iterate, Krylov work vectors, and optionally an AMG hierarchy.
class System {
public:
    ~System();
    Probe probe;
    // Cached solve machinery (launch tables, descriptors, graphs) per method / precision.
    std::unique_ptr<Runner>& runner_slot(cfdg_method method, bool use_amg, cfdg_precision precision);
    [[nodiscard]] long long nnz0() const { return nnz0_; }
    explicit System(const cfdg_ldu_pattern& pattern); // values assembled later from LDU
    explicit System(const cfdg_csr& a);               // values given
    System(const System&) = delete;
    System& operator=(const System&) = delete;

    // fuse = false: no fused restrictions are formed (their host product pays off only over
    // many solves with the hierarchy; the server's requests and the AmgCache, which rebuilds
    // every max_age solves, pass false).
    void set_amg(const cfdg_amg_view& v, bool fuse = true);
    void assemble(const double* diag, const double* upper, const double* lower);
    // The same from device arrays (diag, upper, lower of the system's pattern) plus the
    // right-hand side b, copied within HBM.
    void assemble_device(const double* diag, const double* upper, const double* lower, const double* b);
    void set_rhs(const double* b);
    void set_x(const double* x);
    void get_x(double* x);
    // Device-to-device copies of the right-hand side / iterate, queued on stream s.
    void copy_rhs_from(const double* dev, hipStream_t s);
    void copy_x_to(double* dev, hipStream_t s) const;
    [[nodiscard]] std::size_t footprint() const;
    // residual_norms (ldu_matrix.cpp:74-103) of the resident operator, right-hand side and
    // iterate: out[0] = |b - A x|_1, out[1] = the normalisation, out[2] = |b - A x|_2^2.
    void residual_norms(double out[3]);

    [[nodiscard]] int n() const { return n_; }
    [[nodiscard]] int n_faces() const { return nf_; }
    // Levels of the cycle: the hierarchy's, or for the FP32 cycle up to its dense tail level.
    [[nodiscard]] int n_levels(bool fp32 = false) const {
        if (!has_amg_) {
            return 0;
        }
        return fp32 && tail_ > 0 ? tail_ + 1 : static_cast<int>(levels_.size());
    }
    [[nodiscard]] bool has_amg() const { return has_amg_; }
    // V-cycle sweeps per side: the hierarchy's (FP64, the reference algorithm) or the mixed
    // smoother's (FP32 V-cycle).
    [[nodiscard]] int sweeps(bool fp32 = false) const { return fp32 ? mixed_sweeps() : sweeps_; }
    [[nodiscard]] double weight() const { return weight_; }
    // Ratio of the mixed smoother's damping interval the hierarchy's FP32 weights were set for
    // (0: the hierarchy's own weight).
    [[nodiscard]] double smoother_ratio() const { return mixed_ratio_; }
    [[nodiscard]] hipStream_t stream() const { return stream_.get(); }
    [[nodiscard]] SolveState* state() const { return state_.get(); }
    [[nodiscard]] int level_rows(int l) const { return levels_[static_cast<std::size_t>(l)].n; }
    // Lanes per row of level l's operator (which 0), prolongation (1) or restriction (2).
    [[nodiscard]] int tpr(int l, int which) const {
        const DevLevel& d = levels_[static_cast<std::size_t>(l)];
        if (which == 0) {
            return l == 0 ? 1 : d.a.tpr;
        }
        return which == 1 ? d.p.tpr : (which == 2 ? d.r.tpr : d.rt.tpr);
    }
    // Whether level l has the fused restriction of the one-sweep FP32 cycle.
    [[nodiscard]] bool has_fused_restriction(int l) const {
        return levels_[static_cast<std::size_t>(l)].rt.n_rows > 0;
    }

    // Descriptor of this system for kernels of the given precision (FP64 or FP32 V-cycle).
    [[nodiscard]] SysDev descriptor(bool fp32_vcycle, bool use_amg, bool bf16 = false);

    // Refreshes the BF16 copy of the level-0 operator after a new assembly.
    void ensure_bf16() {
        if (!bf16_ready_) {
            make_bf16();
        }
    }

    // Completes work a call left queued on the system's stream (set_x(nullptr)'s fill), so that
    // work on another stream may follow; a no-op when nothing is pending.
    void finish_pending() {
        if (pending_) {
            hip_check(hipStreamSynchronize(stream_.get()), "system stream sync");
            pending_ = false;
        }
    }

    // Generation counter, bumped when the hierarchy or vector set changes (descriptor caches).
    [[nodiscard]] unsigned generation() const { return generation_; }
    // Process-wide unique identity of this system (addresses of destroyed systems are reused).
    [[nodiscard]] unsigned long long uid() const { return uid_; }

private:
    void init_vectors();
    void finish_assembly(); // slot values and inverse diagonal from ldu_, then synchronise
    void ensure_level_vectors(bool fp32);
    // Symmetric stencil storage of the level-0 operator (EllView::sym_w), prepared with an LDU
    // pattern whose offsets form one small stencil, used while the assembled values are exactly
    // symmetric (upper == lower); the full per-slice layout otherwise (and for BF16).
    void build_symmetric_layout(const cfdg_ldu_pattern& pt, const std::vector<int>& csr_src);
    [[nodiscard]] EllView a0_view(bool fp32) const;

    int n_ = 0;
    int nf_ = 0;
    long long nnz0_ = 0; // stored entries of the level-0 operator (padding excluded)
    Stream stream_;
    DevEll a0_;
    DevBuf<int> a0_src_;      // ELL slot -> index into [diag | upper | lower], -1 padding
    DevBuf<int> sym_src_;     // symmetric storage slot -> LDU index, -1 absent (empty: no such layout)
    DevBuf<int> sym_tab_;     // stencil offsets ascending, then the stored slot of each
    DevBuf<double> sym_v64_;
    DevBuf<float> sym_v32_;
    DevBuf<int> sym_flag_;    // set by the symmetry check of an assembly
    int sym_w_ = 0;
    int sym_n_ = 0;
    bool sym_active_ = false; // the last assembly was symmetric: kernels read the symmetric storage
    bool pending_ = false;    // set_x(nullptr)'s fill is queued on stream_ and not yet waited for
    DevBuf<double> ldu_;      // staging of LDU coefficients
    DevBuf<double> dinv0_;
    DevBuf<double> wd0_64_;
    DevBuf<float> wd0_32_;
    DevBuf<double> b_, x_, r_, p_, p2_, q_, z_, r0_, v_, s_, t_, ph_, sh_;
    DevBuf<float> r32_;
    DevBuf<double> partials_;
    DevBuf<SolveState> state_;
    DevBuf<double> norms_; // sum of x and the three residual norms
    std::vector<DevLevel> levels_;
    bool has_amg_ = false;
    int tail_ = 0; // > 0: the FP32 cycle solves level tail_ exactly (cinv_tail) and stops there
    bool assembled_ = false;
    int sweeps_ = 2;
    double weight_ = 0.7;
    double weight0_32_ = 0.7; // level-0 damping of the mixed smoother
    double mixed_ratio_ = 0.0; // interval ratio of the FP32 weights (set with the hierarchy)
    unsigned generation_ = 0;
    unsigned long long uid_ = next_uid();
    static unsigned long long next_uid();
    std::unique_ptr<Runner> runners_[5];
    bool bf16_ready_ = false;
    void make_bf16();
};

// One solve of a resident system from its resident right-hand side and iterate (the C-ABI's
// single-system path: cached launch tables and graphs per method and precision).
cfdg_solve_stats solve_system(System& s, cfdg_method method, bool use_amg, cfdg_precision precision,
                              const cfdg_solve_control& control);

// Solve options shared by every system of a batch.
struct BatchOptions {
    cfdg_method method = CFDG_METHOD_CG;
    bool use_amg = true;
    cfdg_precision precision = CFDG_PREC_FP64;
};

// Solves a set of resident systems together: one launch per stage covers all of them.
// Every system keeps its own reductions, scalars and convergence test, so each result is
// independent of the other members of the set.
void solve_batch(const std::vector<System*>& systems, const BatchOptions& opt,
                 const std::vector<cfdg_solve_control>& control,
                 std::vector<cfdg_solve_stats>& stats, hipStream_t stream);

// Upper bound of the device bytes a system built from this pattern and hierarchy will hold
// (both precisions of the V-cycle, BF16 level-0 values included), from sizes and row lengths.
std::size_t estimate_footprint(const cfdg_ldu_pattern& pattern, const cfdg_amg_view* amg);

struct SlotMachinery;

// Continuous batching over a fixed number of slots. Solves start and finish at iteration
// boundaries while the others keep iterating: each step() queues one iteration that covers
// every running solve, then reads the convergence flags of the previous iteration and retires
// the solves that ended there. Every solve keeps its own state and reductions, so its result
// is bit-identical to a solve of the same system on its own. One thread drives a SlotRunner.
class SlotRunner {
public:
    struct Done {
        int slot;
        cfdg_solve_stats stats;
    };
    SlotRunner(int slots, const BatchOptions& opt, int sweeps, hipStream_t s);
    ~SlotRunner();
    SlotRunner(const SlotRunner&) = delete;
    SlotRunner& operator=(const SlotRunner&) = delete;

    [[nodiscard]] int slots() const;
    [[nodiscard]] int free_slots() const;
    // Solves running or finishing.
    [[nodiscard]] bool busy() const;
    // Starts a solve of the system's resident right-hand side from its resident iterate; its
    // first iteration is part of the next step(). Returns the slot. The system must stay alive
    // and untouched until the slot is reported done.
    int start(System* sys, const cfdg_solve_control& control);
    // One iteration boundary. Blocks until the previous iteration has finished on the device;
    // appends the solves whose results have reached the host to `done`.
    void step(std::vector<Done>& done);

private:
    std::unique_ptr<SlotMachinery> m_;
};

} // namespace cfdg
