// engine.hpp -- device-resident state of one heat-bath VCI solve.
//
// One Engine per request.  It owns, on its own HIP stream, the basis (packed
// uint8 quanta keys plus an open-addressing hash), the Hamiltonian in CSR form,
// the Ritz block and the heat-bath / PT2 work buffers, so they stay on the GPU
// across heat-bath iterations.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "problem.hpp"

namespace vhci {

enum Form { kFP64 = 0, kMixed = 1 };

// Growable device allocation; contents are not preserved on growth.
// Device bytes held by the DevBufs that this thread allocated and not yet
// released, and the peak since the last reset (serving-loop accounting).
int64_t thread_device_bytes();
int64_t thread_device_peak();
void thread_device_reset_peak();

// page-locked host buffer (hipHostMalloc) for small per-step transfers:
// copies from and to it are DMA, asynchronous to the host
struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf();
    void* ensure(size_t bytes);  // contents are not kept when it grows
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Device memory of one context.  Blocks that the context's buffers give
// back are kept idle and handed out again to later requests of up to twice
// their size, so the phases of a solve call neither hipFree (which waits
// for the whole device) nor, once the context has held blocks of the sizes
// they need, hipMalloc.  Idle blocks go back to the device at heat-bath
// iteration boundaries (trim), when an allocation fails, and when the
// context ends.  Every block is used on the context's stream only, so a
// block handed out again is ordered after its previous use.
struct DevPool {
    hipStream_t stream = nullptr;
    std::multimap<size_t, void*> idle;  // capacity -> block
    int64_t held = 0;                   // bytes taken from the device (in use + idle)
    int64_t idle_bytes = 0;
    int64_t mallocs = 0, frees = 0, reuses = 0;  // hipMalloc / hipFree calls, idle blocks handed out
    DevPool() = default;
    DevPool(const DevPool&) = delete;
    DevPool& operator=(const DevPool&) = delete;
    ~DevPool();
    // a block of at least `bytes` (its capacity in cap); throws on failure
    void* take(size_t bytes, size_t& cap);
    void give(void* p, size_t cap);
    // waits for the stream, then returns every idle block to the device
    void trim();
};

// The pool that buffers first sized on this thread draw from: set for the
// duration of each Engine call (PoolScope), none outside.
DevPool* current_pool();
struct PoolScope {
    DevPool* prev;
    explicit PoolScope(DevPool* p);
    ~PoolScope();
    PoolScope(const PoolScope&) = delete;
    PoolScope& operator=(const PoolScope&) = delete;
};

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    DevPool* pool = nullptr;  // taken from current_pool() at the first ensure; none: plain hipMalloc/hipFree
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void* ensure(size_t bytes);  // throws std::runtime_error on failure
    void release();
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(pool, o.pool);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct HStats {
    int64_t n = 0;
    int64_t nnz_offdiag = 0;     // stored off-diagonal entries (both triangles)
    int64_t nnz_struct = -1;     // reference's structural count, if computed
    double build_ms = 0;
    int64_t probed_rows = 0;     // rows whose connections were probed (all, or the new ones)
    int64_t carried = 0;         // entries carried over from the previous build
};

struct EigStats {
    int iterations = 0;
    int spmm_calls = 0;
    int64_t spmm_vectors = 0;
    int fp64_spmm_vectors = 0;
    int device_ortho_steps = 0;   // Lanczos steps orthonormalised on the device
    int device_ortho_redone = 0;  // of those, repeated on the host's general path
    int graphs_captured = 0;      // cycles' device steps captured as a HIP graph
    int graph_launches = 0;       // and launched from it
    double max_resid = 0;
    double ms = 0;
    bool dense = false;
};

struct Pt2Stats {
    int64_t ns2 = 0;
    int64_t pairs = 0;  // (external, variational) pairs of the push form
    int64_t candidates = 0;
    int chunks = 0;     // hash-tag ranges the external space was summed in
    double select_ms = 0, sum_ms = 0;
};

struct HbStats {
    int64_t candidates = 0;
    int64_t unique = 0;
    double ms = 0;
};

class Engine {
  public:
    Engine(const Problem& P, int form, int device);
    ~Engine();

    const Problem& problem() const { return P_; }
    int form() const { return form_; }
    int64_t n() const { return n_; }

    // BasisSet (VCI_headers.h:149): replace the basis; quanta rows n x M.
    void set_basis(const int32_t* q, int64_t n);
    // the initial product basis of the problem, enumerated on the device in
    // the reference's order (equal to Problem::basis0)
    void set_initial_basis();
    void get_basis(int32_t* q) const;

    // MakeHamSparse (Ham.cpp:328-344) on the current basis.
    HStats build_h();
    // Structural nonzero count of the reference's triplet list (pairs that
    // pass its prefilter, explicit zeros included), for the "% sparse" line.
    int64_t structural_nnz();
    // copy of the device CSR (off-diagonal part, stored order) and diagonal
    void get_h(std::vector<int64_t>& rowptr, std::vector<int32_t>& col, std::vector<double>& val,
               std::vector<double>& diag);

    // the same for the basis rows `rows` only (CSR over the nr rows,
    // columns in basis order ascending; diag[nr])
    void get_h_rows(const int64_t* rows, int64_t nr, std::vector<int64_t>& rowptr, std::vector<int32_t>& col,
                    std::vector<double>& val, std::vector<double>& diag);

    // SparseDiagonalize (Ham.cpp:347-370): nev lowest eigenpairs.  Resident
    // Ritz vectors seed the next call (warm start) after the basis grows.
    EigStats eigsolve(int nev, double tol, int maxit, std::vector<double>& E);
    // DenseDiagonalize (Ham.cpp:372-379): all n eigenpairs.
    EigStats dense_all(std::vector<double>& E);
    // resident eigenvectors, row-major n x ncol
    // products regenerate H (no stored entries): chosen when CSR would not fit
    bool matrix_free() const { return matfree_; }
    // storage of the last build: 0 full CSR, 1 upper triangle, 2 matrix-free
    int h_storage() const { return matfree_ ? 2 : (sym_ ? 1 : 0); }
    int ritz_cols() const { return xcols_; }
    void get_vectors(double* C, int ncol) const;
    // PrintFreqs assignment scan of the first ncol resident eigenvectors on
    // the device: per column, the basis rows with |C| >= max |C| / 2
    // (ascending), their |C| and quanta
    void freq_assignments(int ncol, std::vector<FreqColumn>& out) const;
    void set_vectors(const double* C, int ncol);  // upload (restart)

    // AddStatesHB over every basis state (HB.cpp:221-223).  cmax may be null
    // (then computed from the resident vectors with nopt leading columns).
    HbStats hb_select(const double* cmax_host, int nopt, double eps);
    int64_t new_count() const { return nnew_; }
    void get_new_states(int32_t* q) const;
    void append_new_states();

    // DoPT2 (PT2.cpp:16-154) with the resident vectors and E[0..nopt).
    Pt2Stats pt2(const std::vector<double>& E, int nopt, double eps2, std::vector<double>& dE);

    hipStream_t stream() const { return stream_; }
    double last_spmm_ms() const { return last_spmm_ms_; }
    // Y = H X for b host vectors (row-major n x b, basis order) through the
    // solver's product kernels; fp32 selects the mixed form's FP32 product
    void apply_h(const double* X, double* Y, int b, bool fp32);
    // the same product with every device buffer it uses (staging of X and
    // Y, their solver-order copies) in a fixed per-device scratch array:
    // no device allocation.  n * b is limited (kScratchElems); calls on one
    // GPU are serialised by a lock, since they share that GPU's array
    void apply_h_noalloc(const double* X, double* Y, int b, bool fp32);
    // time one SpMM of b vectors, averaged over reps (for the roofline figure)
    double time_spmm(int b, int reps, bool fp32, double* bytes_per_call);
    // a plain copy of `bytes` bytes on the context stream; returns ms
    double copy_probe(size_t bytes);
    // kernel launches per SpMM of b vectors
    int spmm_launches(int b) const;
    // bases of at least this many states are stored in solver order
    static constexpr int64_t kReorderMin = 16384;
    // entries per call of the row sort of the H build (rocPRIM counts in 32 bits)
    static constexpr int64_t kSortSlice = int64_t(1) << 30;
    // PT2 chunk budgets (VHCI_PT2_CHUNK_PAIRS / VHCI_PT2_CHUNK_STATES
    // override them): (a, i) pairs (24 bytes each with the sort's copies,
    // 51.5 GB at the budget) and external states (the dedup pool) one chunk
    // may hold.  Every chunk probes all connection patterns of every
    // variational state twice (count, fill), so fewer chunks cost less; a
    // chunk whose buffers cannot be allocated is split.
    static constexpr int64_t kPt2ChunkPairs = int64_t(1) << 31;
    static constexpr int64_t kPt2ChunkStates = int64_t(1) << 28;
    // per-launch timing of the b-column products that the solver issues:
    // b > 0 starts (and clears) collection, b = 0 stops it.  The read gives
    // total ms, total algorithmic bytes and the launch count since the start.
    void spmm_profile(int b);
    void spmm_profile_read(double out[3]);
    // the eigensolver's Ritz rotations (two n x m x keep GEMMs per restart)
    // while SpMM profiling is on: total ms, flops, calls
    void gemm_profile_read(double out[3]);
    // last H build: n, stored entries, connection patterns, ms; last
    // heat-bath selection (the PT2 one after pt2()): candidates, unique, ms;
    // last PT2: external states, (external, variational) pairs, sum ms; key
    // bytes; PT2 states
    void phase_stats(double out[16]) const;
    // algorithmic bytes of one product of b vectors: stored entries (value +
    // column), row pointers, the diagonal, X read once and Y written once
    double spmm_bytes(int b, int mode) const;

    // device-memory pool of this context: calls made, bytes held
    void pool_stats(int64_t out[5]) const;

  private:
    // declared first so that it outlives every buffer below
    mutable DevPool pool_;
    void upload_tables();
    void make_solver_order();
    void quanta_order(const uint32_t* keys, int64_t n, int32_t* perm);
    void rebuild_hash(int64_t cap_hint);
    void spmm(const double* X, int ldx, double* Y, int ldy, int b, bool low, const float* X32 = nullptr,
              int ldx32 = 0);
    void dense_solve(int nev, std::vector<double>& E, EigStats& st);
    // column-major block helpers of the iterative eigensolvers
    void apply_cm(const double* X, double* Y, int64_t ld, int k, bool low, EigStats& st, double* Y2 = nullptr);
    // orthonormalise the k columns at V + m0*ld against the first m0 and
    // among themselves; lanczos: the block already had its local Lanczos
    // components removed, so the first projection is applied only when one
    // of its coefficients exceeds kOrthoSkip times the block's smallest norm
    int ortho_cm(double* V, int64_t ld, int m0, int k, bool lanczos = false);
    static constexpr double kOrthoSkip = 1e-13;
    void random_cm(double* V, int64_t ld, int k, uint64_t seed);
    void davidson(int nev, double tol, int maxit, std::vector<double>& E, EigStats& st);
    void krylov_schur(int nev, double tol, int maxit, std::vector<double>& E, EigStats& st);
    int64_t hb_count(double eps);
    template <int NW>
    bool hb_gen_t(int64_t total, uint64_t tlo, uint64_t thi, uint64_t max_pool, HbStats& st);
    template <int NW>
    void hb_select_t(double eps, HbStats& st);
    template <int NW>
    void build_h_t(HStats& st);
    template <int NW>
    void pt2_t(const std::vector<double>& E, int nopt, double eps2, Pt2Stats& st, std::vector<double>& dE);
    template <int NW>
    bool pt2_push_t(const double* Ed, int nopt, int nwaves, int64_t max_pairs, bool carry, uint64_t tlo,
                    uint64_t thi, Pt2Stats& st);
    template <int NW>
    void pt2_pull_t(const double* Ed, int nopt, int nwaves, bool carry);
    void release_solver_work();

    Problem P_;
    Tables T_;
    int form_;
    int device_;
    int KW_;  // uint4 words per key
    hipStream_t stream_ = nullptr;
    rocblas_handle blas_ = nullptr;

    // basis
    int64_t n_ = 0;
    DevBuf keys_, slots_;
    uint64_t hmask_ = 0;
    DevBuf bloom_;        // filter of the basis hash's keys
    uint64_t bmask_ = 0;  // its bit count - 1
    // tables
    DevBuf d_ladder_, d_elem_, d_deltas_, d_groups_, d_fclist_, d_diag_, d_freq_;
    DevBuf d_hbfc_, d_hbvaroff_, d_hbvars_, d_ball_;
    DevBuf d_walk_;           // the pattern walks' 8-byte records of d_deltas_
    DevBuf d_ball_walk_;      // and of d_ball_
    DevBuf d_doff_, d_boff_;  // block offsets of the pattern table and the ball
    int P_deltas_ = 0;
    // Hamiltonian
    bool h_valid_ = false;
    int64_t nnz_ = 0;
    HostBuf hr_, hz_;  // block Gram-Schmidt: projections down, block mix up
    // incremental H build: basis size of the last stored build whose states
    // are still the first ones of the basis (0: none), and the previous
    // order and matrix while the next one is assembled
    int64_t built_n_ = 0;
    DevBuf perm_o_, inv_o_, rowptr_o_, col_o_, val_o_, hi_o_, lo_o_, extra_, oldcnt_, cursor_;
    bool matfree_ = false;  // products regenerate H from the hash (no CSR)
    // half storage: only the strict upper triangle (solver order) is stored;
    // products add the transposed terms through a fixed-point accumulator
    bool sym_ = false;
    double hnorm_ = 0;  // max over rows of sum_j |H_ij| (off-diagonal), the accumulator's bound
    DevBuf rowcnt_up_, rowabs_, hacc_, hcmax_, hpack_x_, hpack_y_;
    size_t hacc_zero_ = 0;  // bytes of hacc_ known to hold zeros
    // 0 full CSR, 1 half storage, 2 matrix-free: the first that fits beside
    // the eigensolver's blocks (VHCI_H_STORAGE=full|half|free forces one);
    // half storage may also shrink the Krylov-Schur basis (ks_mfac_) to fit
    // words: bytes of the build's hit words, held now and counted as free;
    // *keep_words: full storage, and its fill and sort fit beside them
    int choose_storage(int64_t nnz, int64_t nnz_up, size_t words = 0, bool* keep_words = nullptr);
    double ks_mfac_ = 4.0;  // Krylov-Schur basis size / nev
    void spmm_half(const double* X, int ldx, double* Y, int ldy, int b, bool low, const float* X32, int ldx32);
    template <int NW>
    void spmm_free_t(const double* X, int ldx, double* Y, int ldy, int b);
    DevBuf rowcnt_, rowptr_, col_, val_, hi_, lo_, D_;
    // solver order: H, D and the eigensolver's blocks index states in the
    // order perm_ (solver row -> basis state; inv_ the inverse); the
    // resident Ritz block X_ stays in basis order
    DevBuf perm_, inv_;
    bool perm_on_ = false;
    // Ritz block (row-major n x xcols_, ld = xcols_)
    DevBuf X_;
    int64_t xrows_ = 0;
    int xcols_ = 0;
    // eigensolver work
    DevBuf V_, W_, HX_, Xw_, T_b_, G_, S_, theta_, work_, part_, X32_, info_;
    DevBuf stage_in_, stage_in32_, stage_out_, gram_, otmp_, Gw_, tsm_part_, rr_part_, rot_s_;
    DevBuf ostat_, ozd_;  // device-side Lanczos orthonormalisation: step statuses, the block's Z
    // heat-bath work
    DevBuf cmax_, candcnt_, candoff_, nslots_, pool_keys_, pool_rank_, counters_, col_rank_, col_idx_,
        srt_rank_, srt_idx_, newkeys_, scan_tmp_;
    int64_t nnew_ = 0;
    uint64_t nmask_ = 0;  // mask of the new-state hash of the last selection
    DevBuf pool_pos_;     // pool entry -> position in first-discovery order
    // PT2 work
    DevBuf s2keys_, pt2_part_, C32_;
    DevBuf pairoff_;  // PT2 push: prefix of the pair counts per variational state
    DevBuf s2slots_;  // PT2 push: hash of one chunk's external states
    DevBuf s2bloom_;  // and its filter
    int64_t last_pairs_ = 0;  // pairs counted by a PT2 chunk that exceeded the budget
    double last_spmm_ms_ = 0;
    struct SpmmProf {
        int b = 0;
        std::vector<hipEvent_t> ev;  // start, end pairs
        size_t used = 0;
        double bytes = 0;
        std::vector<hipEvent_t> gev;  // Ritz rotation GEMMs
        double gemm_ms = 0, gemm_flops = 0;
        int64_t gemm_calls = 0;
    } prof_;
    HStats last_h_;
    HbStats last_hb_;
    Pt2Stats last_pt2_;
    int64_t last_pt2_n_ = 0;
    int last_pt2_nopt_ = 0;
};

}  // namespace vhci
