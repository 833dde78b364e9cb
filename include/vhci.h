/*
 * vhci.h -- C ABI of libvhci, the MI355X heat-bath VCI engine.
 *
 * The reference program (berkelbach-group/VHCI) has no plugin or FFI layer:
 * its hot path is four C++ functions over process globals, called from
 * Perform_HCI (src/HB.cpp:176-265) and main (src/VHCI.cpp:39-137).  Each
 * entry point below replaces one of them (or the globals they read) and is
 * called by the `vhci` driver in the same order.
 *
 * Conventions
 *   - Quanta are int32 rows, N x M row-major (BasisSet, VCI_headers.h:149).
 *   - Eigenvector blocks are row-major N x ncol (column j = state j).
 *   - Every function returns VHCI_OK (0) or a negative VHCI_E* code;
 *     vhci_last_error(ctx) / vhci_problem_error() give the message, which
 *     for the reference's own failure cases is the reference's text.
 *   - A context owns its device state on its own HIP stream.  Calls on one
 *     context must not overlap; different contexts are independent.
 *   - Precision forms: VHCI_FP64 stores H and runs every product in FP64;
 *     VHCI_MIXED stores H as FP32 hi + FP32 lo, runs the bulk of the
 *     eigensolver and PT2 numerators in FP32, and the final eigen-refinement,
 *     PT2 denominators and sums in FP64.
 */
#ifndef VHCI_H
#define VHCI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VHCI_OK 0
#define VHCI_E_ARG -1      /* bad argument or call order */
#define VHCI_E_INPUT -2    /* input file / checkpoint error */
#define VHCI_E_DEVICE -3   /* HIP / device error */
#define VHCI_E_CONVERGE -4 /* eigensolver did not converge */
#define VHCI_E_LIMIT -5    /* size limit (modes, quanta, states) */

#define VHCI_FP64 0
#define VHCI_MIXED 1

typedef struct vhci_problem vhci_problem;
typedef struct vhci_ctx vhci_ctx;

/* ---- problem (replaces ReadCIInput and its globals) ------------------- */

/* Parse an input file: src/Input_Reader.cpp:170-492 (header, modes, force
 * constants, ScaleFC :494-515, order split :273-291, initial basis
 * :296-364, AnharmHB :366-460).  restart != 0 builds only the ground state,
 * as the reference does when a checkpoint exists. */
int vhci_problem_load(const char* path, int restart, vhci_problem** out);
/* Same, from an in-memory copy of the file text. */
int vhci_problem_parse(const char* text, int restart, vhci_problem** out);
void vhci_problem_free(vhci_problem* p);
const char* vhci_problem_error(void);

/* Integer settings: 0 M, 1 NEig, 2 Ntot, 3 PT2 flag, 4 N0 (initial basis
 * size), 5 highest FC order, 6 number of force constants, 7 heat-bath list
 * length, 8 number of connection patterns, 9 element FCs.
 * Real settings: 0 eps1 (HCI_Eps), 1 eps2 (PT2_Eps). */
int64_t vhci_problem_int(const vhci_problem* p, int what);
double vhci_problem_real(const vhci_problem* p, int what);
int vhci_problem_freqs(const vhci_problem* p, double* freqs);
int vhci_problem_initial_basis(const vhci_problem* p, int32_t* q);
/* Force-constant lists as the reference holds them: which = 0 AnharmFC,
 * 1 Cubic, 2 Quartic, 3 Quintic, 4 Sextic, 5 AnharmHB (construction order),
 * 6 AnharmHB sorted by |f| (HB.cpp:17-22, stable).  sm/mp receive up to 6
 * entries (ShortModes, ModePowers). */
int vhci_problem_fc_count(const vhci_problem* p, int which);
int vhci_problem_fc(const vhci_problem* p, int which, int idx, double* fc, int* order, int* nsm, int* sm,
                    int* mp);
/* Ladder path sum <qe| (a+a^dag)^k |qs> in the reference's accumulation
 * order (CreationLO/AnnihilationLO/AnharmPot, src/Ham.cpp:16-104). */
double vhci_ladder_sum(int k, int qs, int qe);

/* Checkpoint files (host only).  Write replaces WriteCheckpoint
 * (src/Core_functions.cpp:94-117): n M NEig, then the frequencies and the nE
 * energies, then one line per state with its M quanta and ncol coefficients
 * (C row-major n x ncol), the reals in fixed notation with 17 decimals,
 * formatted on `threads` host threads.  binary != 0 writes the opt-in
 * binary form (same content).  Read replaces ReadCheckpoint
 * (src/Core_functions.cpp:119-167) for either form with the reference's
 * error messages (through vhci_problem_error): *n receives the number of
 * states; with q (n*M), E and C (n*NEig) non-null and qcap, ccap (in
 * elements) large enough the basis, the NEig energies and coefficients are
 * filled; with q, E or C null only the header is read, *n is set and
 * VHCI_E_ARG returned (a size query). */
int vhci_checkpoint_write(const vhci_problem* p, const char* path, const int32_t* q, int64_t n, const double* E,
                          int nE, const double* C, int ncol, int threads, int binary);
int vhci_checkpoint_read(const vhci_problem* p, const char* path, int threads, int64_t* n, int32_t* q,
                         int64_t qcap, double* E, double* C, int64_t ccap);

/* ---- device context ---------------------------------------------------- */

int vhci_ctx_create(const vhci_problem* p, int form, int device, vhci_ctx** out);
void vhci_ctx_destroy(vhci_ctx* ctx);
const char* vhci_last_error(const vhci_ctx* ctx);

/* BasisSet (VCI_headers.h:149): replace / read the variational basis. */
int vhci_set_basis(vhci_ctx* ctx, const int32_t* q, int64_t n);
int64_t vhci_basis_size(const vhci_ctx* ctx);
int vhci_get_basis(vhci_ctx* ctx, int32_t* q);

/* MakeHamSparse (src/Ham.cpp:328-344): build H on the device for the
 * current basis.  stats (nullable, 4 entries): n, stored off-diagonal
 * nonzeros, structural nonzeros of the reference's triplet list (-1 if not
 * computed), build time in microseconds. */
int vhci_build_h(vhci_ctx* ctx, int64_t* stats);
/* 1 when the built H keeps no stored entries and every product regenerates
 * them from the basis hash (chosen when the stored matrix would not fit
 * beside the eigensolver's blocks, or forced by VHCI_MATRIX_FREE=1), 0 for
 * the stored CSR, negative on error.  No counterpart in the reference, which
 * always stores H (src/Ham.cpp:328-344). */
int vhci_matrix_free(const vhci_ctx* ctx);
/* Structural nonzeros of the reference's triplet list for the current
 * basis: n plus every ordered pair whose quanta differ by an even L1
 * distance <= the highest FC order or an odd one <= that order - 1
 * (src/Ham.cpp:247,283: explicit zeros included, the count behind the
 * "The Hamiltonian is X% sparse." line, src/Ham.cpp:341-342).  Returns -1
 * when the displacement ball exceeds 500000 vectors (then the driver prints
 * the line from the stored entries), or a negative error code. */
int64_t vhci_structural_nnz(vhci_ctx* ctx);
/* Copy H to the host: rowptr[n+1], col/val[nnz] (off-diagonal entries,
 * rows in basis order, columns ascending), diag[n].  Pass rowptr=NULL to
 * query nnz via *nnz. */
int vhci_get_h(vhci_ctx* ctx, int64_t* nnz, int64_t* rowptr, int32_t* col, double* val, double* diag);
/* The same for the nr basis rows rows[] only: rowptr[nr+1], col/val[nnz]
 * (each row's off-diagonal entries, columns in basis order ascending),
 * diag[nr].  Pass rowptr=NULL to query nnz via *nnz.  Test support for
 * bases whose whole matrix is too large to copy; no counterpart in the
 * reference. */
int vhci_get_h_rows(vhci_ctx* ctx, const int64_t* rows, int64_t nr, int64_t* nnz, int64_t* rowptr, int32_t* col,
                    double* val, double* diag);

/* SparseDiagonalize (src/Ham.cpp:347-370): the nev lowest eigenpairs,
 * converged to ||H x - theta x|| < tol * max(eps^(2/3), |theta|).  E[nev]
 * ascending; C (nullable) row-major n x nev.  The Ritz block stays on the
 * device and warm-starts the next call after the basis grows. */
int vhci_eigsolve(vhci_ctx* ctx, int nev, double tol, int maxit, double* E, double* C);
/* DenseDiagonalize (src/Ham.cpp:372-379) for the first HCI iteration
 * (src/HB.cpp:180-194): all n eigenpairs; C (nullable) n x n. */
int vhci_dense_diag(vhci_ctx* ctx, double* E, double* C);
/* Resident eigenvectors: column count, download, upload (checkpoint). */
int vhci_ritz_cols(const vhci_ctx* ctx);
int vhci_get_vectors(vhci_ctx* ctx, double* C, int ncol);
int vhci_set_vectors(vhci_ctx* ctx, const double* C, int ncol);
/* The assignment scan of PrintFreqs (src/Ham.cpp:448-502) on the device,
 * over the first ncol resident eigenvectors: for column c the basis rows m
 * with |C_mc| >= max_m |C_mc| / 2, ascending.  counts[ncol] receives each
 * column's row count and *total their sum; rows[] and absval[] receive the
 * rows and |C_mc|, column after column, when cap >= *total (else
 * VHCI_E_ARG, with counts and *total set). */
int vhci_freq_assignments(vhci_ctx* ctx, int ncol, int64_t* counts, int64_t* rows, double* absval, int64_t cap,
                          int64_t* total);

/* The AddStatesHB loop of Perform_HCI (src/HB.cpp:204-223): Cmax over the
 * first nopt columns (or the given cmax[n]), then every basis state's
 * heat-bath candidates with |Cmax_n f| >= eps.  New states are kept on the
 * device in first-discovery order; *nnew receives their count. */
int vhci_hb_select(vhci_ctx* ctx, const double* cmax, int nopt, double eps, int64_t* nnew);
int vhci_get_new_states(vhci_ctx* ctx, int32_t* q);
/* BasisSet.push_back of the new states (src/HB.cpp:229-231). */
int vhci_append_new_states(vhci_ctx* ctx);

/* DoPT2 (src/PT2.cpp:16-154): external space by heat bath with eps2 from
 * the resident vectors, then dE[n] = sum_a (sum_i H_ai C_in)^2 / (E_n - E_a)
 * for n < nopt.  *ns2 (nullable) receives the external-space size.  The
 * external space is processed in chunks (ranges of its states' hash) so
 * that at most 2^28 external states and 2^30 (a, i) pairs are held at once
 * (VHCI_PT2_CHUNK_STATES / VHCI_PT2_CHUNK_PAIRS lower the budgets);
 * afterwards vhci_get_new_states returns the last chunk's states, which are
 * the whole external space when it fit one chunk.  The eigensolver's work
 * blocks are released first (the resident eigenvectors stay). */
int vhci_pt2(vhci_ctx* ctx, const double* E, int nopt, double eps2, double* dE, int64_t* ns2);

/* Y = H X with the resident H for b vectors (1 <= b <= 64): X and Y are
 * host row-major n x b in basis order.  Replaces the product operator the
 * reference hands to its eigensolver (SparseSymMatProd, src/Ham.cpp:349,
 * the whole H applied to a block).  fp32 != 0 selects the mixed form's FP32
 * product (FP32 values and vectors); otherwise the product is FP64 (FP64
 * values, or hi + lo in the mixed form).  Runs the solver's own kernels. */
int vhci_apply_h(vhci_ctx* ctx, int b, int fp32, const double* X, double* Y);
/* vhci_apply_h without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results (bit for bit) as vhci_apply_h;
 * the device staging of X and Y and their solver-order copies live in a
 * fixed scratch array of 112 MiB that each GPU the library runs on holds
 * until the process ends, so n * b may be at most 4,194,304 (VHCI_E_LIMIT
 * beyond).  Calls on one GPU share its array and are serialised by a lock
 * (two calls never overlap on it); the array keeps the last call's data. */
int vhci_apply_h_noalloc(vhci_ctx* ctx, int b, int fp32, const double* X, double* Y);
/* Time b-vector SpMMs with the resident H (reps products, HIP events on
 * the context stream).  out[0] = ms per product, out[1] = algorithmic
 * bytes per product, out[2] = kernel launches per product.  fp32 != 0
 * selects the mixed form's FP32 product.  Measurement support; no
 * counterpart in the reference. */
int vhci_time_spmm(vhci_ctx* ctx, int b, int reps, int fp32, double* out);
/* One plain device copy of `bytes` bytes on the context stream (the
 * calibration kernel for memory-counter readings); *ms its duration.
 * Measurement support; no counterpart in the reference. */
int vhci_copy_probe(vhci_ctx* ctx, int64_t bytes, double* ms);
/* Per-launch timing of the b-column SpMMs the eigensolver issues (HIP
 * events around each launch on the context stream).  b > 0 starts and
 * clears collection, b = 0 stops it.  The read gives out[0] = total ms,
 * out[1] = total algorithmic bytes, out[2] = launches since the start.
 * Measurement support; no counterpart in the reference. */
int vhci_spmm_profile(vhci_ctx* ctx, int b);
int vhci_spmm_profile_read(vhci_ctx* ctx, double* out);
/* While SpMM profiling is on, the eigensolver's Ritz rotations (two
 * n x m x keep GEMMs per restart, MFMA through rocBLAS) are timed too:
 * out[0] = total ms, out[1] = flops, out[2] = GEMM calls.  Measurement
 * support; no counterpart in the reference. */
int vhci_gemm_profile_read(vhci_ctx* ctx, double* out);
/* Counters of the last phases (measurement support; no counterpart in the
 * reference), out[16]: last H build n, off-diagonal entries, connection
 * patterns, build ms; last heat-bath selection (after vhci_pt2: the PT2
 * one, all chunks) candidates, unique states, ms; last PT2 external states,
 * (external, variational) pairs, sum ms; key bytes; PT2 states; rows the
 * last H build probed (all, or only the new states' rows of an incremental
 * build); entries it carried over from the previous build; chunks the last
 * PT2 summed its external space in; the last build's storage (0 full CSR,
 * 1 upper triangle, 2 matrix-free products). */
int vhci_phase_stats(vhci_ctx* ctx, double* out);
/* Device-memory pool of the context, out[5]: hipMalloc calls, hipFree
 * calls, blocks handed out again from the idle set, bytes held (in use +
 * idle), idle bytes.  Blocks return to the device only at heat-bath
 * iteration boundaries (vhci_append_new_states, vhci_set_basis), on a
 * failed allocation and at vhci_ctx_destroy.  Measurement support; no
 * counterpart in the reference. */
int vhci_pool_stats(vhci_ctx* ctx, int64_t* out);

/* ---- whole solve (Perform_HCI + DoPT2 of one input) --------------------
 * Runs the reference's control flow (src/VHCI.cpp:39-114) on a device
 * context and returns the final energies: E[0..nE) (after PT2 if enabled),
 * *nE, the final variational basis size *nvar and *ns2.  timings (nullable,
 * 8 entries, ms): total, H build, eigensolve, heat bath, PT2 select,
 * PT2 sum, iterations, SpMM vectors. */
int vhci_solve(const vhci_problem* p, int form, int device, double tol, double* E, int* nE, int64_t* nvar,
               int64_t* ns2, double* timings);

/* Same loop on an existing context whose basis is the problem's initial
 * basis (restart must be 0: resuming from a checkpoint is done by the
 * `vhci` program, which also holds the checkpoint's energies); the final
 * basis, H and Ritz block stay resident on the device.  E receives the
 * final energies (after PT2), Evar (nullable) the variational ones, *nE
 * their count.  info (nullable, 4 entries): final basis size, PT2 space
 * size, heat-bath iterations, basis checksum (sum over the int32 quanta of
 * the final basis, row-major, of q_t * 1099511628211^(t+1) mod 2^64).
 * timings (nullable, 8 entries) as in vhci_solve. */
int vhci_run_ctx(vhci_ctx* ctx, double tol, int restart, double* E, double* Evar, int* nE, int64_t* info,
                 double* timings);

/* Basis size after each heat-bath iteration of the last vhci_run_ctx on
 * this context (the initial size first); returns the count written. */
int vhci_iteration_sizes(const vhci_ctx* ctx, int64_t* sizes, int max);

/* ---- serving loop ----------------------------------------------------
 * No counterpart in the reference (one input per process, src/VHCI.cpp).
 * Requests (input-file text, precision form, tolerance) are queued FIFO and
 * admitted to one of `slots` worker threads when a slot is free and their
 * predicted device bytes fit under mem_limit next to the reservations of
 * the running requests (a request is always admitted when none runs).  A
 * running request owns a device context on its own stream for its whole
 * solve; its reservation is revised after every basis growth.  Each
 * request runs the same loop as vhci_run_ctx, so its result does not
 * depend on what else runs. */
typedef struct vhci_server vhci_server;
int vhci_server_create(int device, int slots, int64_t mem_limit, vhci_server** out);
/* Waits for the running requests; queued ones end as failed. */
void vhci_server_destroy(vhci_server* s);
const char* vhci_server_error(const vhci_server* s);
/* Parses the input (VHCI_E_INPUT with the reference's message when it is
 * bad) and queues it; *id receives the request id. */
int vhci_server_submit(vhci_server* s, const char* text, int form, double tol, int64_t* id);
/* *done = 1 when the request finished within timeout_s (< 0: no limit). */
int vhci_server_wait(vhci_server* s, int64_t id, double timeout_s, int* done);
/* Result of a finished request: E/Evar (nullable, room for Nstate values),
 * *nE; info (nullable, 6): state (0 queued, 1 running, 2 done, 3 failed),
 * final basis size, PT2 space size, heat-bath iterations, basis checksum
 * (as vhci_run_ctx), worker slot; times (nullable, 6): submit, start and
 * end time in seconds since the server started, predicted bytes at
 * submission, measured peak device bytes, final reservation.  A failed
 * request returns VHCI_E_* with its message in vhci_server_error(). */
int vhci_server_result(vhci_server* s, int64_t id, double* E, double* Evar, int* nE, int64_t* info,
                       double* times);
/* out[5]: running, queued, reserved bytes, most requests run at once,
 * finished. */
int vhci_server_stats(vhci_server* s, double* out);

/* The drop-in program: same flags, input, freqs.dat and checkpoint formats
 * and log lines as the reference's `vhci -n -i -o -c`. argv as in main. */
int vhci_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif
#endif
