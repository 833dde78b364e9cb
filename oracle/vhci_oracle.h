/*
 * vhci_oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch CPU restatement of the
 * reference VHCI algorithms (berkelbach-group/VHCI, src/ directory).  It exists to
 * check the GPU library; the product never links, loads or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Parity unpinned (the reference's own files cannot fix these, so the tests
 * compare them order-insensitively or within a tolerance): the row order of
 * the states heat bath adds (the reference iterates a hashed unordered_set;
 * here the first-discovery order of its loop), the bits of eigenvectors and
 * with them of Cmax (Spectra's), and the order-dependent bits of the PT2 sum
 * (the reference sums with omp atomic).
 *
 * Quanta are passed as int32 rows (N x M, row-major).  Eigenvector blocks are
 * row-major N x ncol.  All functions return 0 on success and a negative code
 * on error; orc_error() returns the message of the last error on this thread.
 */
#ifndef VHCI_ORACLE_H
#define VHCI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

const char* orc_error(void);
void orc_set_threads(int n);

/* Parse an input file (Input_Reader.cpp:170-492).  restart!=0 builds only the
 * ground state as the reference does when a checkpoint is present. */
orc_problem* orc_load(const char* path, int restart);
void orc_free(orc_problem* p);

/* scalar settings: 0 M, 1 NEig, 2 Ntot, 3 perturb, 4 N0 (initial basis size),
 * 5 fcmax, 6 number of force constants read */
int64_t orc_info_i(const orc_problem* p, int what);
/* 0 eps1, 1 eps2 */
double orc_info_d(const orc_problem* p, int what);
int orc_freqs(const orc_problem* p, double* out);

/* force-constant lists: which = 0 AnharmFC, 1 Cubic, 2 Quartic, 3 Quintic,
 * 4 Sextic, 5 AnharmHB (construction order), 6 AnharmHB sorted by |f| desc
 * (stable).  fcpow/sm/mp have room for 6 entries. */
int orc_fc_count(const orc_problem* p, int which);
int orc_fc_get(const orc_problem* p, int which, int idx, double* fc,
               int* npow, int* fcpow, int* nsm, int* sm, int* mp);

/* initial basis (N0 x M) */
int orc_initial_basis(const orc_problem* p, int32_t* q);

/* <m| (a+a^dag)^k |n> path sum in the reference's enumeration order
 * (Ham.cpp:38-95) for one mode. */
double orc_coeffsum(int k, int qn, int qm);
/* AnharmPot for one force constant of list `which` between two states */
double orc_anharm_pot(const orc_problem* p, int which, int idx,
                      const int32_t* qn, const int32_t* qm);

/* All-pairs sparse Hamiltonian (Ham.cpp:203-344) on the given basis.
 * Two-call pattern: call with rowptr==NULL to get nnz (structural, including
 * explicit zeros) in *nnz; then call with arrays of that size.
 * Output: full symmetric CSR, columns ascending. */
int orc_build_h(const orc_problem* p, const int32_t* q, int64_t n,
                int64_t* nnz, int64_t* rowptr, int32_t* col, double* val);
/* Selected rows of the all-pairs Hamiltonian over the whole basis (every
 * column, element from the lower index, diagonal included, structural zeros
 * kept, columns ascending).  Two-call pattern as orc_build_h. */
int orc_h_rows(const orc_problem* p, const int32_t* q, int64_t n, const int64_t* rows, int64_t nr, int64_t* nnz,
               int64_t* rowptr, int32_t* col, double* val);
/* Dense H (Ham.cpp:107-200) as used for the first HCI iteration */
int orc_build_h_dense(const orc_problem* p, const int32_t* q, int64_t n, double* H);

/* Dense symmetric eigensolver: w ascending, Z row-major n x n, column j is
 * the eigenvector of w[j]. */
int orc_eigh(int n, const double* A, double* w, double* Z);

/* Thick-restart Lanczos for the nev smallest eigenpairs of a CSR matrix;
 * Spectra-style convergence (tol relative to max(eps^(2/3), |theta|)).
 * C is row-major n x nev. */
int orc_lanczos(int64_t n, const int64_t* rowptr, const int32_t* col,
                const double* val, int nev, double tol, int maxit,
                double* E, double* C);

/* Heat-bath selection (HB.cpp:24-174) over all basis states with the given
 * Cmax.  New states come out in first-discovery order.  Two-call pattern:
 * qnew==NULL returns the count in *nnew. */
int orc_hb_select(const orc_problem* p, const int32_t* q, int64_t n,
                  const double* cmax, double eps, int64_t* nnew, int32_t* qnew);

/* Cmax rule of Perform_HCI (HB.cpp:204-214): C row-major n x ncol, nopt
 * leading columns considered. */
int orc_cmax(int64_t n, int ncol, int nopt, const double* C, double* cmax);

/* Epstein-Nesbet PT2 (PT2.cpp:16-154) for a converged variational space.
 * C row-major n x ncol (ncol >= nopt), E[ncol].  dE[nopt] receives the
 * corrections; *ns2 the external-space size.  Per-state numerators sum over
 * i ascending; the sum over external states runs in a fixed order. */
int orc_pt2(const orc_problem* p, const int32_t* q, int64_t n, const double* C,
            int ncol, const double* E, int nopt, double eps2, double* dE,
            int64_t* ns2);

/* Pair-evaluation rates of the reference algorithms on a bounded sample,
 * for the CPU baseline: the all-pairs H loop (Ham.cpp:237-324) over rows
 * [0, rows) of qV against every j >= i, and the PT2 gather (PT2.cpp:66-149)
 * of the na external states qA against all of qV.  out[4] = H pairs, H
 * seconds, PT2 pairs, PT2 seconds (steady clock, current thread count). */
int orc_sample_rates(const orc_problem* p, const int32_t* qV, int64_t nV, int64_t rows, const int32_t* qA,
                     int64_t na, int nopt, double* out);

/* Hash-based host form, benchmark baseline only: the device engine's
 * generate-and-lookup algorithm on the host cores (each element FC's quanta
 * changes looked up in an open-addressing index of the states).
 * orc_fast_h_rows: rows of H as orc_h_rows, without the stored zeros no FC
 * bridges (two-call pattern).  orc_fast_pt2: DoPT2's corrections over a
 * given external space (as orc_hb_select returns it); with the thread count
 * of orc_pt2 the results are equal bit for bit.  orc_fast_sample: a timed
 * sample on a run's states (16 runs of consecutive rows); out[13] = index
 * build s, H rows, their s, their entries, one 8-column product over them
 * (s), external index build s, PT2 rows, pair generation s, pairs,
 * accumulation s, checksum, external states reached, their E_a s. */
int orc_fast_h_rows(const orc_problem* p, const int32_t* q, int64_t n, const int64_t* rows, int64_t nr, int64_t* nnz,
                    int64_t* rowptr, int32_t* col, double* val);
int orc_fast_pt2(const orc_problem* p, const int32_t* q, int64_t n, const double* C, int ncol, const double* E,
                 int nopt, const int32_t* qA, int64_t nA, double* dE, int64_t* npairs);
int orc_fast_sample(const orc_problem* p, const int32_t* qV, int64_t nV, const int32_t* qA, int64_t nA,
                    int64_t rows_h, int64_t rows_p, int nopt, double* out);

/* Whole program (VHCI.cpp:5-185): writes freqs file, optional checkpoint and
 * a JSON summary (basis size per iteration, PT2 space, energies as %.17g,
 * structural nnz per build).  basis_out (optional) receives the final
 * variational basis, one state per line. */
int orc_run(const char* input, const char* freqs_out, const char* ckpt,
            int nthreads, const char* json_out, const char* basis_out, int quiet);

#ifdef __cplusplus
}
#endif
#endif
