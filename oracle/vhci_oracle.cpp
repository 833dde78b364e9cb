/*
 * vhci_oracle.cpp -- CPU oracle for the heat-bath VCI path.
 *
 * TEST INFRASTRUCTURE ONLY.  A from-scratch restatement, in plain C++ with
 * OpenMP, of the reference program's algorithms (berkelbach-group/VHCI).  It
 * is the checker for the GPU library and the CPU baseline ("port") in the
 * benchmark; the product never links or calls it.
 *
 * Each function names the reference file:line it restates.  Third-party
 * pieces the reference uses (Eigen's SelfAdjointEigenSolver, Spectra's
 * SymEigsSolver, Boost hash) are replaced by textbook methods written here:
 *   - dense symmetric eigensolver: Householder tridiagonalisation followed by
 *     implicit symmetric QR steps with a Wilkinson shift (Golub & Van Loan,
 *     Matrix Computations, Alg. 8.3.1/8.3.2);
 *   - sparse eigensolver: thick-restart (Krylov-Schur) Lanczos with full
 *     re-orthogonalisation and Spectra's convergence test;
 *   - hash sets: std::unordered_set over the quanta, with insertion order
 *     kept alongside, so new states have a defined (first-discovery) order.
 */
#include "vhci_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_err;
int g_threads = 0;

int fail(const std::string& m) {
    g_err = m;
    return -1;
}

/* FConst (VCI_headers.h:121-129) */
struct FConst {
    double fc = 0.0;
    std::vector<int> fcpow;  // mode index per power
    std::vector<int> sm;     // ShortModes: distinct modes, first-seen order
    std::vector<int> mp;     // ModePowers
};

struct Problem {
    double eps1 = 0.0, eps2 = 0.0;
    int neig = 0;
    bool perturb = false;
    int ntot = 0;
    int M = 0;
    int nfc_read = 0;
    bool restart = false;
    std::vector<double> freq;
    std::vector<int> cap;
    std::vector<FConst> anharm, cubic, quartic, quintic, sextic, hb, hb_sorted;
    std::vector<int> basis0;  // N0 x M
    int fcmax = 0;
};

/* Fact (Core_functions.cpp:73-83) */
double fact(int n) {
    double v = 1;
    while (n > 0) {
        v *= n;
        n -= 1;
    }
    return v;
}

/* ScaleFC (Input_Reader.cpp:494-515): exact operation order kept. */
void scale_fc(std::vector<FConst>& fcs) {
    for (auto& f : fcs) {
        double s = 1.;
        s = (double)f.fcpow.size();
        s = std::pow(2, s);
        s = 1 / std::sqrt(s);
        for (int k : f.mp) s /= fact(k);
        f.fc *= s;
    }
}

/* ReadCIInput (Input_Reader.cpp:170-492) */
int parse(std::istream& in, bool restart, Problem& P) {
    std::string dummy;
    std::getline(in, dummy);
    in >> dummy >> P.eps1;
    in >> dummy >> P.neig;
    in >> dummy >> P.perturb;
    in >> dummy >> P.eps2;
    if (!P.perturb) P.eps2 = 0;
    in >> dummy >> P.ntot;
    int nmodes = 0;
    in >> dummy >> nmodes;
    if (!in) return fail("Error: malformed input header");
    for (int i = 0; i < nmodes; i++) {
        int id = 0;
        in >> id;
        if (id != i) {
            std::ostringstream os;
            os << "Error: Expected mode " << i << " but read data for mode " << id;
            return fail(os.str());
        }
        double f;
        int c;
        in >> f >> c;
        P.freq.push_back(f);
        P.cap.push_back(c);
    }
    P.M = nmodes;
    int nfc = 0;
    in >> dummy >> nfc;
    P.nfc_read = nfc;
    for (int i = 0; i < nfc; i++) {
        FConst t;
        int p = 0;
        in >> p;
        for (int j = 0; j < p; j++) {
            int m = 0;
            in >> m;
            t.fcpow.push_back(m);
        }
        in >> t.fc;
        P.anharm.push_back(t);
    }
    // ShortModes / ModePowers (Input_Reader.cpp:248-271)
    for (auto& f : P.anharm) {
        for (size_t i = 0; i < f.fcpow.size(); i++) {
            bool found = false;
            for (size_t j = 0; j < f.sm.size(); j++)
                if (f.fcpow[i] == f.sm[j]) {
                    f.mp[j] += 1;
                    found = true;
                }
            if (!found) {
                f.sm.push_back(f.fcpow[i]);
                f.mp.push_back(1);
            }
        }
        for (int m : f.sm)
            if (m < 0 || m >= P.M) return fail("Error: force constant names a mode outside 0..M-1");
    }
    scale_fc(P.anharm);
    // split by order (Input_Reader.cpp:273-291)
    for (auto& f : P.anharm) {
        switch (f.fcpow.size()) {
            case 3: P.cubic.push_back(f); break;
            case 4: P.quartic.push_back(f); break;
            case 5: P.quintic.push_back(f); break;
            case 6: P.sextic.push_back(f); break;
            default: return fail("Error: Force constants are only supported up to 6th order!");
        }
    }
    P.restart = restart;
    const int M = P.M;
    // initial basis (Input_Reader.cpp:296-364)
    if (restart) {
        P.basis0.assign(M, 0);
    } else {
        std::vector<std::vector<int>> B;
        int nexc = (P.cap[0] + 1 > P.ntot) ? P.ntot + 1 : P.cap[0] + 1;
        for (int i = 0; i < nexc; i++) B.push_back({i});
        for (int i = 1; i < M; i++) {
            std::vector<std::vector<int>> NB;
            nexc = (P.cap[i] + 1 > P.ntot) ? P.ntot + 1 : P.cap[i] + 1;
            for (int j = 0; j < nexc; j++)
                for (auto& s : B) {
                    std::vector<int> t = s;
                    t.push_back(j);
                    int tot = 0;
                    for (int x : t) tot += x;
                    if (tot <= P.ntot) NB.push_back(t);
                }
            B.swap(NB);
        }
        for (auto& s : B) P.basis0.insert(P.basis0.end(), s.begin(), s.end());
    }
    // AnharmHB: copy + effective singles and doubles (Input_Reader.cpp:366-460)
    P.hb = P.anharm;
    for (int m = 0; m < M; m++) {
        FConst t;
        t.fcpow = {m};
        t.sm = {m};
        t.mp = {1};
        t.fc = 0.;
        for (auto& c : P.cubic) {
            if (c.sm.size() == 1 && c.sm[0] == m) t.fc += 3 * c.fc;
            if (c.sm.size() == 2)
                for (int i = 0; i < 2; i++)
                    if (c.sm[i] == m && c.mp[i] == 1) {
                        t.fc += 2 * c.fc;
                        break;
                    }
        }
        if (std::abs(t.fc) > 1e-8) P.hb.push_back(t);
    }
    for (int m = 0; m < M; m++) {
        FConst t;
        t.fcpow = {m, m};
        t.sm = {m};
        t.mp = {2};
        t.fc = 0.;
        for (auto& c : P.quartic) {
            if (c.sm.size() == 1 && c.sm[0] == m) t.fc += 4 * c.fc;
            if (c.sm.size() == 2)
                for (int i = 0; i < 2; i++)
                    if (c.sm[i] == m && c.mp[i] == 2) {
                        t.fc += 2 * c.fc;
                        break;
                    }
        }
        if (std::abs(t.fc) > 1e-8) P.hb.push_back(t);
    }
    for (int m = 0; m < M - 1; m++)
        for (int n = m + 1; n < M; n++) {
            FConst t;
            t.fcpow = {m, n};
            t.sm = {m, n};
            t.mp = {1, 1};
            t.fc = 0.;
            for (auto& c : P.quartic) {
                bool pow3 = std::find(c.mp.begin(), c.mp.end(), 3) != c.mp.end();
                auto fm = std::find(c.sm.begin(), c.sm.end(), m);
                auto fn = std::find(c.sm.begin(), c.sm.end(), n);
                bool hm = fm != c.sm.end(), hn = fn != c.sm.end();
                if (c.sm.size() == 2 && pow3 && hm && hn) t.fc += 3 * c.fc;
                if (c.sm.size() == 3 && hm && hn && c.mp[fm - c.sm.begin()] == 1 &&
                    c.mp[fn - c.sm.begin()] == 1)
                    t.fc += 2 * c.fc;
            }
            if (std::abs(t.fc) > 1e-8) P.hb.push_back(t);
        }
    // HeatBath_Sort_FC (HB.cpp:17-22); stable so the order is defined
    P.hb_sorted = P.hb;
    std::stable_sort(P.hb_sorted.begin(), P.hb_sorted.end(),
                     [](const FConst& a, const FConst& b) { return std::abs(a.fc) > std::abs(b.fc); });
    P.fcmax = 0;
    for (auto& f : P.anharm) P.fcmax = std::max<int>(P.fcmax, (int)f.fcpow.size());
    return 0;
}

/* Ladder path sum for one mode (Ham.cpp:16-36 and :46-95): start at qn, apply
 * k factors (a^dag + a), creation branch first, drop negative states, then sum
 * the coefficients of the paths that end at qm in list order. */
double coeffsum(int k, int qn, int qm) {
    int qs[128], qt[128];
    double cs[128], ct[128];
    int len = 1;
    qs[0] = qn;
    cs[0] = 1.0;
    for (int j = 0; j < k; j++) {
        int nl = 0;
        for (int e = 0; e < len; e++) {
            int q = qs[e];
            double c = cs[e];
            qt[nl] = q + 1;
            ct[nl] = c * std::sqrt(q + 1);
            nl++;
            if (q - 1 >= 0) {
                qt[nl] = q - 1;
                ct[nl] = c * std::sqrt(q);
                nl++;
            }
        }
        len = nl;
        std::memcpy(qs, qt, sizeof(int) * nl);
        std::memcpy(cs, ct, sizeof(double) * nl);
    }
    double s = 0.0;
    for (int e = 0; e < len; e++)
        if (qs[e] == qm) s += cs[e];
    return s;
}

/* AnharmPot (Ham.cpp:38-104): <m|V_fc|n>, start state n */
double anharm_pot(const int* qn, const int* qm, const FConst& f) {
    double v = f.fc;
    for (size_t i = 0; i < f.sm.size(); i++) v *= coeffsum(f.mp[i], qn[f.sm[i]], qm[f.sm[i]]);
    return v;
}

/* QDiffVec (Ham.cpp:381-399) */
inline void qdiffvec(const int* a, const int* b, int M, int& qtot, int& mchange, int* dv) {
    for (int i = 0; i < M; i++) {
        int d = a[i] - b[i];
        dv[i] = std::abs(d);
        qtot += std::abs(d);
        if (d != 0) mchange += 1;
    }
}

/* ScreenState (Ham.cpp:401-446) */
inline bool screen(int qdiff, int mchange, const int* dv, int M, const FConst& f) {
    if (qdiff > (int)f.fcpow.size() || mchange > (int)f.sm.size() || qdiff % 2 != (int)f.fcpow.size() % 2)
        return false;
    for (size_t i = 0; i < f.sm.size(); i++)
        if (dv[f.sm[i]] > f.mp[i] || dv[f.sm[i]] % 2 != f.mp[i] % 2) return false;
    for (int i = 0; i < M; i++) {
        bool in = false;
        for (int m : f.sm)
            if (m == i) {
                in = true;
                break;
            }
        if (!in && dv[i] != 0) return false;
    }
    return true;
}

/* Pair element: the sum of AnharmPot over screened FCs for one pair, in the
 * reference's list order (Ham.cpp:247-274 even: Quartic then Sextic;
 * :283-314 odd: Cubic then Quintic, with the fcmax-1 prefilter).  Returns
 * false when the pair fails the prefilter (no triplet is pushed). */
inline bool pair_element(const Problem& P, const int* qs, const int* qe, int* dv, double& v) {
    int qdiff = 0, mch = 0;
    qdiffvec(qs, qe, P.M, qdiff, mch, dv);
    v = 0;
    const int fm = P.fcmax;
    if (qdiff <= fm && mch <= fm && qdiff % 2 == 0) {
        for (auto& f : P.quartic)
            if (screen(qdiff, mch, dv, P.M, f)) v += anharm_pot(qs, qe, f);
        for (auto& f : P.sextic)
            if (screen(qdiff, mch, dv, P.M, f)) v += anharm_pot(qs, qe, f);
        return true;
    }
    if (qdiff <= fm - 1 && mch <= fm - 1 && qdiff % 2 == 1) {
        for (auto& f : P.cubic)
            if (screen(qdiff, mch, dv, P.M, f)) v += anharm_pot(qs, qe, f);
        for (auto& f : P.quintic)
            if (screen(qdiff, mch, dv, P.M, f)) v += anharm_pot(qs, qe, f);
        return true;
    }
    return false;
}

/* harmonic diagonal (Ham.cpp:203-226) */
inline double harmonic(const Problem& P, const int* q) {
    double e = 0.;
    for (int j = 0; j < P.M; j++) {
        double ej = 0.5;
        ej += q[j];
        ej *= P.freq[j];
        e += ej;
    }
    return e;
}

struct CSR {
    int64_t n = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<double> val;
};

/* MakeHamSparse (Ham.cpp:203-344): all pairs j >= i; diagonal = E_i + V_ii;
 * off-diagonal kept even when zero (structural).  Full symmetric CSR. */
void build_h(const Problem& P, const int* q, int64_t n, CSR& H) {
    const int M = P.M;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    std::vector<std::vector<std::pair<int32_t, double>>> up(n);
    std::vector<double> diag(n);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; i++) {
        std::vector<int> dv(M);
        for (int64_t j = i; j < n; j++) {
            double v;
            if (pair_element(P, q + i * M, q + j * M, dv.data(), v)) {
                if (i == j)
                    diag[i] = v;
                else
                    up[i].push_back({(int32_t)j, v});
            }
        }
    }
    // lower part from the transpose
    std::vector<int64_t> cnt(n, 1);
    for (int64_t i = 0; i < n; i++) {
        cnt[i] += up[i].size();
        for (auto& e : up[i]) cnt[e.first] += 1;
    }
    H.n = n;
    H.rowptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; i++) H.rowptr[i + 1] = H.rowptr[i] + cnt[i];
    H.col.resize(H.rowptr[n]);
    H.val.resize(H.rowptr[n]);
    std::vector<int64_t> pos(H.rowptr.begin(), H.rowptr.end() - 1);
    // rows are filled in ascending column order: lower entries (from rows
    // j < i, visited in order), then the diagonal, then the upper entries.
    for (int64_t j = 0; j < n; j++)
        for (auto& e : up[j]) {
            int64_t i = e.first;
            H.col[pos[i]] = (int32_t)j;
            H.val[pos[i]] = e.second;
            pos[i]++;
        }
    for (int64_t i = 0; i < n; i++) {
        double e = harmonic(P, q + i * M);
        H.col[pos[i]] = (int32_t)i;
        H.val[pos[i]] = e + diag[i];
        pos[i]++;
        for (auto& x : up[i]) {
            H.col[pos[i]] = x.first;
            H.val[pos[i]] = x.second;
            pos[i]++;
        }
    }
}

/* ZerothHam + AnharmHam (Ham.cpp:107-200) */
void build_h_dense(const Problem& P, const int* q, int64_t n, std::vector<double>& A) {
    const int M = P.M;
    A.assign(n * n, 0.0);
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int64_t i = 0; i < n; i++) {
        A[i * n + i] += harmonic(P, q + i * M);
        std::vector<int> dv(M);
        for (int64_t j = i; j < n; j++) {
            double v;
            if (pair_element(P, q + i * M, q + j * M, dv.data(), v)) {
                A[i * n + j] += v;
                if (i != j) A[j * n + i] += v;
            }
        }
    }
}

/* ---------------- dense symmetric eigensolver ---------------- */
/* Householder tridiagonalisation + implicit QR (Wilkinson shift).  A is
 * row-major n x n symmetric (copied).  w ascending, Z row-major with column j
 * the eigenvector of w[j]. */
void eigh(int n, const double* Ain, std::vector<double>& w, std::vector<double>& Z) {
    std::vector<double> A(Ain, Ain + (size_t)n * n);
    std::vector<double> d(n), e(n > 1 ? n - 1 : 1, 0.0);
    std::vector<std::vector<double>> hv(n > 2 ? n - 2 : 0);
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    for (int k = 0; k + 2 < n; k++) {
        int L = n - k - 1;
        const double* x = &A[(size_t)k * n + k + 1];
        double sig = 0;
        for (int i = 0; i < L; i++) sig += x[i] * x[i];
        sig = std::sqrt(sig);
        d[k] = A[(size_t)k * n + k];
        if (sig == 0.0) {
            e[k] = 0.0;
            continue;
        }
        double alpha = x[0] > 0 ? -sig : sig;
        std::vector<double> v(x, x + L);
        v[0] -= alpha;
        double vn = 0;
        for (double t : v) vn += t * t;
        vn = std::sqrt(vn);
        if (vn == 0.0) {
            e[k] = x[0];
            continue;
        }
        for (double& t : v) t /= vn;
        e[k] = alpha;
        // trailing block B = A[k+1:, k+1:]; p = B v; K = v.p; w = p - K v;
        // B -= 2 (v w^T + w v^T)
        std::vector<double> p(L, 0.0);
        int kk = k + 1;
#pragma omp parallel for num_threads(nt) if (L > 256)
        for (int i = 0; i < L; i++) {
            const double* row = &A[(size_t)(kk + i) * n + kk];
            double s = 0;
            for (int j = 0; j < L; j++) s += row[j] * v[j];
            p[i] = s;
        }
        double K = 0;
        for (int i = 0; i < L; i++) K += v[i] * p[i];
        for (int i = 0; i < L; i++) p[i] -= K * v[i];
#pragma omp parallel for num_threads(nt) if (L > 256)
        for (int i = 0; i < L; i++) {
            double* row = &A[(size_t)(kk + i) * n + kk];
            double vi = v[i], pi = p[i];
            for (int j = 0; j < L; j++) row[j] -= 2.0 * (vi * p[j] + pi * v[j]);
        }
        hv[k] = std::move(v);
    }
    if (n >= 2) {
        d[n - 2] = A[(size_t)(n - 2) * n + n - 2];
        d[n - 1] = A[(size_t)(n - 1) * n + n - 1];
        e[n - 2] = A[(size_t)(n - 2) * n + n - 1];
    } else if (n == 1) {
        d[0] = A[0];
    }
    // Qt = Q^T, row r of Qt is column r of Q; Q = H_0 H_1 ... H_{n-3}
    std::vector<double> Qt((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Qt[(size_t)i * n + i] = 1.0;
    // Q = H_k Q backwards; in Qt terms: Qt = Qt H_k (H symmetric), acting on
    // columns k+1.. of Qt.
    for (int k = n - 3; k >= 0; k--) {
        if (hv[k].empty()) continue;
        const std::vector<double>& v = hv[k];
        int kk = k + 1, L = n - kk;
#pragma omp parallel for num_threads(nt) if (n > 256)
        for (int r = kk; r < n; r++) {
            double* row = &Qt[(size_t)r * n + kk];
            double s = 0;
            for (int j = 0; j < L; j++) s += row[j] * v[j];
            s *= 2.0;
            for (int j = 0; j < L; j++) row[j] -= s * v[j];
        }
    }
    // implicit symmetric QR on (d, e), rotations applied to rows of Qt
    const double eps = std::numeric_limits<double>::epsilon();
    int end = n - 1;
    int iter = 0;
    while (end > 0) {
        if (std::abs(e[end - 1]) <= eps * (std::abs(d[end - 1]) + std::abs(d[end]))) {
            e[end - 1] = 0.0;
            end--;
            iter = 0;
            continue;
        }
        int l = end - 1;
        while (l > 0 && std::abs(e[l - 1]) > eps * (std::abs(d[l - 1]) + std::abs(d[l]))) l--;
        if (++iter > 60 * n) break;
        double dd = (d[end - 1] - d[end]) / 2;
        double em = e[end - 1];
        double den = dd + (dd >= 0 ? 1.0 : -1.0) * std::hypot(dd, em);
        double mu = d[end] - em * em / den;
        double x = d[l] - mu, z = e[l];
        for (int k = l; k < end; k++) {
            double r = std::hypot(x, z);
            double c = 1.0, s = 0.0;
            if (r != 0.0) {
                c = x / r;
                s = -z / r;
            }
            if (k > l) e[k - 1] = r;
            double a = d[k], b = e[k], cc = d[k + 1];
            d[k] = c * c * a - 2 * c * s * b + s * s * cc;
            e[k] = c * s * a + (c * c - s * s) * b - c * s * cc;
            d[k + 1] = s * s * a + 2 * c * s * b + c * c * cc;
            if (k + 1 < end) {
                z = -s * e[k + 1];
                e[k + 1] = c * e[k + 1];
                x = e[k];
            }
            double* r0 = &Qt[(size_t)k * n];
            double* r1 = &Qt[(size_t)(k + 1) * n];
            for (int j = 0; j < n; j++) {
                double u = r0[j], t = r1[j];
                r0[j] = c * u - s * t;
                r1[j] = s * u + c * t;
            }
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    w.resize(n);
    Z.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; j++) {
        w[j] = d[idx[j]];
        const double* src = &Qt[(size_t)idx[j] * n];
        for (int i = 0; i < n; i++) Z[(size_t)i * n + j] = src[i];
    }
}

/* ---------------- sparse eigensolver ---------------- */
void spmv(const CSR& H, const double* x, double* y, int nt) {
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < H.n; i++) {
        double s = 0;
        for (int64_t k = H.rowptr[i]; k < H.rowptr[i + 1]; k++) s += H.val[k] * x[H.col[k]];
        y[i] = s;
    }
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Replaces SparseDiagonalize (Ham.cpp:347-370): nev smallest eigenpairs,
 * ncv = max(2 nev + 1, 20), convergence when ||r|| < tol*max(eps^(2/3),|theta|).
 * Columns are stored as contiguous vectors (V[j] has n entries). */
int lanczos(const CSR& H, int nev, double tol, int maxit, std::vector<double>& E, std::vector<double>& C) {
    const int64_t n = H.n;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    if (nev < 1 || nev > n) return fail("lanczos: bad nev");
    int m = std::max(2 * nev + 1, 20);
    if (m > n) m = (int)n;
    if (nev >= m) {
        // tiny: dense path
        std::vector<double> A((size_t)n * n, 0.0);
        for (int64_t i = 0; i < n; i++)
            for (int64_t k = H.rowptr[i]; k < H.rowptr[i + 1]; k++) A[i * n + H.col[k]] = H.val[k];
        std::vector<double> w, Z;
        eigh((int)n, A.data(), w, Z);
        E.assign(w.begin(), w.begin() + nev);
        C.assign((size_t)n * nev, 0.0);
        for (int64_t i = 0; i < n; i++)
            for (int j = 0; j < nev; j++) C[i * nev + j] = Z[i * n + j];
        return 0;
    }
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(n)), W(m, std::vector<double>(n));
    uint64_t seed = 0x5EED1234ull;
    for (int64_t i = 0; i < n; i++) V[0][i] = (double)(splitmix(seed) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    auto dot = [&](const double* a, const double* b) {
        double s = 0;
        for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
        return s;
    };
    auto normalize = [&](double* a) {
        double s = std::sqrt(dot(a, a));
        for (int64_t i = 0; i < n; i++) a[i] /= s;
        return s;
    };
    // h[i] = V_i . r for i < cnt, then r -= sum_i h[i] V_i (one parallel
    // region per step, so the cost does not hinge on region start-up)
    auto project_out = [&](int cnt, std::vector<double>& r) {
        std::vector<double> h(cnt);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < cnt; i++) h[i] = dot(V[i].data(), r.data());
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t t = 0; t < n; t++) {
            double s = r[t];
            for (int i = 0; i < cnt; i++) s -= h[i] * V[i][t];
            r[t] = s;
        }
    };
    normalize(V[0].data());
    std::vector<double> G((size_t)m * m, 0.0);
    int k = 0;  // kept vectors
    const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
    for (int it = 0; it < maxit; it++) {
        for (int j = k; j < m; j++) {
            spmv(H, V[j].data(), W[j].data(), nt);
            std::vector<double> r(W[j]);
            project_out(j + 1, r);
            project_out(j + 1, r);
            double wn = std::sqrt(dot(W[j].data(), W[j].data()));
            double beta = std::sqrt(dot(r.data(), r.data()));
            if (beta <= 1e-13 * wn) {
                // invariant subspace: continue with a random orthogonal vector
                for (int64_t t = 0; t < n; t++) r[t] = (double)(splitmix(seed) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
                project_out(j + 1, r);
                project_out(j + 1, r);
                beta = std::sqrt(dot(r.data(), r.data()));
            }
            for (int64_t t = 0; t < n; t++) V[j + 1][t] = r[t] / beta;
            // projected column j: G[:, j] = V^T W[:, j]
            std::vector<double> g(j + 1);
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int i = 0; i <= j; i++) g[i] = dot(V[i].data(), W[j].data());
            for (int i = 0; i <= j; i++) {
                G[(size_t)i * m + j] = g[i];
                G[(size_t)j * m + i] = g[i];
            }
            for (int i = j + 1; i < m; i++) {
                G[(size_t)i * m + j] = 0;
                G[(size_t)j * m + i] = 0;
            }
        }
        // Rayleigh-Ritz
        std::vector<double> th, Y;
        eigh(m, G.data(), th, Y);
        // residuals of the nev wanted pairs
        bool all = true;
        int keep = std::min(m - 1, nev + (m - nev) / 2);
        std::vector<std::vector<double>> VY(keep, std::vector<double>(n)), WY(keep, std::vector<double>(n));
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t t = 0; t < n; t++) {
            for (int c = 0; c < keep; c++) {
                double sv = 0, sw = 0;
                for (int i = 0; i < m; i++) {
                    sv += V[i][t] * Y[(size_t)i * m + c];
                    sw += W[i][t] * Y[(size_t)i * m + c];
                }
                VY[c][t] = sv;
                WY[c][t] = sw;
            }
        }
        std::vector<double> rr(nev, 0.0);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int c = 0; c < nev; c++) {
            double acc = 0;
            for (int64_t t = 0; t < n; t++) {
                double x = WY[c][t] - th[c] * VY[c][t];
                acc += x * x;
            }
            rr[c] = acc;
        }
        for (int c = 0; c < nev; c++)
            if (std::sqrt(rr[c]) >= tol * std::max(eps23, std::abs(th[c]))) all = false;
        if (all) {
            E.assign(th.begin(), th.begin() + nev);
            C.assign((size_t)n * nev, 0.0);
            for (int64_t t = 0; t < n; t++)
                for (int c = 0; c < nev; c++) C[t * nev + c] = VY[c][t];
            return 0;
        }
        // thick restart: keep the leading Ritz vectors, continue from v_m
        std::vector<double> last = V[m];
        for (int c = 0; c < keep; c++) {
            V[c].swap(VY[c]);
            W[c].swap(WY[c]);
        }
        V[keep] = last;
        std::fill(G.begin(), G.end(), 0.0);
        for (int c = 0; c < keep; c++) G[(size_t)c * m + c] = th[c];
        // the coupling column between kept vectors and v_keep is recomputed
        // exactly when W[keep] is formed (G[:, keep] = V^T W[keep]).
        k = keep;
    }
    return fail("Error: Eigenvalues did not converge.");
}

/* ---------------- heat-bath selection ---------------- */
std::string key_of(const int* q, int M) {
    std::string s((size_t)2 * M, '\0');
    for (int m = 0; m < M; m++) {
        s[2 * m] = (char)(q[m] & 0xff);
        s[2 * m + 1] = (char)((q[m] >> 8) & 0xff);
    }
    return s;
}

struct OrderedSet {
    std::unordered_set<std::string> set;
    std::vector<int> rows;  // insertion order, M ints per state
};

/* AddStatesHB (HB.cpp:24-174) for one basis state n */
int add_states_hb(const Problem& P, const std::unordered_set<std::string>& basis, OrderedSet& out,
                  const int* qn, double cn, double eps) {
    const int M = P.M;
    std::vector<int> tmp(qn, qn + M);
    for (const FConst& f : P.hb_sorted) {
        if (!(std::abs(cn * f.fc) >= eps)) break;
        int d = (int)f.sm.size();
        if (f.fcpow.size() == 0 || f.fcpow.size() > 6 || d < 1 || d > 6)
            return fail("Error: HCI only works with force constants up to 6th order.");
        int dl[6];
        for (int t = 0; t < d; t++) dl[t] = -f.mp[t];
        while (true) {
            int sa = 0;
            for (int t = 0; t < d; t++) sa += std::abs(dl[t]);
            if (sa != 0) {
                bool ok = true;
                for (int t = 0; t < d; t++) {
                    tmp[f.sm[t]] = qn[f.sm[t]] + dl[t];
                    if (tmp[f.sm[t]] < 0) ok = false;
                }
                if (ok) {
                    std::string key = key_of(tmp.data(), M);
                    if (!basis.count(key) && out.set.insert(key).second)
                        out.rows.insert(out.rows.end(), tmp.begin(), tmp.end());
                }
                for (int t = 0; t < d; t++) tmp[f.sm[t]] = qn[f.sm[t]];
            }
            // odometer, last mode fastest (HB.cpp nested loops)
            int t = d - 1;
            while (t >= 0) {
                dl[t] += 2;
                if (dl[t] <= f.mp[t]) break;
                dl[t] = -f.mp[t];
                t--;
            }
            if (t < 0) break;
        }
    }
    return 0;
}

/* Cmax (HB.cpp:204-214, PT2.cpp:29-37) */
void cmax_rule(int64_t n, int ncol, int nopt, const double* C, std::vector<double>& cm) {
    cm.resize(n);
    for (int64_t i = 0; i < n; i++) {
        double c = C[i * ncol + 0];
        for (int j = 1; j < nopt; j++)
            if (std::abs(C[i * ncol + j]) > std::abs(c)) c = C[i * ncol + j];
        cm[i] = c;
    }
}

int hb_select(const Problem& P, const int* q, int64_t n, const double* cmax, double eps, OrderedSet& out) {
    std::unordered_set<std::string> basis;
    basis.reserve(n * 2);
    for (int64_t i = 0; i < n; i++) basis.insert(key_of(q + i * P.M, P.M));
    for (int64_t i = 0; i < n; i++)
        if (add_states_hb(P, basis, out, q + i * P.M, cmax[i], eps)) return -1;
    return 0;
}

/* DoPT2 (PT2.cpp:16-154): numerator over all variational states (ascending
 * i), E_a with the diagonal FC terms added one at a time (PT2.cpp:119-144). */
int pt2(const Problem& P, const int* q, int64_t n, const double* C, int ncol, const double* E, int nopt,
        double eps2, std::vector<double>& dE, int64_t& ns2) {
    const int M = P.M;
    std::vector<double> cm;
    cmax_rule(n, ncol, nopt, C, cm);
    OrderedSet S2;
    if (hb_select(P, q, n, cm.data(), eps2, S2)) return -1;
    ns2 = (int64_t)S2.set.size();
    const int* qa_all = S2.rows.data();
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    std::vector<std::vector<double>> part(nt, std::vector<double>(nopt, 0.0));
    const int fm = P.fcmax;
#pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        int nth = omp_get_num_threads();
        int64_t lo = ns2 * tid / nth, hi = ns2 * (tid + 1) / nth;
        std::vector<double> sum(nopt);
        std::vector<int> dv(M), zero(M, 0);
        for (int64_t a = lo; a < hi; a++) {
            const int* qa = qa_all + a * M;
            std::fill(sum.begin(), sum.end(), 0.0);
            for (int64_t i = 0; i < n; i++) {
                const int* qi = q + i * M;
                double hai = 0;
                int qdiff = 0, mch = 0;
                qdiffvec(qa, qi, M, qdiff, mch, dv.data());
                if (qdiff <= fm && mch <= fm && qdiff % 2 == 0) {
                    for (auto& f : P.quartic)
                        if (screen(qdiff, mch, dv.data(), M, f)) hai += anharm_pot(qa, qi, f);
                    for (auto& f : P.sextic)
                        if (screen(qdiff, mch, dv.data(), M, f)) hai += anharm_pot(qa, qi, f);
                    for (int s = 0; s < nopt; s++) sum[s] += hai * C[i * ncol + s];
                }
                if (qdiff <= fm - 1 && mch <= fm - 1 && qdiff % 2 == 1) {
                    for (auto& f : P.cubic)
                        if (screen(qdiff, mch, dv.data(), M, f)) hai += anharm_pot(qa, qi, f);
                    for (auto& f : P.quintic)
                        if (screen(qdiff, mch, dv.data(), M, f)) hai += anharm_pot(qa, qi, f);
                    for (int s = 0; s < nopt; s++) sum[s] += hai * C[i * ncol + s];
                }
            }
            double ea = harmonic(P, qa);
            for (auto& f : P.quartic)
                if (screen(0, 0, zero.data(), M, f)) ea += anharm_pot(qa, qa, f);
            for (auto& f : P.sextic)
                if (screen(0, 0, zero.data(), M, f)) ea += anharm_pot(qa, qa, f);
            for (int s = 0; s < nopt; s++) part[tid][s] += std::pow(sum[s], 2) / (E[s] - ea);
        }
    }
    dE.assign(nopt, 0.0);
    for (int t = 0; t < nt; t++)
        for (int s = 0; s < nopt; s++) dE[s] += part[t][s];
    return 0;
}

/* PrintFreqs (Ham.cpp:448-502) */
void print_freqs(std::ostream& os, const Problem& P, const std::vector<int>& basis, const std::vector<double>& E,
                 const std::vector<double>& C, int64_t rows, int ncol) {
    const int M = P.M;
    int nf = P.neig;
    if ((int)E.size() < P.neig) nf = (int)E.size();
    auto assign = [&](int col, bool sep) {
        double mx = -std::numeric_limits<double>::infinity(), mn = std::numeric_limits<double>::infinity();
        for (int64_t m = 0; m < rows; m++) {
            mx = std::max(mx, C[m * ncol + col]);
            mn = std::min(mn, C[m * ncol + col]);
        }
        double maxc = std::max(std::abs(mx), std::abs(mn));
        int state = 0;
        for (int64_t m = 0; m < rows; m++) {
            if (std::abs(C[m * ncol + col]) >= maxc / 2) {
                if (sep && state > 0) os << "  ";
                os << std::fixed << std::setprecision(2) << std::abs(C[m * ncol + col]) << " (";
                state++;
                int nm = 0;
                for (int l = 0; l < M; l++)
                    if (basis[m * M + l] > 0) {
                        if (nm > 0) os << "+";
                        os << basis[m * M + l] << "w" << l;
                        nm++;
                    }
                os << ")";
            }
        }
    };
    os << std::fixed << std::setprecision(10) << E[0] << "   ";
    assign(0, false);
    os << '\n';
    for (int s = 1; s < nf; s++) {
        os << std::fixed << std::setprecision(12) << E[s] - E[0] << "   ";
        assign(s, true);
        os << '\n';
    }
    os.flush();
}

/* WriteCheckpoint (Core_functions.cpp:94-117) */
void write_ckpt(std::ostream& os, const Problem& P, const std::vector<int>& basis, const std::vector<double>& E,
                const std::vector<double>& C, int ncol) {
    const int M = P.M;
    int64_t n = (int64_t)basis.size() / M;
    os << n << " " << M << " " << P.neig << '\n';
    for (int j = 0; j < M; j++) os << std::fixed << std::setprecision(17) << P.freq[j] << " ";
    for (double e : E) os << std::fixed << std::setprecision(17) << e << " ";
    os << '\n';
    for (int64_t i = 0; i < n; i++) {
        for (int j = 0; j < M; j++) os << basis[i * M + j] << " ";
        for (int j = 0; j < ncol; j++) os << std::fixed << std::setprecision(17) << C[i * ncol + j] << " ";
        os << '\n';
    }
    os.flush();
}

/* ReadCheckpoint (Core_functions.cpp:119-167) */
int read_ckpt(std::istream& is, const Problem& P, std::vector<int>& basis, std::vector<double>& E,
              std::vector<double>& C) {
    int64_t nb;
    int par;
    is >> nb >> par;
    if (par != P.M) return fail("Error: Checkpoint file contains wrong number of fundamental modes!");
    is >> par;
    if (par != P.neig) return fail("Error: Checkpoint file contains wrong number of eigenstates!");
    for (int i = 0; i < P.M; i++) {
        double f;
        is >> f;
        if (f != P.freq[i]) return fail("Error: Checkpoint file contains wrong fundamental frequencies!");
    }
    E.assign(P.neig, 0.0);
    for (int i = 0; i < P.neig; i++) is >> E[i];
    basis.assign(nb * P.M, 0);
    C.assign(nb * P.neig, 0.0);
    for (int64_t i = 0; i < nb; i++) {
        for (int j = 0; j < P.M; j++) is >> basis[i * P.M + j];
        for (int j = 0; j < P.neig; j++) is >> C[i * P.neig + j];
    }
    if (!is) return fail("Error: truncated checkpoint file");
    return 0;
}

/* eigen-solve for the sparse stages; small bases go through the dense solver */
int sparse_diag(const Problem& P, const std::vector<int>& basis, std::vector<double>& E, std::vector<double>& C,
                int& ncol, int64_t& nnz_struct) {
    const int M = P.M;
    int64_t n = (int64_t)basis.size() / M;
    CSR H;
    build_h(P, basis.data(), n, H);
    nnz_struct = H.rowptr[n];
    int nev = P.neig;
    if (nev > n) nev = (int)n;
    if (n <= 2500) {
        std::vector<double> A((size_t)n * n, 0.0), w, Z;
        for (int64_t i = 0; i < n; i++)
            for (int64_t k = H.rowptr[i]; k < H.rowptr[i + 1]; k++) A[i * n + H.col[k]] = H.val[k];
        eigh((int)n, A.data(), w, Z);
        E.assign(w.begin(), w.begin() + nev);
        C.assign((size_t)n * nev, 0.0);
        for (int64_t i = 0; i < n; i++)
            for (int j = 0; j < nev; j++) C[i * nev + j] = Z[i * n + j];
    } else {
        if (lanczos(H, nev, 1e-10, 1000, E, C)) return -1;
    }
    ncol = nev;
    return 0;
}


/* ---- Hash-based host form ("fast") -------------------------------------
 * Benchmark baseline only: the generate-and-lookup algorithm of the device
 * engine, run on the host cores, so that the CPU figure is not set by the
 * reference's all-pairs loops alone.  A row's connections are generated from
 * each element FC's quanta changes (every change that the reference's
 * ScreenState, Ham.cpp:401-446, accepts for that FC, with the odd-pair
 * prefilter of Ham.cpp:283) and looked up in an open-addressing index of the
 * quanta; the FC terms of one pair are added in the reference's list order
 * (Quartic, Sextic / Cubic, Quintic; Ham.cpp:247-314), so the values equal
 * the all-pairs form's bit for bit.  Pairs that no FC bridges (the
 * reference's stored zeros) are not generated. */

struct QIndex {
    int M = 0;
    int64_t n = 0;
    std::vector<uint8_t> q;       // n x M quanta
    std::vector<uint64_t> coef;   // per-mode multipliers of the linear key sum
    std::vector<int32_t> slot;    // open addressing, linear probing, -1 empty
    uint64_t mask = 0;
    std::vector<uint64_t> bloom;  // two bits per state
    uint64_t bmask = 0;
};

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t qlin(const QIndex& X, const uint8_t* q) {
    uint64_t h = 0;
    for (int m = 0; m < X.M; m++) h += X.coef[m] * q[m];
    return h;
}

int qindex_build(QIndex& X, const int32_t* q, int64_t n, int M) {
    X.M = M;
    X.n = n;
    X.q.resize((size_t)n * M);
    for (int64_t i = 0; i < n * M; i++) {
        if (q[i] < 0 || q[i] > 255) return fail("fast form: quanta must lie in 0..255");
        X.q[i] = (uint8_t)q[i];
    }
    uint64_t s = 0x51ED270B27A1C3D5ull;
    X.coef.resize(M);
    for (auto& c : X.coef) c = splitmix(s) | 1ull;
    size_t cap = 16;
    while (cap < (size_t)(2 * n)) cap <<= 1;
    X.slot.assign(cap, -1);
    X.mask = cap - 1;
    size_t bits = 1024;
    while (bits < (size_t)(16 * n)) bits <<= 1;
    X.bloom.assign(bits / 64, 0);
    X.bmask = bits - 1;
    for (int64_t i = 0; i < n; i++) {
        const uint64_t h = mix64(qlin(X, &X.q[(size_t)i * M]));
        const uint64_t b1 = h & X.bmask, b2 = (h >> 32) & X.bmask;
        X.bloom[b1 >> 6] |= 1ull << (b1 & 63);
        X.bloom[b2 >> 6] |= 1ull << (b2 & 63);
        for (uint64_t p = h & X.mask;; p = (p + 1) & X.mask)
            if (X.slot[p] < 0) {
                X.slot[p] = (int32_t)i;
                break;
            }
    }
    return 0;
}

// index of the state with quanta t (hash h), or -1
inline int64_t qindex_find(const QIndex& X, uint64_t h, const uint8_t* t) {
    const uint64_t b1 = h & X.bmask, b2 = (h >> 32) & X.bmask;
    if (!((X.bloom[b1 >> 6] >> (b1 & 63)) & 1ull) || !((X.bloom[b2 >> 6] >> (b2 & 63)) & 1ull)) return -1;
    for (uint64_t p = h & X.mask;; p = (p + 1) & X.mask) {
        const int32_t s = X.slot[p];
        if (s < 0) return -1;
        if (std::memcmp(&X.q[(size_t)s * X.M], t, X.M) == 0) return s;
    }
}

/* AnharmPot (Ham.cpp:38-104) on byte quanta, start state qn */
inline double anharm_pot8(const uint8_t* qn, const uint8_t* qm, const FConst& f) {
    double v = f.fc;
    for (size_t i = 0; i < f.sm.size(); i++) v *= coeffsum(f.mp[i], qn[f.sm[i]], qm[f.sm[i]]);
    return v;
}

inline double harmonic8(const Problem& P, const uint8_t* q) {
    double e = 0.;
    for (int j = 0; j < P.M; j++) {
        double ej = 0.5;
        ej += q[j];
        ej *= P.freq[j];
        e += ej;
    }
    return e;
}

/* every state of X that one element FC connects to the start quanta qs
 * (lin = its linear key sum under X's multipliers): hit(index, fc) in
 * element order; tmp holds a copy of qs and is restored */
template <class F>
void walk_connections(const Problem& P, const QIndex& X, const uint8_t* qs, uint64_t lin, uint8_t* tmp, F&& hit) {
    const int fm = P.fcmax;
    for (const auto* lst : {&P.quartic, &P.sextic, &P.cubic, &P.quintic})
        for (const FConst& f : *lst) {
            const int d = (int)f.sm.size();
            int dl[6];
            for (int t = 0; t < d; t++) dl[t] = -f.mp[t];
            while (true) {
                int qd = 0;
                bool ok = true;
                uint64_t h = lin;
                for (int t = 0; t < d; t++) {
                    qd += std::abs(dl[t]);
                    const int v = qs[f.sm[t]] + dl[t];
                    ok = ok && v >= 0 && v <= 255;
                    h += (uint64_t)(int64_t)dl[t] * X.coef[f.sm[t]];
                }
                if (qd != 0 && ok && ((qd & 1) == 0 || qd <= fm - 1)) {
                    const uint64_t hh = mix64(h);
                    const uint64_t b1 = hh & X.bmask, b2 = (hh >> 32) & X.bmask;
                    if (((X.bloom[b1 >> 6] >> (b1 & 63)) & 1ull) && ((X.bloom[b2 >> 6] >> (b2 & 63)) & 1ull)) {
                        for (int t = 0; t < d; t++) tmp[f.sm[t]] = (uint8_t)(qs[f.sm[t]] + dl[t]);
                        const int64_t j = qindex_find(X, hh, tmp);
                        for (int t = 0; t < d; t++) tmp[f.sm[t]] = qs[f.sm[t]];
                        if (j >= 0) hit(j, f);
                    }
                }
                int t = d - 1;
                while (t >= 0) {
                    dl[t] += 2;
                    if (dl[t] <= f.mp[t]) break;
                    dl[t] = -f.mp[t];
                    t--;
                }
                if (t < 0) break;
            }
        }
}

/* the FCs that couple a state to itself (every mode power even: the ones
 * ScreenState accepts at zero quanta change), Quartic then Sextic in list
 * order */
std::vector<const FConst*> diag_fcs(const Problem& P) {
    std::vector<const FConst*> d;
    for (const auto* lst : {&P.quartic, &P.sextic})
        for (const FConst& f : *lst) {
            bool even = true;
            for (int k : f.mp) even = even && (k % 2 == 0);
            if (even) d.push_back(&f);
        }
    return d;
}

/* row r of H over the states of X: (column, value) ascending, the diagonal
 * (E_r + V_rr, Ham.cpp:203-226) included; dg = diag_fcs(P) */
void fast_row(const Problem& P, const std::vector<const FConst*>& dg, const QIndex& X, int64_t r,
              std::vector<std::pair<int32_t, double>>& out, std::vector<std::pair<int32_t, double>>& raw,
              std::vector<uint8_t>& tmp) {
    const int M = X.M;
    const uint8_t* qr = &X.q[(size_t)r * M];
    tmp.assign(qr, qr + M);
    raw.clear();
    walk_connections(P, X, qr, qlin(X, qr), tmp.data(), [&](int64_t j, const FConst& f) {
        const uint8_t* qj = &X.q[(size_t)j * M];
        // the element is evaluated from the lower-indexed state, as in the
        // reference's upper-triangle loop
        raw.push_back({(int32_t)j, j > r ? anharm_pot8(qr, qj, f) : anharm_pot8(qj, qr, f)});
    });
    double vd = 0;
    for (const FConst* f : dg) vd += anharm_pot8(qr, qr, *f);
    raw.push_back({(int32_t)r, harmonic8(P, qr) + vd});
    // one FC reaches a column at most once, so a stable sort keeps each
    // pair's terms in element order
    std::stable_sort(raw.begin(), raw.end(),
                     [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                         return a.first < b.first;
                     });
    out.clear();
    for (size_t k = 0; k < raw.size();) {
        const int32_t c = raw[k].first;
        if (c == (int32_t)r) {
            out.push_back(raw[k]);
            k++;
            continue;
        }
        double v = 0;
        for (; k < raw.size() && raw[k].first == c; k++) v += raw[k].second;
        out.push_back({c, v});
    }
}

/* PT2 pairs (a, i, term) of variational row i against the external index A,
 * terms AnharmPot(a -> i) in element order (PT2.cpp:80-118) */
struct Pt2Term {
    uint64_t key;  // a << 32 | i
    double v;
};

void fast_pt2_row(const Problem& P, const QIndex& V, const QIndex& A, int64_t i, std::vector<Pt2Term>& out,
                  std::vector<uint8_t>& tmp) {
    const int M = V.M;
    const uint8_t* qi = &V.q[(size_t)i * M];
    tmp.assign(qi, qi + M);
    walk_connections(P, A, qi, qlin(A, qi), tmp.data(), [&](int64_t a, const FConst& f) {
        out.push_back({((uint64_t)a << 32) | (uint64_t)i, anharm_pot8(&A.q[(size_t)a * M], qi, f)});
    });
}

/* E_a of an external state: harmonic, then the diagonal FC terms one at a
 * time (PT2.cpp:119-144); dg = diag_fcs(P) */
inline double pt2_ea(const Problem& P, const std::vector<const FConst*>& dg, const uint8_t* qa) {
    double ea = harmonic8(P, qa);
    for (const FConst* f : dg) ea += anharm_pot8(qa, qa, *f);
    return ea;
}

}  // namespace

/* =================== C API =================== */
struct orc_problem {
    Problem P;
};

extern "C" {

const char* orc_error(void) { return g_err.c_str(); }
void orc_set_threads(int n) { g_threads = n; }

orc_problem* orc_load(const char* path, int restart) {
    std::ifstream in(path);
    if (!in.good()) {
        fail("Error: Could not open input file.");
        return nullptr;
    }
    orc_problem* p = new orc_problem();
    if (parse(in, restart != 0, p->P)) {
        delete p;
        return nullptr;
    }
    return p;
}

void orc_free(orc_problem* p) { delete p; }

int64_t orc_info_i(const orc_problem* p, int what) {
    const Problem& P = p->P;
    switch (what) {
        case 0: return P.M;
        case 1: return P.neig;
        case 2: return P.ntot;
        case 3: return P.perturb ? 1 : 0;
        case 4: return (int64_t)P.basis0.size() / std::max(1, P.M);
        case 5: return P.fcmax;
        case 6: return P.nfc_read;
    }
    return -1;
}

double orc_info_d(const orc_problem* p, int what) { return what == 0 ? p->P.eps1 : p->P.eps2; }

int orc_freqs(const orc_problem* p, double* out) {
    for (int i = 0; i < p->P.M; i++) out[i] = p->P.freq[i];
    return 0;
}

static const std::vector<FConst>* fc_list(const orc_problem* p, int which) {
    const Problem& P = p->P;
    switch (which) {
        case 0: return &P.anharm;
        case 1: return &P.cubic;
        case 2: return &P.quartic;
        case 3: return &P.quintic;
        case 4: return &P.sextic;
        case 5: return &P.hb;
        case 6: return &P.hb_sorted;
    }
    return nullptr;
}

int orc_fc_count(const orc_problem* p, int which) {
    auto* l = fc_list(p, which);
    return l ? (int)l->size() : fail("bad list");
}

int orc_fc_get(const orc_problem* p, int which, int idx, double* fc, int* npow, int* fcpow, int* nsm, int* sm,
               int* mp) {
    auto* l = fc_list(p, which);
    if (!l || idx < 0 || idx >= (int)l->size()) return fail("bad index");
    const FConst& f = (*l)[idx];
    *fc = f.fc;
    *npow = (int)f.fcpow.size();
    for (int i = 0; i < *npow && i < 6; i++) fcpow[i] = f.fcpow[i];
    *nsm = (int)f.sm.size();
    for (int i = 0; i < *nsm && i < 6; i++) {
        sm[i] = f.sm[i];
        mp[i] = f.mp[i];
    }
    return 0;
}

int orc_initial_basis(const orc_problem* p, int32_t* q) {
    for (size_t i = 0; i < p->P.basis0.size(); i++) q[i] = p->P.basis0[i];
    return 0;
}

double orc_coeffsum(int k, int qn, int qm) { return coeffsum(k, qn, qm); }

double orc_anharm_pot(const orc_problem* p, int which, int idx, const int32_t* qn, const int32_t* qm) {
    auto* l = fc_list(p, which);
    return anharm_pot(qn, qm, (*l)[idx]);
}

int orc_build_h(const orc_problem* p, const int32_t* q, int64_t n, int64_t* nnz, int64_t* rowptr, int32_t* col,
                double* val) {
    CSR H;
    build_h(p->P, q, n, H);
    *nnz = H.rowptr[n];
    if (rowptr) {
        std::memcpy(rowptr, H.rowptr.data(), sizeof(int64_t) * (n + 1));
        std::memcpy(col, H.col.data(), sizeof(int32_t) * H.col.size());
        std::memcpy(val, H.val.data(), sizeof(double) * H.val.size());
    }
    return 0;
}

/* Selected rows of the all-pairs H (Ham.cpp:237-326): for row r every
 * column j of the basis, the element computed from the lower index as the
 * reference's i <= j loop does, diagonal = harmonic + V_rr; structural
 * entries (explicit zeros) included, columns ascending. */
int orc_h_rows(const orc_problem* p, const int32_t* q, int64_t n, const int64_t* rows, int64_t nr, int64_t* nnz,
               int64_t* rowptr, int32_t* col, double* val) {
    const Problem& P = p->P;
    const int M = P.M;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    std::vector<std::vector<std::pair<int32_t, double>>> out(nr);
    for (int64_t k = 0; k < nr; k++) {
        const int64_t r = rows[k];
        if (r < 0 || r >= n) return fail("orc_h_rows: row out of range");
        const int64_t nb = (n + 4095) / 4096;
        std::vector<std::vector<std::pair<int32_t, double>>> part(nb);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
        for (int64_t b = 0; b < nb; b++) {
            std::vector<int> dv(M);
            for (int64_t j = b * 4096; j < std::min(n, (b + 1) * 4096); j++) {
                const int64_t lo = std::min(r, j), hi = std::max(r, j);
                double v;
                if (pair_element(P, q + lo * M, q + hi * M, dv.data(), v)) {
                    if (j == r) v += harmonic(P, q + r * M);
                    part[b].push_back({(int32_t)j, v});
                }
            }
        }
        for (auto& x : part) out[k].insert(out[k].end(), x.begin(), x.end());
    }
    int64_t tot = 0;
    for (auto& x : out) tot += (int64_t)x.size();
    *nnz = tot;
    if (rowptr) {
        rowptr[0] = 0;
        for (int64_t k = 0; k < nr; k++) {
            rowptr[k + 1] = rowptr[k] + (int64_t)out[k].size();
            for (size_t t = 0; t < out[k].size(); t++) {
                col[rowptr[k] + t] = out[k][t].first;
                val[rowptr[k] + t] = out[k][t].second;
            }
        }
    }
    return 0;
}

int orc_build_h_dense(const orc_problem* p, const int32_t* q, int64_t n, double* H) {
    std::vector<double> A;
    build_h_dense(p->P, q, n, A);
    std::memcpy(H, A.data(), sizeof(double) * A.size());
    return 0;
}

int orc_eigh(int n, const double* A, double* w, double* Z) {
    std::vector<double> ww, zz;
    eigh(n, A, ww, zz);
    std::memcpy(w, ww.data(), sizeof(double) * n);
    std::memcpy(Z, zz.data(), sizeof(double) * zz.size());
    return 0;
}

int orc_lanczos(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val, int nev, double tol,
                int maxit, double* E, double* C) {
    CSR H;
    H.n = n;
    H.rowptr.assign(rowptr, rowptr + n + 1);
    H.col.assign(col, col + rowptr[n]);
    H.val.assign(val, val + rowptr[n]);
    std::vector<double> e, c;
    if (lanczos(H, nev, tol, maxit, e, c)) return -1;
    std::memcpy(E, e.data(), sizeof(double) * nev);
    std::memcpy(C, c.data(), sizeof(double) * c.size());
    return 0;
}

int orc_hb_select(const orc_problem* p, const int32_t* q, int64_t n, const double* cmax, double eps, int64_t* nnew,
                  int32_t* qnew) {
    OrderedSet out;
    if (hb_select(p->P, q, n, cmax, eps, out)) return -1;
    *nnew = (int64_t)out.set.size();
    if (qnew) std::memcpy(qnew, out.rows.data(), sizeof(int32_t) * out.rows.size());
    return 0;
}

int orc_cmax(int64_t n, int ncol, int nopt, const double* C, double* cmax) {
    std::vector<double> cm;
    cmax_rule(n, ncol, nopt, C, cm);
    std::memcpy(cmax, cm.data(), sizeof(double) * n);
    return 0;
}

int orc_pt2(const orc_problem* p, const int32_t* q, int64_t n, const double* C, int ncol, const double* E, int nopt,
            double eps2, double* dE, int64_t* ns2) {
    std::vector<double> d;
    if (pt2(p->P, q, n, C, ncol, E, nopt, eps2, d, *ns2)) return -1;
    std::memcpy(dE, d.data(), sizeof(double) * nopt);
    return 0;
}

int orc_sample_rates(const orc_problem* p, const int32_t* qV, int64_t nV, int64_t rows, const int32_t* qA,
                     int64_t na, int nopt, double* out) {
    using clk = std::chrono::steady_clock;
    const Problem& P = p->P;
    const int M = P.M;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    rows = std::min(rows, nV);
    // H pair loop (no triplet storage: the sums are reduced into a checksum
    // so the work is not optimised away)
    double chk = 0;
    auto t0 = clk::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : chk)
    for (int64_t i = 0; i < rows; i++) {
        std::vector<int> dv(M);
        for (int64_t j = i; j < nV; j++) {
            double v;
            if (pair_element(P, qV + i * M, qV + j * M, dv.data(), v)) chk += v;
        }
    }
    double th = std::chrono::duration<double>(clk::now() - t0).count();
    double hpairs = 0;
    for (int64_t i = 0; i < rows; i++) hpairs += (double)(nV - i);
    // PT2 gather with unit coefficients
    std::vector<double> C((size_t)nV * nopt, 1e-3);
    t0 = clk::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : chk)
    for (int64_t a = 0; a < na; a++) {
        std::vector<int> dv(M);
        std::vector<double> sum(nopt, 0.0);
        const int* qa = qA + a * M;
        for (int64_t i = 0; i < nV; i++) {
            double v;
            if (pair_element(P, qa, qV + i * M, dv.data(), v))
                for (int s2 = 0; s2 < nopt; s2++) sum[s2] += v * C[i * nopt + s2];
        }
        for (int s2 = 0; s2 < nopt; s2++) chk += sum[s2] * sum[s2];
    }
    double tp = std::chrono::duration<double>(clk::now() - t0).count();
    out[0] = hpairs;
    out[1] = th;
    out[2] = (double)na * (double)nV;
    out[3] = tp;
    if (chk == 12345.6789) out[0] += 1;  // keeps chk live
    return 0;
}

/* Hash-based host form (benchmark baseline).  Rows of H over the states q as
 * orc_h_rows gives them, without the stored zeros that no FC bridges.
 * Two-call pattern as orc_h_rows. */
int orc_fast_h_rows(const orc_problem* p, const int32_t* q, int64_t n, const int64_t* rows, int64_t nr, int64_t* nnz,
                    int64_t* rowptr, int32_t* col, double* val) {
    const Problem& P = p->P;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    QIndex X;
    if (qindex_build(X, q, n, P.M)) return -1;
    for (int64_t k = 0; k < nr; k++)
        if (rows[k] < 0 || rows[k] >= n) return fail("orc_fast_h_rows: row out of range");
    std::vector<std::vector<std::pair<int32_t, double>>> out(nr);
    const auto dg = diag_fcs(P);
#pragma omp parallel num_threads(nt)
    {
        std::vector<std::pair<int32_t, double>> raw;
        std::vector<uint8_t> tmp;
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < nr; k++) fast_row(P, dg, X, rows[k], out[k], raw, tmp);
    }
    int64_t tot = 0;
    for (auto& x : out) tot += (int64_t)x.size();
    *nnz = tot;
    if (rowptr) {
        rowptr[0] = 0;
        for (int64_t k = 0; k < nr; k++) {
            rowptr[k + 1] = rowptr[k] + (int64_t)out[k].size();
            for (size_t t = 0; t < out[k].size(); t++) {
                col[rowptr[k] + t] = out[k][t].first;
                val[rowptr[k] + t] = out[k][t].second;
            }
        }
    }
    return 0;
}

namespace {

/* pairs of the given variational rows against A, grouped by owner thread of
 * their external state (contiguous ranges of A, as orc_pt2 splits them), each
 * group sorted by (a, i) with every pair's terms in element order */
void fast_pt2_pairs(const Problem& P, const QIndex& V, const QIndex& A, const std::vector<int64_t>& rows, int nt,
                    std::vector<std::vector<Pt2Term>>& own) {
    std::vector<std::vector<Pt2Term>> gen(nt);
#pragma omp parallel num_threads(nt)
    {
        std::vector<uint8_t> tmp;
        std::vector<Pt2Term>& g = gen[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1)
        for (size_t k = 0; k < rows.size(); k++) fast_pt2_row(P, V, A, rows[k], g, tmp);
    }
    const int64_t na = A.n;
    own.assign(nt, {});
    // a thread's row blocks come in ascending order of its generation list;
    // the stable sort below keeps one pair's terms (all from one row, one
    // thread) in element order
    for (int t = 0; t < nt; t++)
        for (const Pt2Term& x : gen[t]) {
            const int64_t a = (int64_t)(x.key >> 32);
            int o = (int)((a * nt) / std::max<int64_t>(1, na));
            while (o + 1 < nt && a >= na * (o + 1) / nt) o++;
            while (o > 0 && a < na * o / nt) o--;
            own[o].push_back(x);
        }
#pragma omp parallel for schedule(static, 1) num_threads(nt)
    for (int t = 0; t < nt; t++)
        std::stable_sort(own[t].begin(), own[t].end(),
                         [](const Pt2Term& x, const Pt2Term& y) { return x.key < y.key; });
}

}  // namespace

/* PT2 corrections (DoPT2, PT2.cpp:16-154) over a given external space qA of
 * nA states (as orc_hb_select returns it) by the hash-based form.  The
 * numerators run over ascending i and the corrections over the same thread
 * ranges of A as orc_pt2, so with the same thread count the results are
 * equal bit for bit. */
int orc_fast_pt2(const orc_problem* p, const int32_t* q, int64_t n, const double* C, int ncol, const double* E,
                 int nopt, const int32_t* qA, int64_t nA, double* dE, int64_t* npairs) {
    const Problem& P = p->P;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    QIndex V, A;
    if (qindex_build(V, q, n, P.M) || qindex_build(A, qA, nA, P.M)) return -1;
    std::vector<int64_t> rows(n);
    for (int64_t i = 0; i < n; i++) rows[i] = i;
    std::vector<std::vector<Pt2Term>> own;
    fast_pt2_pairs(P, V, A, rows, nt, own);
    std::vector<std::vector<double>> part(nt, std::vector<double>(nopt, 0.0));
    std::vector<int64_t> pc(nt, 0);
    const auto dg = diag_fcs(P);
#pragma omp parallel for schedule(static, 1) num_threads(nt)
    for (int t = 0; t < nt; t++) {
        const auto& w = own[t];
        std::vector<double> sum(nopt);
        for (size_t k = 0; k < w.size();) {
            const uint64_t a = w[k].key >> 32;
            std::fill(sum.begin(), sum.end(), 0.0);
            while (k < w.size() && (w[k].key >> 32) == a) {
                const uint64_t key = w[k].key;
                const int64_t i = (int64_t)(key & 0xffffffffull);
                double hai = 0;
                for (; k < w.size() && w[k].key == key; k++) hai += w[k].v;
                for (int s = 0; s < nopt; s++) sum[s] += hai * C[i * ncol + s];
                pc[t]++;
            }
            const double ea = pt2_ea(P, dg, &A.q[(size_t)a * P.M]);
            for (int s = 0; s < nopt; s++) part[t][s] += std::pow(sum[s], 2) / (E[s] - ea);
        }
    }
    for (int s = 0; s < nopt; s++) dE[s] = 0.0;
    int64_t tot = 0;
    for (int t = 0; t < nt; t++) {
        for (int s = 0; s < nopt; s++) dE[s] += part[t][s];
        tot += pc[t];
    }
    if (npairs) *npairs = tot;
    return 0;
}

/* Timed sample of the hash-based form on a run's states (benchmark
 * baseline).  H rows and PT2 rows are taken as 16 runs of consecutive rows
 * spread over the basis.  out: [0] index build s (nV states), [1] H rows
 * timed, [2] their s, [3] their entries (diagonal included), [4] one
 * 8-column product over those rows, s (best of 3), [5] external index build
 * s (nA states), [6] PT2 rows timed, [7] their pair generation s, [8] pairs
 * found, [9] numerator accumulation over those pairs (nopt columns) and the
 * corrections of their external states, s, [10] a checksum, [11] external
 * states those pairs reach, [12] their E_a, s */
int orc_fast_sample(const orc_problem* p, const int32_t* qV, int64_t nV, const int32_t* qA, int64_t nA,
                    int64_t rows_h, int64_t rows_p, int nopt, double* out) {
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); };
    const Problem& P = p->P;
    int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
    auto runs = [&](int64_t want) {
        std::vector<int64_t> r;
        const int nrun = 16;
        const int64_t len = std::max<int64_t>(1, std::min(want / nrun, nV / nrun));
        for (int c = 0; c < nrun; c++) {
            const int64_t s0 = (nV - len) * c / (nrun - 1);
            for (int64_t i = s0; i < s0 + len && i < nV; i++)
                if (r.empty() || i > r.back()) r.push_back(i);
        }
        return r;
    };
    double chk = 0;
    auto t0 = clk::now();
    QIndex V;
    if (qindex_build(V, qV, nV, P.M)) return -1;
    out[0] = secs(t0);
    // H rows
    const std::vector<int64_t> hr = runs(rows_h);
    const int64_t nh = (int64_t)hr.size();
    std::vector<std::vector<std::pair<int32_t, double>>> rowv(nh);
    const auto dg = diag_fcs(P);
    t0 = clk::now();
#pragma omp parallel num_threads(nt)
    {
        std::vector<std::pair<int32_t, double>> raw;
        std::vector<uint8_t> tmp;
#pragma omp for schedule(dynamic, 1)
        for (int64_t k = 0; k < nh; k++) fast_row(P, dg, V, hr[k], rowv[k], raw, tmp);
    }
    out[1] = (double)nh;
    out[2] = secs(t0);
    std::vector<int64_t> rp(nh + 1, 0);
    for (int64_t k = 0; k < nh; k++) rp[k + 1] = rp[k] + (int64_t)rowv[k].size();
    out[3] = (double)rp[nh];
    std::vector<int32_t> cl(rp[nh]);
    std::vector<double> vl(rp[nh]);
    for (int64_t k = 0; k < nh; k++)
        for (size_t t = 0; t < rowv[k].size(); t++) {
            cl[rp[k] + t] = rowv[k][t].first;
            vl[rp[k] + t] = rowv[k][t].second;
        }
    rowv.clear();
    rowv.shrink_to_fit();
    // one 8-column product over the sampled rows, X over the whole basis
    std::vector<double> X((size_t)nV * 8), Y((size_t)nh * 8);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < nV * 8; i++) X[i] = 1e-3 * (double)((i * 2654435761ull) % 1000);
    double best = 1e300;
    for (int rep = 0; rep < 3; rep++) {
        t0 = clk::now();
#pragma omp parallel for schedule(static, 64) num_threads(nt)
        for (int64_t k = 0; k < nh; k++) {
            double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t e = rp[k]; e < rp[k + 1]; e++) {
                const double v = vl[e];
                const double* x = &X[(size_t)cl[e] * 8];
                for (int c = 0; c < 8; c++) acc[c] += v * x[c];
            }
            for (int c = 0; c < 8; c++) Y[k * 8 + c] = acc[c];
        }
        best = std::min(best, secs(t0));
    }
    out[4] = best;
    for (int64_t k = 0; k < nh * 8; k += 97) chk += Y[k];
    // PT2
    t0 = clk::now();
    QIndex A;
    if (qindex_build(A, qA, nA, P.M)) return -1;
    out[5] = secs(t0);
    const std::vector<int64_t> pr = runs(rows_p);
    std::vector<int32_t> slot_of(nV, -1);
    for (size_t k = 0; k < pr.size(); k++) slot_of[pr[k]] = (int32_t)k;
    std::vector<double> Cs(pr.size() * (size_t)nopt);
    for (size_t k = 0; k < Cs.size(); k++) Cs[k] = 1e-3 * (double)((k * 40503ull) % 997);
    std::vector<double> Es(nopt, -1e4);
    out[6] = (double)pr.size();
    t0 = clk::now();
    std::vector<std::vector<Pt2Term>> own;
    fast_pt2_pairs(P, V, A, pr, nt, own);
    out[7] = secs(t0);
    // E_a of every external state the sample reaches (a cost per external
    // state), then the numerators and corrections (a cost per pair)
    std::vector<std::vector<double>> eas(nt);
    int64_t seen = 0;
    t0 = clk::now();
#pragma omp parallel for schedule(static, 1) num_threads(nt) reduction(+ : seen)
    for (int t = 0; t < nt; t++) {
        const auto& w = own[t];
        for (size_t k = 0; k < w.size(); k++)
            if (k == 0 || (w[k].key >> 32) != (w[k - 1].key >> 32)) {
                eas[t].push_back(pt2_ea(P, dg, &A.q[(size_t)(w[k].key >> 32) * P.M]));
                seen++;
            }
    }
    out[12] = secs(t0);
    out[11] = (double)seen;
    int64_t pairs = 0;
    t0 = clk::now();
#pragma omp parallel for schedule(static, 1) num_threads(nt) reduction(+ : chk, pairs)
    for (int t = 0; t < nt; t++) {
        const auto& w = own[t];
        std::vector<double> sum(nopt);
        size_t g = 0;
        for (size_t k = 0; k < w.size(); g++) {
            const uint64_t a = w[k].key >> 32;
            std::fill(sum.begin(), sum.end(), 0.0);
            while (k < w.size() && (w[k].key >> 32) == a) {
                const uint64_t key = w[k].key;
                const int64_t sl = slot_of[(int64_t)(key & 0xffffffffull)];
                double hai = 0;
                for (; k < w.size() && w[k].key == key; k++) hai += w[k].v;
                for (int s = 0; s < nopt; s++) sum[s] += hai * Cs[(size_t)sl * nopt + s];
                pairs++;
            }
            const double ea = eas[t][g];
            for (int s = 0; s < nopt; s++) chk += sum[s] * sum[s] / (Es[s] - ea);
        }
    }
    out[9] = secs(t0);
    out[8] = (double)pairs;
    out[10] = chk;
    return 0;
}

/* Perform_HCI + DoPT2 + PrintFreqs + checkpoint policy (VHCI.cpp:5-185,
 * HB.cpp:176-265). */
int orc_run(const char* input, const char* freqs_out, const char* ckpt, int nthreads, const char* json_out,
            const char* basis_out, int quiet) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    if (nthreads > 0) g_threads = nthreads;
    std::string cpname = ckpt ? ckpt : "none";
    bool restart = false;
    if (cpname != "none") {
        std::ifstream cf(cpname);
        restart = cf.good() && cf.peek() != std::ifstream::traits_type::eof();
    }
    std::ifstream in(input);
    if (!in.good()) return fail("Error: Could not open input file.");
    Problem P;
    if (parse(in, restart, P)) return -1;
    const int M = P.M;
    std::vector<int> basis = P.basis0;
    std::vector<double> E, C;
    int ncol = 0;
    std::ostringstream js;
    js << "{\"M\": " << M << ", \"N0\": " << basis.size() / M << ", \"iters\": [";
    bool first_iter = true;
    auto log = [&](const std::string& s) {
        if (!quiet) std::cout << s << std::flush;
    };
    auto save_ckpt = [&]() {
        std::ofstream cf(cpname, std::ios::out | std::ios::trunc);
        if (!cf.good() || cpname == "none")
            log(" Error: checkpoint file could not be found or no name was given.\n");
        else
            write_ckpt(cf, P, basis, E, C, ncol);
    };
    if (restart) {
        std::ifstream cf(cpname);
        if (read_ckpt(cf, P, basis, E, C)) return -1;
        ncol = P.neig;
    }
    if (P.eps1 == 0.) {
        if (!restart) {
            int64_t nnz;
            if (sparse_diag(P, basis, E, C, ncol, nnz)) return -1;
            js << "{\"N\": " << basis.size() / M << ", \"nnz\": " << nnz << "}";
            first_iter = false;
            if ((int64_t)basis.size() / M > 100000) save_ckpt();
        }
    } else {
        if (!restart) {
            int64_t n = (int64_t)basis.size() / M;
            std::vector<double> A, w, Z;
            build_h_dense(P, basis.data(), n, A);
            eigh((int)n, A.data(), w, Z);
            E = w;
            C = Z;
            ncol = (int)n;
            std::ostringstream os;
            os << " Initial basis size is: " << n << "\n ZPE is: " << std::fixed << std::setprecision(2) << E[0]
               << "\n";
            log(os.str());
        }
        int iter = 1;
        while (true) {
            int64_t n = (int64_t)basis.size() / M;
            int nopt = P.neig > n ? (int)n : P.neig;
            std::vector<double> cm;
            cmax_rule(n, ncol, nopt, C.data(), cm);
            OrderedSet ns;
            if (hb_select(P, basis.data(), n, cm.data(), P.eps1, ns)) return -1;
            if ((double)ns.set.size() / (double)n <= 0.01) {
                std::ostringstream os;
                os << " Basis converged after " << iter - 1 << " iterations and contains " << n
                   << " product states.\n";
                log(os.str());
                break;
            }
            basis.insert(basis.end(), ns.rows.begin(), ns.rows.end());
            int64_t nnz;
            if (sparse_diag(P, basis, E, C, ncol, nnz)) return -1;
            std::ostringstream os;
            os << " Basis size after iteration " << iter << " is: " << basis.size() / M << "   ZPE "
               << std::fixed << std::setprecision(6) << E[0] << "\n";
            log(os.str());
            if (!first_iter) js << ", ";
            first_iter = false;
            js << "{\"N\": " << basis.size() / M << ", \"nnz\": " << nnz << ", \"new\": " << ns.set.size() << "}";
            iter++;
            if ((int64_t)basis.size() / M > 100000) save_ckpt();
        }
    }
    js << "]";
    std::vector<double> Evar = E;
    int64_t ns2 = -1;
    if (P.perturb) {
        int64_t n = (int64_t)basis.size() / M;
        int nopt = P.neig > n ? (int)n : P.neig;
        std::vector<double> dE;
        if (pt2(P, basis.data(), n, C.data(), ncol, E.data(), nopt, P.eps2, dE, ns2)) return -1;
        for (int s = 0; s < nopt; s++) E[s] += dE[s];
        std::ostringstream os;
        os << " Perturbative space contains " << ns2 << " states.\n";
        log(os.str());
    }
    {
        std::ofstream fo(freqs_out);
        if (!fo.good()) return fail("Error: Could not output file.");
        print_freqs(fo, P, basis, E, C, (int64_t)basis.size() / M, ncol);
    }
    double secs = std::chrono::duration<double>(clk::now() - t0).count();
    js << ", \"N_final\": " << basis.size() / M << ", \"pt2_space\": " << ns2 << ", \"seconds\": " << secs
       << ", \"threads\": " << (g_threads > 0 ? g_threads : omp_get_max_threads());
    auto arr = [&](const char* name, const std::vector<double>& v) {
        js << ", \"" << name << "\": [";
        char buf[64];
        for (size_t i = 0; i < v.size(); i++) {
            std::snprintf(buf, sizeof buf, "%.17g", v[i]);
            js << (i ? ", " : "") << buf;
        }
        js << "]";
    };
    arr("E_var", Evar);
    arr("E", E);
    js << "}\n";
    if (json_out) {
        std::ofstream jo(json_out);
        jo << js.str();
    }
    if (basis_out) {
        std::ofstream bo(basis_out);
        int64_t n = (int64_t)basis.size() / M;
        for (int64_t i = 0; i < n; i++) {
            for (int j = 0; j < M; j++) bo << basis[i * M + j] << (j + 1 < M ? " " : "\n");
        }
    }
    return 0;
}

}  // extern "C"
