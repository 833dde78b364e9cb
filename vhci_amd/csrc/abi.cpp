// abi.cpp -- C ABI of libvhci and the drop-in `vhci` driver.
//
// The driver keeps the reference's control flow and log lines
// (src/VHCI.cpp:5-185, src/HB.cpp:176-265, src/PT2.cpp:16-154) and replaces
// its numerical kernels by the device engine.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vhci.h"
#include "driver.hpp"
#include "server.hpp"
#include "engine.hpp"

using vhci::Engine;
using vhci::Problem;

struct vhci_problem {
    Problem P;
};

struct vhci_server {
    std::unique_ptr<vhci::srv::Server> s;
    std::string err;
};

struct vhci_ctx {
    std::unique_ptr<Engine> eng;
    std::string err;
    std::vector<int64_t> sizes;
};

namespace {

thread_local std::string g_problem_err;

int classify(const std::string& m) {
    if (m.find("did not converge") != std::string::npos) return VHCI_E_CONVERGE;
    if (m.find("HIP error") != std::string::npos || m.find("rocBLAS") != std::string::npos ||
        m.find("device") != std::string::npos)
        return VHCI_E_DEVICE;
    if (m.find("255 quanta") != std::string::npos || m.find("at most") != std::string::npos ||
        m.find("larger than") != std::string::npos)
        return VHCI_E_LIMIT;
    return VHCI_E_ARG;
}

template <class F>
int sguard(vhci_server* s, F&& f) {
    if (!s || !s->s) return VHCI_E_ARG;
    try {
        f();
        s->err.clear();
        return VHCI_OK;
    } catch (const std::invalid_argument& e) {
        s->err = e.what();
        return VHCI_E_INPUT;
    } catch (const std::exception& e) {
        s->err = e.what();
        return classify(s->err);
    }
}

template <class F>
int guard(vhci_ctx* ctx, F&& f) {
    if (!ctx || !ctx->eng) return VHCI_E_ARG;
    try {
        f();
        ctx->err.clear();
        return VHCI_OK;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return classify(ctx->err);
    }
}

using vhci::drv::Logger;
using vhci::drv::RunOpts;
using vhci::drv::RunResult;
using vhci::drv::fmt_fixed;
using vhci::drv::run;

std::string time_str(double secs, std::string& units) {
    double t = secs;
    if (t >= 3600) {
        t /= 3600;
        units = "hours";
    } else if (t >= 60) {
        t /= 60;
        units = "minutes";
    } else {
        units = "seconds";
    }
    return fmt_fixed(t, 2);
}

void banner() {
    std::cout << '\n'
              << "#############################################################################\n"
              << "#                                                                           #\n"
              << "#              Vibrational Heat-Bath Configuration Interaction              #\n"
              << "#                         By: Jonathan H. Fetherolf                         #\n"
              << "#                                                                           #\n"
              << "#                      Based on LOVCI by Eric G. Kratz                      #\n"
              << "#                                                                           #\n"
              << "#############################################################################\n\n";
}

void usage_short(const char* why) {
    std::cout << '\n' << why << "\n\nUsage: vhci -n Ncpus -i Modes.inp -o Freqs.dat -c Checkpoint.txt\n\n"
              << "Use -h or --help for detailed instructions.\n\n";
    std::cout.flush();
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- problem
const char* vhci_problem_error(void) { return g_problem_err.c_str(); }

int vhci_problem_load(const char* path, int restart, vhci_problem** out) {
    if (!path || !out) return VHCI_E_ARG;
    auto p = std::make_unique<vhci_problem>();
    std::string e = vhci::load_input(path, restart != 0, p->P);
    if (!e.empty()) {
        g_problem_err = e;
        return VHCI_E_INPUT;
    }
    *out = p.release();
    return VHCI_OK;
}

int vhci_problem_parse(const char* text, int restart, vhci_problem** out) {
    if (!text || !out) return VHCI_E_ARG;
    auto p = std::make_unique<vhci_problem>();
    std::istringstream in(text);
    std::string e = vhci::parse_input(in, restart != 0, p->P);
    if (!e.empty()) {
        g_problem_err = e;
        return VHCI_E_INPUT;
    }
    *out = p.release();
    return VHCI_OK;
}

void vhci_problem_free(vhci_problem* p) { delete p; }

int64_t vhci_problem_int(const vhci_problem* p, int what) {
    if (!p) return -1;
    const Problem& P = p->P;
    switch (what) {
        case 0: return P.M;
        case 1: return P.neig;
        case 2: return P.ntot;
        case 3: return P.perturb ? 1 : 0;
        case 4: return P.n0();
        case 5: return P.fcmax;
        case 6: return P.nfc;
        case 7: return (int64_t)P.hb.size();
        case 8:
        case 9: {
            vhci::Tables T;
            vhci::build_tables(P, T);
            return what == 8 ? (int64_t)T.deltas.size() : (int64_t)T.elem.size();
        }
    }
    return -1;
}

double vhci_problem_real(const vhci_problem* p, int what) {
    if (!p) return NAN;
    return what == 0 ? p->P.eps1 : p->P.eps2;
}

int vhci_problem_freqs(const vhci_problem* p, double* f) {
    if (!p || !f) return VHCI_E_ARG;
    std::memcpy(f, p->P.freq.data(), sizeof(double) * p->P.M);
    return VHCI_OK;
}

int vhci_problem_initial_basis(const vhci_problem* p, int32_t* q) {
    if (!p || !q) return VHCI_E_ARG;
    std::memcpy(q, p->P.basis0.data(), sizeof(int32_t) * p->P.basis0.size());
    return VHCI_OK;
}

int vhci_checkpoint_write(const vhci_problem* p, const char* path, const int32_t* q, int64_t n, const double* E,
                          int nE, const double* C, int ncol, int threads, int binary) {
    if (!p || !path || !q || !E || !C || n < 0 || nE < 0 || ncol < 0) return VHCI_E_ARG;
    try {
        std::ofstream cf(path, std::ios::out | std::ios::trunc | std::ios::binary);
        if (!cf.good()) {
            g_problem_err = "Error: checkpoint file could not be found or no name was given.";
            return VHCI_E_INPUT;
        }
        std::vector<double> ev(E, E + nE);
        vhci::write_checkpoint(cf, p->P, q, n, ev, C, ncol, std::max(1, threads), binary != 0);
        return cf.good() ? VHCI_OK : VHCI_E_INPUT;
    } catch (const std::exception& e) {
        g_problem_err = e.what();
        return VHCI_E_INPUT;
    }
}

int vhci_checkpoint_read(const vhci_problem* p, const char* path, int threads, int64_t* n, int32_t* q,
                         int64_t qcap, double* E, double* C, int64_t ccap) {
    if (!p || !path || !n) return VHCI_E_ARG;
    try {
        std::ifstream cf(path, std::ios::in | std::ios::binary);
        if (!cf.good()) {
            g_problem_err = "Error: checkpoint file could not be found or no name was given.";
            return VHCI_E_INPUT;
        }
        if (!q || !E || !C) {
            // size query: the state count from the header alone
            char head[8 + sizeof(int64_t)] = {};
            cf.read(head, sizeof(head));
            if (cf.gcount() == (std::streamsize)sizeof(head) &&
                std::memcmp(head, vhci::kCheckpointMagic, sizeof(vhci::kCheckpointMagic)) == 0) {
                std::memcpy(n, head + 8, sizeof(int64_t));
            } else {
                cf.clear();
                cf.seekg(0);
                if (!(cf >> *n)) {
                    g_problem_err = "Error: checkpoint file is truncated.";
                    return VHCI_E_INPUT;
                }
            }
            return VHCI_E_ARG;
        }
        std::vector<int32_t> qv;
        std::vector<double> ev, cv;
        std::string err = vhci::read_checkpoint(cf, p->P, qv, ev, cv, std::max(1, threads));
        if (!err.empty()) {
            g_problem_err = err;
            return VHCI_E_INPUT;
        }
        *n = (int64_t)qv.size() / std::max(1, p->P.M);
        if (qcap < (int64_t)qv.size() || ccap < (int64_t)cv.size()) return VHCI_E_ARG;
        std::copy(qv.begin(), qv.end(), q);
        std::copy(ev.begin(), ev.end(), E);
        std::copy(cv.begin(), cv.end(), C);
        return VHCI_OK;
    } catch (const std::exception& e) {
        g_problem_err = e.what();
        return VHCI_E_INPUT;
    }
}

static const vhci::ForceConst* fc_at(const Problem& P, int which, int idx, int& count) {
    const std::vector<int>* lst = nullptr;
    switch (which) {
        case 0: count = (int)P.fcs.size(); return (idx >= 0 && idx < count) ? &P.fcs[idx] : nullptr;
        case 1: lst = &P.cubic; break;
        case 2: lst = &P.quartic; break;
        case 3: lst = &P.quintic; break;
        case 4: lst = &P.sextic; break;
        case 5: count = (int)P.hb.size(); return (idx >= 0 && idx < count) ? &P.hb[idx] : nullptr;
        case 6:
            count = (int)P.hb_sorted.size();
            return (idx >= 0 && idx < count) ? &P.hb[P.hb_sorted[idx]] : nullptr;
        default: count = -1; return nullptr;
    }
    count = (int)lst->size();
    return (idx >= 0 && idx < count) ? &P.fcs[(*lst)[idx]] : nullptr;
}

int vhci_problem_fc_count(const vhci_problem* p, int which) {
    if (!p) return VHCI_E_ARG;
    int c = -1;
    fc_at(p->P, which, -1, c);
    return c < 0 ? VHCI_E_ARG : c;
}

int vhci_problem_fc(const vhci_problem* p, int which, int idx, double* fc, int* order, int* nsm, int* sm, int* mp) {
    if (!p) return VHCI_E_ARG;
    int c = 0;
    const vhci::ForceConst* f = fc_at(p->P, which, idx, c);
    if (!f) return VHCI_E_ARG;
    *fc = f->fc;
    *order = f->order();
    *nsm = (int)f->sm.size();
    for (int i = 0; i < *nsm && i < 6; i++) {
        sm[i] = f->sm[i];
        mp[i] = f->mp[i];
    }
    return VHCI_OK;
}

double vhci_ladder_sum(int k, int qs, int qe) { return vhci::ladder_sum(k, qs, qe); }

// ---------------------------------------------------------------- context
int vhci_ctx_create(const vhci_problem* p, int form, int device, vhci_ctx** out) {
    if (!p || !out) return VHCI_E_ARG;
    auto c = std::make_unique<vhci_ctx>();
    try {
        c->eng = std::make_unique<Engine>(p->P, form, device);
        c->eng->set_initial_basis();
    } catch (const std::exception& e) {
        g_problem_err = e.what();
        return classify(g_problem_err);
    }
    *out = c.release();
    return VHCI_OK;
}

void vhci_ctx_destroy(vhci_ctx* ctx) { delete ctx; }

const char* vhci_last_error(const vhci_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int vhci_set_basis(vhci_ctx* ctx, const int32_t* q, int64_t n) {
    return guard(ctx, [&] { ctx->eng->set_basis(q, n); });
}

int64_t vhci_basis_size(const vhci_ctx* ctx) { return ctx && ctx->eng ? ctx->eng->n() : -1; }

int vhci_get_basis(vhci_ctx* ctx, int32_t* q) {
    return guard(ctx, [&] { ctx->eng->get_basis(q); });
}

int vhci_build_h(vhci_ctx* ctx, int64_t* stats) {
    return guard(ctx, [&] {
        auto h = ctx->eng->build_h();
        if (stats) {
            stats[0] = h.n;
            stats[1] = h.nnz_offdiag;
            stats[2] = h.nnz_struct;
            stats[3] = (int64_t)(h.build_ms * 1000);
        }
    });
}

int vhci_matrix_free(const vhci_ctx* ctx) {
    if (!ctx || !ctx->eng) return VHCI_E_ARG;
    return ctx->eng->matrix_free() ? 1 : 0;
}

int64_t vhci_structural_nnz(vhci_ctx* ctx) {
    if (!ctx || !ctx->eng) return VHCI_E_ARG;
    int64_t r = -1;
    int rc = guard(ctx, [&] { r = ctx->eng->structural_nnz(); });
    return rc ? rc : r;
}

int vhci_get_h(vhci_ctx* ctx, int64_t* nnz, int64_t* rowptr, int32_t* col, double* val, double* diag) {
    return guard(ctx, [&] {
        std::vector<int64_t> rp;
        std::vector<int32_t> cl;
        std::vector<double> vl, dg;
        ctx->eng->get_h(rp, cl, vl, dg);
        if (nnz) *nnz = (int64_t)cl.size();
        if (rowptr) {
            std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int64_t));
            std::memcpy(col, cl.data(), cl.size() * sizeof(int32_t));
            std::memcpy(val, vl.data(), vl.size() * sizeof(double));
            std::memcpy(diag, dg.data(), dg.size() * sizeof(double));
        }
    });
}

int vhci_get_h_rows(vhci_ctx* ctx, const int64_t* rows, int64_t nr, int64_t* nnz, int64_t* rowptr, int32_t* col,
                    double* val, double* diag) {
    if (!rows || nr < 0 || !nnz) return VHCI_E_ARG;
    return guard(ctx, [&] {
        std::vector<int64_t> rp;
        std::vector<int32_t> cl;
        std::vector<double> vl, dg;
        ctx->eng->get_h_rows(rows, nr, rp, cl, vl, dg);
        *nnz = (int64_t)cl.size();
        if (rowptr) {
            std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int64_t));
            std::memcpy(col, cl.data(), cl.size() * sizeof(int32_t));
            std::memcpy(val, vl.data(), vl.size() * sizeof(double));
            std::memcpy(diag, dg.data(), dg.size() * sizeof(double));
        }
    });
}

int vhci_eigsolve(vhci_ctx* ctx, int nev, double tol, int maxit, double* E, double* C) {
    return guard(ctx, [&] {
        std::vector<double> e;
        ctx->eng->eigsolve(nev, tol, maxit, e);
        std::memcpy(E, e.data(), e.size() * sizeof(double));
        if (C) ctx->eng->get_vectors(C, (int)e.size());
    });
}

int vhci_dense_diag(vhci_ctx* ctx, double* E, double* C) {
    return guard(ctx, [&] {
        std::vector<double> e;
        ctx->eng->dense_all(e);
        std::memcpy(E, e.data(), e.size() * sizeof(double));
        if (C) ctx->eng->get_vectors(C, (int)e.size());
    });
}

int vhci_ritz_cols(const vhci_ctx* ctx) { return ctx && ctx->eng ? ctx->eng->ritz_cols() : -1; }

int vhci_get_vectors(vhci_ctx* ctx, double* C, int ncol) {
    return guard(ctx, [&] { ctx->eng->get_vectors(C, ncol); });
}

int vhci_set_vectors(vhci_ctx* ctx, const double* C, int ncol) {
    return guard(ctx, [&] { ctx->eng->set_vectors(C, ncol); });
}

int vhci_freq_assignments(vhci_ctx* ctx, int ncol, int64_t* counts, int64_t* rows, double* absval, int64_t cap,
                          int64_t* total) {
    if (!counts || !total) return VHCI_E_ARG;
    bool small = false;
    int rc = guard(ctx, [&] {
        std::vector<vhci::FreqColumn> cols;
        ctx->eng->freq_assignments(ncol, cols);
        int64_t t = 0;
        for (int c = 0; c < ncol; c++) {
            counts[c] = (int64_t)cols[c].row.size();
            t += counts[c];
        }
        *total = t;
        if (!rows || !absval || cap < t) {
            small = true;
            return;
        }
        int64_t w = 0;
        for (int c = 0; c < ncol; c++)
            for (size_t k = 0; k < cols[c].row.size(); k++, w++) {
                rows[w] = cols[c].row[k];
                absval[w] = cols[c].absval[k];
            }
    });
    return rc != VHCI_OK ? rc : (small ? VHCI_E_ARG : VHCI_OK);
}

int vhci_hb_select(vhci_ctx* ctx, const double* cmax, int nopt, double eps, int64_t* nnew) {
    return guard(ctx, [&] {
        ctx->eng->hb_select(cmax, nopt, eps);
        if (nnew) *nnew = ctx->eng->new_count();
    });
}

int vhci_get_new_states(vhci_ctx* ctx, int32_t* q) {
    return guard(ctx, [&] { ctx->eng->get_new_states(q); });
}

int vhci_append_new_states(vhci_ctx* ctx) {
    return guard(ctx, [&] { ctx->eng->append_new_states(); });
}

int vhci_pt2(vhci_ctx* ctx, const double* E, int nopt, double eps2, double* dE, int64_t* ns2) {
    return guard(ctx, [&] {
        std::vector<double> e(E, E + nopt), d;
        auto st = ctx->eng->pt2(e, nopt, eps2, d);
        std::memcpy(dE, d.data(), nopt * sizeof(double));
        if (ns2) *ns2 = st.ns2;
    });
}

int vhci_copy_probe(vhci_ctx* ctx, int64_t bytes, double* ms) {
    return guard(ctx, [&] { *ms = ctx->eng->copy_probe((size_t)bytes); });
}

int vhci_spmm_profile(vhci_ctx* ctx, int b) {
    return guard(ctx, [&] {
        if (b < 0) throw std::invalid_argument("vhci_spmm_profile: block width must be >= 0");
        ctx->eng->spmm_profile(b);
    });
}

int vhci_spmm_profile_read(vhci_ctx* ctx, double* out) {
    return guard(ctx, [&] { ctx->eng->spmm_profile_read(out); });
}

int vhci_gemm_profile_read(vhci_ctx* ctx, double* out) {
    return guard(ctx, [&] { ctx->eng->gemm_profile_read(out); });
}

int vhci_phase_stats(vhci_ctx* ctx, double* out) {
    return guard(ctx, [&] { ctx->eng->phase_stats(out); });
}

int vhci_pool_stats(vhci_ctx* ctx, int64_t* out) {
    if (!out) return VHCI_E_ARG;
    return guard(ctx, [&] { ctx->eng->pool_stats(out); });
}

int vhci_apply_h(vhci_ctx* ctx, int b, int fp32, const double* X, double* Y) {
    return guard(ctx, [&] {
        if (!X || !Y) throw std::runtime_error("vhci_apply_h: null vector");
        ctx->eng->apply_h(X, Y, b, fp32 != 0);
    });
}

int vhci_apply_h_noalloc(vhci_ctx* ctx, int b, int fp32, const double* X, double* Y) {
    return guard(ctx, [&] {
        if (!X || !Y) throw std::runtime_error("vhci_apply_h_noalloc: null vector");
        ctx->eng->apply_h_noalloc(X, Y, b, fp32 != 0);
    });
}

int vhci_time_spmm(vhci_ctx* ctx, int b, int reps, int fp32, double* out) {
    return guard(ctx, [&] {
        double bytes = 0;
        out[0] = ctx->eng->time_spmm(b, reps, fp32 != 0, &bytes);
        out[1] = bytes;
        out[2] = ctx->eng->spmm_launches(b);
    });
}

static uint64_t basis_checksum(Engine& eng) { return vhci::drv::basis_checksum(eng); }

static void fill_timings(double* t, const RunResult& R, double total) {
    if (!t) return;
    t[0] = total;
    t[1] = R.t_build;
    t[2] = R.t_eig;
    t[3] = R.t_hb;
    t[4] = R.t_pt2sel;
    t[5] = R.t_pt2sum;
    t[6] = R.iters;
    t[7] = (double)R.spmm_vectors;
}

int vhci_run_ctx(vhci_ctx* ctx, double tol, int restart, double* E, double* Evar, int* nE, int64_t* info,
                 double* timings) {
    return guard(ctx, [&] {
        if (restart) throw std::runtime_error("vhci_run_ctx: restart is handled by the vhci program");
        auto t0 = std::chrono::steady_clock::now();
        Engine& eng = *ctx->eng;
        RunOpts o;
        if (tol > 0) o.tol = tol;
        o.restart = restart != 0;
        RunResult R;
        run(eng, eng.problem(), o, R, Logger{false});
        ctx->sizes = R.sizes;
        if (E) std::memcpy(E, R.E.data(), R.E.size() * sizeof(double));
        if (Evar) std::memcpy(Evar, R.Evar.data(), R.Evar.size() * sizeof(double));
        if (nE) *nE = (int)R.E.size();
        if (info) {
            info[0] = R.nvar;
            info[1] = R.ns2;
            info[2] = R.hb_iters;
            info[3] = (int64_t)basis_checksum(eng);
        }
        fill_timings(timings, R,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    });
}

int vhci_iteration_sizes(const vhci_ctx* ctx, int64_t* sizes, int max) {
    if (!ctx) return VHCI_E_ARG;
    int n = std::min<int>(max, (int)ctx->sizes.size());
    for (int i = 0; i < n; i++) sizes[i] = ctx->sizes[i];
    return n;
}

int vhci_solve(const vhci_problem* p, int form, int device, double tol, double* E, int* nE, int64_t* nvar,
               int64_t* ns2, double* timings) {
    if (!p) return VHCI_E_ARG;
    auto t0 = std::chrono::steady_clock::now();
    try {
        Engine eng(p->P, form, device);
        eng.set_initial_basis();
        RunOpts o;
        if (tol > 0) o.tol = tol;
        RunResult R;
        run(eng, p->P, o, R, Logger{false});
        if (E) std::memcpy(E, R.E.data(), R.E.size() * sizeof(double));
        if (nE) *nE = (int)R.E.size();
        if (nvar) *nvar = R.nvar;
        if (ns2) *ns2 = R.ns2;
        fill_timings(timings, R,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    } catch (const std::exception& e) {
        g_problem_err = e.what();
        return classify(g_problem_err);
    }
    return VHCI_OK;
}

// ---------------------------------------------------------------- driver
// ---------------------------------------------------------------- serving
int vhci_server_create(int device, int slots, int64_t mem_limit, vhci_server** out) {
    if (!out || slots < 1 || mem_limit <= 0) return VHCI_E_ARG;
    *out = nullptr;
    try {
        auto s = new vhci_server;
        s->s = std::make_unique<vhci::srv::Server>(device, slots, mem_limit);
        *out = s;
        return VHCI_OK;
    } catch (const std::exception& e) {
        g_problem_err = e.what();
        return VHCI_E_DEVICE;
    }
}

void vhci_server_destroy(vhci_server* s) { delete s; }

const char* vhci_server_error(const vhci_server* s) { return s ? s->err.c_str() : ""; }

int vhci_server_submit(vhci_server* s, const char* text, int form, double tol, int64_t* id) {
    if (!text || !id) return VHCI_E_ARG;
    return sguard(s, [&] { *id = s->s->submit(text, form, tol); });
}

int vhci_server_wait(vhci_server* s, int64_t id, double timeout_s, int* done) {
    if (!done) return VHCI_E_ARG;
    return sguard(s, [&] { *done = s->s->wait(id, timeout_s) ? 1 : 0; });
}

int vhci_server_result(vhci_server* s, int64_t id, double* E, double* Evar, int* nE, int64_t* info,
                       double* times) {
    if (!s || !s->s) return VHCI_E_ARG;
    vhci::srv::Request r;
    int rc = sguard(s, [&] { r = s->s->get(id); });
    if (rc != VHCI_OK) return rc;
    if (info) {
        info[0] = (int64_t)r.state;
        info[1] = r.R.nvar;
        info[2] = r.R.ns2;
        info[3] = r.R.hb_iters;
        info[4] = (int64_t)r.checksum;
        info[5] = r.slot;
    }
    if (times) {
        times[0] = r.t_submit;
        times[1] = r.t_start;
        times[2] = r.t_end;
        times[3] = r.predicted;
        times[4] = (double)r.peak;
        times[5] = r.reserved;
    }
    if (nE) *nE = (int)r.R.E.size();
    if (E) std::copy(r.R.E.begin(), r.R.E.end(), E);
    if (Evar) std::copy(r.R.Evar.begin(), r.R.Evar.end(), Evar);
    if (r.state == vhci::srv::State::Failed) {
        s->err = r.err;
        return classify(r.err);
    }
    return VHCI_OK;
}

int vhci_server_stats(vhci_server* s, double* out) {
    if (!out) return VHCI_E_ARG;
    return sguard(s, [&] { s->s->stats(out); });
}

static int main_body(int argc, char** argv) {
    const auto start = std::chrono::steady_clock::now();
    std::cout << std::fixed << std::setprecision(12);
    banner();
    std::cout << "Last modification: " << __TIME__ << " on " << __DATE__ << "\n\n";
    std::cout << "Reading input...\n";
    // ReadCIArgs (Input_Reader.cpp:17-168)
    if (argc == 1) {
        usage_short("Missing arguments...");
        return 1;
    }
    if ((argc % 2) != 1) {
        std::string a1 = argv[1];
        if (a1 != "-h" && a1 != "--help") {
            usage_short("Wrong number of arguments...");
            return 1;
        }
    }
    int ncpus = 0;
    std::string in_path, out_path, cpname = "none";
    bool have_in = false, have_out = false;
    for (int i = 0; i < argc; i++) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << "\nUsage: vhci -n Ncpus -i Modes.inp -o Freqs.txt\n\nCommand line arguments:\n\n"
                      << "  -n    Number of CPUs to use for the calculation.\n\n"
                      << "  -i    Input file.\n\n"
                      << "  -o    Output file for the first NEig VCI frequencies.\n\n"
                      << "  -c    Checkpoint file for HCI iterations on large systems.\n\n";
            std::cout.flush();
            return 0;
        }
        if (i + 1 >= argc) continue;
        if (a == "-n") ncpus = std::atoi(argv[i + 1]);
        if (a == "-i") {
            in_path = argv[i + 1];
            have_in = true;
        }
        if (a == "-o") {
            out_path = argv[i + 1];
            have_out = true;
        }
        if (a == "-c") cpname = argv[i + 1];
    }
    const int maxthr = std::max(1u, std::thread::hardware_concurrency());
    if (ncpus < 1) {
        std::cout << " Warning: Calculations cannot run with " << ncpus << " CPUs.\n"
                  << "  Do you know how computers work? Ncpus set to 1\n";
        ncpus = 1;
    }
    if (ncpus > maxthr) {
        std::cout << " Warning: Too many threads requested.\n  Requested: " << ncpus << ", Available: " << maxthr
                  << "\n  Ncpus set to " << maxthr << "\n";
        ncpus = maxthr;
    }
    std::ifstream vcidata;
    std::ofstream freqfile;
    if (have_in) vcidata.open(in_path);
    if (have_out) freqfile.open(out_path);
    bool quit = false;
    if (!have_in || !vcidata.good()) {
        std::cout << " Error: Could not open input file.\n";
        quit = true;
    }
    if (!have_out || !freqfile.good()) {
        std::cout << " Error: Could not output file.\n";
        quit = true;
    }
    if (quit) {
        std::cout << '\n';
        std::cout.flush();
        return 1;
    }
    std::cout << "\nNo fatal errors detected.\n And there was much rejoicing. Yay...\n\n";
    bool restart = false;
    if (cpname == "none") {
        std::cout << "No checkpoint filename specified. Checkpoints will not be saved.\n\n";
    } else {
        std::cout << "Locating checkpoint file...\n";
        std::ifstream cf(cpname);
        restart = cf.good() && cf.peek() != std::ifstream::traits_type::eof();
        if (restart)
            std::cout << "Checkpoint file found. Calculation will be resumed.\n\n";
        else
            std::cout << "No checkpoint file found. Calculation will start from the beginning.\n\n";
    }
    std::cout.flush();
    // ReadCIInput
    Problem P;
    std::string err = vhci::parse_input(vcidata, restart, P);
    if (!err.empty()) {
        std::cout << err << "\n\n";
        return 1;
    }
    std::cout << "General settings:\n  CPU threads: " << ncpus << "\n\n  Normal modes: " << P.M
              << "\n  Max quanta per state: " << P.ntot << "\n  Basis functions: " << P.n0()
              << "\n  Force constants: ";
    if (P.nfc == 0)
        std::cout << "N/A\n\n";
    else
        std::cout << P.fcs.size() << "\n\n";
    if (P.eps1 != 0.) {
        std::cout << std::setprecision(3) << "  Using HCI with energy cutoff: " << P.eps1 << '\n';
        if (P.neig == 1)
            std::cout << "  Optimizing just the ground states (ZPE).\n";
        else
            std::cout << "  Optimizing the first " << P.neig << " eigenstates.\n";
    }
    std::cout << '\n' << std::setprecision(12);
    std::cout.flush();
    int form = VHCI_FP64;
    if (const char* f = std::getenv("VHCI_FORM")) form = (std::string(f) == "mixed") ? VHCI_MIXED : VHCI_FP64;
    int device = 0;
    if (const char* d = std::getenv("VHCI_DEVICE")) device = std::atoi(d);
    RunOpts o;
    if (const char* t = std::getenv("VHCI_TOL")) o.tol = std::atof(t);
    if (const char* t = std::getenv("VHCI_STOP_AFTER_ITERATION")) o.stop_after = std::atoi(t);
    if (const char* t = std::getenv("VHCI_MAX_STATES")) o.max_states = std::atoll(t);
    if (const char* t = std::getenv("VHCI_CHECKPOINT_FORMAT")) o.binary_checkpoint = std::string(t) == "binary";
    o.threads = ncpus;
    o.cpname = cpname;
    o.restart = restart;
    o.log = true;
    Logger log{true};
    RunResult R;
    try {
        Engine eng(P, form, device);
        if (restart) {
            std::cout << "\nLoading checkpoint file...";
            std::cout.flush();
            std::ifstream cf(cpname, std::ios::in | std::ios::binary);
            std::vector<int32_t> q;
            std::vector<double> C;
            std::string ce = vhci::read_checkpoint(cf, P, q, R.E, C, ncpus);
            if (!ce.empty()) {
                std::cout << ce << "\n";
                return 1;
            }
            const int64_t nb = (int64_t)q.size() / P.M;
            eng.set_basis(q.data(), nb);
            eng.set_vectors(C.data(), P.neig);
            std::cout << "done.\n  Basis functions: " << nb << "\n" << std::endl;
        } else {
            eng.set_initial_basis();
        }
        const bool pt = P.perturb;
        if (pt) {
            // the "Variational stage took" line is printed between the stages
            RunOpts ov = o;
            Problem Pv = P;
            Pv.perturb = false;
            run(eng, Pv, ov, R, log);
            std::string units;
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
            std::string t = time_str(secs, units);
            std::cout << " Variational stage took " << t << " " << units << "\n\n";
            std::cout << "Performing Epstein-Nesbet perturbation theory...\n";
            std::cout.flush();
            // PT2 on the converged space
            const int64_t n = eng.n();
            const int nopt = (int)std::min<int64_t>(P.neig, n);
            std::vector<double> E = R.E, dE;
            if (P.eps2 > 0) {
                std::ostringstream os;
                os << " Using heat bath sorting of perturbative states with cutoff: " << std::scientific
                   << std::setprecision(2) << P.eps2 << "\n";
                log(os.str());
            } else {
                log(" Including all connected states in PT2 calculation.\n");
            }
            log(" Finding connected states...\n");
            auto ps = eng.pt2(E, nopt, P.eps2, dE);
            log(" Perturbative space contains " + std::to_string(ps.ns2) + " states.\n");
            if (P.neig == 1)
                log(" Calculating the 2nd-order perturbative correction on the ground state energy\n");
            else
                log(" Calculating the 2nd-order perturbative correction for on first " + std::to_string(nopt) +
                    " eigenvalues.\n");
            for (int s = 0; s < nopt; s++) E[s] += dE[s];
            log(" New ZPE after PT2 correction is: " + fmt_fixed(E[0], 2) + "\n");
            R.E = E;
        } else {
            run(eng, P, o, R, log);
        }
        // PrintFreqs (VHCI.cpp:116-119)
        const int ncol = std::min<int>(eng.ritz_cols(), std::max(1, std::min<int>(P.neig, (int)R.E.size())));
        std::vector<vhci::FreqColumn> cols;
        eng.freq_assignments(ncol, cols);
        vhci::write_freqs(freqfile, P, R.E, cols);
    } catch (const std::exception& e) {
        std::cout << "\n" << e.what() << "\n";
        return 1;
    }
    double ezpe = R.E[0];
    for (double e : R.E) ezpe = std::min(ezpe, e);
    if (ezpe < 0) {
        std::cout << "Error: The zero-point energy is negative. Something is wrong.\n"
                  << "Try reducing the number of quanta or removing large anharmonic\n"
                  << "force constants.\n\n";
        return 1;
    }
    std::string units;
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::string t = time_str(secs, units);
    std::cout << "Printing the results...\n\nResults:\n  Zero-point energy: " << fmt_fixed(ezpe, 2)
              << " (1/cm)\n  Run time: " << t << " " << units << "\n\nDone.\n\n";
    std::cout.flush();
    return 0;
}

// The reference prints its error messages and exits with status 0 in every
// case (exit(0) after each message, src/Input_Reader.cpp:17-168, src/VHCI.cpp);
// this program returns 1 after the same messages so that scripts can tell a
// failed run, unless VHCI_ERROR_STATUS=0 asks for the reference's status.
int vhci_main(int argc, char** argv) {
    const int rc = main_body(argc, argv);
    if (rc != 0) {
        const char* e = std::getenv("VHCI_ERROR_STATUS");
        if (e && std::string(e) == "0") return 0;
    }
    return rc;
}

}  // extern "C"
