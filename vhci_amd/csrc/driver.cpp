// driver.cpp -- the select -> build -> diagonalize -> PT2 loop of one solve
// with the reference's log lines (src/VHCI.cpp:39-114, src/HB.cpp:176-265,
// src/PT2.cpp:16-154).
#include "driver.hpp"

#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace vhci {
namespace drv {

void Logger::operator()(const std::string& s) const {
    if (on) {
        std::cout << s;
        std::cout.flush();
    }
}

namespace {

std::string sparse_line(const vhci::HStats& h) {
    // "% sparse" (Ham.cpp:341-342) from the structural count when it was
    // computed, else from the stored entries plus the diagonal.
    double nnz = h.nnz_struct >= 0 ? (double)h.nnz_struct : (double)(h.nnz_offdiag + h.n);
    double pct = 100 * (1. - nnz / ((double)h.n * (double)h.n));
    return "The Hamiltonian is " + fmt_fixed(pct, 2) + "% sparse.\n";
}

void save_checkpoint(Engine& eng, const Problem& P, const std::vector<double>& E, const RunOpts& o,
                     const Logger& log) {
    // The reference opens the file before it tests for the name "none"
    // (src/HB.cpp:253-256), which leaves an empty file called
    // "none" in the working directory; the name is tested first here.
    if (o.cpname == "none") {
        log("\n Error: checkpoint file could not be found or no name was given.\n\n");
        return;
    }
    std::ofstream cf(o.cpname, std::ios::out | std::ios::trunc | std::ios::binary);
    if (!cf.good()) {
        log("\n Error: checkpoint file could not be found or no name was given.\n\n");
        return;
    }
    const int64_t n = eng.n();
    const int nc = eng.ritz_cols();
    std::vector<int32_t> q((size_t)n * P.M);
    std::vector<double> C((size_t)n * nc);
    eng.get_basis(q.data());
    eng.get_vectors(C.data(), nc);
    vhci::write_checkpoint(cf, P, q.data(), n, E, C.data(), nc, o.threads, o.binary_checkpoint);
}

}  // namespace

std::string fmt_fixed(double v, int prec) {
    std::ostringstream os;
    os << std::fixed << std::setprecision(prec) << v;
    return os.str();
}

uint64_t basis_checksum(Engine& eng) {
    const int64_t n = eng.n();
    const int M = eng.problem().M;
    std::vector<int32_t> q((size_t)n * M);
    eng.get_basis(q.data());
    const uint64_t P = 1099511628211ull;
    uint64_t h = 0, pw = 1;
    for (int32_t v : q) {
        pw *= P;
        h += (uint64_t)(int64_t)v * pw;
    }
    return h;
}

void run(Engine& eng, const Problem& P, const RunOpts& o, RunResult& R, const Logger& log) {
    std::vector<double> E;
    R.sizes.push_back(eng.n());
    if (o.on_grow) o.on_grow(eng.n());
    if (P.eps1 == 0.) {
        // one-shot VCI (VHCI.cpp:49-76)
        if (!o.restart) {
            log("Constructing the Vibrational CI Hamiltonian...");
            auto hs = eng.build_h();
            R.t_build += hs.build_ms;
            if (log.on) hs.nnz_struct = eng.structural_nnz();
            log(sparse_line(hs));
            if (P.neig == 1)
                log("done.\n\nSolving for the ground state...\n\n");
            else
                log("done.\n\nSolving for the lowest " + std::to_string(P.neig) + " eigenstates...\n\n");
            auto es = eng.eigsolve(P.neig, o.tol, o.maxit, E);
            R.t_eig += es.ms;
            R.spmm_vectors += es.spmm_vectors;
            if (eng.n() > 100000) {
                log("Writing checkpoint file...");
                save_checkpoint(eng, P, E, o, log);
                log("done.\n\n");
            }
        } else {
            E = R.E;
        }
    } else {
        // Perform_HCI (HB.cpp:176-265)
        log("Beginning Heat Bath CI procedure...\n");
        int iter = 1;
        if (!o.restart) {
            log(" Initial basis size is: " + std::to_string(eng.n()) + "\n");
            log(" Constructing Hamiltonian...");
            auto hs = eng.build_h();
            R.t_build += hs.build_ms;
            log(" Diagonalizing...\n");
            auto es = eng.dense_all(E);
            R.t_eig += es.ms;
            log(" ZPE is: " + fmt_fixed(E[0], 2) + "\n\n");
        } else {
            E = R.E;
        }
        while (true) {
            log(" Performing HCI iteration: " + std::to_string(iter) + "\n");
            const int64_t n = eng.n();
            const int nopt = (int)std::min<int64_t>(P.neig, n);
            auto hb = eng.hb_select(nullptr, nopt, P.eps1);
            R.t_hb += hb.ms;
            R.hb_iters++;
            if ((double)eng.new_count() / (double)n <= 0.01) {
                log(" Basis converged after " + std::to_string(iter - 1) + " iterations and contains " +
                    std::to_string(n) + " product states.\n");
                break;
            }
            if (o.max_states > 0 && n + eng.new_count() > o.max_states)
                throw std::runtime_error("heat-bath iteration " + std::to_string(iter) + " would grow the basis to " +
                                         std::to_string(n + eng.new_count()) + " states, beyond the limit of " +
                                         std::to_string(o.max_states) + " (VHCI_MAX_STATES)");
            eng.append_new_states();
            R.sizes.push_back(eng.n());
            if (o.on_grow) o.on_grow(eng.n());
            log(" Basis size after iteration " + std::to_string(iter) + " is: " + std::to_string(eng.n()) + "\n");
            ++iter;
            log(" Constructing Hamiltonian...");
            auto hs = eng.build_h();
            R.t_build += hs.build_ms;
            if (log.on) hs.nnz_struct = eng.structural_nnz();
            log(sparse_line(hs));
            if (P.neig == 1)
                log(" Finding the ground state...\n");
            else
                log(" Finding the first " + std::to_string(P.neig) + " eigenstates...\n");
            auto es = eng.eigsolve(P.neig, o.tol, o.maxit, E);
            R.t_eig += es.ms;
            R.spmm_vectors += es.spmm_vectors;
            R.iters++;
            log(" ZPE is: " + fmt_fixed(E[0], 2) + "\n\n");
            if (eng.n() > 100000) {
                log(" Writing checkpoint file...");
                save_checkpoint(eng, P, E, o, log);
                log("done.\n\n");
            }
            if (o.stop_after > 0 && iter - 1 >= o.stop_after)
                throw std::runtime_error("stopped after heat-bath iteration " + std::to_string(iter - 1) +
                                         " (VHCI_STOP_AFTER_ITERATION)");
        }
    }
    R.Evar = E;
    R.nvar = eng.n();
    if (P.perturb) {
        const int64_t n = eng.n();
        const int nopt = (int)std::min<int64_t>(P.neig, n);
        std::vector<double> dE;
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
        R.ns2 = ps.ns2;
        R.t_pt2sel += ps.select_ms;
        R.t_pt2sum += ps.sum_ms;
        log(" Perturbative space contains " + std::to_string(ps.ns2) + " states.\n");
        if (P.neig == 1)
            log(" Calculating the 2nd-order perturbative correction on the ground state energy\n");
        else
            log(" Calculating the 2nd-order perturbative correction for on first " + std::to_string(nopt) +
                " eigenvalues.\n");
        for (int s = 0; s < nopt; s++) E[s] += dE[s];
        log(" New ZPE after PT2 correction is: " + fmt_fixed(E[0], 2) + "\n");
    }
    R.E = E;
}


}  // namespace drv
}  // namespace vhci
