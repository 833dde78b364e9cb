// driver.hpp -- the iterative select -> build -> diagonalize -> PT2 loop of
// one solve on a device context (Perform_HCI, src/HB.cpp:176-265, and
// DoPT2, src/PT2.cpp:16-154, as called from main, src/VHCI.cpp:39-114).
// Shared by the C ABI, the `vhci` program and the serving loop.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "engine.hpp"

namespace vhci {
namespace drv {

struct RunOpts {
    double tol = 1e-10;  // as the reference's Spectra call (src/Ham.cpp:363)
    int maxit = 500;
    std::string cpname = "none";
    bool restart = false;
    bool log = false;
    int stop_after = 0;  // failure injection: end the run after this HCI iteration
    int threads = 1;     // host threads for checkpoint formatting (-n)
    bool binary_checkpoint = false;  // opt-in binary checkpoint form
    // end the run with an error when a heat-bath iteration would grow the
    // basis beyond this many states (0: no limit)
    int64_t max_states = 0;
    // called after every basis growth with the new basis size (serving
    // loop: memory reservation updates)
    std::function<void(int64_t)> on_grow;
};

struct RunResult {
    std::vector<double> E;  // final energies (after PT2)
    std::vector<double> Evar;
    int64_t nvar = 0;
    int64_t ns2 = 0;
    double t_build = 0, t_eig = 0, t_hb = 0, t_pt2sel = 0, t_pt2sum = 0;
    int iters = 0;
    int hb_iters = 0;
    int64_t spmm_vectors = 0;
    std::vector<int64_t> sizes;
};

struct Logger {
    bool on;
    void operator()(const std::string& s) const;
};

std::string fmt_fixed(double v, int prec);
// sum over the int32 quanta of the basis, row-major, of
// q_t * 1099511628211^(t+1) mod 2^64
uint64_t basis_checksum(Engine& eng);
// Runs the loop on eng, whose basis is P's initial basis (or, with
// o.restart, the checkpoint's basis with R.E holding its energies).
void run(Engine& eng, const Problem& P, const RunOpts& o, RunResult& R, const Logger& log);

}  // namespace drv
}  // namespace vhci
