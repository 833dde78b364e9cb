// engine_stub.cpp -- a CPU stand-in for the device engine, for a
// ThreadSanitizer build of the serving loop (server.cpp) and the driver
// (driver.cpp).  It makes no HIP call.  Each phase sleeps for a short,
// request-dependent time and moves the thread's device-byte counters as the
// real engine's pool does, so several worker threads interleave through
// admission, reservation updates and completion the way they do on the GPU.
// The basis grows 4x, 2x, then converges; energies are a function of the
// request, so every result is reproducible and can be checked.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <thread>

#include "engine.hpp"

namespace vhci {
namespace {
thread_local int64_t t_bytes = 0;
thread_local int64_t t_peak = 0;
thread_local DevPool* t_pool = nullptr;

// requests of STUB_SLOW_MODES modes (environment) take 200 times as long
thread_local bool t_slow = false;
void nap(int64_t us) { std::this_thread::sleep_for(std::chrono::microseconds(t_slow ? 200 * us : us)); }
}  // namespace

int64_t thread_device_bytes() { return t_bytes; }
int64_t thread_device_peak() { return t_peak; }
void thread_device_reset_peak() { t_peak = t_bytes; }

HostBuf::~HostBuf() {}
DevPool::~DevPool() {}
void* DevPool::take(size_t, size_t&) { return nullptr; }
void DevPool::give(void*, size_t) {}
void DevPool::trim() {}
DevPool* current_pool() { return t_pool; }
PoolScope::PoolScope(DevPool* p) : prev(t_pool) { t_pool = p; }
PoolScope::~PoolScope() { t_pool = prev; }
DevBuf::~DevBuf() {}
void DevBuf::release() {}
void* DevBuf::ensure(size_t) { return nullptr; }

Engine::Engine(const Problem& P, int form, int device) : P_(P), form_(form), device_(device) {
    KW_ = (P_.M + 15) / 16;
    const char* slow = std::getenv("STUB_SLOW_MODES");
    t_slow = slow && std::atoi(slow) == P_.M;
    t_bytes += 1 << 20;
    t_peak = std::max(t_peak, t_bytes);
}
Engine::~Engine() { t_bytes -= 1 << 20; }

void Engine::set_initial_basis() { n_ = P_.n0(); }
void Engine::get_basis(int32_t* q) const {
    for (int64_t i = 0; i < n_; i++)
        for (int m = 0; m < P_.M; m++) q[i * P_.M + m] = (int32_t)((i >> (m % 16)) & 3);
}
HStats Engine::build_h() {
    nap(200 + n_ % 300);
    t_bytes += n_ * 64;
    t_peak = std::max(t_peak, t_bytes);
    HStats st;
    st.n = n_;
    st.nnz_offdiag = n_ * 8;
    return st;
}
int64_t Engine::structural_nnz() { return n_ * 8; }
EigStats Engine::dense_all(std::vector<double>& E) {
    nap(300);
    E.assign(n_, 0.0);
    for (int64_t i = 0; i < n_; i++) E[i] = 1000.0 + 10.0 * (double)i + 0.5 * P_.M;
    xcols_ = (int)std::min<int64_t>(n_, std::max(1, P_.neig));
    xrows_ = n_;
    EigStats st;
    st.dense = true;
    return st;
}
EigStats Engine::eigsolve(int nev, double, int, std::vector<double>& E) {
    nap(500 + (n_ * 7) % 900);
    if (nev > n_) nev = (int)n_;
    E.assign(nev, 0.0);
    for (int i = 0; i < nev; i++) E[i] = 1000.0 + 10.0 * (double)i + 0.5 * P_.M - 1e-3 * (double)n_;
    xcols_ = nev;
    xrows_ = n_;
    EigStats st;
    st.spmm_vectors = 8 * nev;
    return st;
}
HbStats Engine::hb_select(const double*, int, double) {
    nap(150);
    // 4x, then 2x, then converged (at most 1 % new)
    const int64_t n0 = P_.n0();
    nnew_ = n_ == n0 ? 3 * n_ : (n_ < 8 * n0 ? n_ : n_ / 200);
    HbStats st;
    st.candidates = 10 * n_;
    st.unique = nnew_;
    return st;
}
void Engine::append_new_states() {
    n_ += nnew_;
    nnew_ = 0;
}
Pt2Stats Engine::pt2(const std::vector<double>&, int nopt, double eps2, std::vector<double>& dE) {
    nap(400);
    dE.assign(nopt, 0.0);
    for (int s = 0; s < nopt; s++) dE[s] = -1e-3 * (double)(s + 1) * (1.0 + eps2);
    Pt2Stats st;
    st.ns2 = 5 * n_;
    return st;
}
// not reached by the driver's loop without checkpoints
void Engine::get_vectors(double* C, int ncol) const {
    for (int64_t i = 0; i < n_ * ncol; i++) C[i] = 0.0;
}
void Engine::set_vectors(const double*, int) {}
void Engine::set_basis(const int32_t*, int64_t n) { n_ = n; }

}  // namespace vhci
