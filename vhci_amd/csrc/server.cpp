// server.cpp -- the serving loop (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>

namespace vhci {
namespace srv {

double predict_bytes(const Problem& P, const Tables& T, int form, double n) {
    const double K = 16.0 * ((P.M + 15) / 16);
    const double nev = std::max(1, P.neig);
    // stored entries per row: a fraction of the connection patterns find a
    // basis state (about 6 % at the config-1 field's final basis; 10 % here)
    const double row = std::min(0.1 * (double)T.deltas.size(), n);
    const double sv = 8.0;  // FP64 value, or FP32 hi + FP32 lo (both forms)
    (void)form;
    const double keys = K + 16.0 + 8.0 + 8.0;  // key, two hash slots, row pointer, diagonal
    const double hmat = row * (sv + 4.0) * 2.0;  // entries, plus the build's sort buffers
    const double m = 3.0 * nev;
    const double krylov = (2.0 * m + 8.0 + 2.0 * (m - nev / 2.0)) * 8.0 + nev * 8.0;
    const double pt2 = P.perturb ? 40.0 * (K + 24.0) : 0.0;  // external states per basis state
    return n * (keys + hmat + krylov + pt2);
}

Server::Server(int device, int slots, int64_t mem_limit)
    : device_(device), slots_(std::max(1, slots)), limit_((double)mem_limit), t0_(std::chrono::steady_clock::now()) {
    for (int s = 0; s < slots_; s++) threads_.emplace_back([this, s] { worker(s); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        for (int64_t id : queue_) {
            Request& r = *reqs_[id];
            r.state = State::Failed;
            r.err = "server stopped before the request was admitted";
            r.t_end = now();
        }
        queue_.clear();
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    for (auto& t : threads_) t.join();
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

bool Server::fits(const Request& r) const { return running_ == 0 || reserved_ + r.predicted <= limit_; }

// Admission: the first queued request that fits beside the running ones, so
// that a large request waiting for memory does not hold back smaller ones
// queued after it; but a request passed over kMaxPass times is not passed
// again (the scan stops there), so it runs once the running ones finish.
int Server::pick() const {
    for (size_t i = 0; i < queue_.size(); i++) {
        const Request& q = *reqs_.at(queue_[i]);
        if (fits(q)) return (int)i;
        if (q.passed >= kMaxPass) return -1;
    }
    return -1;
}

int64_t Server::submit(const std::string& text, int form, double tol) {
    auto r = std::make_unique<Request>();
    std::istringstream in(text);
    const std::string err = parse_input(in, false, r->P);
    if (!err.empty()) throw std::invalid_argument(err);
    if (form != 0 && form != 1) throw std::invalid_argument("unknown precision form");
    r->form = form;
    r->tol = tol > 0 ? tol : 1e-10;
    Tables T;
    build_tables(r->P, T);
    // first heat-bath iteration: assume 64-fold growth of the initial basis
    const double n0 = (double)r->P.n0();
    r->predicted = predict_bytes(r->P, T, form, r->P.eps1 > 0 ? 64.0 * n0 : n0);
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("server is stopping");
    r->id = next_id_++;
    r->t_submit = now();
    const int64_t id = r->id;
    reqs_[id] = std::move(r);
    queue_.push_back(id);
    cv_work_.notify_all();
    return id;
}

bool Server::wait(int64_t id, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw std::invalid_argument("unknown request id");
    Request* r = it->second.get();
    auto done = [&] { return r->state == State::Done || r->state == State::Failed; };
    if (timeout_s < 0) {
        cv_done_.wait(lk, done);
        return true;
    }
    return cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
}

Request Server::get(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw std::invalid_argument("unknown request id");
    return *it->second;
}

void Server::stats(double out[5]) {
    std::lock_guard<std::mutex> lk(mu_);
    out[0] = running_;
    out[1] = (double)queue_.size();
    out[2] = reserved_;
    out[3] = max_running_;
    out[4] = (double)finished_;
}

void Server::worker(int slot) {
    while (true) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] { return stop_ || pick() >= 0; });
        if (stop_) return;
        const int qi = pick();
        const int64_t id = queue_[qi];
        for (int i = 0; i < qi; i++) reqs_[queue_[i]]->passed++;
        queue_.erase(queue_.begin() + qi);
        Request& r = *reqs_[id];
        r.state = State::Running;
        r.slot = slot;
        r.t_start = now();
        r.reserved = r.predicted;
        reserved_ += r.reserved;
        running_++;
        max_running_ = std::max(max_running_, running_);
        // copies of what the solve needs, so the lock is not held meanwhile
        const Problem P = r.P;
        const int form = r.form;
        const double tol = r.tol;
        lk.unlock();

        drv::RunResult R;
        std::string err;
        int64_t peak = 0;
        uint64_t checksum = 0;
        const int64_t base = thread_device_bytes();
        thread_device_reset_peak();
        try {
            Tables T;
            build_tables(P, T);
            Engine eng(P, form, device_);
            eng.set_initial_basis();
            drv::RunOpts o;
            o.tol = tol;
            int64_t nprev = 0;
            o.on_grow = [&](int64_t n) {
                // the next iteration is assumed to grow the basis by the
                // ratio of the last one (at least 5 %, at most 64-fold)
                const double g = nprev > 0 ? std::min(64.0, std::max(1.05, (double)n / (double)nprev)) : 64.0;
                nprev = n;
                const double next = P.eps1 > 0 ? g * (double)n : (double)n;
                const double res = std::max((double)(thread_device_bytes() - base), predict_bytes(P, T, form, next));
                std::lock_guard<std::mutex> g2(mu_);
                reserved_ += res - r.reserved;
                r.reserved = res;
                cv_work_.notify_all();
            };
            drv::run(eng, P, o, R, drv::Logger{false});
            peak = thread_device_peak() - base;
            checksum = drv::basis_checksum(eng);
        } catch (const std::exception& e) {
            err = e.what();
        }

        lk.lock();
        r.R = std::move(R);
        r.err = err;
        r.peak = peak;
        r.checksum = checksum;
        r.state = err.empty() ? State::Done : State::Failed;
        r.t_end = now();
        reserved_ -= r.reserved;
        running_--;
        if (running_ == 0) reserved_ = 0;  // drop the rounding left by the updates
        finished_++;
        lk.unlock();
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

}  // namespace srv
}  // namespace vhci
