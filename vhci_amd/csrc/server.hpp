// server.hpp -- the serving loop: solve requests of different sizes arrive
// at any time, wait in a FIFO queue, and are admitted to one of `slots`
// worker threads when a slot is free and their predicted device memory fits
// under the limit.  Each admitted request owns a device context (its basis
// hash, H and Ritz block stay resident across its heat-bath iterations) on
// its own HIP stream, so the work of concurrent requests overlaps on the
// GPU.  A request's reservation starts from a size prediction and is
// revised after every basis growth; new requests are admitted against the
// sum of the current reservations.
//
// There is no counterpart in the reference, which solves one input per
// process (src/VHCI.cpp:39-137); each request runs the same loop as the
// `vhci` program (src/HB.cpp:176-265, src/PT2.cpp:16-154).
#pragma once
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "driver.hpp"
#include "problem.hpp"

namespace vhci {
namespace srv {

enum class State { Queued, Running, Done, Failed };

struct Request {
    int64_t id = 0;
    Problem P;
    int form = 0;
    double tol = 1e-10;
    State state = State::Queued;
    std::string err;
    drv::RunResult R;
    double predicted = 0;  // predicted device bytes at submission
    double reserved = 0;   // current reservation (revised while running)
    int64_t peak = 0;      // measured peak device bytes
    uint64_t checksum = 0; // basis checksum of the final variational basis
    double t_submit = 0, t_start = 0, t_end = 0;  // seconds since server start
    int slot = -1;
    int passed = 0;  // requests queued after it that were admitted before it
};

// Device bytes predicted for a solve whose basis has n states (the state
// keys and hash, the Hamiltonian rows, the eigensolver's Krylov blocks, the
// Ritz block and the external space of PT2).
double predict_bytes(const Problem& P, const Tables& T, int form, double n);

class Server {
  public:
    Server(int device, int slots, int64_t mem_limit);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // parses the input (throws on a bad input) and queues the request
    int64_t submit(const std::string& text, int form, double tol);
    // true when the request has finished (done or failed) within timeout_s
    // (timeout_s < 0: wait without limit)
    bool wait(int64_t id, double timeout_s);
    // copy of a request's record (throws for an unknown id)
    Request get(int64_t id);
    // running, queued, reserved bytes, max concurrent so far, finished
    void stats(double out[5]);

    // a queued request is passed over by later ones that fit at most this
    // many times; after that the requests behind it wait for it
    static constexpr int kMaxPass = 16;

  private:
    void worker(int slot);
    double now() const;
    bool fits(const Request& r) const;
    // position in the queue of the request to admit next, or -1 (mu_ held)
    int pick() const;

    int device_;
    int slots_;
    double limit_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<int64_t> queue_;
    std::map<int64_t, std::unique_ptr<Request>> reqs_;
    double reserved_ = 0;
    int running_ = 0;
    int max_running_ = 0;
    int64_t finished_ = 0;
    int64_t next_id_ = 1;
    bool stop_ = false;
    std::vector<std::thread> threads_;
};

}  // namespace srv
}  // namespace vhci
