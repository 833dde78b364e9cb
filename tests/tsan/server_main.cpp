// server_main.cpp -- drives the serving loop (vhci::srv::Server) from two
// client threads under ThreadSanitizer: 40 requests of different sizes are
// submitted while others run on four worker slots; a third thread polls the
// statistics; every request is waited for and checked against a run of the
// same input alone.  Exit status 0 when every result matches.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "server.hpp"

namespace {
// an input of the reference's format with M modes, eps1 1, PT2 on
std::string sized(int M, int tag) {
    std::ostringstream os;
    os << "sized " << M << " " << tag << "\n";
    os << "HCI_Eps: 1.0\nNstate: 4\nPT2: 1\nPT2_Eps: 0.01\nMax_quanta: 2\n";
    os << "Modes: " << M << "\n";
    for (int m = 0; m < M; m++) os << m << " " << 1000 + 100 * m << " 6\n";
    os << "Force_constants: 1\n";
    os << "3 0 0 0 " << 10.0 + tag << "\n";
    return os.str();
}

// device bytes the server predicts for an input when its basis has
// `growth` times the initial states
double predicted(const std::string& text, double growth) {
    vhci::Problem P;
    std::istringstream in(text);
    if (!vhci::parse_input(in, false, P).empty()) return -1;
    vhci::Tables T;
    vhci::build_tables(P, T);
    return vhci::srv::predict_bytes(P, T, 0, growth * (double)P.n0());
}

// a small input of the reference's format: M modes, eps1 > 0, PT2 on
std::string request(int r) {
    const int M = 3 + r % 5;
    std::ostringstream os;
    os << "request " << r << "\n";
    os << "HCI_Eps: " << (r % 3 == 0 ? 0.5 : 1.0) << "\n";
    os << "Nstate: " << 4 + r % 7 << "\n";
    os << "PT2: 1\n";
    os << "PT2_Eps: 0.01\n";
    os << "Max_quanta: 2\n";
    os << "Modes: " << M << "\n";
    for (int m = 0; m < M; m++) os << m << " " << 1000 + 100 * m << " 6\n";
    os << "Force_constants: 1\n";
    os << "3 0 0 0 " << 10.0 + r << "\n";
    return os.str();
}
}  // namespace

int main() {
    using vhci::srv::Server;
    using vhci::srv::State;
    const int nreq = 40;
    std::vector<double> alone_e(nreq);
    std::vector<uint64_t> alone_ck(nreq);
    {
        Server one(0, 1, int64_t(1) << 40);
        for (int r = 0; r < nreq; r++) {
            const int64_t id = one.submit(request(r), 0, 0.0);
            one.wait(id, -1.0);
            const auto q = one.get(id);
            if (q.state != State::Done) {
                std::printf("request %d failed alone: %s\n", r, q.err.c_str());
                return 1;
            }
            alone_e[r] = q.R.E[0];
            alone_ck[r] = q.checksum;
        }
    }
    Server srv(0, 4, int64_t(1) << 40);
    std::vector<int64_t> ids(nreq, 0);
    std::thread a([&] {
        for (int r = 0; r < nreq; r += 2) ids[r] = srv.submit(request(r), 0, 0.0);
    });
    std::thread b([&] {
        for (int r = 1; r < nreq; r += 2) ids[r] = srv.submit(request(r), 0, 0.0);
    });
    bool stop = false;
    std::thread poll([&] {
        double st[5];
        while (!__atomic_load_n(&stop, __ATOMIC_ACQUIRE)) {
            srv.stats(st);
            std::this_thread::yield();
        }
    });
    a.join();
    b.join();
    int bad = 0;
    for (int r = 0; r < nreq; r++) {
        srv.wait(ids[r], -1.0);
        const auto q = srv.get(ids[r]);
        if (q.state != State::Done || q.R.E.empty() || q.R.E[0] != alone_e[r] || q.checksum != alone_ck[r]) {
            std::printf("request %d differs in the mix (%s)\n", r, q.err.c_str());
            bad++;
        }
    }
    __atomic_store_n(&stop, true, __ATOMIC_RELEASE);
    poll.join();
    double st[5];
    srv.stats(st);
    std::printf("served %d requests, at most %d at once, %d differing from alone\n", nreq, (int)st[3], bad);
    if (bad != 0 || st[3] <= 1) return 1;

    // Admission: A (5 modes, slow) runs; B is too large to fit beside it and
    // waits; C, small, is queued after B and must start while A runs instead
    // of waiting behind B.  The memory limit is B's prediction, so B fits
    // alone, and at least A's largest reservation plus C's prediction.
    const std::string A = sized(5, 1), C = sized(3, 3);
    const double a_max = 1.5 * predicted(A, 256.0), c_pred = predicted(C, 64.0);
    std::string B;
    double b_pred = 0;
    for (int M = 8; M <= 40 && b_pred < a_max + c_pred; M++) {
        B = sized(M, 2);
        b_pred = predicted(B, 64.0);
    }
    if (!(b_pred >= a_max + c_pred)) {
        std::printf("admission: no large request found\n");
        return 1;
    }
    setenv("STUB_SLOW_MODES", "5", 1);
    Server adm(0, 2, (int64_t)std::ceil(b_pred));
    const int64_t ia = adm.submit(A, 0, 0.0);
    for (int t = 0; t < 20000 && adm.get(ia).state == State::Queued; t++)
        std::this_thread::sleep_for(std::chrono::microseconds(500));
    const int64_t ib = adm.submit(B, 0, 0.0);
    const int64_t ic = adm.submit(C, 0, 0.0);
    for (int64_t id : {ia, ib, ic}) adm.wait(id, -1.0);
    const auto qa = adm.get(ia), qb = adm.get(ib), qc = adm.get(ic);
    const bool done = qa.state == State::Done && qb.state == State::Done && qc.state == State::Done;
    const bool early = qc.t_start < qb.t_start && qc.t_start < qa.t_end;
    std::printf("admission: small request started %.3f s before the large one, while the first ran: %s; "
                "large request passed over %d time(s); all done: %s\n",
                qb.t_start - qc.t_start, early ? "yes" : "no", qb.passed, done ? "yes" : "no");
    return done && early && qb.passed == 1 ? 0 : 1;
}
