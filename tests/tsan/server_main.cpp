// server_main.cpp -- drives the serving loop (vhci::srv::Server) from two
// client threads under ThreadSanitizer: 40 requests of different sizes are
// submitted while others run on four worker slots; a third thread polls the
// statistics; every request is waited for and checked against a run of the
// same input alone.  Exit status 0 when every result matches.
#include <cstdio>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "server.hpp"

namespace {
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
    return bad == 0 && st[3] > 1 ? 0 : 1;
}
