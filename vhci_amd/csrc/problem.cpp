// problem.cpp -- input parsing, force-constant preprocessing, device tables
// and the text output formats of the heat-bath VCI engine.
#include "problem.hpp"

#include <algorithm>
#include <charconv>
#include <cstring>
#include <functional>
#include <thread>
#include <numeric>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <istream>
#include <limits>
#include <map>
#include <ostream>
#include <sstream>

namespace vhci {

namespace {

// Fact (Core_functions.cpp:73-83): descending product in double
double factorial(int n) {
    double v = 1;
    for (; n > 0; --n) v *= n;
    return v;
}

// ScaleFC (Input_Reader.cpp:494-515): f *= 1/sqrt(2^p) / prod_j k_j!, in that
// operation order so the scaled values are bit-identical to the reference.
void scale(ForceConst& f) {
    double s = (double)f.fcpow.size();
    s = std::pow(2, s);
    s = 1 / std::sqrt(s);
    for (int k : f.mp) s /= factorial(k);
    f.fc *= s;
}

}  // namespace

std::string parse_input(std::istream& in, bool restart, Problem& P) {
    std::string tok;
    std::getline(in, tok);  // comment line
    in >> tok >> P.eps1;
    in >> tok >> P.neig;
    in >> tok >> P.perturb;
    in >> tok >> P.eps2;
    if (!P.perturb) P.eps2 = 0;  // Input_Reader.cpp:199-201
    in >> tok >> P.ntot;
    int nm = 0;
    in >> tok >> nm;
    if (!in || nm < 1) return "Error: could not read the input header.";
    if (nm > 64) return "Error: at most 64 modes are supported.";
    P.M = nm;
    for (int i = 0; i < nm; i++) {
        int id = -1;
        in >> id;
        if (id != i) {
            std::ostringstream os;
            os << "Error: Expected mode " << i << " but read data for mode " << id;
            return os.str();
        }
        double f = 0;
        int c = 0;
        in >> f >> c;
        P.freq.push_back(f);
        P.cap.push_back(c);
    }
    in >> tok >> P.nfc;
    for (int i = 0; i < P.nfc; i++) {
        ForceConst f;
        int p = 0;
        in >> p;
        for (int j = 0; j < p; j++) {
            int m = 0;
            in >> m;
            f.fcpow.push_back(m);
        }
        in >> f.fc;
        if (!in) return "Error: could not read force constant " + std::to_string(i) + ".";
        // ShortModes / ModePowers in first-seen order (Input_Reader.cpp:248-271)
        for (int m : f.fcpow) {
            auto it = std::find(f.sm.begin(), f.sm.end(), m);
            if (it == f.sm.end()) {
                f.sm.push_back(m);
                f.mp.push_back(1);
            } else {
                f.mp[it - f.sm.begin()] += 1;
            }
            if (m < 0 || m >= nm) return "Error: force constant " + std::to_string(i) + " names a mode outside 0..M-1.";
        }
        P.fcs.push_back(f);
    }
    for (auto& f : P.fcs) scale(f);
    for (int i = 0; i < (int)P.fcs.size(); i++) {
        switch (P.fcs[i].order()) {  // Input_Reader.cpp:273-291
            case 3: P.cubic.push_back(i); break;
            case 4: P.quartic.push_back(i); break;
            case 5: P.quintic.push_back(i); break;
            case 6: P.sextic.push_back(i); break;
            default: return "Error: Force constants are only supported up to 6th order!";
        }
    }
    P.restart = restart;
    const int M = P.M;
    // initial basis (Input_Reader.cpp:296-364): mode 0 varies fastest, the
    // last mode is the slowest key; total quanta <= Ntot, per-mode cap.
    P.basis0.clear();
    if (restart) {
        P.basis0.assign(M, 0);
    } else {
        std::vector<int32_t> cur;
        int lim0 = std::min(P.cap[0], P.ntot);
        for (int q = 0; q <= lim0; q++) cur.push_back(q);
        int width = 1;
        for (int m = 1; m < M; m++) {
            std::vector<int32_t> nxt;
            int lim = std::min(P.cap[m], P.ntot);
            int64_t ncur = (int64_t)cur.size() / width;
            for (int q = 0; q <= lim; q++)
                for (int64_t s = 0; s < ncur; s++) {
                    int tot = q;
                    for (int t = 0; t < width; t++) tot += cur[s * width + t];
                    if (tot <= P.ntot) {
                        nxt.insert(nxt.end(), cur.begin() + s * width, cur.begin() + (s + 1) * width);
                        nxt.push_back(q);
                    }
                }
            cur.swap(nxt);
            width++;
        }
        P.basis0 = cur;
    }
    // AnharmHB = AnharmFC + effective single and double terms
    // (Input_Reader.cpp:366-460), summed in list order, kept if |f| > 1e-8.
    P.hb = P.fcs;
    auto has = [](const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    for (int m = 0; m < M; m++) {
        ForceConst t;
        t.fcpow = {m};
        t.sm = {m};
        t.mp = {1};
        for (int id : P.cubic) {
            const ForceConst& c = P.fcs[id];
            if (c.sm.size() == 1 && c.sm[0] == m) t.fc += 3 * c.fc;
            if (c.sm.size() == 2) {
                if (c.sm[0] == m && c.mp[0] == 1)
                    t.fc += 2 * c.fc;
                else if (c.sm[1] == m && c.mp[1] == 1)
                    t.fc += 2 * c.fc;
            }
        }
        if (std::abs(t.fc) > 1e-8) P.hb.push_back(t);
    }
    for (int m = 0; m < M; m++) {
        ForceConst t;
        t.fcpow = {m, m};
        t.sm = {m};
        t.mp = {2};
        for (int id : P.quartic) {
            const ForceConst& c = P.fcs[id];
            if (c.sm.size() == 1 && c.sm[0] == m) t.fc += 4 * c.fc;
            if (c.sm.size() == 2) {
                if (c.sm[0] == m && c.mp[0] == 2)
                    t.fc += 2 * c.fc;
                else if (c.sm[1] == m && c.mp[1] == 2)
                    t.fc += 2 * c.fc;
            }
        }
        if (std::abs(t.fc) > 1e-8) P.hb.push_back(t);
    }
    // double terms: a quartic with a cubed mode adds 3 f to the pair of its
    // two modes, one with two single modes 2 f to the pair of those.  Each
    // quartic adds to its one pair, in list order, so every pair's sum has
    // the order of a scan of the list per pair.
    std::vector<double> pair_fc((size_t)M * M, 0.0);
    for (int id : P.quartic) {
        const ForceConst& c = P.fcs[id];
        if (c.sm.size() == 2 && has(c.mp, 3)) {
            const int a = std::min(c.sm[0], c.sm[1]), b = std::max(c.sm[0], c.sm[1]);
            pair_fc[(size_t)a * M + b] += 3 * c.fc;
        } else if (c.sm.size() == 3) {
            int ones[3], k = 0;
            for (int t = 0; t < 3; t++)
                if (c.mp[t] == 1) ones[k++] = c.sm[t];
            if (k == 2) {
                const int a = std::min(ones[0], ones[1]), b = std::max(ones[0], ones[1]);
                pair_fc[(size_t)a * M + b] += 2 * c.fc;
            }
        }
    }
    for (int m = 0; m + 1 < M; m++)
        for (int n = m + 1; n < M; n++) {
            ForceConst t;
            t.fcpow = {m, n};
            t.sm = {m, n};
            t.mp = {1, 1};
            t.fc = pair_fc[(size_t)m * M + n];
            if (std::abs(t.fc) > 1e-8) P.hb.push_back(t);
        }
    // HeatBath_Sort_FC (HB.cpp:17-22), stable so ties keep construction order
    P.hb_sorted.resize(P.hb.size());
    for (size_t i = 0; i < P.hb.size(); i++) P.hb_sorted[i] = (int)i;
    std::stable_sort(P.hb_sorted.begin(), P.hb_sorted.end(),
                     [&](int a, int b) { return std::abs(P.hb[a].fc) > std::abs(P.hb[b].fc); });
    P.fcmax = 0;
    for (auto& f : P.fcs) P.fcmax = std::max(P.fcmax, f.order());
    return "";
}

std::string load_input(const std::string& path, bool restart, Problem& P) {
    std::ifstream in(path);
    if (!in.good()) return "Error: Could not open input file.";
    return parse_input(in, restart, P);
}

double ladder_sum(int k, int qs, int qe) {
    // Expand the path list level by level: every entry first spawns its
    // creation branch (coefficient *= sqrt(q+1)), then, if q >= 1, its
    // annihilation branch (coefficient *= sqrt(q)).
    std::vector<int> q{qs}, nq;
    std::vector<double> c{1.0}, nc;
    for (int s = 0; s < k; s++) {
        nq.clear();
        nc.clear();
        for (size_t e = 0; e < q.size(); e++) {
            nq.push_back(q[e] + 1);
            nc.push_back(c[e] * std::sqrt(q[e] + 1));
            if (q[e] >= 1) {
                nq.push_back(q[e] - 1);
                nc.push_back(c[e] * std::sqrt(q[e]));
            }
        }
        q.swap(nq);
        c.swap(nc);
    }
    double sum = 0.0;
    for (size_t e = 0; e < q.size(); e++)
        if (q[e] == qe) sum += c[e];
    return sum;
}

void build_ladder(int qmax, std::vector<double>& L) {
    const int nq = qmax + 1;
    L.assign((size_t)7 * nq * kLadderW, 0.0);
    for (int k = 0; k <= 6; k++)
        for (int qs = 0; qs < nq; qs++)
            for (int d = -k; d <= k; d++) {
                int qe = qs + d;
                if (qe < 0 || ((d + k) & 1)) continue;
                L[((size_t)k * nq + qs) * kLadderW + (d + k)] = ladder_sum(k, qs, qe);
            }
}

void order_patterns(std::vector<DeltaEnt>& v, int M, std::vector<uint32_t>& off) {
    // block of a pattern: 0 when no delta is negative, else 1 + the lowest
    // mode with a negative delta
    auto block = [](const DeltaEnt& e) {
        const int nz = (int)(e.meta >> 28);
        int lo = -1;
        for (int t = 0; t < nz; t++)
            if (e.delta[t] < 0 && (lo < 0 || e.mode[t] < lo)) lo = e.mode[t];
        return lo + 1;
    };
    std::vector<std::vector<DeltaEnt>> b((size_t)M + 1);
    for (const DeltaEnt& e : v) b[block(e)].push_back(e);
    off.assign((size_t)M + 2, 0);
    v.clear();
    for (int k = 0; k <= M; k++) {
        off[k] = (uint32_t)v.size();
        v.insert(v.end(), b[k].begin(), b[k].end());
    }
    off[(size_t)M + 1] = (uint32_t)v.size();
}

void build_tables(const Problem& P, Tables& T) {
    T = Tables();
    // element FCs: Quartic, Sextic, Cubic, Quintic
    std::vector<int> order;
    for (auto* lst : {&P.quartic, &P.sextic, &P.cubic, &P.quintic}) order.insert(order.end(), lst->begin(), lst->end());
    for (int id : order) {
        const ForceConst& f = P.fcs[id];
        ElemFC e{};
        e.fc = f.fc;
        e.d = (uint8_t)f.sm.size();
        for (size_t t = 0; t < f.sm.size(); t++) {
            e.mode[t] = (uint8_t)f.sm[t];
            e.pow[t] = (uint8_t)f.mp[t];
        }
        T.elem.push_back(e);
    }
    // connection groups: the set of |delta| patterns every element FC can
    // bridge (ScreenState, Ham.cpp:401-446, is true exactly on these), with
    // each group's FCs in element order.  Odd patterns with qdiff > fcmax-1
    // fail the reference's odd prefilter (Ham.cpp:283) and are dropped.
    std::map<std::vector<std::pair<int, int>>, std::vector<uint32_t>> gm;
    for (uint32_t e = 0; e < T.elem.size(); e++) {
        const ElemFC& f = T.elem[e];
        int d = f.d;
        int a[6];
        for (int t = 0; t < d; t++) a[t] = f.pow[t] & 1;
        while (true) {
            std::vector<std::pair<int, int>> pat;
            int qdiff = 0;
            for (int t = 0; t < d; t++)
                if (a[t] > 0) {
                    pat.push_back({f.mode[t], a[t]});
                    qdiff += a[t];
                }
            std::sort(pat.begin(), pat.end());
            bool odd = qdiff & 1;
            if (pat.empty())
                T.diag_fcs.push_back(e);
            else if (!odd || qdiff <= P.fcmax - 1)
                gm[pat].push_back(e);
            int t = d - 1;
            while (t >= 0) {
                a[t] += 2;
                if (a[t] <= f.pow[t]) break;
                a[t] = f.pow[t] & 1;
                t--;
            }
            if (t < 0) break;
        }
    }
    uint32_t gid = 0;
    for (auto& kv : gm) {
        const auto& pat = kv.first;
        GroupEnt g{(uint32_t)T.fclist.size(), (uint32_t)kv.second.size()};
        T.fclist.insert(T.fclist.end(), kv.second.begin(), kv.second.end());
        T.groups.push_back(g);
        int nz = (int)pat.size();
        for (int s = 0; s < (1 << nz); s++) {
            DeltaEnt de{};
            for (int t = 0; t < nz; t++) {
                de.mode[t] = (uint8_t)pat[t].first;
                int v = pat[t].second;
                de.delta[t] = (int8_t)(((s >> t) & 1) ? -v : v);
                T.max_delta = std::max(T.max_delta, v);
            }
            de.meta = gid | ((uint32_t)nz << 28);
            T.deltas.push_back(de);
        }
        gid++;
    }
    // L1 ball of the pair prefilter, enumerated as (sorted modes, signed
    // deltas); counted first so an oversized ball is skipped
    {
        const int R = P.fcmax;
        const int M = P.M;
        // number of vectors with exactly k nonzero modes and L1 norm s:
        // C(M,k) * C(s-1,k-1) * 2^k
        auto binom = [](int nn, int kk) {
            if (kk < 0 || kk > nn) return 0.0;
            double r = 1;
            for (int i = 1; i <= kk; i++) r = r * (nn - kk + i) / i;
            return r;
        };
        double total = 0;
        for (int sN = 1; sN <= R; sN++) {
            if ((sN & 1) && sN > R - 1) continue;
            for (int k = 1; k <= std::min(sN, 6); k++) total += binom(M, k) * binom(sN - 1, k - 1) * std::pow(2.0, k);
        }
        if (R >= 1 && total <= (double)kMaxBall) {
            std::vector<int> modes, mags;
            std::function<void(int, int)> rec = [&](int start, int left) {
                const int sN = std::accumulate(mags.begin(), mags.end(), 0);
                if (!modes.empty() && (!(sN & 1) || sN <= R - 1)) {
                    const int nz = (int)modes.size();
                    for (int sg = 0; sg < (1 << nz); sg++) {
                        DeltaEnt de{};
                        for (int t = 0; t < nz; t++) {
                            de.mode[t] = (uint8_t)modes[t];
                            de.delta[t] = (int8_t)(((sg >> t) & 1) ? -mags[t] : mags[t]);
                        }
                        de.meta = (uint32_t)nz << 28;
                        T.ball.push_back(de);
                    }
                }
                if ((int)modes.size() == 6) return;
                for (int m = start; m < M; m++)
                    for (int v = 1; v <= left; v++) {
                        modes.push_back(m);
                        mags.push_back(v);
                        rec(m + 1, left - v);
                        modes.pop_back();
                        mags.pop_back();
                    }
            };
            rec(0, R);
        }
    }
    order_patterns(T.deltas, P.M, T.delta_off);
    order_patterns(T.ball, P.M, T.ball_off);
    // heat-bath variants in sorted-FC order (HB.cpp:27-164): delta_t runs
    // -k..k in steps of 2 with the last mode fastest; all-zero is skipped.
    T.hb_varoff.push_back(0);
    for (int idx : P.hb_sorted) {
        const ForceConst& f = P.hb[idx];
        T.hb_fc.push_back(f.fc);
        int d = (int)f.sm.size();
        int dl[6];
        for (int t = 0; t < d; t++) dl[t] = -f.mp[t];
        int64_t cnt = 0;
        while (d > 0) {
            bool zero = true;
            for (int t = 0; t < d; t++) zero = zero && dl[t] == 0;
            if (!zero) {
                HbVar v{};
                v.d = (uint8_t)d;
                for (int t = 0; t < d; t++) {
                    v.mode[t] = (uint8_t)f.sm[t];
                    v.delta[t] = (int8_t)dl[t];
                    T.max_hb_delta = std::max(T.max_hb_delta, std::abs(dl[t]));
                }
                T.hb_vars.push_back(v);
                cnt++;
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
        T.hb_varoff.push_back(T.hb_varoff.back() + cnt);
    }

}

void write_freqs(std::ostream& os, const Problem& P, const std::vector<double>& E,
                 const std::vector<FreqColumn>& cols) {
    const int M = P.M;
    const int nf = std::min<int>(std::min<int>(P.neig, (int)E.size()), (int)cols.size());
    auto labels = [&](int col, const char* sep) {
        const FreqColumn& fc = cols[col];
        for (size_t t = 0; t < fc.row.size(); t++) {
            if (t > 0) os << sep;
            os << std::fixed << std::setprecision(2) << fc.absval[t] << " (";
            int nm = 0;
            for (int l = 0; l < M; l++) {
                const int q = fc.quanta[t][l];
                if (q > 0) os << (nm++ ? "+" : "") << q << "w" << l;
            }
            os << ")";
        }
    };
    // line 0: ZPE, assignments written back to back (Ham.cpp:454-472)
    os << std::fixed << std::setprecision(10) << E[0] << "   ";
    labels(0, "");
    os << '\n';
    for (int n = 1; n < nf; n++) {
        os << std::fixed << std::setprecision(12) << E[n] - E[0] << "   ";
        labels(n, "  ");
        os << '\n';
    }
    os.flush();
}

namespace {

// run f(t, lo, hi) over [0, n) split into `threads` contiguous ranges
void parallel_ranges(int64_t n, int threads, const std::function<void(int, int64_t, int64_t)>& f) {
    threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n / 2048 + 1));
    if (threads == 1) {
        f(0, 0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back(f, t, n * t / threads, n * (t + 1) / threads);
    for (auto& th : pool) th.join();
}

// "%.17f" as std::fixed << std::setprecision(17) writes it (to_chars with a
// precision follows printf, in the C locale)
inline char* put_fixed17(char* p, char* end, double v) {
    return std::to_chars(p, end, v, std::chars_format::fixed, 17).ptr;
}

template <class T>
void put_raw(std::ostream& os, const T* p, size_t count) {
    os.write(reinterpret_cast<const char*>(p), (std::streamsize)(sizeof(T) * count));
}

// largest fixed 17-decimal text of a double: sign, 309 integer digits, point, 17 decimals
constexpr size_t kFixedMax = 1 + 309 + 1 + 17;

}  // namespace

void write_checkpoint(std::ostream& os, const Problem& P, const int32_t* basis, int64_t n,
                      const std::vector<double>& E, const double* C, int ncol, int threads, bool binary) {
    const int M = P.M;
    if (binary) {
        os.write(kCheckpointMagic, sizeof(kCheckpointMagic));
        const int64_t head[4] = {n, M, P.neig, (int64_t)E.size()};
        put_raw(os, head, 4);
        put_raw(os, P.freq.data(), (size_t)M);
        put_raw(os, E.data(), E.size());
        const int64_t cols = ncol;
        put_raw(os, &cols, 1);
        std::vector<uint8_t> q((size_t)n * M);
        for (size_t i = 0; i < q.size(); i++) q[i] = (uint8_t)basis[i];
        put_raw(os, q.data(), q.size());
        put_raw(os, C, (size_t)n * ncol);
        os.flush();
        return;
    }
    {
        std::ostringstream h;
        h << n << " " << M << " " << P.neig << '\n';
        os << h.str();
        std::string line;
        char buf[kFixedMax + 1];
        for (int j = 0; j < M; j++) {
            line.append(buf, put_fixed17(buf, buf + kFixedMax, P.freq[j]));
            line += ' ';
        }
        for (double e : E) {
            line.append(buf, put_fixed17(buf, buf + kFixedMax, e));
            line += ' ';
        }
        line += '\n';
        os << line;
    }
    // rows: "q_0 .. q_(M-1) C_0 .. C_(ncol-1) \n", formatted in parallel ranges
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n / 2048 + 1));
    std::vector<std::string> out(T);
    parallel_ranges(n, T, [&](int t, int64_t lo, int64_t hi) {
        std::string& s = out[t];
        s.reserve((size_t)(hi - lo) * (M * 4 + ncol * 21 + 1));
        char buf[kFixedMax + 1];
        for (int64_t i = lo; i < hi; i++) {
            for (int j = 0; j < M; j++) {
                s.append(buf, std::to_chars(buf, buf + sizeof(buf), basis[i * M + j]).ptr);
                s += ' ';
            }
            for (int j = 0; j < ncol; j++) {
                s.append(buf, put_fixed17(buf, buf + kFixedMax, C[i * ncol + j]));
                s += ' ';
            }
            s += '\n';
        }
    });
    for (auto& s : out) os.write(s.data(), (std::streamsize)s.size());
    os.flush();
}

namespace {

const char* kTruncated = "Error: checkpoint file is truncated.";

// whitespace-separated number at p (the reference's operator>> reading)
template <class T>
bool next_num(const char*& p, const char* end, T& v) {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) p++;
    if (p >= end) return false;
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
}

template <class T>
bool get_raw(const char*& p, const char* end, T* v, size_t count) {
    const size_t b = sizeof(T) * count;
    if ((size_t)(end - p) < b) return false;
    std::memcpy(v, p, b);
    p += b;
    return true;
}

std::string read_binary(const char* p, const char* end, const Problem& P, std::vector<int32_t>& basis,
                        std::vector<double>& E, std::vector<double>& C) {
    p += sizeof(kCheckpointMagic);
    int64_t head[4];
    if (!get_raw(p, end, head, 4)) return kTruncated;
    const int64_t nb = head[0];
    if (head[1] != P.M) return "Error: Checkpoint file contains wrong number of fundamental modes!";
    if (head[2] != P.neig) return "Error: Checkpoint file contains wrong number of eigenstates!";
    std::vector<double> f((size_t)P.M);
    if (!get_raw(p, end, f.data(), f.size())) return kTruncated;
    for (int i = 0; i < P.M; i++)
        if (f[i] != P.freq[i]) return "Error: Checkpoint file contains wrong fundamental frequencies!";
    if (head[3] < P.neig || nb < 1) return kTruncated;
    std::vector<double> ew((size_t)head[3]);
    if (!get_raw(p, end, ew.data(), ew.size())) return kTruncated;
    E.assign(ew.begin(), ew.begin() + P.neig);
    int64_t cols = 0;
    if (!get_raw(p, end, &cols, 1) || cols < P.neig) return kTruncated;
    std::vector<uint8_t> q((size_t)nb * P.M);
    if (!get_raw(p, end, q.data(), q.size())) return kTruncated;
    basis.assign(q.begin(), q.end());
    std::vector<double> Cw((size_t)nb * cols);
    if (!get_raw(p, end, Cw.data(), Cw.size())) return kTruncated;
    C.assign((size_t)nb * P.neig, 0.0);
    for (int64_t i = 0; i < nb; i++)
        std::copy(Cw.begin() + i * cols, Cw.begin() + i * cols + P.neig, C.begin() + i * P.neig);
    return "";
}

}  // namespace

std::string read_checkpoint(std::istream& is, const Problem& P, std::vector<int32_t>& basis,
                            std::vector<double>& E, std::vector<double>& C, int threads) {
    std::string text;
    {
        // whole file in one read where the stream can seek, else in blocks
        const std::streampos at = is.tellg();
        is.seekg(0, std::ios::end);
        const std::streampos len = is.tellg();
        if (at != std::streampos(-1) && len != std::streampos(-1) && len >= at) {
            is.seekg(at);
            text.resize((size_t)(len - at));
            is.read(&text[0], (std::streamsize)text.size());
            text.resize((size_t)is.gcount());
        } else {
            is.clear();
            char buf[1 << 16];
            while (is.read(buf, sizeof(buf)) || is.gcount() > 0) text.append(buf, (size_t)is.gcount());
        }
    }
    const char* p = text.data();
    const char* end = p + text.size();
    if (text.size() >= sizeof(kCheckpointMagic) && std::memcmp(p, kCheckpointMagic, sizeof(kCheckpointMagic)) == 0)
        return read_binary(p, end, P, basis, E, C);
    int64_t nb = 0;
    int par = 0;
    if (!next_num(p, end, nb) || !next_num(p, end, par)) return kTruncated;
    if (par != P.M) return "Error: Checkpoint file contains wrong number of fundamental modes!";
    if (!next_num(p, end, par)) return kTruncated;
    if (par != P.neig) return "Error: Checkpoint file contains wrong number of eigenstates!";
    for (int i = 0; i < P.M; i++) {
        double f = 0;
        if (!next_num(p, end, f)) return kTruncated;
        if (f != P.freq[i]) return "Error: Checkpoint file contains wrong fundamental frequencies!";
    }
    E.assign(P.neig, 0.0);
    for (int i = 0; i < P.neig; i++)
        if (!next_num(p, end, E[i])) return kTruncated;
    if (nb < 1) return kTruncated;
    basis.assign((size_t)nb * P.M, 0);
    C.assign((size_t)nb * P.neig, 0.0);
    // Rows are read one line per basis state when the energy line ends after
    // NEig values (the writer's output: Freqs.size() = NEig); otherwise the
    // values are taken one after another, as the reference's stream reader
    // does.
    const char* eol = p;
    while (eol < end && (*eol == ' ' || *eol == '\t' || *eol == '\r')) eol++;
    if (eol >= end || *eol != '\n') {
        for (int64_t i = 0; i < nb; i++) {
            for (int j = 0; j < P.M; j++)
                if (!next_num(p, end, basis[i * P.M + j])) return kTruncated;
            for (int j = 0; j < P.neig; j++)
                if (!next_num(p, end, C[i * P.neig + j])) return kTruncated;
        }
        return "";
    }
    p = eol;
    // one line per basis state: find the line starts, then parse in ranges
    std::vector<const char*> starts;
    starts.reserve((size_t)nb + 1);
    for (const char* s = p; s < end && (int64_t)starts.size() < nb;) {
        const char* nl = (const char*)std::memchr(s, '\n', (size_t)(end - s));
        if (!nl) break;
        if (nl + 1 < end) starts.push_back(nl + 1);
        s = nl + 1;
    }
    if ((int64_t)starts.size() < nb) return kTruncated;
    starts.push_back(end);
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(threads, nb / 2048 + 1));
    std::vector<char> ok(T, 1);
    parallel_ranges(nb, T, [&](int t, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi && ok[t]; i++) {
            const char* q = starts[i];
            const char* qe = (i + 1 < nb) ? starts[i + 1] : end;
            for (int j = 0; j < P.M && ok[t]; j++) ok[t] = next_num(q, qe, basis[i * P.M + j]);
            for (int j = 0; j < P.neig && ok[t]; j++) ok[t] = next_num(q, qe, C[i * P.neig + j]);
        }
    });
    for (char o : ok)
        if (!o) return kTruncated;
    return "";
}

}  // namespace vhci
