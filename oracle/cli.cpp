/*
 * cli.cpp -- command line of the CPU oracle (vhci_cpu).
 *
 * TEST INFRASTRUCTURE ONLY: the oracle's own driver, used by the tests and as
 * the CPU baseline of the benchmark.  Flags follow the reference program
 * (Input_Reader.cpp:17-168): -n threads, -i input, -o freqs, -c checkpoint;
 * extra: --json FILE (summary), --basis FILE (final basis), --quiet.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "vhci_oracle.h"

int main(int argc, char** argv) {
    const char *in = nullptr, *out = nullptr, *ck = nullptr, *js = nullptr, *bs = nullptr;
    int n = 0, quiet = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--quiet") {
            quiet = 1;
            continue;
        }
        if (i + 1 >= argc) {
            std::fprintf(stderr, "missing value for %s\n", a.c_str());
            return 2;
        }
        const char* v = argv[++i];
        if (a == "-n") n = std::atoi(v);
        else if (a == "-i") in = v;
        else if (a == "-o") out = v;
        else if (a == "-c") ck = v;
        else if (a == "--json") js = v;
        else if (a == "--basis") bs = v;
        else {
            std::fprintf(stderr, "unknown flag %s\n", a.c_str());
            return 2;
        }
    }
    if (!in || !out) {
        std::fprintf(stderr, "usage: vhci_cpu -n N -i input -o freqs [-c ckpt] [--json f] [--basis f] [--quiet]\n");
        return 2;
    }
    if (n < 1) n = 1;
    int rc = orc_run(in, out, ck, n, js, bs, quiet);
    if (rc) {
        std::printf("%s\n", orc_error());
        return 1;
    }
    return 0;
}
