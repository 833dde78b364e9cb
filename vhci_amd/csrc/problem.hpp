// problem.hpp -- host-side problem model of the heat-bath VCI engine.
//
// Holds what the reference keeps in process globals (VCI_headers.h:147-160):
// frequencies, force constants split by order, the heat-bath list, the
// initial basis and the run settings, plus the tables the device kernels use
// (connection patterns, per-mode ladder tables, heat-bath variants).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace vhci {

// One force constant in the reference's FConst layout (VCI_headers.h:121-129):
// fcpow lists the mode of every power; sm/mp are the distinct modes in
// first-seen order and their powers.
struct ForceConst {
    double fc = 0.0;
    std::vector<int> fcpow;
    std::vector<int> sm;
    std::vector<int> mp;
    int order() const { return (int)fcpow.size(); }
};

struct Problem {
    double eps1 = 0.0, eps2 = 0.0;
    int neig = 0;
    bool perturb = false;
    int ntot = 0;
    int M = 0;
    int nfc = 0;
    bool restart = false;
    std::vector<double> freq;
    std::vector<int> cap;
    std::vector<ForceConst> fcs;     // AnharmFC, file order, scaled
    std::vector<int> cubic, quartic, quintic, sextic;  // indices into fcs
    std::vector<ForceConst> hb;      // AnharmHB in construction order
    std::vector<int> hb_sorted;      // indices into hb, |fc| descending, stable
    std::vector<int32_t> basis0;     // initial basis, N0 x M
    int fcmax = 0;
    int64_t n0() const { return M ? (int64_t)basis0.size() / M : 0; }
};

// Parse an input file with the reference's token order (Input_Reader.cpp:170-492).
// Returns an empty string on success, else the reference's error message.
std::string parse_input(std::istream& in, bool restart, Problem& P);
std::string load_input(const std::string& path, bool restart, Problem& P);

// Sum over ladder paths of (a + a^dag)^k from quanta qs to qe, enumerated and
// accumulated in the reference's order (Ham.cpp:16-95).
double ladder_sum(int k, int qs, int qe);

// ---------------------------------------------------------------------------
// Device-side tables.
//
// Element FCs: the four order lists concatenated as Quartic, Sextic, Cubic,
// Quintic, so that every connection group's FC list is ascending in this
// numbering and matches the reference's summation order (Ham.cpp:251-313).
struct ElemFC {
    double fc;
    uint8_t d;
    uint8_t mode[6];
    uint8_t pow[6];
    uint8_t pad[3];
};
static_assert(sizeof(ElemFC) == 24, "ElemFC layout");

// One signed connection pattern: up to 6 modes change by delta.  meta packs
// the group id (low 26 bits) and the number of changed modes (bits 28..30).
struct DeltaEnt {
    uint8_t mode[6];
    int8_t delta[6];
    uint32_t meta;
};
static_assert(sizeof(DeltaEnt) == 16, "DeltaEnt layout");

struct GroupEnt {
    uint32_t fc_start;
    uint32_t fc_count;
};

// One heat-bath candidate displacement of one HB force constant, in the
// reference's enumeration order (HB.cpp:27-164).
struct HbVar {
    uint8_t d;
    uint8_t mode[6];
    int8_t delta[6];
    uint8_t pad[3];
};
static_assert(sizeof(HbVar) == 16, "HbVar layout");

struct Tables {
    std::vector<ElemFC> elem;          // element FC table
    std::vector<DeltaEnt> deltas;      // signed patterns, grouped
    std::vector<GroupEnt> groups;      // per group: range in fclist
    std::vector<uint32_t> fclist;      // element FC ids per group
    std::vector<uint32_t> diag_fcs;    // element FC ids of the zero pattern
    // heat bath: sorted FC values and the flattened variant table
    std::vector<double> hb_fc;         // |fc| descending (signed values)
    std::vector<int64_t> hb_varoff;    // prefix of variant counts, size nhb+1
    std::vector<HbVar> hb_vars;
    int max_hb_delta = 0;              // largest |delta| of any HB variant
    // every nonzero displacement with even L1 norm <= fcmax or odd L1 norm
    // <= fcmax-1 (the reference's pair prefilter); empty when it would
    // exceed kMaxBall entries
    std::vector<DeltaEnt> ball;
    int max_delta = 0;                 // largest |delta| of any pattern
    // deltas and ball are ordered in blocks: block 0 holds the patterns that
    // lower no mode, block 1 + m those whose lowest lowered mode is m (a
    // row can use them only when it has quanta in mode m).  Block b is
    // [off[b], off[b + 1]); M + 2 offsets each.
    std::vector<uint32_t> delta_off, ball_off;
};

// Coefficient of mode m in the linear key sum of the device hash
// (sum_m q_m * key_coef(m) mod 2^64, then a 64-bit finalizer): pseudo-random
// and odd (splitmix64 of m), so that distinct small-integer quanta vectors
// give distinct sums except with probability about 2^-64.  The sum is linear,
// so the hash of a key moved by a pattern is that of the sum plus
// sum_t delta_t * key_coef(mode_t): no rehash of the whole key.
#if defined(__HIP__)
__host__ __device__
#endif
inline constexpr uint64_t key_coef(int m) {
    uint64_t z = 0x56484349ull * 0x9E3779B97F4A7C15ull + (uint64_t)(m + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (z ^ (z >> 31)) | 1ull;
}

// Stable reorder of patterns into the blocks described above; off receives
// the M + 2 block offsets.
void order_patterns(std::vector<DeltaEnt>& v, int M, std::vector<uint32_t>& off);

void build_tables(const Problem& P, Tables& T);
constexpr size_t kMaxBall = 500000;

// Ladder table T[k][qs][qe-qs+k], k = 0..6, qs = 0..qmax, 13 entries per row.
void build_ladder(int qmax, std::vector<double>& L);
constexpr int kLadderW = 13;

// Output files -----------------------------------------------------------
// One eigenvector's assignment for PrintFreqs: the basis rows with
// |C| >= max |C| / 2 in ascending order, their |C| and their quanta.
struct FreqColumn {
    std::vector<int64_t> row;
    std::vector<double> absval;
    std::vector<std::vector<int32_t>> quanta;
};
// PrintFreqs (Ham.cpp:448-502) from the per-eigenvector assignments
void write_freqs(std::ostream& os, const Problem& P, const std::vector<double>& E,
                 const std::vector<FreqColumn>& cols);
// WriteCheckpoint (Core_functions.cpp:94-117): the reference's text, byte for
// byte (fixed, 17 decimals), with the rows formatted on `threads` host
// threads.  binary = true writes the opt-in binary form instead (same
// content: header, frequencies, energies, quanta as bytes, coefficients as
// raw doubles).
void write_checkpoint(std::ostream& os, const Problem& P, const int32_t* basis, int64_t n,
                      const std::vector<double>& E, const double* C, int ncol, int threads = 1,
                      bool binary = false);
// ReadCheckpoint (Core_functions.cpp:119-167) of either form (the binary one
// is recognised by its first bytes); returns "" or the error text
std::string read_checkpoint(std::istream& is, const Problem& P, std::vector<int32_t>& basis,
                            std::vector<double>& E, std::vector<double>& C, int threads = 1);
// first bytes of a binary checkpoint
constexpr char kCheckpointMagic[8] = {'V', 'H', 'C', 'I', 'C', 'K', 'B', '1'};

}  // namespace vhci
