// engine.hip -- gfx950 kernels and device orchestration of the heat-bath VCI
// engine: basis hash, Hamiltonian build, SpMM, block-Davidson eigensolver,
// heat-bath selection and Epstein-Nesbet PT2.
//
// Compiled with -ffp-contract=off: every floating-point expression that must
// reproduce the reference's rounding (matrix elements, diagonal energies,
// selection thresholds) is written as separate multiplies and adds; fused
// multiply-adds are requested explicitly with fma() where rounding is free.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>

#include "engine.hpp"

namespace vhci {
static double* noalloc_scratch(int device, std::mutex** lock);
}

namespace vhci {

namespace {

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +  \
                                     __FILE__ + ":" + std::to_string(__LINE__));                    \
    } while (0)
#define BLASCHK(x)                                                                                   \
    do {                                                                                             \
        rocblas_status s_ = (x);                                                                     \
        if (s_ != rocblas_status_success)                                                            \
            throw std::runtime_error(std::string("rocBLAS/rocSOLVER status ") + std::to_string(s_) + \
                                     " at " + __FILE__ + ":" + std::to_string(__LINE__));           \
    } while (0)

constexpr uint64_t kEmpty = ~0ull;
constexpr int kWave = 64;
// entries of padding after the last stored entry of col/val/hi/lo, so that a
// 16-byte load of 4 entries that starts inside the matrix stays in bounds
constexpr int64_t kCsrPad = 4;
typedef int32_t sp_i4 __attribute__((ext_vector_type(4)));
typedef double sp_d2 __attribute__((ext_vector_type(2)));
typedef float sp_f4 __attribute__((ext_vector_type(4)));

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (1ll << 30)) g = 1ll << 30;
    return (unsigned)g;
}

uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// ------------------------------------------------------------------------
// key helpers: a key is NW 32-bit words (4 uint8 quanta per word)
// ------------------------------------------------------------------------
__device__ __forceinline__ uint64_t fmix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// Key hash: fmix64 of the linear sum  sum_m q_m * key_coef(m)  (mod 2^64).
// The sum is linear in the quanta, so the hash of a key moved by a
// connection pattern is fmix64(sum + the pattern's shift): the kernels that
// probe many patterns per state add a precomputed shift instead of
// rehashing the whole key.
template <int NW>
__device__ __forceinline__ uint64_t key_lin(const uint32_t (&w)[NW]) {
    uint64_t h = 0;
#pragma unroll
    for (int i = 0; i < NW; i++)
#pragma unroll
        for (int b = 0; b < 4; b++) h += (uint64_t)((w[i] >> (8 * b)) & 0xffu) * key_coef(4 * i + b);
    return h;
}

template <int NW>
__device__ __forceinline__ uint64_t key_hash(const uint32_t (&w)[NW]) {
    return fmix64(key_lin<NW>(w));
}

template <int NW>
__device__ __forceinline__ void key_load(const uint4* __restrict__ keys, int64_t i, uint32_t (&w)[NW]) {
    const uint4* k = keys + i * (NW / 4);
#pragma unroll
    for (int t = 0; t < NW / 4; t++) {
        uint4 v = k[t];
        w[4 * t + 0] = v.x;
        w[4 * t + 1] = v.y;
        w[4 * t + 2] = v.z;
        w[4 * t + 3] = v.w;
    }
}

template <int NW>
__device__ __forceinline__ bool key_eq(const uint4* __restrict__ keys, uint32_t idx, const uint32_t (&w)[NW]) {
    const uint4* k = keys + (size_t)idx * (NW / 4);
    bool eq = true;
#pragma unroll
    for (int t = 0; t < NW / 4; t++) {
        uint4 v = k[t];
        eq = eq && v.x == w[4 * t] && v.y == w[4 * t + 1] && v.z == w[4 * t + 2] && v.w == w[4 * t + 3];
    }
    return eq;
}

// Word i of a key held in registers.  Every word is masked and or-ed in, so
// the array is only ever indexed by constants and stays out of scratch.
template <int NW>
__device__ __forceinline__ uint32_t key_word(const uint32_t (&w)[NW], int i) {
    uint32_t r = 0;
#pragma unroll
    for (int t = 0; t < NW; t++) r |= w[t] & (0u - (uint32_t)(i == t));
    return r;
}

template <int NW>
__device__ __forceinline__ int key_q(const uint32_t (&w)[NW], int m) {
    return (int)((key_word<NW>(w, m >> 2) >> ((m & 3) * 8)) & 0xffu);
}

// add d to the quanta of mode m; the caller has checked 0 <= q+d <= 255,
// so the byte add never carries or borrows into a neighbour.
template <int NW>
__device__ __forceinline__ void key_add(uint32_t (&w)[NW], int m, int d) {
    const uint32_t add = ((uint32_t)d) << ((m & 3) * 8);
    const int wi = m >> 2;
#pragma unroll
    for (int t = 0; t < NW; t++) w[t] += add & (0u - (uint32_t)(t == wi));
}

// probe with the key's hash h already computed
template <int NW>
__device__ __forceinline__ int64_t hash_find_h(const uint64_t* __restrict__ slots, uint64_t mask,
                                               const uint4* __restrict__ keys, const uint32_t (&w)[NW], uint64_t h) {
    uint64_t pos = h & mask;
    const uint32_t tag = (uint32_t)(h >> 32);
    while (true) {
        const uint64_t s = slots[pos];
        if (s == kEmpty) return -1;
        if ((uint32_t)(s >> 32) == tag && key_eq<NW>(keys, (uint32_t)s, w)) return (int64_t)(uint32_t)s;
        pos = (pos + 1) & mask;
    }
}

template <int NW>
__device__ __forceinline__ int64_t hash_find(const uint64_t* __restrict__ slots, uint64_t mask,
                                             const uint4* __restrict__ keys, const uint32_t (&w)[NW]) {
    return hash_find_h<NW>(slots, mask, keys, w, key_hash<NW>(w));
}

// Bloom filter in front of a hash (one bit array, two bits set per key,
// 4 to 32 bits per key and at most 4 MB, small enough to stay in a die's L2
// while the hash itself does not).  Most probes of the pattern walks find
// nothing; the filter answers them without touching the hash.
// Blocked: both bits of a key lie in one 64-bit word, so a test is one read
// (48-mode H builds 2.19 -> 2.08 s and PT2 pairs 1.97 -> 1.78 s against two
// independent bits; the 4 MB size is still the fastest).
__device__ __forceinline__ uint64_t bloom_second(uint64_t H) {
    return ((H ^ (H >> 29)) * 0xBF58476D1CE4E5B9ULL) >> 32;
}
// the key's word (index into 64-bit words; fm + 1 bits in all, a multiple of 64) and its two bits
__device__ __forceinline__ uint64_t bloom_word(uint64_t H, uint64_t fm) { return (H & fm) >> 6; }
__device__ __forceinline__ uint64_t bloom_bits2(uint64_t H) {
    const uint64_t s = bloom_second(H);
    return (1ull << (s & 63)) | (1ull << ((s >> 6) & 63));
}
__device__ __forceinline__ void bloom_set(uint32_t* f, uint64_t fm, uint64_t H) {
    atomicOr(reinterpret_cast<unsigned long long*>(f) + bloom_word(H, fm), (unsigned long long)bloom_bits2(H));
}
__device__ __forceinline__ bool bloom_maybe(const uint32_t* __restrict__ f, uint64_t fm, uint64_t H) {
    const uint64_t b = bloom_bits2(H);
    return (reinterpret_cast<const uint64_t*>(f)[bloom_word(H, fm)] & b) == b;
}

// Probe for the key w moved by pattern e (nz modes; the caller has checked
// that every moved quantum stays in 0..255), with the moved key's hash H
// given: the moved key is built only when a slot's tag matches.
template <int NW>
__device__ __forceinline__ int64_t probe_shift(const uint64_t* __restrict__ slots, uint64_t mask,
                                               const uint4* __restrict__ keys, const uint32_t (&w)[NW],
                                               const uint4 e, int nz, uint64_t H);

// ------------------------------------------------------------------------
// table records in registers.  DeltaEnt and HbVar (16 bytes) are read as one
// uint4, ElemFC as its FC plus the 16 bytes after it; their byte fields are
// taken with constant shifts in fully unrolled loops, since a runtime index
// into a local struct would place it in scratch memory.
//   DeltaEnt: mode[s] = byte s, delta[s] = byte 6 + s, meta = word 3
//   HbVar:    d = byte 0, mode[t] = byte 1 + t, delta[t] = byte 7 + t
//   ElemFC:   d = byte 0, mode[t] = byte 1 + t, pow[t] = byte 7 + t
// ------------------------------------------------------------------------
constexpr int kMaxD = 6;  // modes per pattern, variant or FC

__device__ __forceinline__ uint32_t rec_u8(const uint4& r, int i) {
    const uint32_t w = i < 4 ? r.x : i < 8 ? r.y : i < 12 ? r.z : r.w;
    return (w >> ((i & 3) * 8)) & 0xffu;
}
__device__ __forceinline__ int rec_i8(const uint4& r, int i) { return (int)(int8_t)(uint8_t)rec_u8(r, i); }

template <class T>
__device__ __forceinline__ uint4 rec_load(const T* base, int64_t i) {
    static_assert(sizeof(T) == 16, "16-byte record");
    return reinterpret_cast<const uint4*>(base)[i];
}

__device__ __forceinline__ int delta_nz(const uint4& e) { return (int)(e.w >> 28); }
__device__ __forceinline__ uint32_t delta_group(const uint4& e) { return e.w & 0x03ffffffu; }

// Walk record: the part of a connection pattern that the pattern walks test
// (modes and quanta changes), 8 bytes instead of the 16 of DeltaEnt, so the
// walks stream half the bytes; the full record is read only for a hit.
// Bits 10 s .. 10 s + 5: mode s (< 64, as every key holds at most 64
// modes); 10 s + 6 .. 10 s + 9: its change, signed (|change| <= 6); bits
// 60..63: the number of modes.
__device__ __forceinline__ int delta_nz(uint64_t e) { return (int)(e >> 60); }
__device__ __forceinline__ int walk_mode(uint64_t e, int s) { return (int)((e >> (10 * s)) & 63u); }
__device__ __forceinline__ int walk_delta(uint64_t e, int s) {
    return (int)((int64_t)(e << (54 - 10 * s)) >> 60);  // bits 10 s + 6 .. 10 s + 9, sign-extended
}

__device__ __forceinline__ double elem_load(const ElemFC* __restrict__ elem, uint32_t i, uint4& b) {
    const uint64_t* q = reinterpret_cast<const uint64_t*>(elem + i);
    const uint64_t u = q[1], v = q[2];
    b = make_uint4((uint32_t)u, (uint32_t)(u >> 32), (uint32_t)v, (uint32_t)(v >> 32));
    return __longlong_as_double((long long)q[0]);
}

// ------------------------------------------------------------------------
// small utility kernels
// ------------------------------------------------------------------------
template <class T>
__global__ void k_fill(T* p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Initial product basis (Input_Reader.cpp:308-364): every q with
// q_m <= lim_m and sum q <= ntot, ordered with the last mode most
// significant and mode 0 fastest.  Thread r unranks state r: cnt[m][s] =
// number of choices for modes 0..m-1 with sum <= s, so taking the modes from
// the last down, q_m is the first value whose block of completions holds r.
__global__ void k_initial_basis(int64_t n0, int M, int ntot, int K, const int64_t* __restrict__ cnt,
                                const int32_t* __restrict__ lim, uint8_t* __restrict__ keys) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n0) return;
    uint4* k4 = reinterpret_cast<uint4*>(keys + r * K);
    for (int t = 0; t < K / 16; t++) k4[t] = make_uint4(0u, 0u, 0u, 0u);
    int64_t rr = r;
    int b = ntot;
    for (int m = M - 1; m >= 0; m--) {
        const int top = min(lim[m], b);
        int q = 0;
        for (; q < top; q++) {
            const int64_t c = cnt[(int64_t)m * (ntot + 1) + (b - q)];
            if (rr < c) break;
            rr -= c;
        }
        keys[r * K + m] = (uint8_t)q;
        b -= q;
    }
}

// row-major n x ld block of identity columns c0 .. c0+bc-1 (zero beyond bc)
__global__ void k_identity_block(double* __restrict__ X, int64_t n, int ld, int64_t c0, int bc) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * ld; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / ld;
        const int c = (int)(t - r * ld);
        X[t] = (c < bc && r == c0 + c) ? 1.0 : 0.0;
    }
}

// copy an r x c block between row-major arrays with different leading dims
__global__ void k_copy2d(const double* __restrict__ src, int lds, double* __restrict__ dst, int ldd, int64_t rows,
                         int cols) {
    int64_t total = rows * cols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = t / cols;
        int c = (int)(t - r * cols);
        dst[r * ldd + c] = src[r * lds + c];
    }
}

__global__ void k_to_f32(const double* __restrict__ src, int lds, float* __restrict__ dst, int ldd, int64_t rows,
                         int cols) {
    int64_t total = rows * cols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = t / cols;
        int c = (int)(t - r * cols);
        dst[r * ldd + c] = (float)src[r * lds + c];
    }
}

template <int NW>
__global__ void k_hash_insert(const uint4* __restrict__ keys, int64_t i0, int64_t i1, uint64_t* slots, uint64_t mask,
                              uint32_t* bloom, uint64_t bmask) {
    int64_t i = i0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= i1) return;
    uint32_t w[NW];
    key_load<NW>(keys, i, w);
    const uint64_t h = key_hash<NW>(w);
    if (bloom) bloom_set(bloom, bmask, h);
    uint64_t pos = h & mask;
    const uint64_t val = (h & 0xffffffff00000000ULL) | (uint64_t)(uint32_t)i;
    while (true) {
        unsigned long long prev = atomicCAS((unsigned long long*)&slots[pos], (unsigned long long)kEmpty,
                                            (unsigned long long)val);
        if (prev == kEmpty) return;
        pos = (pos + 1) & mask;
    }
}

// Sum over the 64 lanes of a wave of each lane's acc[0..B): a transposing
// butterfly over lane bits 5, 4, ... (each step keeps half of the columns and
// trades the other half with the partner lane), then a sum over the R = 64/B
// lanes that share a column.  Lane l (l % R == 0) returns column l / R.
// Fixed order, so the result is reproducible bit for bit.
template <int B, class AT>
__device__ __forceinline__ AT lane_colsum(AT (&acc)[B], int lane) {
#pragma unroll
    for (int h = B / 2, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
        const bool up = (lane & bit) != 0;
#pragma unroll
        for (int i = 0; i < h; i++) {
            const AT send = up ? acc[i] : acc[i + h];
            const AT keep = up ? acc[i + h] : acc[i];
            acc[i] = keep + __shfl_xor(send, bit);
        }
    }
    constexpr int R = 64 / B;  // lanes that share one column after the butterfly
    AT s = acc[0];
#pragma unroll
    for (int off = R / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    return s;
}

// ------------------------------------------------------------------------
// Hamiltonian build (generate-and-lookup)
// ------------------------------------------------------------------------
struct HBuildArgs {
    int64_t n;
    int M;
    int P;  // number of signed patterns
    int nq;  // ladder table rows per power
    const uint4* keys;
    const uint64_t* slots;
    uint64_t mask;
    const uint32_t* bloom;  // filter of the hash's keys (bits bmask + 1)
    uint64_t bmask;
    const DeltaEnt* deltas;
    const uint64_t* walk;  // the walk records of deltas
    const uint32_t* poff;     // M + 2 block offsets of deltas
    const GroupEnt* groups;
    const uint32_t* fclist;
    const ElemFC* elem;
    const double* ladder;
    const int32_t* perm;  // solver row -> basis state (null: identity)
    const int32_t* inv;   // basis state -> solver row (null: identity)
    int64_t* rowcnt;
    int64_t* rowcnt_up;  // count pass: entries with a higher stored column (null: not counted)
    int half;            // fill pass: store only entries with a higher stored column
    double* rowabs;      // half fill pass: sum of |H_ij| over the whole row
    const int64_t* rowptr;
    int32_t* col;
    double* val;
    float* hi;
    float* lo;
    int form;
    // incremental build (n_old > 0): block b builds the row of basis state
    // first + b; a hit on an old state j < n_old also counts (extra[j]) or
    // stores (after oldcnt[j] entries, at cursor[j]) the transposed entry in
    // j's row
    int64_t first;
    int64_t n_old;
    unsigned long long* extra;
    const int64_t* oldcnt;
    unsigned long long* cursor;
    // hit words (null: none): the count pass stores the ballot of every
    // 64-pattern step of a row's walk at hitw[woff[b] + step] (b the block
    // index of the launch), and the fill pass then visits only those hits
    uint64_t* hitw;
    const int64_t* woff;
};

// steps of 64 patterns in the walk of each launched row (the count pass's
// hit words per row); row mapping as k_hbuild's
template <int NW>
__global__ void k_walk_words(HBuildArgs a, int64_t nb, int64_t* __restrict__ words) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b > nb) return;
    if (b == nb) {
        words[nb] = 0;
        return;
    }
    int64_t src;
    if (a.n_old > 0) {
        src = a.first + b;
    } else {
        src = a.perm ? a.perm[b] : b;
    }
    uint32_t w[NW];
    key_load<NW>(a.keys, src, w);
    int64_t cnt = 0;
    for (int m = -1; m < a.M; m++) {
        if (m >= 0 && key_q<NW>(w, m) == 0) continue;
        const int blk = m + 1;
        cnt += ((int64_t)a.poff[blk + 1] - (int64_t)a.poff[blk] + kWave - 1) / kWave;
    }
    words[b] = cnt;
}

// sum over the group's FCs of fc * prod_t L[k][q_start][q_end], in the
// element order (= the reference's Quartic/Sextic or Cubic/Quintic order)
// (sgn = -1 applies the pattern's deltas negated)
__device__ __forceinline__ double group_element(const GroupEnt ge, const uint32_t* __restrict__ fclist,
                                                const ElemFC* __restrict__ elem, const double* __restrict__ L,
                                                int nq, const uint8_t* __restrict__ qrow, const uint4 e, int nz,
                                                bool row_is_start, int sgn = 1) {
    double v = 0.0;
    for (uint32_t f = 0; f < ge.fc_count; f++) {
        uint4 b;
        double x = elem_load(elem, fclist[ge.fc_start + f], b);
        const int d = (int)rec_u8(b, 0);
#pragma unroll
        for (int t = 0; t < kMaxD; t++) {
            if (t < d) {
                const int m = (int)rec_u8(b, 1 + t);
                const int k = (int)rec_u8(b, 7 + t);
                const int qr = qrow[m];
                int dq = 0;
#pragma unroll
                for (int s = 0; s < kMaxD; s++) dq = (s < nz && (int)rec_u8(e, s) == m) ? sgn * rec_i8(e, 6 + s) : dq;
                const int qs = row_is_start ? qr : qr + dq;
                const int qe = row_is_start ? qr + dq : qr;
                x = x * L[((size_t)k * nq + qs) * kLadderW + (qe - qs + k)];
            }
        }
        v = v + x;
    }
    return v;
}

// apply pattern e (nz modes) to key t; false when a quantum leaves 0..255
template <int NW>
__device__ __forceinline__ bool apply_pattern(const uint4 e, int nz, const uint8_t* __restrict__ sq,
                                              uint32_t (&t)[NW]) {
    bool ok = true;
#pragma unroll
    for (int s = 0; s < kMaxD; s++) {
        if (s < nz) {
            const int m = (int)rec_u8(e, s);
            const int dl = rec_i8(e, 6 + s);
            const int nqv = (int)sq[m] + dl;
            ok = ok && nqv >= 0 && nqv <= 255;
            key_add<NW>(t, m, dl);
        }
    }
    return ok;
}

// true when pattern e (nz modes) keeps every quantum of the row in 0..255
__device__ __forceinline__ bool pattern_ok(const uint4 e, int nz, const uint8_t* __restrict__ sq) {
    bool ok = true;
#pragma unroll
    for (int s = 0; s < kMaxD; s++) {
        if (s < nz) {
            const int nqv = (int)sq[rec_u8(e, s)] + rec_i8(e, 6 + s);
            ok = ok && nqv >= 0 && nqv <= 255;
        }
    }
    return ok;
}
__device__ __forceinline__ bool pattern_ok(uint64_t e, int nz, const uint8_t* __restrict__ sq) {
    bool ok = true;
#pragma unroll
    for (int s = 0; s < kMaxD; s++) {
        if (s < nz) {
            const int nqv = (int)sq[walk_mode(e, s)] + walk_delta(e, s);
            ok = ok && nqv >= 0 && nqv <= 255;
        }
    }
    return ok;
}

template <int NW>
__device__ __forceinline__ int64_t probe_shift(const uint64_t* __restrict__ slots, uint64_t mask,
                                               const uint4* __restrict__ keys, const uint32_t (&w)[NW],
                                               const uint4 e, int nz, uint64_t H) {
    uint64_t pos = H & mask;
    const uint32_t tag = (uint32_t)(H >> 32);
    // most probes end at an empty slot before any tag matches: the moved key
    // is built only after the first match
    uint64_t s;
    while (true) {
        s = slots[pos];
        if (s == kEmpty) return -1;
        if ((uint32_t)(s >> 32) == tag) break;
        pos = (pos + 1) & mask;
    }
    uint32_t t[NW];
#pragma unroll
    for (int i = 0; i < NW; i++) t[i] = w[i];
#pragma unroll
    for (int k = 0; k < kMaxD; k++)
        if (k < nz) key_add<NW>(t, (int)rec_u8(e, k), rec_i8(e, 6 + k));
    while (true) {
        if ((uint32_t)(s >> 32) == tag && key_eq<NW>(keys, (uint32_t)s, t)) return (int64_t)(uint32_t)s;
        pos = (pos + 1) & mask;
        s = slots[pos];
        if (s == kEmpty) return -1;
    }
}

template <int NW>
__device__ __forceinline__ int64_t probe_shift(const uint64_t* __restrict__ slots, uint64_t mask,
                                               const uint4* __restrict__ keys, const uint32_t (&w)[NW],
                                               uint64_t e, int nz, uint64_t H) {
    uint64_t pos = H & mask;
    const uint32_t tag = (uint32_t)(H >> 32);
    // most probes end at an empty slot before any tag matches: the moved key
    // is built only after the first match
    uint64_t s;
    while (true) {
        s = slots[pos];
        if (s == kEmpty) return -1;
        if ((uint32_t)(s >> 32) == tag) break;
        pos = (pos + 1) & mask;
    }
    uint32_t t[NW];
#pragma unroll
    for (int i = 0; i < NW; i++) t[i] = w[i];
#pragma unroll
    for (int k = 0; k < kMaxD; k++)
        if (k < nz) key_add<NW>(t, walk_mode(e, k), walk_delta(e, k));
    while (true) {
        if ((uint32_t)(s >> 32) == tag && key_eq<NW>(keys, (uint32_t)s, t)) return (int64_t)(uint32_t)s;
        pos = (pos + 1) & mask;
        s = slots[pos];
        if (s == kEmpty) return -1;
    }
}

// Walk of the pattern blocks a row can use (the table is ordered so that
// block 0 lowers no mode and block 1 + m has m as its lowest lowered mode):
// block 0, then the block of every mode with quanta.  b = -1 starts the
// walk; false when no block is left.  sq is the row's quanta in LDS, so the
// walk is the same for every lane.
__device__ __forceinline__ bool next_block(const uint32_t* __restrict__ off, int M, const uint8_t* __restrict__ sq,
                                           int& b, int& p0, int& p1) {
    if (b < 0) {
        b = 0;
    } else {
        do {
            ++b;
        } while (b <= M && sq[b - 1] == 0);
        if (b > M) return false;
    }
    p0 = (int)off[b];
    p1 = (int)off[b + 1];
    return true;
}

// sum over a list of zero-pattern FCs of fc * prod_t L[k][q][q] (in list order)
__device__ __forceinline__ double diag_fc_term(const ElemFC* __restrict__ elem, uint32_t i,
                                               const double* __restrict__ L, int nq,
                                               const uint8_t* __restrict__ sq) {
    uint4 b;
    double x = elem_load(elem, i, b);
    const int d = (int)rec_u8(b, 0);
#pragma unroll
    for (int t = 0; t < kMaxD; t++) {
        if (t < d) {
            const int q = sq[rec_u8(b, 1 + t)];
            const int k = (int)rec_u8(b, 7 + t);
            x = x * L[((size_t)k * nq + q) * kLadderW + k];
        }
    }
    return x;
}

// the key-sum coefficients of the 64 modes, for a block's LDS table
__device__ __forceinline__ void load_coefs(uint64_t* cf) {
    for (int m = threadIdx.x; m < 64; m += blockDim.x) cf[m] = key_coef(m);
}

// linear key sum of a key held in registers, from the coefficient table in
// LDS (bytes beyond the modes are zero)
template <int NW>
__device__ __forceinline__ uint64_t key_lin_cf(const uint32_t (&w)[NW], const uint64_t* __restrict__ cf) {
    uint64_t h = 0;
#pragma unroll
    for (int i = 0; i < NW; i++)
#pragma unroll
        for (int b = 0; b < 4; b++) h += (uint64_t)((w[i] >> (8 * b)) & 0xffu) * cf[4 * i + b];
    return h;
}

// linear key sum of a row whose quanta are in LDS (sq, M modes), from the
// coefficient table in LDS (a loop over the modes: unrolled compile-time
// coefficients would hold every one in registers)
__device__ __forceinline__ uint64_t row_lin(const uint8_t* __restrict__ sq, int M, const uint64_t* __restrict__ cf) {
    uint64_t h = 0;
    for (int m = 0; m < M; m++) h += (uint64_t)sq[m] * cf[m];
    return h;
}

// change of the linear key sum when pattern e (nz modes at bytes 0..5,
// deltas at bytes 6..11) is applied; cf is the coefficient table in LDS
__device__ __forceinline__ uint64_t pattern_shift(const uint4 e, int nz, const uint64_t* __restrict__ cf) {
    uint64_t h = 0;
#pragma unroll
    for (int s = 0; s < kMaxD; s++)
        if (s < nz) h += (uint64_t)(int64_t)rec_i8(e, 6 + s) * cf[rec_u8(e, s)];
    return h;
}
__device__ __forceinline__ uint64_t pattern_shift(uint64_t e, int nz, const uint64_t* __restrict__ cf) {
    uint64_t h = 0;
#pragma unroll
    for (int s = 0; s < kMaxD; s++)
        if (s < nz) h += (uint64_t)(int64_t)walk_delta(e, s) * cf[walk_mode(e, s)];
    return h;
}

// One walk step's candidate, prepared a step ahead so that its Bloom word is
// loading while the step before is tested and probed: ok when the pattern
// keeps every quantum in 0..255 (and, with a tag range, its hash tag lies in
// it), H its moved key's hash, bw the Bloom word to test.
struct WalkCand {
    uint64_t e, H, bw;
    bool ok;
};
__device__ __forceinline__ WalkCand walk_cand(uint64_t e, bool in, const uint8_t* __restrict__ sq, uint64_t h0,
                                              const uint64_t* __restrict__ cf, const uint32_t* __restrict__ bloom,
                                              uint64_t bmask, uint64_t tlo = 0, uint64_t thi = ~0ull) {
    WalkCand c{e, 0ull, 0ull, false};
    if (in) {
        const int nz = delta_nz(e);
        if (pattern_ok(e, nz, sq)) {
            c.H = fmix64(h0 + pattern_shift(e, nz, cf));
            const uint64_t tag = c.H >> 32;
            if (tag >= tlo && tag < thi) {
                c.ok = true;
                c.bw = reinterpret_cast<const uint64_t*>(bloom)[bloom_word(c.H, bmask)];
            }
        }
    }
    return c;
}
__device__ __forceinline__ bool walk_maybe(const WalkCand& c) {
    const uint64_t b = bloom_bits2(c.H);
    return c.ok && (c.bw & b) == b;
}


// the same for a heat-bath variant (d modes at bytes 1..6, deltas at 7..12)
__device__ __forceinline__ uint64_t variant_shift(const uint4 v, int d, const uint64_t* __restrict__ cf) {
    uint64_t h = 0;
#pragma unroll
    for (int t = 0; t < kMaxD; t++)
        if (t < d) h += (uint64_t)(int64_t)rec_i8(v, 7 + t) * cf[rec_u8(v, 1 + t)];
    return h;
}

template <int NW, bool FILL>
__global__ __launch_bounds__(64) void k_hbuild(HBuildArgs a) {
    __shared__ uint8_t sq[64];
    __shared__ uint64_t cf[64];
    // fill pass: hits wait here (in pattern order) until 64 of them can be
    // evaluated at once, one per lane, instead of a few lanes per chunk
    __shared__ int32_t qj[2 * kWave];
    __shared__ int32_t qp[2 * kWave];
    const int lane = threadIdx.x;
    int64_t row, src;  // stored row and its basis state
    if (a.n_old > 0) {
        src = a.first + blockIdx.x;
        row = a.inv ? a.inv[src] : src;
    } else {
        row = blockIdx.x;
        src = a.perm ? a.perm[row] : row;
    }
    uint32_t w[NW];
    key_load<NW>(a.keys, src, w);
    if (lane < a.M) sq[lane] = (uint8_t)key_q<NW>(w, lane);
    load_coefs(cf);
    __syncthreads();
    const uint64_t h0 = row_lin(sq, a.M, cf);
    int64_t base = FILL ? a.rowptr[row] : 0;
    int64_t cnt = 0, cnt_up = 0;
    int nq = 0;
    double absum = 0.0;  // half fill: this lane's share of the row's sum of |H_ij|
    // evaluate and store the first `count` (<= 64) queued hits (half
    // storage: evaluate all, store those above the diagonal)
    auto drain = [&](int count) {
        const int64_t cj = lane < count ? (a.inv ? a.inv[qj[lane]] : qj[lane]) : -1;
        const bool keep = lane < count && (!a.half || cj > row);
        const uint64_t km = __ballot(keep);
        if (lane < count) {
            const int p = qp[lane];
            const int64_t j = qj[lane];
            const uint4 e = rec_load(a.deltas, p);
            const GroupEnt ge = a.groups[delta_group(e)];
            // the element starts from the lower basis index, as AnharmPot(min, max)
            const double v = group_element(ge, a.fclist, a.elem, a.ladder, a.nq, sq, e, delta_nz(e), src < j);
            absum += fabs(v);
            if (keep) {
                const int64_t pos = base + __popcll(km & ((1ull << lane) - 1ull));
                a.col[pos] = (int32_t)cj;
                // an old state's row gets the transposed entry (same value:
                // the element starts from the lower basis index either way)
                const bool tr = j < a.n_old;
                const int64_t pt =
                    tr ? a.rowptr[a.inv ? a.inv[j] : j] + a.oldcnt[j] + (int64_t)atomicAdd(&a.cursor[j], 1ull) : 0;
                if (tr) a.col[pt] = (int32_t)row;
                if (a.form == kFP64) {
                    a.val[pos] = v;
                    if (tr) a.val[pt] = v;
                } else {
                    const float h = (float)v;
                    const float l = (float)(v - (double)h);
                    a.hi[pos] = h;
                    a.lo[pos] = l;
                    if (tr) {
                        a.hi[pt] = h;
                        a.lo[pt] = l;
                    }
                }
            }
        }
        base += __popcll(km);
    };
    int blk = -1, pb0 = 0, pb1 = 0;
    uint64_t* hw = a.hitw ? a.hitw + a.woff[blockIdx.x] : nullptr;
    const bool known = FILL && hw;  // fill pass over recorded hits
    while (next_block(a.poff, a.M, sq, blk, pb0, pb1)) {
    // each lane's next pattern record is loaded one step ahead
    // the current step's candidate and the next step's record
    WalkCand cur{};
    uint64_t en = 0ull;
    if (!known) {
        cur = walk_cand(pb0 + lane < pb1 ? a.walk[pb0 + lane] : 0ull, pb0 + lane < pb1, sq, h0, cf, a.bloom, a.bmask);
        en = pb0 + kWave + lane < pb1 ? a.walk[pb0 + kWave + lane] : 0ull;
    }
    for (int p0 = pb0; p0 < pb1; p0 += kWave) {
        const int p = p0 + lane;
        bool hit = false;
        int64_t j = -1;
        if (known) {
            const uint64_t word = *hw++;
            if (word == 0ull) continue;
            if ((word >> lane) & 1ull) {
                const uint64_t e = a.walk[p];
                const int nz = delta_nz(e);
                const uint64_t H = fmix64(h0 + pattern_shift(e, nz, cf));
                j = probe_shift<NW>(a.slots, a.mask, a.keys, w, e, nz, H);
                hit = j >= 0;
            }
        } else {
            // the next step's candidate first, so its Bloom word loads while
            // this one is probed
            const WalkCand nxt = walk_cand(en, p + kWave < pb1, sq, h0, cf, a.bloom, a.bmask);
            en = p + 2 * kWave < pb1 ? a.walk[p + 2 * kWave] : 0ull;
            if (walk_maybe(cur)) {
                j = probe_shift<NW>(a.slots, a.mask, a.keys, w, cur.e, delta_nz(cur.e), cur.H);
                hit = j >= 0;
            }
            cur = nxt;
        }
        const uint64_t bal = __ballot(hit);
        if (!FILL && hw) {
            if (lane == 0) *hw = bal;
            hw++;
        }
        if (!FILL) {
            cnt += __popcll(bal);
            if (a.rowcnt_up) cnt_up += __popcll(__ballot(hit && (a.inv ? a.inv[j] : j) > row));
            if (hit && j < a.n_old) atomicAdd(&a.extra[j], 1ull);
        } else {
            if (hit) {
                const int slot = nq + __popcll(bal & ((1ull << lane) - 1ull));
                qj[slot] = (int32_t)j;
                qp[slot] = p;
            }
            nq += __popcll(bal);
            __syncthreads();
            if (nq >= kWave) {
                drain(kWave);
                __syncthreads();
                if (lane < nq - kWave) {
                    qj[lane] = qj[kWave + lane];
                    qp[lane] = qp[kWave + lane];
                }
                nq -= kWave;
                __syncthreads();
            }
        }
    }
    }
    if (FILL && nq > 0) drain(nq);
    if (!FILL && lane == 0) {
        a.rowcnt[row] = cnt;
        if (a.rowcnt_up) a.rowcnt_up[row] = cnt_up;
    }
    if (FILL && a.half) {
        // fixed-order wave sum (the row's bound for the fixed-point
        // accumulation of the transposed products)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) absum += __shfl_xor(absum, off);
        if (lane == 0) a.rowabs[row] = absum;
    }
}

// Matrix-free product Y = (D + H_off) X for B-column blocks (bcols <= B
// used): each row regenerates its connections as the H build does (every
// pattern probes the basis hash; hits queue in LDS and are evaluated 64 at
// a time, one per lane) and multiplies them into the gathered rows of X
// instead of storing them.  X and Y are row-major in stored-row order.
struct HFreeArgs {
    HBuildArgs h;
    const double* D;
    const double* X;
    int ldx;
    double* Y;
    int ldy;
    int bcols;
};

template <int NW, int B>
__global__ __launch_bounds__(64) void k_hfree(HFreeArgs f) {
    const HBuildArgs& a = f.h;
    __shared__ uint8_t sq[64];
    __shared__ uint64_t cf[64];
    __shared__ int32_t qj[2 * kWave];
    __shared__ int32_t qp[2 * kWave];
    const int lane = threadIdx.x;
    const int64_t row = blockIdx.x;
    const int64_t src = a.perm ? a.perm[row] : row;
    uint32_t w[NW];
    key_load<NW>(a.keys, src, w);
    if (lane < a.M) sq[lane] = (uint8_t)key_q<NW>(w, lane);
    load_coefs(cf);
    __syncthreads();
    double acc[B];
#pragma unroll
    for (int c = 0; c < B; c++) acc[c] = 0.0;
    int nq = 0;
    auto drain = [&](int count) {
        if (lane < count) {
            const int p = qp[lane];
            const int64_t j = qj[lane];
            const uint4 e = rec_load(a.deltas, p);
            const GroupEnt ge = a.groups[delta_group(e)];
            const double v = group_element(ge, a.fclist, a.elem, a.ladder, a.nq, sq, e, delta_nz(e), src < j);
            const int64_t jr = a.inv ? (int64_t)a.inv[j] : j;
            const double* xr = f.X + jr * f.ldx;
#pragma unroll
            for (int c = 0; c < B; c++)
                if (c < f.bcols) acc[c] = acc[c] + v * xr[c];
        }
    };
    const uint64_t h0 = row_lin(sq, a.M, cf);
    int blk = -1, pb0 = 0, pb1 = 0;
    while (next_block(a.poff, a.M, sq, blk, pb0, pb1)) {
    // each lane's next pattern record is loaded one step ahead
    uint64_t en = pb0 + lane < pb1 ? a.walk[pb0 + lane] : 0ull;
    for (int p0 = pb0; p0 < pb1; p0 += kWave) {
        const int p = p0 + lane;
        const uint64_t e = en;
        if (p + kWave < pb1) en = a.walk[p + kWave];
        bool hit = false;
        int64_t j = -1;
        if (p < pb1) {
            const int nz = delta_nz(e);
            if (pattern_ok(e, nz, sq)) {
                const uint64_t H = fmix64(h0 + pattern_shift(e, nz, cf));
                if (bloom_maybe(a.bloom, a.bmask, H)) {
                    j = probe_shift<NW>(a.slots, a.mask, a.keys, w, e, nz, H);
                    hit = j >= 0;
                }
            }
        }
        const uint64_t bal = __ballot(hit);
        if (hit) {
            const int slot = nq + __popcll(bal & ((1ull << lane) - 1ull));
            qj[slot] = (int32_t)j;
            qp[slot] = p;
        }
        nq += __popcll(bal);
        __syncthreads();
        if (nq >= kWave) {
            drain(kWave);
            __syncthreads();
            if (lane < nq - kWave) {
                qj[lane] = qj[kWave + lane];
                qp[lane] = qp[kWave + lane];
            }
            nq -= kWave;
            __syncthreads();
        }
    }
    }
    if (nq > 0) drain(nq);
    const double s = lane_colsum<B>(acc, lane);
    constexpr int R = 64 / B;
    if ((lane & (R - 1)) == 0) {
        const int c = lane / R;
        if (c < f.bcols) f.Y[row * f.ldy + c] = f.D[row] * f.X[row * f.ldx + c] + s;
    }
}

// Incremental build: rows of the n_old states of the previous build.  Each
// keeps its old entries (old solver order -> basis state -> new solver
// order; both orders sort the same keys, so the columns stay ascending) and
// gains extra[x] entries for the new states.
__global__ void k_old_counts(int64_t n_old, const int32_t* __restrict__ inv, const int32_t* __restrict__ inv_o,
                             const int64_t* __restrict__ rowptr_o, const unsigned long long* __restrict__ extra,
                             int64_t* __restrict__ oldcnt, int64_t* __restrict__ rowcnt) {
    const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (x >= n_old) return;
    const int64_t ro = inv_o ? inv_o[x] : x;
    const int64_t oc = rowptr_o[ro + 1] - rowptr_o[ro];
    oldcnt[x] = oc;
    rowcnt[inv ? inv[x] : x] = oc + (int64_t)extra[x];
}

// one wave per old state
__global__ __launch_bounds__(256) void k_copy_old_rows(int64_t n_old, const int32_t* __restrict__ inv,
                                                       const int32_t* __restrict__ inv_o,
                                                       const int32_t* __restrict__ perm_o,
                                                       const int64_t* __restrict__ rowptr_o,
                                                       const int32_t* __restrict__ col_o,
                                                       const double* __restrict__ val_o,
                                                       const float* __restrict__ hi_o, const float* __restrict__ lo_o,
                                                       const int64_t* __restrict__ rowptr, int32_t* __restrict__ col,
                                                       double* __restrict__ val, float* __restrict__ hi,
                                                       float* __restrict__ lo, int form) {
    const int lane = threadIdx.x & 63;
    const int64_t x = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (x >= n_old) return;
    const int64_t ro = inv_o ? inv_o[x] : x;
    const int64_t k0 = rowptr_o[ro], k1 = rowptr_o[ro + 1];
    const int64_t d = rowptr[inv ? inv[x] : x] - k0;
    for (int64_t k = k0 + lane; k < k1; k += 64) {
        const int64_t c = col_o[k];
        const int64_t b = perm_o ? perm_o[c] : c;
        col[k + d] = inv ? inv[b] : (int32_t)b;
        if (form == kFP64) {
            val[k + d] = val_o[k];
        } else {
            hi[k + d] = hi_o[k];
            lo[k + d] = lo_o[k];
        }
    }
}

// H_ii = E_i + V_ii (Ham.cpp:203-226 plus the zero-pattern FCs), in the
// reference's order: E_i summed over modes, V_ii summed over FCs, then added.
template <int NW>
__global__ void k_diag(int64_t n, int M, const uint4* __restrict__ keys, const double* __restrict__ freq,
                       const uint32_t* __restrict__ dfc, int ndfc, const ElemFC* __restrict__ elem,
                       const double* __restrict__ L, int nq, const int32_t* __restrict__ perm,
                       double* __restrict__ D) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[NW];
    key_load<NW>(keys, perm ? (int64_t)perm[i] : i, w);
    double e = 0.;
    for (int j = 0; j < M; j++) {
        double ej = 0.5;
        ej = ej + (double)key_q<NW>(w, j);
        ej = ej * freq[j];
        e = e + ej;
    }
    double v = 0.0;
    for (int f = 0; f < ndfc; f++) {
        uint4 b;
        double x = elem_load(elem, dfc[f], b);
        const int d = (int)rec_u8(b, 0);
#pragma unroll
        for (int t = 0; t < kMaxD; t++) {
            if (t < d) {
                const int q = key_q<NW>(w, (int)rec_u8(b, 1 + t));
                const int k = (int)rec_u8(b, 7 + t);
                x = x * L[((size_t)k * nq + q) * kLadderW + k];
            }
        }
        v = v + x;
    }
    D[i] = e + v;
}

// ------------------------------------------------------------------------
// SpMM: Y = (D + H_off) X, X/Y row-major.  The general kernel k_spmm takes
// one wave per row with the lanes split into 64/B entry groups x B columns;
// Workgroups are handed to the 8 XCDs in turn, so consecutive workgroups
// land on different dies and their neighbouring rows gather from different
// L2 caches.  With xc > 0 a kernel renumbers its workgroups so that each XCD
// takes runs of xc consecutive ones (the grid is a multiple of 8 xc): rows
// adjacent in the solver order then share their die's L2 (on the 48-mode
// H, 1024-row runs cut the 8-column product from 14.9 to 12.8 ms).
__device__ __forceinline__ int64_t xcd_block(int xc) {
    int64_t blk = blockIdx.x;
    if (xc > 0) {
        const int64_t t = blk >> 3, x = blk & 7;
        blk = ((t / xc) * 8 + x) * xc + t % xc;
    }
    return blk;
}

// k_spmm_lane (below) is the form for exactly 4, 8 or 16 contiguous columns.
// MODE 0: FP64 values; 1: FP32 values and FP32 X (FP32 accumulation, the
// diagonal term in FP64); 2: FP64 value = hi + lo.
// ------------------------------------------------------------------------
template <int MODE>
struct SpAcc {
    using T = double;
};
template <>
struct SpAcc<1> {
    using T = float;
};

// value of entry k in FP64 (MODE 0, 2) or FP32 (MODE 1); the matrix arrays
// are streamed once per product and loaded non-temporally so the gathered
// vector stays in cache
template <int MODE>
__device__ __forceinline__ typename SpAcc<MODE>::T sp_val(const double* __restrict__ val,
                                                          const float* __restrict__ hi,
                                                          const float* __restrict__ lo, int64_t k) {
    if (MODE == 0) return __builtin_nontemporal_load(val + k);
    if (MODE == 1) return __builtin_nontemporal_load(hi + k);
    return (double)__builtin_nontemporal_load(hi + k) + (double)__builtin_nontemporal_load(lo + k);
}

__device__ __forceinline__ float sp_fma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double sp_fma(double a, double b, double c) { return fma(a, b, c); }

template <int MODE>
__device__ __forceinline__ typename SpAcc<MODE>::T sp_x(const double* __restrict__ X,
                                                        const float* __restrict__ X32, int64_t idx32, int64_t idx) {
    if (MODE == 1) return X32[idx32];
    return X[idx];
}

template <int B, int MODE>
__global__ __launch_bounds__(256) void k_spmm(int64_t n, const int64_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col, const double* __restrict__ val,
                                              const float* __restrict__ hi, const float* __restrict__ lo,
                                              const double* __restrict__ D, const float* __restrict__ X32, int ldx32,
                                              const double* __restrict__ X, int ldx, double* __restrict__ Y, int ldy,
                                              int b) {
    using AT = typename SpAcc<MODE>::T;
    constexpr int G = kWave / B;
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const int grp = lane / B, c = lane % B;
    const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    const double d = D[row];
    for (int c0 = 0; c0 < b; c0 += B) {
        const int cc = c0 + c;
        const bool act = cc < b;
        AT a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        if (act) {
            int64_t k = k0 + grp;
            for (; k + 3 * G < k1; k += 4 * G) {
                const int64_t j0 = __builtin_nontemporal_load(col + k);
                const int64_t j1 = __builtin_nontemporal_load(col + k + G);
                const int64_t j2 = __builtin_nontemporal_load(col + k + 2 * G);
                const int64_t j3 = __builtin_nontemporal_load(col + k + 3 * G);
                const AT v0 = sp_val<MODE>(val, hi, lo, k), v1 = sp_val<MODE>(val, hi, lo, k + G);
                const AT v2 = sp_val<MODE>(val, hi, lo, k + 2 * G), v3 = sp_val<MODE>(val, hi, lo, k + 3 * G);
                a0 = sp_fma(v0, sp_x<MODE>(X, X32, j0 * ldx32 + cc, j0 * ldx + cc), a0);
                a1 = sp_fma(v1, sp_x<MODE>(X, X32, j1 * ldx32 + cc, j1 * ldx + cc), a1);
                a2 = sp_fma(v2, sp_x<MODE>(X, X32, j2 * ldx32 + cc, j2 * ldx + cc), a2);
                a3 = sp_fma(v3, sp_x<MODE>(X, X32, j3 * ldx32 + cc, j3 * ldx + cc), a3);
            }
            for (; k < k1; k += G) {
                const int64_t j0 = __builtin_nontemporal_load(col + k);
                a0 = sp_fma(sp_val<MODE>(val, hi, lo, k), sp_x<MODE>(X, X32, j0 * ldx32 + cc, j0 * ldx + cc), a0);
            }
        }
        AT acc = (a0 + a1) + (a2 + a3);
#pragma unroll
        for (int off = B; off < kWave; off <<= 1) acc += __shfl_xor(acc, off);
        if (grp == 0 && act) Y[row * ldy + cc] = d * X[row * ldx + cc] + (double)acc;
    }
}

// Block product for exactly B contiguous row-major columns (ldx == B, rows
// 16-byte aligned): one wave per row.  The gathered row of X for a stored
// entry is fetched cooperatively: P = B * (element bytes) / 16 lanes take
// one 16-byte piece each, so one wave instruction brings E = 64 / P whole
// rows (one cache line each) instead of a quarter of 64 rows.  Each lane
// accumulates its W = B / P columns; the wave then sums over the lanes that
// hold the same piece (fixed order, so products are reproducible bit for
// bit).
template <int B, int MODE>
struct SpCoop {
    static constexpr int XB = MODE == 1 ? 4 : 8;  // bytes per X element
    static constexpr int W = 16 / XB;             // elements per lane piece
    static constexpr int P = B / W;               // lanes per entry
    static constexpr int E = 64 / P;              // entries per wave step
};

// the W-element piece q of row j of X (FP64, or the FP32 copy for MODE 1)
template <int B, int MODE>
__device__ __forceinline__ void sp_piece(const double* __restrict__ X, const float* __restrict__ X32, int64_t j, int q,
                                         typename SpAcc<MODE>::T (&x)[SpCoop<B, MODE>::W]) {
    if constexpr (MODE == 1) {
        const float4 t = *reinterpret_cast<const float4*>(X32 + j * B + 4 * q);
        x[0] = t.x;
        x[1] = t.y;
        x[2] = t.z;
        x[3] = t.w;
    } else {
        const double2 t = *reinterpret_cast<const double2*>(X + j * B + 2 * q);
        x[0] = t.x;
        x[1] = t.y;
    }
}

// A row is summed by a group of G lanes (the wave, or half of it for short
// rows).  Lane (q = lane % P, e = lane / P) of the group sums entries
// k0 + e, k0 + e + E, ... (E = G / P) of its row against piece q; after the
// reduction lanes 0..P-1 of the group hold the row's products for columns
// q*W .. q*W+W-1.  The column and value streams are read once, one entry
// per lane (fully coalesced), in chunks of G entries (C chunks in flight),
// and handed to the P lanes that gather each entry's row by shuffles within
// the group.
template <int B, int MODE, int G>
__device__ __forceinline__ void sp_coop_dot(int64_t k0, int64_t k1, int lane, const int32_t* __restrict__ col,
                                            const double* __restrict__ val, const float* __restrict__ hi,
                                            const float* __restrict__ lo, const float* __restrict__ X32,
                                            const double* __restrict__ X,
                                            typename SpAcc<MODE>::T (&acc)[SpCoop<B, MODE>::W]) {
    using AT = typename SpAcc<MODE>::T;
    using Cp = SpCoop<B, MODE>;
    constexpr int W = Cp::W, P = Cp::P, E = G / P;
    constexpr int C = P <= 4 ? 4 : 2;  // chunks of G entries in flight (fewer for wide rows: registers)
    const int q = lane % P, e = lane / P;
#pragma unroll
    for (int c = 0; c < W; c++) acc[c] = 0;
    int64_t k = k0;
    for (; k + G * C <= k1; k += G * C) {
        int32_t jl[C];
        AT vl[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            jl[c] = __builtin_nontemporal_load(col + k + G * c + lane);
            vl[c] = sp_val<MODE>(val, hi, lo, k + G * c + lane);
        }
        AT v[C][P];
        AT x[C][P][W];
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int u = 0; u < P; u++) {
                const int32_t j = __shfl(jl[c], u * E + e, G);
                v[c][u] = __shfl(vl[c], u * E + e, G);
                sp_piece<B, MODE>(X, X32, j, q, x[c][u]);
            }
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int u = 0; u < P; u++)
#pragma unroll
                for (int w = 0; w < W; w++) acc[w] = sp_fma(v[c][u], x[c][u][w], acc[w]);
    }
    // the rest of the row: G-entry chunks, masked, handed out the same way
    for (; k < k1; k += G) {
        const bool in = k + lane < k1;
        const int32_t jl = in ? __builtin_nontemporal_load(col + k + lane) : 0;
        const AT vl = in ? sp_val<MODE>(val, hi, lo, k + lane) : AT(0);
#pragma unroll
        for (int u = 0; u < P; u++) {
            const int32_t j = __shfl(jl, u * E + e, G);
            const AT v = __shfl(vl, u * E + e, G);
            if (k + u * E + e < k1) {
                AT x[W];
                sp_piece<B, MODE>(X, X32, j, q, x);
#pragma unroll
                for (int w = 0; w < W; w++) acc[w] = sp_fma(v, x[w], acc[w]);
            }
        }
    }
#pragma unroll
    for (int off = P; off < G; off <<= 1)
#pragma unroll
        for (int w = 0; w < W; w++) acc[w] += __shfl_xor(acc[w], off, G);
}

template <int B, int MODE, int G = 64>
__global__ __launch_bounds__(256) void k_spmm_lane(int64_t n, const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double* __restrict__ val,
                                                   const float* __restrict__ hi, const float* __restrict__ lo,
                                                   const double* __restrict__ D, const float* __restrict__ X32,
                                                   const double* __restrict__ X, double* __restrict__ Y, int ldy,
                                                   int xc) {
    const int lane = threadIdx.x & (G - 1);
    const int64_t row = xcd_block(xc) * (256 / G) + threadIdx.x / G;
    if (row >= n) return;
    using C = SpCoop<B, MODE>;
    typename SpAcc<MODE>::T acc[C::W];
    sp_coop_dot<B, MODE, G>(rowptr[row], rowptr[row + 1], lane, col, val, hi, lo, X32, X, acc);
    if (lane < C::P) {
        const double d = D[row];
#pragma unroll
        for (int c = 0; c < C::W; c++) {
            const int cc = lane * C::W + c;
            Y[row * ldy + cc] = d * X[row * B + cc] + (double)acc[c];
        }
    }
}

// entries kk .. kk+3 (kk a multiple of 4) as one or two 16-byte loads
template <int MODE>
__device__ __forceinline__ void sp_val4(const double* __restrict__ val, const float* __restrict__ hi,
                                        const float* __restrict__ lo, int64_t kk,
                                        typename SpAcc<MODE>::T (&v)[4]) {
    if constexpr (MODE == 0) {
        const sp_d2 a = __builtin_nontemporal_load(reinterpret_cast<const sp_d2*>(val + kk));
        const sp_d2 b = __builtin_nontemporal_load(reinterpret_cast<const sp_d2*>(val + kk + 2));
        v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y;
    } else if constexpr (MODE == 1) {
        const sp_f4 a = __builtin_nontemporal_load(reinterpret_cast<const sp_f4*>(hi + kk));
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
    } else {
        const sp_f4 a = __builtin_nontemporal_load(reinterpret_cast<const sp_f4*>(hi + kk));
        const sp_f4 b = __builtin_nontemporal_load(reinterpret_cast<const sp_f4*>(lo + kk));
        v[0] = (double)a.x + (double)b.x, v[1] = (double)a.y + (double)b.y;
        v[2] = (double)a.z + (double)b.z, v[3] = (double)a.w + (double)b.w;
    }
}

// single vector: 16 lanes per row, 16 rows per 256-thread block.  Each lane
// takes groups of 4 consecutive entries (a 16-byte column load and 16-byte
// value loads), 8 groups in flight.  The groups start at the row's first
// entry rounded down to a multiple of 4; entries outside the row are loaded
// (the arrays carry kCsrPad entries of padding) but not used.  The gathers
// of x bound this kernel, not the streams: one 8-byte element per cache
// line fetched from L2.
template <int MODE>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int64_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col, const double* __restrict__ val,
                                              const float* __restrict__ hi, const float* __restrict__ lo,
                                              const double* __restrict__ D, const float* __restrict__ x32,
                                              const double* __restrict__ x, double* __restrict__ y, int xc) {
    using AT = typename SpAcc<MODE>::T;
    constexpr int G = 16, U = 8;
    const int sub = threadIdx.x % G;
    const int64_t row = xcd_block(xc) * (256 / G) + threadIdx.x / G;
    const bool live = row < n;
    AT acc = 0;
    if (live) {
        const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
        for (int64_t k = (k0 & ~int64_t(3)) + 4 * sub; k < k1; k += 4 * G * U) {
            sp_i4 c[U];
            AT v[U][4];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t kk = k + 4 * G * u;
                if (kk < k1) {
                    c[u] = __builtin_nontemporal_load(reinterpret_cast<const sp_i4*>(col + kk));
                    sp_val4<MODE>(val, hi, lo, kk, v[u]);
                } else {
                    c[u] = sp_i4{0, 0, 0, 0};
#pragma unroll
                    for (int t = 0; t < 4; t++) v[u][t] = 0;
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t kk = k + 4 * G * u;
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (kk + t >= k0 && kk + t < k1) {
                        const int64_t j = c[u][t];
                        acc = sp_fma(v[u][t], sp_x<MODE>(x, x32, j, j), acc);
                    }
            }
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (live && sub == 0) y[row] = D[row] * x[row] + (double)acc;
}

// ------------------------------------------------------------------------
// Half (upper-triangle) storage: Y = D X + U X + U^T X, U the strict upper
// triangle in CSR.  The row's wave forms (D X + U X)_i in FP64 in a fixed
// order and stores it.  The transposed products U_ij X_i go to row j of an
// int64 fixed-point accumulator by integer atomic adds: integer addition is
// associative, so the accumulated sums, and the product, are the same bit
// for bit whatever order the adds land in.  Column c is scaled by
// s_c = 2^e_c with hnorm * max_i |X_ic| * s_c < 2^62 (hnorm = max_i sum_j
// |H_ij| over the whole row), which bounds every partial sum of every row
// below 2^62; each term is rounded once, to a multiple of 1/s_c.
// k_half_finish adds the accumulator to Y and clears it for the next
// product.  The adds bound the product: about 1e11 atomic adds per second
// (the same with one accumulator copy per XCD added at workgroup scope in
// that XCD's L2), against the gathers' 1.5e11 entries per second of the
// full-storage product.
// ------------------------------------------------------------------------
__device__ __forceinline__ double half_scale(double hnorm, uint64_t cmax_bits) {
    // the margin covers FP32 rounding of values and X in the mixed form
    const double bound = hnorm * __longlong_as_double((long long)cmax_bits) * 1.001;
    if (!(bound > 0.0)) return 1.0;
    int ex;
    (void)frexp(bound, &ex);  // bound < 2^ex
    return ldexp(1.0, max(-1000, min(1000, 62 - ex)));
}

// per-column max |X_rc| as the bit pattern of a non-negative double (such
// patterns order as the values do), atomically maxed into cmax[c]
__global__ __launch_bounds__(256) void k_colmax(int64_t n, int b, const double* __restrict__ X, int ldx,
                                                unsigned long long* __restrict__ cmax) {
    const int c = threadIdx.x % b;
    const int rpb = 256 / b;  // rows per block step
    if (threadIdx.x >= rpb * b) return;
    double m = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * rpb + threadIdx.x / b; r < n; r += (int64_t)gridDim.x * rpb)
        m = fmax(m, fabs(X[r * ldx + c]));
    atomicMax(&cmax[c], (unsigned long long)__double_as_longlong(m));
}

template <int B, int MODE, int G = 64>
__global__ __launch_bounds__(256) void k_spmm_half(int64_t n, const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double* __restrict__ val,
                                                   const float* __restrict__ hi, const float* __restrict__ lo,
                                                   const double* __restrict__ D, const float* __restrict__ X32,
                                                   const double* __restrict__ X, double* __restrict__ Y, int ldy,
                                                   double hnorm, const unsigned long long* __restrict__ cmax,
                                                   unsigned long long* __restrict__ acc_t) {
    using AT = typename SpAcc<MODE>::T;
    using Cp = SpCoop<B, MODE>;
    constexpr int W = Cp::W, P = Cp::P, E = G / P;
    const int lane = threadIdx.x & (G - 1);
    const int64_t row = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
    if (row >= n) return;
    const int q = lane % P, e = lane / P;
    // this lane's columns q*W .. q*W+W-1: scales and the row's own X
    double sc[W];
    AT xi[W];
#pragma unroll
    for (int w = 0; w < W; w++) sc[w] = half_scale(hnorm, cmax[q * W + w]);
    sp_piece<B, MODE>(X, X32, row, q, xi);
    AT acc[W];
#pragma unroll
    for (int w = 0; w < W; w++) acc[w] = 0;
    const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    for (int64_t k = k0; k < k1; k += G) {
        const bool in = k + lane < k1;
        const int32_t jl = in ? __builtin_nontemporal_load(col + k + lane) : 0;
        const AT vl = in ? sp_val<MODE>(val, hi, lo, k + lane) : AT(0);
#pragma unroll
        for (int u = 0; u < P; u++) {
            const int32_t j = __shfl(jl, u * E + e, G);
            const AT v = __shfl(vl, u * E + e, G);
            if (k + u * E + e < k1) {
                AT x[W];
                sp_piece<B, MODE>(X, X32, j, q, x);
#pragma unroll
                for (int w = 0; w < W; w++) {
                    acc[w] = sp_fma(v, x[w], acc[w]);
                    const unsigned long long t =
                        (unsigned long long)__double2ll_rn((double)(v * xi[w]) * sc[w]);
                    (void)__hip_atomic_fetch_add(&acc_t[(int64_t)j * B + q * W + w], t, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
#pragma unroll
    for (int off = P; off < G; off <<= 1)
#pragma unroll
        for (int w = 0; w < W; w++) acc[w] += __shfl_xor(acc[w], off, G);
    if (lane < P) {
        const double d = D[row];
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int cc = lane * W + w;
            Y[row * ldy + cc] = d * X[row * B + cc] + (double)acc[w];
        }
    }
}

// Y[r, c] += acc_t[r, c] / s_c, and acc_t back to zero
template <int B>
__global__ void k_half_finish(int64_t n, unsigned long long* __restrict__ acc_t, double hnorm,
                              const unsigned long long* __restrict__ cmax, double* __restrict__ Y, int ldy) {
    const int64_t total = n * B;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / B;
        const int c = (int)(t - r * B);
        const double inv = 1.0 / half_scale(hnorm, cmax[c]);  // a power of two: exact
        Y[r * ldy + c] += (double)(long long)acc_t[t] * inv;
        acc_t[t] = 0ull;
    }
}

// columns c0 .. c0+bc-1 of X (row-major, ldx) into a zero-padded n x B
// block, and back
__global__ void k_pack_cols(const double* __restrict__ X, int ldx, int64_t n, int bc, int B, double* __restrict__ out) {
    const int64_t total = n * B;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / B;
        const int c = (int)(t - r * B);
        out[t] = c < bc ? X[r * ldx + c] : 0.0;
    }
}
__global__ void k_unpack_cols(const double* __restrict__ in, int B, int64_t n, int bc, double* __restrict__ Y,
                              int ldy) {
    const int64_t total = n * bc;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / bc;
        const int c = (int)(t - r * bc);
        Y[r * ldy + c] = in[r * B + c];
    }
}

// Entries of U in the columns of requested rows (mark[col] >= 0): the lower
// half of those rows.  count pass (out_req null): their number in
// counter[0]; fill pass: (request, source row, value) at counter[0]'s
// successive positions.
__global__ void k_lower_of(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                           const double* __restrict__ val, const float* __restrict__ hi, const float* __restrict__ lo,
                           const int32_t* __restrict__ mark, unsigned long long* __restrict__ counter,
                           int32_t* __restrict__ out_req, int32_t* __restrict__ out_row, double* __restrict__ out_val,
                           unsigned long long cap) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    for (int64_t k = rowptr[r] + (threadIdx.x & 63); k < rowptr[r + 1]; k += 64) {
        const int32_t q = mark[col[k]];
        if (q < 0) continue;
        if (!out_req) {
            atomicAdd(&counter[0], 1ull);
            continue;
        }
        const unsigned long long pos = atomicAdd(&counter[0], 1ull);
        if (pos >= cap) continue;
        out_req[pos] = q;
        out_row[pos] = (int32_t)r;
        out_val[pos] = val ? val[k] : (double)hi[k] + (double)lo[k];
    }
}

// pack / unpack the mixed form's hi, lo pair into one 64-bit word (sort payload)
__global__ void k_pack_hilo(int64_t nnz, const float* __restrict__ hi, const float* __restrict__ lo,
                            uint64_t* __restrict__ out) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
        out[k] = (uint64_t)__float_as_uint(hi[k]) | ((uint64_t)__float_as_uint(lo[k]) << 32);
}
__global__ void k_unpack_hilo(int64_t nnz, const uint64_t* __restrict__ in, float* __restrict__ hi,
                              float* __restrict__ lo) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        hi[k] = __uint_as_float((uint32_t)in[k]);
        lo[k] = __uint_as_float((uint32_t)(in[k] >> 32));
    }
}

// row offsets of a slice of rows relative to the slice's first entry
__global__ void k_rel_offsets(const int64_t* __restrict__ rp, int64_t m, int64_t base, uint32_t* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) out[i] = (uint32_t)(rp[i] - base);
}

// ------------------------------------------------------------------------
// solver order of the basis: states sorted by their quanta read as one
// integer with the highest mode most significant, so that states adjacent
// in the order share most quanta and their Hamiltonian rows gather nearby
// rows of X in the block products
// ------------------------------------------------------------------------
__global__ void k_iota32(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// out[r] = 64-bit chunk c (quanta of modes 8c..8c+7) of the key of state perm[r]
__global__ void k_key_chunk(const uint32_t* __restrict__ keys, int wpk, const int32_t* __restrict__ perm, int64_t n,
                            int c, uint64_t* __restrict__ out) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint32_t* k = keys + (int64_t)perm[r] * wpk;
    out[r] = ((uint64_t)k[2 * c + 1] << 32) | (uint64_t)k[2 * c];
}

__global__ void k_invert32(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ inv) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r < n) inv[perm[r]] = (int32_t)r;
}

// V[r, c] += T[perm[r], c] for the rows whose state existed in the previous
// basis (perm[r] < rows); T column-major with leading dimension rows
__global__ void k_add_perm_rows(double* __restrict__ V, int64_t ldv, int64_t n, int k, const double* __restrict__ T,
                                int64_t rows, const int32_t* __restrict__ perm) {
    const int c = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t src = perm[r];
        if (src < rows) V[r + (int64_t)c * ldv] += T[src + (int64_t)c * rows];
    }
}

// ------------------------------------------------------------------------
// Davidson helpers
// ------------------------------------------------------------------------
// partial column sums of (HX - X theta)^2 over row blocks; cols <= 256
__global__ __launch_bounds__(256) void k_resid_partial(int64_t n, int b, const double* __restrict__ HX,
                                                       const double* __restrict__ X, int ld,
                                                       const double* __restrict__ theta, int64_t rows_per_blk,
                                                       double* __restrict__ part) {
    const int c = threadIdx.x;
    if (c >= b) return;
    const int64_t r0 = blockIdx.x * rows_per_blk;
    const int64_t r1 = min(n, r0 + rows_per_blk);
    const double th = theta[c];
    double s = 0.0;
    for (int64_t r = r0; r < r1; r++) {
        const double x = HX[r * ld + c] - th * X[r * ld + c];
        s = fma(x, x, s);
    }
    part[blockIdx.x * (int64_t)b + c] = s;
}

__global__ void k_reduce_cols(const double* __restrict__ part, int nblk, int b, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b) return;
    double s = 0.0;
    for (int i = 0; i < nblk; i++) s += part[(int64_t)i * b + c];
    out[c] = s;
}

// T[:, jj] = (HX[:, J[jj]] - theta_J X[:, J[jj]]) / (theta_J - D)
__global__ void k_precond(int64_t n, int k, const int* __restrict__ J, const double* __restrict__ HX,
                          const double* __restrict__ X, int ldx, const double* __restrict__ theta,
                          const double* __restrict__ D, double* __restrict__ T, int ldt) {
    int64_t total = n * k;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / k;
        const int jj = (int)(t - r * k);
        const int j = J[jj];
        const double th = theta[j];
        const double res = HX[r * ldx + j] - th * X[r * ldx + j];
        double den = th - D[r];
        if (fabs(den) < 1e-6) den = den < 0 ? -1e-6 : 1e-6;
        T[r * ldt + jj] = res / den;
    }
}

__global__ void k_symmetrize(double* G, int m, int ld) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i >= m || j >= m || i <= j) return;
    double a = G[(int64_t)j * ld + i], b = G[(int64_t)i * ld + j];
    double s = 0.5 * (a + b);
    G[(int64_t)j * ld + i] = s;
    G[(int64_t)i * ld + j] = s;
}

// dense symmetric matrix (column-major n x n) from the CSR
// (half storage, sym: each stored entry also goes to its mirror position)
__global__ void k_densify(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                          const double* __restrict__ val, const float* __restrict__ hi, const float* __restrict__ lo,
                          const double* __restrict__ D, double* __restrict__ A, int sym) {
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    for (int64_t k = rowptr[i] + threadIdx.x; k < rowptr[i + 1]; k += blockDim.x) {
        const double v = val ? val[k] : (double)hi[k] + (double)lo[k];
        A[(int64_t)col[k] * n + i] = v;
        if (sym) A[i * n + col[k]] = v;
    }
    if (threadIdx.x == 0) A[i * n + i] = D[i];
}

// column-major (ld) eigenvector block -> row-major n x ncol
__global__ void k_colmajor_to_rows(const double* __restrict__ A, int64_t lda, int64_t n, int ncol,
                                   double* __restrict__ X, const int32_t* __restrict__ rowmap) {
    int64_t total = n * ncol;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / ncol;
        const int c = (int)(t - r * ncol);
        X[(rowmap ? (int64_t)rowmap[r] : r) * ncol + c] = A[(int64_t)c * lda + r];
    }
}

// ------------------------------------------------------------------------
// column-major block helpers of the eigensolver
// ------------------------------------------------------------------------
// B (row-major, ldb) <- A (column-major, lda) for an n x k block; 32x32 LDS
// tiles so both sides are read and written along contiguous addresses.
template <class TO>
__global__ __launch_bounds__(256) void k_cm_to_rm(const double* __restrict__ A, int64_t lda, int64_t n, int k,
                                                  TO* __restrict__ B, int ldb, const int32_t* __restrict__ rowmap) {
    __shared__ double tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.x * 32;
    const int c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + tx;
        const int c = c0 + j;
        tile[j][tx] = (r < n && c < k) ? A[r + (int64_t)c * lda] : 0.0;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + j;
        const int c = c0 + tx;
        if (r < n && c < k) B[(rowmap ? (int64_t)rowmap[r] : r) * ldb + c] = (TO)tile[tx][j];
    }
}

// A (column-major, lda) <- B (row-major, ldb)
// (A2, when not null, receives a second copy with the same leading dim)
__global__ __launch_bounds__(256) void k_rm_to_cm(const double* __restrict__ B, int ldb, int64_t n, int k,
                                                  double* __restrict__ A, int64_t lda, double* __restrict__ A2) {
    __shared__ double tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.x * 32;
    const int c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + j;
        const int c = c0 + tx;
        tile[j][tx] = (r < n && c < k) ? B[r * ldb + c] : 0.0;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + tx;
        const int c = c0 + j;
        if (r < n && c < k) {
            A[r + (int64_t)c * lda] = tile[tx][j];
            if (A2) A2[r + (int64_t)c * lda] = tile[tx][j];
        }
    }
}

// partial sums over row chunks of (HX - theta X)^2 for column blockIdx.y
__global__ __launch_bounds__(256) void k_resid_cm(int64_t n, const double* __restrict__ HX,
                                                  const double* __restrict__ X, int64_t ld,
                                                  const double* __restrict__ theta, int64_t rows_per_blk,
                                                  double* __restrict__ part) {
    __shared__ double red[256];
    const int c = blockIdx.y;
    const int64_t r0 = blockIdx.x * rows_per_blk;
    const int64_t r1 = min(n, r0 + rows_per_blk);
    const double th = theta[c];
    double s = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const double x = HX[r + c * ld] - th * X[r + c * ld];
        s = fma(x, x, s);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)c * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_sum_parts(const double* __restrict__ part, int nblk, int b, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= b) return;
    double s = 0.0;
    for (int i = 0; i < nblk; i++) s += part[(int64_t)c * nblk + i];
    out[c] = s;
}

// T[:, jj] = (HX[:, J[jj]] - theta_J X[:, J[jj]]) / (theta_J - D), column-major
__global__ void k_precond_cm(int64_t n, int k, const int* __restrict__ J, const double* __restrict__ HX,
                             const double* __restrict__ X, int64_t ldx, const double* __restrict__ theta,
                             const double* __restrict__ D, double* __restrict__ T, int64_t ldt) {
    const int jj = blockIdx.y;
    const int j = J[jj];
    const double th = theta[j];
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double res = HX[r + j * ldx] - th * X[r + j * ldx];
        double den = th - D[r];
        if (fabs(den) < 1e-6) den = den < 0 ? -1e-6 : 1e-6;
        T[r + jj * ldt] = res / den;
    }
}

// G (column-major, ld) = diag(theta[0..m)) on an m x m block
__global__ void k_diag_block(double* G, int m, int ld, const double* __restrict__ theta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < m && j < m) G[i + (int64_t)j * ld] = (i == j) ? theta[i] : 0.0;
}

// C[c * ldc + a] = sum over nb partial p x q products (each column-major,
// ld p, consecutive), in a fixed order
__global__ void k_sum_batches(const double* __restrict__ part, int nb, int p, int q, double* __restrict__ C,
                              int64_t ldc) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)p * q) return;
    const int c = (int)(t / p), a = (int)(t - (int64_t)c * p);
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[(int64_t)b * p * q + t];
    C[c * ldc + a] = s;
}

// Ritz residual norms from the Krylov-Schur relation: out[j] =
// || B s_j(last bs rows) ||^2 with B[c * bs + a] = (V_m^T W_last)[a][c] and
// s_j column j of the eigenvector matrix Gw (ld m)
__global__ void k_relation_resid(const double* __restrict__ B, int bs, const double* __restrict__ Gw, int m, int nev,
                                 double* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nev) return;
    const double* sj = Gw + (int64_t)j * m + (m - bs);
    double sq = 0.0;
    for (int a = 0; a < bs; a++) {
        double v = 0.0;
        for (int c = 0; c < bs; c++) v += B[c * bs + a] * sj[c];
        sq += v * v;
    }
    out[j] = sq;
}

// Davidson start block: columns c < k0 from the resident row-major vectors
// Xr (basis order; stored row r is basis state perm[r] when perm is given),
// the others unit vectors on the stored rows idx.
__global__ void k_start_block(int64_t n, int b, int k0, const double* __restrict__ Xr, int64_t xrows, int xcols,
                              const int64_t* __restrict__ idx, double* __restrict__ V, int64_t ldv,
                              const int32_t* __restrict__ perm) {
    const int c = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        const int64_t br = perm ? perm[r] : r;
        if (c < k0)
            v = Xr && br < xrows ? Xr[br * xcols + c] : 0.0;
        else
            v = (idx[c - k0] == r) ? 1.0 : 0.0;
        V[r + c * ldv] = v;
    }
}

// ------------------------------------------------------------------------
// tall-skinny products of the Krylov basis (n >> p, q)
// ------------------------------------------------------------------------
// partial C = A^T B over one row chunk: A n x p, B n x q, column-major.
// Block: AW columns of A x QW columns of B (AW * QW = 1024: 64 x 16, or
// 128 x 8 when q <= 8); 32-row tiles staged in LDS with loads along the rows
// (coalesced), 4 outputs per thread.
// part layout: [chunk][c][a] (c < q, a < p)
template <int AW, int QW>
__global__ __launch_bounds__(256) void k_tsm_tn(int64_t n, int p, int q, const double* __restrict__ A, int64_t lda,
                                                const double* __restrict__ B, int64_t ldb, int64_t rows_per_chunk,
                                                double* __restrict__ part) {
    static_assert(AW * QW == 1024, "256 threads x 4 outputs");
    constexpr int NA = AW / 8, NB = QW / 8;  // staged values per thread
    __shared__ double As[32][AW + 1];
    __shared__ double Bs[32][QW + 1];
    const int t = threadIdx.x;
    const int a0 = blockIdx.y * AW, c0 = blockIdx.z * QW;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = min(n, r0 + rows_per_chunk);
    const int aa = t % AW, cg = t / AW;  // outputs (aa, 4*cg .. 4*cg+3)
    // each thread stages NA values of A and NB of B per 32-row tile; the
    // next tile is fetched into registers while the current one is consumed
    const int ai = t & 31, aj = t >> 5;  // rows ai, columns aj + 8 u
    double ra[NA], rb[NB];
    auto fetch = [&](int64_t rt) {
        const int64_t r = rt + ai;
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int j = aj + 8 * u;
            ra[u] = (r < r1 && a0 + j < p) ? A[r + (int64_t)(a0 + j) * lda] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int j = aj + 8 * u;
            rb[u] = (r < r1 && c0 + j < q) ? B[r + (int64_t)(c0 + j) * ldb] : 0.0;
        }
    };
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (r0 < r1) fetch(r0);
    for (int64_t rt = r0; rt < r1; rt += 32) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NA; u++) As[ai][aj + 8 * u] = ra[u];
#pragma unroll
        for (int u = 0; u < NB; u++) Bs[ai][aj + 8 * u] = rb[u];
        __syncthreads();
        if (rt + 32 < r1) fetch(rt + 32);
#pragma unroll 8
        for (int i = 0; i < 32; i++) {
            const double x = As[i][aa];
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] = fma(x, Bs[i][4 * cg + u], acc[u]);
        }
    }
    if (a0 + aa < p)
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int c = c0 + 4 * cg + u;
            if (c < q) part[((int64_t)blockIdx.x * q + c) * p + a0 + aa] = acc[u];
        }
}

// The same partial products for wide B (q > 16): a 64 x 64 output tile per
// block, 4 x 4 outputs per thread, so A is read q/64 times instead of q/16.
__global__ __launch_bounds__(256) void k_tsm_tn_wide(int64_t n, int p, int q, const double* __restrict__ A,
                                                     int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                     int64_t rows_per_chunk, double* __restrict__ part) {
    __shared__ double As[32][64 + 2];
    __shared__ double Bs[32][64 + 2];
    const int t = threadIdx.x;
    const int a0 = blockIdx.y * 64, c0 = blockIdx.z * 64;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
    const int64_t r1 = min(n, r0 + rows_per_chunk);
    const int ta = t & 15, tc = t >> 4;  // outputs (4 ta + u, 4 tc + v)
    const int ai = t & 31, aj = t >> 5;  // staging: rows ai, columns aj + 8 s
    double ra[8], rb[8];
    auto fetch = [&](int64_t rt) {
        const int64_t r = rt + ai;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = aj + 8 * u;
            ra[u] = (r < r1 && a0 + j < p) ? A[r + (int64_t)(a0 + j) * lda] : 0.0;
            rb[u] = (r < r1 && c0 + j < q) ? B[r + (int64_t)(c0 + j) * ldb] : 0.0;
        }
    };
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] = 0.0;
    if (r0 < r1) fetch(r0);
    for (int64_t rt = r0; rt < r1; rt += 32) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; u++) {
            As[ai][aj + 8 * u] = ra[u];
            Bs[ai][aj + 8 * u] = rb[u];
        }
        __syncthreads();
        if (rt + 32 < r1) fetch(rt + 32);
#pragma unroll 4
        for (int i = 0; i < 32; i++) {
            double x[4], y[4];
#pragma unroll
            for (int u = 0; u < 4; u++) x[u] = As[i][4 * ta + u];
#pragma unroll
            for (int v = 0; v < 4; v++) y[v] = Bs[i][4 * tc + v];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] = fma(x[u], y[v], acc[u][v]);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int a = a0 + 4 * ta + u;
        if (a >= p) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int c = c0 + 4 * tc + v;
            if (c < q) part[((int64_t)blockIdx.x * q + c) * p + a] = acc[u][v];
        }
    }
}

// C[a + c*ldc] = sum over chunks: 16 outputs per block, the chunks split
// over 16 lanes groups (chunk w, w + 16, ...), combined in a fixed order
// (deterministic)
__global__ __launch_bounds__(256) void k_tsm_reduce(const double* __restrict__ part, int nchunk, int p, int q,
                                                    double* __restrict__ C, int ldc) {
    __shared__ double red[16][17];
    const int o = threadIdx.x & 15, w = threadIdx.x >> 4;
    const int64_t t = (int64_t)blockIdx.x * 16 + o;
    const int64_t pq = (int64_t)p * q;
    double s0 = 0.0, s1 = 0.0;
    if (t < pq) {
        int k = w;
        for (; k + 16 < nchunk; k += 32) {
            s0 += part[(int64_t)k * pq + t];
            s1 += part[(int64_t)(k + 16) * pq + t];
        }
        if (k < nchunk) s0 += part[(int64_t)k * pq + t];
    }
    red[w][o] = s0 + s1;
    __syncthreads();
    if (w == 0 && t < pq) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 16; i++) s += red[i][o];
        const int c = (int)(t / p), a = (int)(t - (int64_t)c * p);
        C[a + (int64_t)c * ldc] = s;
    }
}

// Y = beta Y + alpha A S: A n x p (lda), S p x q (lds, col-major) staged in
// LDS (p*q <= 4096), one row per thread, q <= 16 per launch tile.  gate
// (null: none): nothing is done when *gate is 0.
template <int QW, int U>
__global__ __launch_bounds__(256) void k_tsm_nn(int64_t n, int p, int q, double alpha, const double* __restrict__ A,
                                                int64_t lda, const double* __restrict__ S, int lds, double beta,
                                                double* __restrict__ Y, int64_t ldy,
                                                const double* __restrict__ gate) {
    // QW output columns per workgroup (p <= 256 rows of S in LDS); U columns
    // of A loaded before they are used, so that each wave keeps U loads in
    // flight.  Every output sums a = 0, 1, ... in order, whatever U.
    __shared__ double Ss[256 * QW];
    if (gate && *gate == 0.0) return;
    const int c0 = blockIdx.y * QW;
    const int qq = min(QW, q - c0);
    for (int idx = threadIdx.x; idx < p * QW; idx += blockDim.x) {
        const int a = idx / QW, c = idx - a * QW;
        Ss[idx] = c < qq ? S[a + (int64_t)(c0 + c) * lds] : 0.0;
    }
    __syncthreads();
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double acc[QW];
#pragma unroll
        for (int c = 0; c < QW; c++) acc[c] = 0.0;
        const double* Ar = A + r;
        int a = 0;
        for (; a + U <= p; a += U) {
            double x[U];
#pragma unroll
            for (int u = 0; u < U; u++) x[u] = Ar[(int64_t)(a + u) * lda];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int c = 0; c < QW; c++) acc[c] = fma(x[u], Ss[(a + u) * QW + c], acc[c]);
        }
        for (; a < p; a++) {
            const double x = Ar[(int64_t)a * lda];
#pragma unroll
            for (int c = 0; c < QW; c++) acc[c] = fma(x, Ss[a * QW + c], acc[c]);
        }
        for (int c = 0; c < qq; c++) {
            double* y = Y + r + (int64_t)(c0 + c) * ldy;
            *y = (beta == 0.0 ? 0.0 : beta * *y) + alpha * acc[c];
        }
    }
}

// ------------------------------------------------------------------------
// Ritz rotation X = A S on the matrix cores (v_mfma_f64_16x16x4f64): A is
// n x p (column-major, lda; the Krylov basis or its product block), S is
// p x q (column-major, lds; the kept eigenvectors of the projected matrix),
// X is n x q (column-major, ldx).  A workgroup of four waves takes 64 rows
// and CT column tiles of 16 (all q columns for q <= 16 CT), so A is read from
// HBM once; the wave's 16 rows x 16 CT columns accumulate in 4 CT registers
// per lane.  S is staged through LDS in chunks of 32 of its rows.  MFMA
// operands: lane l holds A[i][k] and S[k][j] with i = j = l % 16, k = l / 16,
// and gets X[l / 16 + 4 r][l % 16] in register r.  Each product is summed
// in a fixed order (k ascending in steps of 4 inside the MFMA), so the
// rotation is reproducible bit for bit.
// ------------------------------------------------------------------------
typedef double rot_d4 __attribute__((ext_vector_type(4)));
constexpr int kRotKC = 16;  // rows of S per LDS chunk

// S (p x q, column-major, lds) as launch tiles of TW = 16 CT columns, each
// pk x TW with zeros beyond p rows and q columns (pk = p rounded up to
// kRotKC): every LDS chunk is then one contiguous block.  Within a row k of
// a tile, column 16 t + j sits at position CT j + t, so that the lane of
// column j reads its CT operands of the row as CT / 2 16-byte words.
__global__ void k_rot_stage(const double* __restrict__ S, int lds, int p, int q, int pk, int CT, int ntile,
                            double* __restrict__ Sr) {
    const int TW = 16 * CT;
    const int64_t total = (int64_t)ntile * pk * TW;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t tile = t / ((int64_t)pk * TW);
        const int64_t rem = t - tile * pk * TW;
        const int k = (int)(rem / TW), pos = (int)(rem - (int64_t)k * TW);
        const int j = pos / CT, tt = pos - j * CT;
        const int c = (int)tile * TW + 16 * tt + j;
        Sr[t] = k < p && c < q ? S[k + (int64_t)c * lds] : 0.0;
    }
}

template <int CT>
__global__ __launch_bounds__(256) void k_rot_mfma(int64_t n, int p, int pk, int q, const double* __restrict__ A,
                                                  int64_t lda, const double* __restrict__ Sr, double* __restrict__ X,
                                                  int64_t ldx) {
    constexpr int TW = 16 * CT;
    constexpr int SPT = kRotKC * TW / 256;  // S elements each thread stages per chunk
    __shared__ __align__(16) double Ss[kRotKC * TW];
    const int l = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int c0 = blockIdx.y * TW;  // first column of this launch tile
    const double* St = Sr + (int64_t)blockIdx.y * pk * TW;
    const int64_t rbase = (int64_t)blockIdx.x * 64 + w * 16;
    const int64_t arow = rbase + (l & 15);
    const bool rin = arow < n;
    const int kl = l >> 4;
    rot_d4 acc[CT];
#pragma unroll
    for (int t = 0; t < CT; t++) acc[t] = rot_d4{0.0, 0.0, 0.0, 0.0};
    // chunk 0 in registers; each step stages the next one while the
    // current one is multiplied
    double a_cur[kRotKC / 4], a_nxt[kRotKC / 4], s_nxt[SPT];
#pragma unroll
    for (int s = 0; s < kRotKC / 4; s++) {
        const int gk = 4 * s + kl;
        a_cur[s] = rin && gk < p ? A[arow + (int64_t)gk * lda] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < SPT; i++) s_nxt[i] = St[threadIdx.x + 256 * i];
    for (int k0 = 0; k0 < pk; k0 += kRotKC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SPT; i++) Ss[threadIdx.x + 256 * i] = s_nxt[i];
        __syncthreads();
        const int k1 = k0 + kRotKC;
        if (k1 < pk) {
#pragma unroll
            for (int s = 0; s < kRotKC / 4; s++) {
                const int gk = k1 + 4 * s + kl;
                a_nxt[s] = rin && gk < p ? A[arow + (int64_t)gk * lda] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < SPT; i++) s_nxt[i] = St[(int64_t)k1 * TW + threadIdx.x + 256 * i];
        }
#pragma unroll
        for (int s = 0; s < kRotKC / 4; s++) {
            const double2* srow = reinterpret_cast<const double2*>(Ss + (4 * s + kl) * TW + (l & 15) * CT);
#pragma unroll
            for (int t = 0; t < CT; t += 2) {
                const double2 b = srow[t / 2];
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_cur[s], b.x, acc[t], 0, 0, 0);
                acc[t + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_cur[s], b.y, acc[t + 1], 0, 0, 0);
            }
        }
#pragma unroll
        for (int s = 0; s < kRotKC / 4; s++) a_cur[s] = a_nxt[s];
    }
#pragma unroll
    for (int t = 0; t < CT; t++) {
        const int col = c0 + 16 * t + (l & 15);
        if (col >= q) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t row = rbase + kl + 4 * r;
            if (row < n) X[row + (int64_t)col * ldx] = acc[t][r];
        }
    }
}

// X = A S (see k_rot_mfma); work holds the staged S
static void rot_mfma(hipStream_t s, int64_t n, int p, int q, const double* A, int64_t lda, const double* S, int lds,
                     double* X, int64_t ldx, DevBuf& work) {
    const unsigned gx = grid_for(n, 64);
    const int pk = (p + kRotKC - 1) / kRotKC * kRotKC;
    const int TW = q <= 128 ? 128 : 256;
    const int ntile = (q + TW - 1) / TW;
    work.ensure((size_t)ntile * pk * TW * sizeof(double));
    k_rot_stage<<<grid_for((int64_t)ntile * pk * TW, 256), 256, 0, s>>>(S, lds, p, q, pk, TW / 16, ntile,
                                                                        work.as<double>());
    HIPCHK(hipGetLastError());
    if (TW == 128)
        k_rot_mfma<8><<<dim3(gx, ntile), 256, 0, s>>>(n, p, pk, q, A, lda, work.as<double>(), X, ldx);
    else
        k_rot_mfma<16><<<dim3(gx, ntile), 256, 0, s>>>(n, p, pk, q, A, lda, work.as<double>(), X, ldx);
    HIPCHK(hipGetLastError());
}

// R (p x k, column-major, ld p) holds V^T T in rows 0..m0 and T^T T in rows
// m0..m0+k.  *gate = 1 if any projection coefficient exceeds rel times the
// smallest column norm of T, else 0 (one workgroup)
__global__ __launch_bounds__(256) void k_gate_proj(const double* __restrict__ R, int m0, int k, int p, double rel,
                                                   double* __restrict__ gate) {
    __shared__ int any;
    __shared__ double thr;
    if (threadIdx.x == 0) {
        double mn = R[m0];
        for (int c = 1; c < k; c++) mn = fmin(mn, R[(int64_t)c * p + m0 + c]);
        thr = rel * sqrt(fmax(mn, 0.0));
        any = 0;
    }
    __syncthreads();
    int mine = 0;
    for (int64_t i = threadIdx.x; i < (int64_t)m0 * k; i += blockDim.x) {
        const int c = (int)(i / m0), r = (int)(i - (int64_t)c * m0);
        mine |= !(fabs(R[(int64_t)c * p + r]) <= thr);
    }
    if (mine) atomicOr(&any, 1);
    __syncthreads();
    if (threadIdx.x == 0) *gate = any ? 1.0 : 0.0;
}

// T[:, 0..kk) = T[:, 0..k) Z in place (T column-major n x k, ld; Z k x kk
// column-major, kk <= k <= KMAX): each row is independent.  status (null:
// none): nothing is done when *status is not 0.
template <int KMAX>
__global__ __launch_bounds__(256) void k_rowmix(int64_t n, double* __restrict__ T, int64_t ld, int k, int kk,
                                                const double* __restrict__ Z, const int* __restrict__ status) {
    __shared__ double Zs[KMAX * KMAX];
    if (status && *status != 0) return;
    for (int i = threadIdx.x; i < k * kk; i += blockDim.x) Zs[i] = Z[i];
    __syncthreads();
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double t[KMAX];
#pragma unroll
        for (int a = 0; a < KMAX; a++) t[a] = a < k ? T[r + a * ld] : 0.0;
#pragma unroll
        for (int c = 0; c < KMAX; c++) {
            if (c < kk) {
                double acc = 0.0;
#pragma unroll
                for (int a = 0; a < KMAX; a++)
                    if (a < k) acc = fma(t[a], Zs[c * k + a], acc);
                T[r + c * ld] = acc;
            }
        }
    }
}

// the largest basis offset m0 that k_orth_block stages in LDS (8 x 1024
// doubles, 64 KB); steps beyond it take the host path
constexpr int kOrthRows = 1024;

// One Lanczos block's orthonormalisation finished on the device: the common
// case of Engine::ortho_cm's Lanczos path, so that a cycle's steps need no
// host round trip.  R (p x k, ld p, k <= 8) = [V^T T; T^T T] with the
// projection gate at R[p k] (k_gate_proj); when the gate is set, V S has
// already been taken from T (k_tsm_nn), so the corrected block's Gram matrix
// is G = T^T T - S^T S.  With D = diag(G)^-1/2 and the Cholesky factor
// D G D = L L^T, Z = D L^-T makes T Z orthonormal (Cholesky QR), written to
// Z (k x k, upper triangular).  *status = 0 when one pass finishes the
// block; 1 when it does not (a column lost more than half its squared norm
// to the projection, or the Frobenius condition bound |DGD| |(DGD)^-1| of
// the scaled Gram matrix reaches 1e4, which covers the host's test that its
// eigenvalues lie within a factor 1e4): then Z is not written, and the host
// repeats the step on its general path (SVQB).  decline != 0 declines every
// block (a test of that path).
__global__ __launch_bounds__(256) void k_orth_block(const double* __restrict__ R, int m0, int k, int p,
                                                    double* __restrict__ Z, int* __restrict__ status,
                                                    int decline) {
    __shared__ double A[8][8];  // A[c][i]: column c, row i
    __shared__ double Li[8][8];  // L^-1 (lower), Li[r][c]: row r, column c
    __shared__ double before[8], dsc[8];
    __shared__ double inv2[64];
    __shared__ double gf2;  // squared Frobenius norm of the scaled Gram matrix
    __shared__ int stop;
    __shared__ double Rs[8 * kOrthRows];  // the projection coefficients S (m0 x k), column by column
    const int t = threadIdx.x;
    if (decline) {  // every step to the host's general path (a test of that path)
        if (t == 0) *status = 1;
        return;
    }
    const bool gated = R[(int64_t)p * k] != 0.0;
    if (gated) {
        // S^T S: S staged in LDS, four lanes per entry (quarters of the rows,
        // summed in a fixed order)
        for (int idx = t; idx < k * m0; idx += blockDim.x) {
            const int c = idx / m0, r = idx - c * m0;
            Rs[c * m0 + r] = R[(int64_t)c * p + r];
        }
        __syncthreads();
        const int e = t >> 2, part = t & 3;
        double d = 0.0;
        if (e < k * k) {
            const int c = e / k, i = e - c * k;
            const int r0 = (m0 * part) / 4, r1 = (m0 * (part + 1)) / 4;
            for (int r = r0; r < r1; r++) d += Rs[c * m0 + r] * Rs[i * m0 + r];
        }
        d += __shfl_xor(d, 1, 4);
        d += __shfl_xor(d, 2, 4);
        if (e < k * k && part == 0) {
            const int c = e / k, i = e - c * k;
            A[c][i] = R[(int64_t)c * p + m0 + i] - d;
        }
    } else if (t < k * k) {
        const int c = t / k, i = t - c * k;
        A[c][i] = R[(int64_t)c * p + m0 + i];
    }
    if (t < k) before[t] = R[(int64_t)t * p + m0 + t];
    __syncthreads();
    if (t == 0) {
        int bad = 0;
        for (int i = 0; i < k; i++) {
            if (A[i][i] < 0.5 * before[i]) bad = 1;
            if (!(A[i][i] > 0.0)) bad = 1;
        }
        stop = bad;
    }
    __syncthreads();
    if (stop) {
        if (t == 0) *status = 1;
        return;
    }
    if (t < k) dsc[t] = 1.0 / sqrt(A[t][t]);
    __syncthreads();
    if (t < k * k) {
        const int c = t / k, i = t - c * k;
        A[c][i] *= dsc[i] * dsc[c];
    }
    __syncthreads();
    if (k == 8) {
        // The eigensolver's block size: the same steps as below, in the same
        // order, by one lane on registers (fully unrolled), since the serial
        // chain through LDS cost most of the kernel's time.
        if (t == 0) {
            double a[8][8], li[8][8];
#pragma unroll
            for (int c = 0; c < 8; c++)
#pragma unroll
                for (int i = 0; i < 8; i++) a[c][i] = A[c][i];
            double f2 = 0.0;
#pragma unroll
            for (int c = 0; c < 8; c++)
#pragma unroll
                for (int i = 0; i < 8; i++) f2 += a[c][i] * a[c][i];
            int bad = 0;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                double dd = a[c][c];
#pragma unroll
                for (int r = 0; r < c; r++) dd -= a[r][c] * a[r][c];
                if (!(dd > 0.0)) bad = 1;  // nothing computed after this is used
                const double l = sqrt(dd);
                a[c][c] = l;
#pragma unroll
                for (int i = c + 1; i < 8; i++) {
                    double v = a[c][i];
#pragma unroll
                    for (int r = 0; r < c; r++) v -= a[r][i] * a[r][c];
                    a[c][i] = v / l;
                }
            }
#pragma unroll
            for (int c = 0; c < 8; c++)
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    if (r < c) {
                        li[r][c] = 0.0;
                        continue;
                    }
                    double v = r == c ? 1.0 : 0.0;
#pragma unroll
                    for (int q = c; q < r; q++) v -= a[q][r] * li[q][c];
                    li[r][c] = v / a[r][r];
                }
            double i2 = 0.0;
#pragma unroll
            for (int i = 0; i < 8; i++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    double v = 0.0;
#pragma unroll
                    for (int r = 0; r < 8; r++) v += li[r][i] * li[r][j];
                    i2 += v * v;
                }
            // cond_2 <= |DGD|_F |(DGD)^-1|_F
            const int st = bad || !(sqrt(f2 * i2) < 1e4);
            *status = st;
            if (!st)
#pragma unroll
                for (int c = 0; c < 8; c++)
#pragma unroll
                    for (int i = 0; i < 8; i++) Z[c * 8 + i] = dsc[i] * li[c][i];
        }
        return;
    }
    if (t == 0) {
        double f2 = 0.0;
        for (int c = 0; c < k; c++)
            for (int i = 0; i < k; i++) f2 += A[c][i] * A[c][i];
        gf2 = f2;
        // Cholesky of the scaled Gram matrix (lower L in A's lower triangle,
        // A[c][i] = L[i][c] for i >= c), then L^-1 by forward substitution
        int bad = 0;
        for (int c = 0; c < k && !bad; c++) {
            double dd = A[c][c];
            for (int r = 0; r < c; r++) dd -= A[r][c] * A[r][c];
            if (!(dd > 0.0)) {
                bad = 1;
                break;
            }
            const double l = sqrt(dd);
            A[c][c] = l;
            for (int i = c + 1; i < k; i++) {
                double v = A[c][i];
                for (int r = 0; r < c; r++) v -= A[r][i] * A[r][c];
                A[c][i] = v / l;
            }
        }
        if (!bad)
            for (int c = 0; c < k; c++)
                for (int r = 0; r < k; r++) {
                    if (r < c) {
                        Li[r][c] = 0.0;
                        continue;
                    }
                    double v = r == c ? 1.0 : 0.0;
                    for (int q = c; q < r; q++) v -= A[q][r] * Li[q][c];
                    Li[r][c] = v / A[r][r];
                }
        stop = bad;
    }
    __syncthreads();
    if (stop) {
        if (t == 0) *status = 1;
        return;
    }
    // (D G D)^-1 = L^-T L^-1: its squared entries, one per lane
    if (t < k * k) {
        const int i = t / k, j = t - i * k;
        double v = 0.0;
        for (int r = 0; r < k; r++) v += Li[r][i] * Li[r][j];
        inv2[t] = v * v;
    }
    __syncthreads();
    if (t == 0) {
        double i2 = 0.0;
        for (int e = 0; e < k * k; e++) i2 += inv2[e];
        // cond_2 <= |DGD|_F |(DGD)^-1|_F
        stop = !(sqrt(gf2 * i2) < 1e4);
        *status = stop ? 1 : 0;
    }
    __syncthreads();
    if (stop) return;
    if (t < k * k) {
        // Z = D L^-T: Z[i][c] = dsc[i] * Li[c][i]
        const int c = t / k, i = t - c * k;
        Z[(int64_t)c * k + i] = dsc[i] * Li[c][i];
    }
}

// ------------------------------------------------------------------------
// heat-bath selection
// ------------------------------------------------------------------------
// Cmax (HB.cpp:204-214): signed entry of largest |C| over the first nopt
// columns, earlier column wins ties
__global__ void k_cmax(int64_t n, const double* __restrict__ X, int ld, int nopt, double* __restrict__ cm) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c = X[i * ld];
    for (int j = 1; j < nopt; j++) {
        const double x = X[i * ld + j];
        if (fabs(x) > fabs(c)) c = x;
    }
    cm[i] = c;
}

// number of candidates of state n: the FCs with |C_n f| >= eps form a prefix
// of the |f|-sorted list (HB.cpp:25-26, 172); candidates = variants of that
// prefix.
__global__ void k_hb_count(int64_t n, const double* __restrict__ cm, const double* __restrict__ fcs, int nfc,
                           const int64_t* __restrict__ varoff, double eps, int64_t* __restrict__ cnt) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c = cm[i];
    int lo = 0, hi = nfc;  // first index where the test fails
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (fabs(c * fcs[mid]) >= eps)
            lo = mid + 1;
        else
            hi = mid;
    }
    cnt[i] = varoff[lo];
}

struct HbGenArgs {
    int64_t n;
    int64_t total;
    const int64_t* candoff;  // n+1
    const uint4* keys;
    const uint64_t* bslots;
    uint64_t bmask;
    const uint32_t* bbloom;  // filter of the basis hash's keys
    uint64_t bbmask;
    const HbVar* vars;
    uint64_t* nslots;
    uint64_t nmask;
    uint4* pool_keys;
    uint64_t* pool_rank;
    uint32_t* counters;  // [0] pool fill, [1] overflow flags
    uint32_t pool_cap;
    uint64_t tlo, thi;   // keep only candidates whose hash tag lies in [tlo, thi)
};

template <int NW>
__device__ __forceinline__ bool pool_key_eq(const uint4* pool, uint32_t idx, const uint32_t (&w)[NW]) {
    const uint64_t* p = reinterpret_cast<const uint64_t*>(pool + (size_t)idx * (NW / 4));
    bool eq = true;
#pragma unroll
    for (int t = 0; t < NW / 2; t++) {
        const uint64_t v = __hip_atomic_load(p + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        eq = eq && v == ((uint64_t)w[2 * t] | ((uint64_t)w[2 * t + 1] << 32));
    }
    return eq;
}

// insert a candidate into the new-state hash; keeps the smallest rank per key
template <int NW>
__device__ void new_insert(const HbGenArgs& a, const uint32_t (&w)[NW], uint64_t h, uint64_t rank) {
    uint64_t pos = h & a.nmask;
    const uint32_t tag = (uint32_t)(h >> 32);
    int64_t mine = -1;
    while (true) {
        uint64_t s = __hip_atomic_load(&a.nslots[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s == kEmpty) {
            if (mine < 0) {
                const uint32_t id = atomicAdd(&a.counters[0], 1u);
                if (id >= a.pool_cap) {
                    atomicOr(&a.counters[1], 2u);
                    return;
                }
                mine = id;
                uint4* dst = a.pool_keys + (size_t)id * (NW / 4);
#pragma unroll
                for (int t = 0; t < NW / 4; t++) dst[t] = make_uint4(w[4 * t], w[4 * t + 1], w[4 * t + 2], w[4 * t + 3]);
                a.pool_rank[id] = rank;
                __threadfence();
            }
            const uint64_t want = ((uint64_t)tag << 32) | (uint64_t)(uint32_t)mine;
            const unsigned long long prev = atomicCAS((unsigned long long*)&a.nslots[pos],
                                                      (unsigned long long)kEmpty, (unsigned long long)want);
            if (prev == kEmpty) return;
            s = prev;
        }
        if ((uint32_t)(s >> 32) == tag && pool_key_eq<NW>(a.pool_keys, (uint32_t)s, w)) {
            // most candidates repeat a state found with a smaller rank: skip
            // the atomic then (within one selection the stored rank only
            // decreases, from the fill value down)
            uint64_t* pr = &a.pool_rank[(uint32_t)s];
            if (rank < __hip_atomic_load(pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMin((unsigned long long*)pr, (unsigned long long)rank);
            return;
        }
        pos = (pos + 1) & a.nmask;
    }
}

// candidate g = (state n, sorted FC, variant) flattened in the reference's
// loop order, so g = candoff[n] + r is the first-discovery rank.  One wave
// per basis state (grid-stride over states): the lanes walk its candidates.
template <int NW>
__global__ __launch_bounds__(256) void k_hb_gen(HbGenArgs a) {
    __shared__ uint64_t cf[64];
    load_coefs(cf);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t n = wave; n < a.n; n += nwaves) {
        const int64_t g0 = a.candoff[n], cnt = a.candoff[n + 1] - g0;
        if (cnt == 0) continue;
        uint32_t w0[NW];
        key_load<NW>(a.keys, n, w0);
        const uint64_t h0 = key_lin_cf<NW>(w0, cf);
        for (int64_t r = lane; r < cnt; r += 64) {
            const uint4 v = rec_load(a.vars, r);
            const int d = (int)rec_u8(v, 0);
            uint32_t w[NW];
#pragma unroll
            for (int i = 0; i < NW; i++) w[i] = w0[i];
            bool ok = true;
            bool over = false;
#pragma unroll
            for (int t = 0; t < kMaxD; t++) {
                if (t < d) {
                    const int q = key_q<NW>(w, (int)rec_u8(v, 1 + t)) + rec_i8(v, 7 + t);
                    ok = ok && q >= 0;
                    over = over || q > 255;
                }
            }
            if (!ok) continue;
            if (over) {
                atomicOr(&a.counters[1], 1u);
                continue;
            }
            const uint64_t h = fmix64(h0 + variant_shift(v, d, cf));
            const uint64_t tag = h >> 32;
            if (tag < a.tlo || tag >= a.thi) continue;
#pragma unroll
            for (int t = 0; t < kMaxD; t++)
                if (t < d) key_add<NW>(w, (int)rec_u8(v, 1 + t), rec_i8(v, 7 + t));
            if (bloom_maybe(a.bbloom, a.bbmask, h) && hash_find_h<NW>(a.bslots, a.bmask, a.keys, w, h) >= 0) continue;
            new_insert<NW>(a, w, h, (uint64_t)(g0 + r));
        }
    }
}

__global__ void k_collect(const uint64_t* __restrict__ slots, uint64_t cap, const uint64_t* __restrict__ pool_rank,
                          uint32_t* counter, uint64_t* __restrict__ rank_out, uint32_t* __restrict__ idx_out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = slots[i];
        if (s == kEmpty) continue;
        const uint32_t idx = (uint32_t)s;
        const uint32_t o = atomicAdd(counter, 1u);
        rank_out[o] = pool_rank[idx];
        idx_out[o] = idx;
    }
}

__global__ void k_gather_keys(const uint4* __restrict__ pool, const uint32_t* __restrict__ idx, int64_t u, int kw,
                              uint4* __restrict__ out) {
    int64_t total = u * kw;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / kw;
        const int c = (int)(t - r * kw);
        out[t] = pool[(size_t)idx[r] * kw + c];
    }
}

// ------------------------------------------------------------------------
// PT2: one wave per external state a (grid-stride); numerators over all
// variational states i connected to a, E_a, and Kahan-compensated sums of
// num^2 / (E_n - E_a) per column.
// ------------------------------------------------------------------------
struct Pt2Args {
    int64_t ns2;
    int M;
    int P;
    int nq;
    int nopt;
    const uint4* akeys;
    const uint4* vkeys;
    const uint64_t* vslots;
    uint64_t vmask;
    const uint32_t* vbloom;  // filter of the variational hash's keys
    uint64_t vbmask;
    const DeltaEnt* deltas;
    const uint32_t* poff;    // M + 2 block offsets of deltas
    const GroupEnt* groups;
    const uint32_t* fclist;
    const ElemFC* elem;
    const double* ladder;
    const uint32_t* dfc;
    int ndfc;
    const double* freq;
    const double* C;
    const float* C32;
    int ldc;
    const double* E;
    double* part;  // nwaves x nopt x 2
    int carry;     // 1: continue the sums already in part (a later chunk)
};

constexpr int kHitBuf = 512;

template <int NW, bool MIXED>
__global__ __launch_bounds__(64) void k_pt2(Pt2Args a) {
    __shared__ uint8_t sq[64];
    __shared__ int32_t hidx[kHitBuf];
    __shared__ double hval[kHitBuf];
    __shared__ double sdiag[64];
    __shared__ uint64_t cf[64];
    const int lane = threadIdx.x;
    load_coefs(cf);
    constexpr int R = 4;  // up to 256 columns
    double ks[R], kc[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int c = lane + 64 * r;
        const bool on = a.carry && c < a.nopt;
        ks[r] = on ? a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 0] : 0.0;
        kc[r] = on ? a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 1] : 0.0;
    }
    for (int64_t ia = blockIdx.x; ia < a.ns2; ia += gridDim.x) {
        uint32_t w[NW];
        key_load<NW>(a.akeys, ia, w);
        if (lane < a.M) sq[lane] = (uint8_t)key_q<NW>(w, lane);
        __syncthreads();
        double num[R];
        float num32[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            num[r] = 0.0;
            num32[r] = 0.0f;
        }
        const uint64_t h0 = row_lin(sq, a.M, cf);
        int nh = 0;
        auto flush = [&]() {
            __syncthreads();
            for (int h = 0; h < nh; h++) {
                const int64_t i = hidx[h];
                const double hv = hval[h];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int c = lane + 64 * r;
                    if (c < a.nopt) {
                        if (MIXED)
                            num32[r] = fmaf((float)hv, a.C32[i * a.ldc + c], num32[r]);
                        else
                            num[r] = fma(hv, a.C[i * a.ldc + c], num[r]);
                    }
                }
            }
            __syncthreads();
            nh = 0;
        };
        int blk = -1, pb0 = 0, pb1 = 0;
        while (next_block(a.poff, a.M, sq, blk, pb0, pb1))
        for (int p0 = pb0; p0 < pb1; p0 += kWave) {
            const int p = p0 + lane;
            bool hit = false;
            int64_t j = -1;
            uint4 e = make_uint4(0u, 0u, 0u, 0u);
            int nz = 0;
            if (p < pb1) {
                e = rec_load(a.deltas, p);
                nz = delta_nz(e);
                if (pattern_ok(e, nz, sq)) {
                    const uint64_t H = fmix64(h0 + pattern_shift(e, nz, cf));
                    if (bloom_maybe(a.vbloom, a.vbmask, H)) {
                        j = probe_shift<NW>(a.vslots, a.vmask, a.vkeys, w, e, nz, H);
                        hit = j >= 0;
                    }
                }
            }
            const uint64_t bal = __ballot(hit);
            if (hit) {
                const int pos = nh + __popcll(bal & ((1ull << lane) - 1ull));
                const GroupEnt ge = a.groups[delta_group(e)];
                // PT2.cpp:70-118: AnharmPot(a, i) starts at the external state
                hidx[pos] = (int32_t)j;
                hval[pos] = group_element(ge, a.fclist, a.elem, a.ladder, a.nq, sq, e, nz, true);
            }
            nh += __popcll(bal);
            if (nh > kHitBuf - kWave) flush();
        }
        flush();
        // E_a (PT2.cpp:119-144): harmonic sum, then each zero-pattern FC
        // term added one at a time in list order
        double ea = 0.;
        for (int f0 = 0; f0 < a.ndfc || f0 == 0; f0 += kWave) {
            const int f = f0 + lane;
            double x = 0.0;
            if (f < a.ndfc) x = diag_fc_term(a.elem, a.dfc[f], a.ladder, a.nq, sq);
            sdiag[lane] = x;
            __syncthreads();
            if (lane == 0) {
                if (f0 == 0)
                    for (int m = 0; m < a.M; m++) {
                        double ej = 0.5;
                        ej = ej + (double)sq[m];
                        ej = ej * a.freq[m];
                        ea = ea + ej;
                    }
                const int lim = min(kWave, a.ndfc - f0);
                for (int t = 0; t < lim; t++) ea = ea + sdiag[t];
            }
            __syncthreads();
            if (a.ndfc == 0) break;
        }
        ea = __shfl(ea, 0);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int c = lane + 64 * r;
            if (c < a.nopt) {
                const double nm = MIXED ? (double)num32[r] : num[r];
                const double term = (nm * nm) / (a.E[c] - ea);
                const double y = term - kc[r];
                const double t = ks[r] + y;
                kc[r] = (t - ks[r]) - y;
                ks[r] = t;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int c = lane + 64 * r;
        if (c < a.nopt) {
            a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 0] = ks[r];
            a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 1] = kc[r];
        }
    }
}

// Neumaier sum of the per-wave (sum, -compensation) pairs in wave order
__global__ void k_pt2_reduce(const double* __restrict__ part, int nw, int nopt, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nopt) return;
    double s = 0.0, cmp = 0.0;
    for (int w = 0; w < nw; w++) {
        for (int h = 0; h < 2; h++) {
            const double v = h == 0 ? part[((int64_t)w * nopt + c) * 2] : -part[((int64_t)w * nopt + c) * 2 + 1];
            const double t = s + v;
            if (fabs(s) >= fabs(v))
                cmp += (s - t) + v;
            else
                cmp += (v - t) + s;
            s = t;
        }
    }
    out[c] = s + cmp;
}

// ------------------------------------------------------------------------
// PrintFreqs assignment scan (Ham.cpp:448-502) on the resident Ritz block X
// (row-major, ld columns): per column c, half = max_m |X_mc| / 2 and the rows
// with |X_mc| >= half.  Blocks take chunks of rows, threads take columns.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_col_absmax(int64_t n, const double* __restrict__ X, int ld, int ncol,
                                                    int64_t rpc, double* __restrict__ part) {
    const int64_t r0 = (int64_t)blockIdx.x * rpc, r1 = min(n, r0 + rpc);
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
        double m = 0.0;
        for (int64_t r = r0; r < r1; r++) m = fmax(m, fabs(X[r * ld + c]));
        part[(int64_t)blockIdx.x * ncol + c] = m;
    }
}

__global__ void k_col_half(const double* __restrict__ part, int nchunk, int ncol, double* __restrict__ half) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    double m = 0.0;
    for (int k = 0; k < nchunk; k++) m = fmax(m, part[(int64_t)k * ncol + c]);
    half[c] = m / 2;
}

__global__ __launch_bounds__(256) void k_col_select(int64_t n, const double* __restrict__ X, int ld, int ncol,
                                                    int64_t rpc, const double* __restrict__ half,
                                                    uint32_t* __restrict__ cnt, uint32_t cap,
                                                    int64_t* __restrict__ rows, double* __restrict__ vals) {
    const int64_t r0 = (int64_t)blockIdx.x * rpc, r1 = min(n, r0 + rpc);
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
        const double h = half[c];
        for (int64_t r = r0; r < r1; r++) {
            const double v = fabs(X[r * ld + c]);
            if (v >= h) {
                const uint32_t slot = atomicAdd(&cnt[c], 1u);
                if (slot < cap) {
                    rows[(int64_t)c * cap + slot] = r;
                    vals[(int64_t)c * cap + slot] = v;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// PT2, push form: every variational state i probes its connection patterns
// in the external-space hash; each hit emits (a, i, pattern) at offsets fixed
// by a prefix sum over i and the lane order, so after a stable sort by a the
// pairs of every external state come in ascending i.
// ------------------------------------------------------------------------
struct Pt2PushArgs {
    int64_t n;
    int M;
    int P;
    const uint4* vkeys;
    const uint64_t* aslots;
    uint64_t amask;
    const uint32_t* abloom;  // filter of the chunk's external states
    uint64_t abmask;
    const uint4* apool;
    const uint32_t* apos;
    const DeltaEnt* deltas;
    const uint64_t* walk;    // the walk records of deltas
    const uint32_t* poff;    // M + 2 block offsets of deltas
    uint64_t tlo, thi;       // external states whose hash tag lies in [tlo, thi)
    int64_t rstride;       // count pass over every rstride-th state only (a sample)
    int64_t* cnt;
    const int64_t* off;
    uint32_t* pkey;
    uint64_t* pval;
};

template <int NW, bool FILL>
__global__ __launch_bounds__(64) void k_pt2_push(Pt2PushArgs a) {
    __shared__ uint8_t sq[64];
    __shared__ uint64_t cf[64];
    const int lane = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * a.rstride;
    uint32_t w[NW];
    key_load<NW>(a.vkeys, row, w);
    if (lane < a.M) sq[lane] = (uint8_t)key_q<NW>(w, lane);
    load_coefs(cf);
    __syncthreads();
    const uint64_t h0 = row_lin(sq, a.M, cf);
    int64_t base = FILL ? a.off[row] : 0;
    int64_t cnt = 0;
    int blk = -1, pb0 = 0, pb1 = 0;
    while (next_block(a.poff, a.M, sq, blk, pb0, pb1)) {
    // each lane's next pattern record is loaded one step ahead
    uint64_t en = pb0 + lane < pb1 ? a.walk[pb0 + lane] : 0ull;
    for (int p0 = pb0; p0 < pb1; p0 += kWave) {
        const int p = p0 + lane;
        const uint64_t e = en;
        if (p + kWave < pb1) en = a.walk[p + kWave];
        bool hit = false;
        int64_t idx = -1;
        if (p < pb1) {
            const int nz = delta_nz(e);
            if (pattern_ok(e, nz, sq)) {
                const uint64_t h = fmix64(h0 + pattern_shift(e, nz, cf));
                const uint64_t tag = h >> 32;
                if (tag >= a.tlo && tag < a.thi && bloom_maybe(a.abloom, a.abmask, h)) {
                    idx = probe_shift<NW>(a.aslots, a.amask, a.apool, w, e, nz, h);
                    hit = idx >= 0;
                }
            }
        }
        const uint64_t bal = __ballot(hit);
        if (!FILL) {
            cnt += __popcll(bal);
        } else {
            if (hit) {
                const int64_t pos = base + __popcll(bal & ((1ull << lane) - 1ull));
                a.pkey[pos] = a.apos ? a.apos[idx] : (uint32_t)idx;
                a.pval[pos] = ((uint64_t)(uint32_t)row << 32) | (uint32_t)p;
            }
            base += __popcll(bal);
        }
    }
    }
    if (!FILL && lane == 0) a.cnt[blockIdx.x] = cnt;
}

// PT2 pairs in one pass: each wave queues its row's hits in LDS and, when
// the queue fills or the row ends, reserves space for them with one atomic
// add on a global cursor.  Rows interleave in the output, so a pair is
// written as the key (a << 32 | i) with its pattern as the value; sorting by
// the key (unique: one pattern connects a given pair) restores the order
// of the counted form, a-major and ascending i, on every run.  Pairs beyond
// the capacity are counted but not written (the caller then counts exactly
// and fills again).
constexpr int kPushQ = 256;  // LDS queue of a wave, in pairs
template <int NW>
__global__ __launch_bounds__(64) void k_pt2_push1(Pt2PushArgs a, unsigned long long* __restrict__ cursor,
                                                  uint64_t* __restrict__ okey, uint32_t* __restrict__ oval,
                                                  unsigned long long cap) {
    __shared__ uint8_t sq[64];
    __shared__ uint64_t cf[64];
    __shared__ uint64_t qk[kPushQ];
    __shared__ uint32_t qv[kPushQ];
    const int lane = threadIdx.x;
    const int64_t row = blockIdx.x;
    uint32_t w[NW];
    key_load<NW>(a.vkeys, row, w);
    if (lane < a.M) sq[lane] = (uint8_t)key_q<NW>(w, lane);
    load_coefs(cf);
    __syncthreads();
    const uint64_t h0 = row_lin(sq, a.M, cf);
    int nq = 0;  // queued pairs (wave-uniform)
    auto flush = [&]() {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(cursor, (unsigned long long)nq);
        base = __shfl(base, 0);
        for (int t = lane; t < nq; t += 64) {
            const unsigned long long pos = base + (unsigned long long)t;
            if (pos < cap) {
                okey[pos] = qk[t];
                oval[pos] = qv[t];
            }
        }
        nq = 0;
    };
    int blk = -1, pb0 = 0, pb1 = 0;
    while (next_block(a.poff, a.M, sq, blk, pb0, pb1)) {
    // each lane's next pattern record is loaded one step ahead
    // the current step's candidate (its Bloom word loading a step ahead)
    // and the next step's record
    WalkCand cur = walk_cand(pb0 + lane < pb1 ? a.walk[pb0 + lane] : 0ull, pb0 + lane < pb1, sq, h0, cf, a.abloom,
                             a.abmask, a.tlo, a.thi);
    uint64_t en = pb0 + kWave + lane < pb1 ? a.walk[pb0 + kWave + lane] : 0ull;
    for (int p0 = pb0; p0 < pb1; p0 += kWave) {
        const int p = p0 + lane;
        const WalkCand nxt = walk_cand(en, p + kWave < pb1, sq, h0, cf, a.abloom, a.abmask, a.tlo, a.thi);
        en = p + 2 * kWave < pb1 ? a.walk[p + 2 * kWave] : 0ull;
        bool hit = false;
        int64_t idx = -1;
        if (walk_maybe(cur)) {
            idx = probe_shift<NW>(a.aslots, a.amask, a.apool, w, cur.e, delta_nz(cur.e), cur.H);
            hit = idx >= 0;
        }
        cur = nxt;
        const uint64_t bal = __ballot(hit);
        const int nh = __popcll(bal);
        if (nh == 0) continue;
        if (nq + nh > kPushQ) {
            __syncthreads();
            flush();
            __syncthreads();
        }
        if (hit) {
            const int t = nq + __popcll(bal & ((1ull << lane) - 1ull));
            const uint32_t ai = a.apos ? a.apos[idx] : (uint32_t)idx;
            qk[t] = ((uint64_t)ai << 32) | (uint32_t)row;
            qv[t] = (uint32_t)p;
        }
        nq += nh;
    }
    }
    __syncthreads();
    if (nq > 0) flush();
}

// (a << 32 | i, pattern) pairs sorted by key into the counted form's
// layout: key a, value (i << 32 | pattern)
__global__ void k_pair_split(const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sval, int64_t npairs,
                             uint32_t* __restrict__ akey, uint64_t* __restrict__ ival) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < npairs; t += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = skey[t];
        akey[t] = (uint32_t)(k >> 32);
        ival[t] = (k << 32) | sval[t];
    }
}

// canonical position of every pool entry (inverse of the sort by rank)
__global__ void k_pool_pos(const uint32_t* __restrict__ sorted_idx, int64_t u, uint32_t* __restrict__ pos) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < u) pos[sorted_idx[t]] = (uint32_t)t;
}

// first pair of every external state in the sorted pair list
__global__ void k_seg_start(const uint32_t* __restrict__ skey, int64_t npairs, int64_t ns2, int64_t* __restrict__ seg) {
    const int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (a > ns2) return;
    int64_t lo = 0, hi = npairs;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)skey[mid] < a)
            lo = mid + 1;
        else
            hi = mid;
    }
    seg[a] = lo;
}

struct Pt2NumArgs {
    int64_t ns2;
    int M;
    int nq;
    int nopt;
    const uint4* akeys;
    const int64_t* seg;
    const uint64_t* sval;
    const DeltaEnt* deltas;
    const GroupEnt* groups;
    const uint32_t* fclist;
    const ElemFC* elem;
    const double* ladder;
    const uint32_t* dfc;
    int ndfc;
    const double* freq;
    const double* C;
    const float* C32;
    int ldc;
    const double* E;
    double* part;
    int carry;  // 1: continue the sums already in part (a later chunk)
};

// one wave per external state: numerators sum_i H_ai C_in over its pairs in
// ascending i (FP64: product then sum, as the reference's SumHaiCi), E_a,
// then Kahan sums of num^2 / (E_n - E_a) per column.  R = columns / 64
// (nopt <= 64 R); the C rows of U pairs are loaded before they are summed
// in order, so U row reads are in flight at once.
template <int NW, bool MIXED, int R>
__global__ __launch_bounds__(64) void k_pt2_num(Pt2NumArgs a) {
    __shared__ uint8_t sq[64];
    __shared__ int32_t hidx[64];
    __shared__ double hval[64];
    __shared__ double sdiag[64];
    const int lane = threadIdx.x;
    constexpr int U = 8;
    double ks[R], kc[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int c = lane + 64 * r;
        const bool on = a.carry && c < a.nopt;
        ks[r] = on ? a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 0] : 0.0;
        kc[r] = on ? a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 1] : 0.0;
    }
    for (int64_t ia = blockIdx.x; ia < a.ns2; ia += gridDim.x) {
        uint32_t w[NW];
        key_load<NW>(a.akeys, ia, w);
        if (lane < a.M) sq[lane] = (uint8_t)key_q<NW>(w, lane);
        __syncthreads();
        double num[R];
        float num32[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            num[r] = 0.0;
            num32[r] = 0.0f;
        }
        const int64_t s0 = a.seg[ia], s1 = a.seg[ia + 1];
        for (int64_t h0 = s0; h0 < s1; h0 += kWave) {
            const int64_t h = h0 + lane;
            if (h < s1) {
                const uint64_t v = a.sval[h];
                const uint4 e = rec_load(a.deltas, (uint32_t)v);
                const GroupEnt ge = a.groups[delta_group(e)];
                hidx[lane] = (int32_t)(v >> 32);
                // i = a - delta: the pattern applied negated from a
                hval[lane] = group_element(ge, a.fclist, a.elem, a.ladder, a.nq, sq, e, delta_nz(e), true, -1);
            }
            __syncthreads();
            const int lim = (int)min((int64_t)kWave, s1 - h0);
            for (int l0 = 0; l0 < lim; l0 += U) {
                double cv[U][R];
                float cv32[U][R];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int64_t i = hidx[min(l0 + u, lim - 1)];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int c = lane + 64 * r;
                        const bool on = c < a.nopt && l0 + u < lim;
                        if (MIXED)
                            cv32[u][r] = on ? a.C32[i * a.ldc + c] : 0.0f;
                        else
                            cv[u][r] = on ? a.C[i * a.ldc + c] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (l0 + u < lim) {
                        const double hv = hval[l0 + u];
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            if (MIXED)
                                num32[r] = fmaf((float)hv, cv32[u][r], num32[r]);
                            else
                                num[r] = num[r] + hv * cv[u][r];
                        }
                    }
                }
            }
            __syncthreads();
        }
        double ea = 0.;
        for (int f0 = 0; f0 < a.ndfc || f0 == 0; f0 += kWave) {
            const int f = f0 + lane;
            double x = 0.0;
            if (f < a.ndfc) x = diag_fc_term(a.elem, a.dfc[f], a.ladder, a.nq, sq);
            sdiag[lane] = x;
            __syncthreads();
            if (lane == 0) {
                if (f0 == 0)
                    for (int m = 0; m < a.M; m++) {
                        double ej = 0.5;
                        ej = ej + (double)sq[m];
                        ej = ej * a.freq[m];
                        ea = ea + ej;
                    }
                const int lim = min(kWave, a.ndfc - f0);
                for (int t = 0; t < lim; t++) ea = ea + sdiag[t];
            }
            __syncthreads();
            if (a.ndfc == 0) break;
        }
        ea = __shfl(ea, 0);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int c = lane + 64 * r;
            if (c < a.nopt) {
                const double nm = MIXED ? (double)num32[r] : num[r];
                const double term = (nm * nm) / (a.E[c] - ea);
                const double y = term - kc[r];
                const double t = ks[r] + y;
                kc[r] = (t - ks[r]) - y;
                ks[r] = t;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int c = lane + 64 * r;
        if (c < a.nopt) {
            a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 0] = ks[r];
            a.part[((int64_t)blockIdx.x * a.nopt + c) * 2 + 1] = kc[r];
        }
    }
}

}  // namespace

// ==========================================================================
// DevBuf
// ==========================================================================
namespace {
thread_local int64_t t_dev_bytes = 0;
thread_local int64_t t_dev_peak = 0;
}  // namespace

int64_t thread_device_bytes() { return t_dev_bytes; }
int64_t thread_device_peak() { return t_dev_peak; }
void thread_device_reset_peak() { t_dev_peak = t_dev_bytes; }

HostBuf::~HostBuf() {
    if (p) (void)hipHostFree(p);
}
void* HostBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return p;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 4096);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess || !p)
        throw std::runtime_error("page-locked host allocation failed");
    cap = want;
    return p;
}

namespace {
thread_local DevPool* t_pool = nullptr;

// capacity class of a request: multiples of 256 B below 64 KiB, above that
// eight classes per power of two (at most 12.5 % over the request)
size_t size_class(size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes <= (size_t(1) << 16)) return (bytes + 255) & ~size_t(255);
    int top = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = size_t(1) << (top - 3);
    return (bytes + step - 1) & ~(step - 1);
}
}  // namespace

DevPool* current_pool() { return t_pool; }
PoolScope::PoolScope(DevPool* p) : prev(t_pool) { t_pool = p; }
PoolScope::~PoolScope() { t_pool = prev; }

void* DevPool::take(size_t bytes, size_t& cap) {
    const size_t want = size_class(bytes);
    auto it = idle.lower_bound(want);
    if (it != idle.end() && it->first <= 2 * want) {
        void* p = it->second;
        cap = it->first;
        idle_bytes -= (int64_t)cap;
        idle.erase(it);
        reuses++;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess && !idle.empty()) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw std::runtime_error("device allocation of " + std::to_string(want) + " bytes failed: " +
                                 hipGetErrorString(e));
    }
    mallocs++;
    cap = want;
    held += (int64_t)want;
    t_dev_bytes += (int64_t)want;
    t_dev_peak = std::max(t_dev_peak, t_dev_bytes);
    return p;
}

void DevPool::give(void* p, size_t cap) {
    idle.emplace(cap, p);
    idle_bytes += (int64_t)cap;
}

void DevPool::trim() {
    if (idle.empty()) return;
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& b : idle) {
        (void)hipFree(b.second);
        frees++;
        held -= (int64_t)b.first;
        t_dev_bytes -= (int64_t)b.first;
    }
    idle.clear();
    idle_bytes = 0;
}

DevPool::~DevPool() { trim(); }

DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
    if (p) {
        if (pool) {
            pool->give(p, cap);
        } else {
            (void)hipFree(p);
            t_dev_bytes -= (int64_t)cap;
        }
    }
    p = nullptr;
    cap = 0;
}
void* DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return p;
    release();
    if (!pool) pool = t_pool;
    if (pool) {
        size_t c = 0;
        p = pool->take(bytes, c);
        cap = c;
        return p;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        throw std::runtime_error("device allocation of " + std::to_string(want) + " bytes failed: " +
                                 hipGetErrorString(e));
    }
    cap = want;
    t_dev_bytes += (int64_t)want;
    t_dev_peak = std::max(t_dev_peak, t_dev_bytes);
    return p;
}

// ==========================================================================
// Engine
// ==========================================================================
Engine::Engine(const Problem& P, int form, int device) : P_(P), form_(form), device_(device) {
    if (form != kFP64 && form != kMixed) throw std::runtime_error("unknown precision form");
    if (P_.M < 1 || P_.M > 64) throw std::runtime_error("number of modes must be 1..64");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) throw std::runtime_error("no HIP device available");
    if (device < 0 || device >= ndev) throw std::runtime_error("device index out of range");
    HIPCHK(hipSetDevice(device));
    PoolScope ps(&pool_);
    KW_ = (P_.M + 15) / 16;
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    pool_.stream = stream_;
    // set-up of the no-allocation product: the scratch address of this device
    (void)noalloc_scratch(device, nullptr);
    BLASCHK(rocblas_create_handle(&blas_));
    BLASCHK(rocblas_set_stream(blas_, stream_));
    // no atomics-based algorithms: the same inputs give the same bits on every run
    BLASCHK(rocblas_set_atomics_mode(blas_, rocblas_atomics_not_allowed));
    build_tables(P_, T_);
    upload_tables();
}

Engine::~Engine() {
    // the buffers go back to the pool after this body, and the pool frees
    // them without waiting for the stream, which is finished here
    if (stream_) (void)hipStreamSynchronize(stream_);
    pool_.stream = nullptr;
    for (hipEvent_t e : prof_.ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : prof_.gev) (void)hipEventDestroy(e);
    if (blas_) rocblas_destroy_handle(blas_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

template <class T>
static void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// the pattern walks' 8-byte records of a pattern table (modes < 64,
// changes within -8..7; layout at walk_mode / walk_delta)
static std::vector<uint64_t> walk_records(const std::vector<DeltaEnt>& table) {
    std::vector<uint64_t> walk(table.size());
    for (size_t i = 0; i < walk.size(); i++) {
        const DeltaEnt& d = table[i];
        const int nz = (int)(d.meta >> 28);
        if (nz > kMaxD) throw std::runtime_error("connection pattern of more than 6 modes");
        uint64_t r = (uint64_t)nz << 60;
        for (int t = 0; t < nz; t++) {
            if (d.mode[t] >= 64 || d.delta[t] < -8 || d.delta[t] > 7)
                throw std::runtime_error("connection pattern outside the walk record's range");
            r |= (uint64_t)d.mode[t] << (10 * t);
            r |= (uint64_t)((unsigned)d.delta[t] & 15u) << (10 * t + 6);
        }
        walk[i] = r;
    }
    return walk;
}

void Engine::upload_tables() {
    std::vector<double> L;
    build_ladder(255, L);
    upload(d_ladder_, L, stream_);
    upload(d_elem_, T_.elem, stream_);
    upload(d_deltas_, T_.deltas, stream_);
    upload(d_walk_, walk_records(T_.deltas), stream_);
    upload(d_doff_, T_.delta_off, stream_);
    upload(d_groups_, T_.groups, stream_);
    upload(d_fclist_, T_.fclist, stream_);
    upload(d_diag_, T_.diag_fcs, stream_);
    upload(d_freq_, P_.freq, stream_);
    upload(d_hbfc_, T_.hb_fc, stream_);
    upload(d_hbvaroff_, T_.hb_varoff, stream_);
    upload(d_hbvars_, T_.hb_vars, stream_);
    P_deltas_ = (int)T_.deltas.size();
    HIPCHK(hipStreamSynchronize(stream_));
}

void Engine::set_basis(const int32_t* q, int64_t n) {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    if (n < 1) throw std::runtime_error("empty basis");
    if (n >= (int64_t)INT32_MAX) throw std::runtime_error("basis larger than 2^31-1 states");
    const int M = P_.M;
    const int K = KW_ * 16;
    std::vector<uint8_t> packed((size_t)n * K, 0);
    for (int64_t i = 0; i < n; i++)
        for (int m = 0; m < M; m++) {
            int v = q[i * M + m];
            if (v < 0 || v > 255) throw std::runtime_error("quanta must lie in 0..255");
            packed[(size_t)i * K + m] = (uint8_t)v;
        }
    keys_.ensure(packed.size());
    HIPCHK(hipMemcpyAsync(keys_.p, packed.data(), packed.size(), hipMemcpyHostToDevice, stream_));
    n_ = n;
    built_n_ = 0;
    h_valid_ = false;
    xrows_ = 0;
    xcols_ = 0;
    rebuild_hash(n);
    HIPCHK(hipStreamSynchronize(stream_));
    pool_.trim();
}

void Engine::set_initial_basis() {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    const int M = P_.M;
    const int K = KW_ * 16;
    if (M < 1) throw std::runtime_error("empty basis");
    if (P_.restart) {
        // the restart placeholder: one state of zero quanta (its basis comes from the checkpoint)
        std::vector<int32_t> z(M, 0);
        set_basis(z.data(), 1);
        return;
    }
    const int nt = P_.ntot;
    if (nt < 0) throw std::runtime_error("negative total quanta");
    std::vector<int32_t> lim(M);
    for (int m = 0; m < M; m++) lim[m] = std::min(P_.cap[m], nt);  // a negative cap leaves no state
    // cnt[m][s]: choices for modes 0..m-1 with sum <= s
    std::vector<int64_t> cnt((size_t)(M + 1) * (nt + 1), 0);
    for (int s = 0; s <= nt; s++) cnt[s] = 1;
    for (int m = 0; m < M; m++)
        for (int s = 0; s <= nt; s++) {
            int64_t c = 0;
            for (int q = 0; q <= std::min(lim[m], s); q++) c += cnt[(size_t)m * (nt + 1) + (s - q)];
            if (c > (int64_t)INT32_MAX) throw std::runtime_error("basis larger than 2^31-1 states");
            cnt[(size_t)(m + 1) * (nt + 1) + s] = c;
        }
    const int64_t n0 = cnt[(size_t)M * (nt + 1) + nt];
    if (n0 < 1) throw std::runtime_error("empty basis");
    for (int m = 0; m < M; m++)
        if (lim[m] > 255) throw std::runtime_error("quanta must lie in 0..255");
    DevBuf dcnt, dlim;
    dcnt.ensure(cnt.size() * sizeof(int64_t));
    dlim.ensure(lim.size() * sizeof(int32_t));
    HIPCHK(hipMemcpyAsync(dcnt.p, cnt.data(), cnt.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(dlim.p, lim.data(), lim.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    keys_.ensure((size_t)n0 * K);
    k_initial_basis<<<grid_for(n0, 256), 256, 0, stream_>>>(n0, M, nt, K, dcnt.as<int64_t>(), dlim.as<int32_t>(),
                                                           keys_.as<uint8_t>());
    HIPCHK(hipGetLastError());
    n_ = n0;
    built_n_ = 0;
    h_valid_ = false;
    xrows_ = 0;
    xcols_ = 0;
    rebuild_hash(n0);
    HIPCHK(hipStreamSynchronize(stream_));
}

void Engine::get_basis(int32_t* q) const {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    const int M = P_.M, K = KW_ * 16;
    std::vector<uint8_t> packed((size_t)n_ * K);
    HIPCHK(hipMemcpyAsync(packed.data(), keys_.p, packed.size(), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int64_t i = 0; i < n_; i++)
        for (int m = 0; m < M; m++) q[i * M + m] = packed[(size_t)i * K + m];
}

// Bloom filter bits for a hash of cap slots: as many as one die's 4 MB L2
// holds, between 2 and 16 per slot (VHCI_BLOOM_BITS: bits per slot).  On the
// 48-mode workload a 4 MB filter gives the fastest pattern walks: fewer
// false positives send fewer probes to the hash, and a larger filter no
// longer stays in L2 (H builds 2.70 s at 2 bits per slot, 2.18 s at 8,
// 2.38 s at 16; PT2 pairs 2.09 s at 2, 1.95 s at 4, 2.40 s at 8).
static uint64_t bloom_bits(uint64_t cap) {
    if (const char* e = std::getenv("VHCI_BLOOM_BITS"))
        return next_pow2((uint64_t)std::max(1, std::min(64, std::atoi(e)))) * cap;
    return std::max<uint64_t>(2 * cap, std::min<uint64_t>(1ull << 25, 16 * cap));
}

void Engine::rebuild_hash(int64_t n) {
    uint64_t cap = next_pow2(std::max<uint64_t>(1024, (uint64_t)n * 2));
    slots_.ensure(cap * sizeof(uint64_t));
    hmask_ = cap - 1;
    k_fill<uint64_t><<<grid_for(cap, 256), 256, 0, stream_>>>(slots_.as<uint64_t>(), (int64_t)cap, kEmpty);
    // at least 4 filter bits per key while the hash holds at most half its slots
    const uint64_t fbits = bloom_bits(cap);
    bloom_.ensure(fbits / 8);
    bmask_ = fbits - 1;
    k_fill<uint32_t><<<grid_for((int64_t)(fbits / 32), 256), 256, 0, stream_>>>(bloom_.as<uint32_t>(),
                                                                                (int64_t)(fbits / 32), 0u);
    const uint4* k = keys_.as<uint4>();
    switch (KW_) {
        case 1: k_hash_insert<4><<<grid_for(n, 256), 256, 0, stream_>>>(k, 0, n, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
        case 2: k_hash_insert<8><<<grid_for(n, 256), 256, 0, stream_>>>(k, 0, n, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
        case 3: k_hash_insert<12><<<grid_for(n, 256), 256, 0, stream_>>>(k, 0, n, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
        default: k_hash_insert<16><<<grid_for(n, 256), 256, 0, stream_>>>(k, 0, n, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
    }
    HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------- H build
// perm = the n keys' indices sorted by their quanta read as one integer,
// the highest mode most significant (stable LSD radix sorts over the packed
// keys, 8 modes at a time)
void Engine::quanta_order(const uint32_t* keys, int64_t n, int32_t* perm) {
    DevBuf kin, kout, pout;
    kin.ensure(n * sizeof(uint64_t));
    kout.ensure(n * sizeof(uint64_t));
    pout.ensure(n * sizeof(int32_t));
    k_iota32<<<grid_for(n, 256), 256, 0, stream_>>>(perm, n);
    const int nchunk = (P_.M + 7) / 8;
    for (int c = 0; c < nchunk; c++) {
        k_key_chunk<<<grid_for(n, 256), 256, 0, stream_>>>(keys, 4 * KW_, perm, n, c, kin.as<uint64_t>());
        const unsigned endbit = 8u * (unsigned)std::min(8, P_.M - 8 * c);
        size_t tmp = 0;
        HIPCHK(rocprim::radix_sort_pairs(nullptr, tmp, kin.as<uint64_t>(), kout.as<uint64_t>(), perm,
                                         pout.as<int32_t>(), (size_t)n, 0, endbit, stream_));
        scan_tmp_.ensure(tmp);
        HIPCHK(rocprim::radix_sort_pairs(scan_tmp_.p, tmp, kin.as<uint64_t>(), kout.as<uint64_t>(), perm,
                                         pout.as<int32_t>(), (size_t)n, 0, endbit, stream_));
        HIPCHK(hipMemcpyAsync(perm, pout.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, stream_));
    }
    HIPCHK(hipGetLastError());
}

void Engine::make_solver_order() {
    const int64_t n = n_;
    const char* e = std::getenv("VHCI_REORDER");
    const char* alg = std::getenv("VHCI_EIGSOLVER");
    const bool off = (e && std::string(e) == "0") || (alg && std::string(alg) == "davidson");
    perm_on_ = !off && n >= kReorderMin;
    if (!perm_on_) return;
    perm_.ensure(n * sizeof(int32_t));
    inv_.ensure(n * sizeof(int32_t));
    quanta_order(keys_.as<uint32_t>(), n, perm_.as<int32_t>());
    k_invert32<<<grid_for(n, 256), 256, 0, stream_>>>(perm_.as<int32_t>(), n, inv_.as<int32_t>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
}

template <int NW>
void Engine::build_h_t(HStats& st) {
    const int64_t n = n_;
    HBuildArgs a{};
    a.n = n;
    a.M = P_.M;
    a.P = P_deltas_;
    a.nq = 256;
    a.keys = keys_.as<uint4>();
    a.slots = slots_.as<uint64_t>();
    a.mask = hmask_;
    a.bloom = bloom_.as<uint32_t>();
    a.bmask = bmask_;
    a.deltas = d_deltas_.as<DeltaEnt>();
    a.walk = d_walk_.as<uint64_t>();
    if (!a.walk) throw std::runtime_error("pattern walk records missing");
    a.poff = d_doff_.as<uint32_t>();
    a.groups = d_groups_.as<GroupEnt>();
    a.fclist = d_fclist_.as<uint32_t>();
    a.elem = d_elem_.as<ElemFC>();
    a.ladder = d_ladder_.as<double>();
    a.form = form_;
    // Incremental build when the previous stored H covers the first n_old
    // states (the basis only grew since) and fewer than half are new: only
    // the new states' rows are probed; old rows keep their entries and gain
    // the transposes of the new rows' entries to them.
    const int64_t n_old = built_n_;
    const char* fb = std::getenv("VHCI_FULL_BUILD");
    bool incr = n_old > 0 && n_old < n && 2 * (n - n_old) < n && a.P > 0 && !matfree_ && !sym_ &&
                !(fb && std::string(fb) == "1");
    built_n_ = 0;
    const bool old_perm_on = perm_on_;
    auto swap_buf = [](DevBuf& x, DevBuf& y) { x.swap(y); };
    if (incr) {
        swap_buf(perm_, perm_o_);
        swap_buf(inv_, inv_o_);
        swap_buf(rowptr_, rowptr_o_);
        swap_buf(col_, col_o_);
        swap_buf(val_, val_o_);
        swap_buf(hi_, hi_o_);
        swap_buf(lo_, lo_o_);
    }
    make_solver_order();
    a.perm = perm_on_ ? perm_.as<int32_t>() : nullptr;
    a.inv = perm_on_ ? inv_.as<int32_t>() : nullptr;
    rowcnt_.ensure((n + 1) * sizeof(int64_t));
    rowptr_.ensure((n + 1) * sizeof(int64_t));
    a.rowcnt = rowcnt_.as<int64_t>();
    // from-scratch count passes also count the upper triangle (half storage)
    rowcnt_up_.ensure((n + 1) * sizeof(int64_t));
    a.rowcnt_up = incr ? nullptr : rowcnt_up_.as<int64_t>();
    // Hit words: the count pass records the ballot of every 64-pattern step
    // of each row's walk, so the fill pass probes only the hits instead of
    // walking every pattern again; made when the words take at most a
    // quarter of the free device memory (VHCI_HIT_WORDS=0: never).
    DevBuf wcnt, woff, hitw;
    auto prepare_words = [&](int64_t nb) {
        a.hitw = nullptr;
        a.woff = nullptr;
        hitw.release();
        const char* he = std::getenv("VHCI_HIT_WORDS");
        if (a.P <= 0 || nb <= 0 || (he && std::string(he) == "0")) return;
        wcnt.ensure((nb + 1) * sizeof(int64_t));
        woff.ensure((nb + 1) * sizeof(int64_t));
        k_walk_words<NW><<<grid_for(nb + 1, 256), 256, 0, stream_>>>(a, nb, wcnt.as<int64_t>());
        HIPCHK(hipGetLastError());
        size_t tmp = 0;
        HIPCHK(rocprim::exclusive_scan(nullptr, tmp, wcnt.as<int64_t>(), woff.as<int64_t>(), (int64_t)0,
                                       (size_t)(nb + 1), rocprim::plus<int64_t>(), stream_));
        scan_tmp_.ensure(tmp);
        HIPCHK(rocprim::exclusive_scan(scan_tmp_.p, tmp, wcnt.as<int64_t>(), woff.as<int64_t>(), (int64_t)0,
                                       (size_t)(nb + 1), rocprim::plus<int64_t>(), stream_));
        int64_t words = 0;
        HIPCHK(hipMemcpyAsync(&words, woff.as<int64_t>() + nb, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        if (std::getenv("VHCI_DEBUG"))
            std::fprintf(stderr, "[h-build] %lld rows walk %lld steps of 64 patterns (%.0f per row; table %lld)\n",
                         (long long)nb, (long long)words, (double)words / (double)nb, (long long)a.P);
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        if ((double)words * 8.0 > 0.25 * ((double)free_b + (double)pool_.idle_bytes)) return;
        try {
            hitw.ensure((size_t)std::max<int64_t>(1, words) * sizeof(uint64_t));
        } catch (const std::runtime_error&) {
            return;
        }
        a.hitw = hitw.as<uint64_t>();
        a.woff = woff.as<int64_t>();
    };
    if (incr) {
        extra_.ensure(n_old * sizeof(uint64_t));
        oldcnt_.ensure(n_old * sizeof(int64_t));
        cursor_.ensure(n_old * sizeof(uint64_t));
        k_fill<uint64_t><<<grid_for(n_old, 256), 256, 0, stream_>>>(extra_.as<uint64_t>(), n_old, 0ull);
        k_fill<uint64_t><<<grid_for(n_old, 256), 256, 0, stream_>>>(cursor_.as<uint64_t>(), n_old, 0ull);
        a.first = n_old;
        a.n_old = n_old;
        a.extra = extra_.as<unsigned long long>();
        a.oldcnt = oldcnt_.as<int64_t>();
        a.cursor = cursor_.as<unsigned long long>();
        prepare_words(n - n_old);
        k_hbuild<NW, false><<<(unsigned)(n - n_old), 64, 0, stream_>>>(a);
        HIPCHK(hipGetLastError());
        k_old_counts<<<grid_for(n_old, 256), 256, 0, stream_>>>(
            n_old, a.inv, old_perm_on ? inv_o_.as<int32_t>() : nullptr, rowptr_o_.as<int64_t>(),
            extra_.as<unsigned long long>(), oldcnt_.as<int64_t>(), a.rowcnt);
        HIPCHK(hipGetLastError());
    } else if (a.P > 0) {
        prepare_words(n);
        k_hbuild<NW, false><<<(unsigned)n, 64, 0, stream_>>>(a);
        HIPCHK(hipGetLastError());
    } else {
        k_fill<int64_t><<<grid_for(n, 256), 256, 0, stream_>>>(a.rowcnt, n, 0);
        k_fill<int64_t><<<grid_for(n, 256), 256, 0, stream_>>>(a.rowcnt_up, n, 0);
    }
    // row offsets (exclusive scan of n + 1 counts, the last one zero); returns the total
    auto scan_counts = [&](int64_t* cnt) {
        k_fill<int64_t><<<1, 1, 0, stream_>>>(cnt + n, 1, 0);
        size_t tmp = 0;
        HIPCHK(rocprim::exclusive_scan(nullptr, tmp, cnt, rowptr_.as<int64_t>(), (int64_t)0, (size_t)(n + 1),
                                       rocprim::plus<int64_t>(), stream_));
        scan_tmp_.ensure(tmp);
        HIPCHK(rocprim::exclusive_scan(scan_tmp_.p, tmp, cnt, rowptr_.as<int64_t>(), (int64_t)0, (size_t)(n + 1),
                                       rocprim::plus<int64_t>(), stream_));
        int64_t tot = 0;
        HIPCHK(hipMemcpyAsync(&tot, rowptr_.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        return tot;
    };
    int64_t nnz_up = a.rowcnt_up ? scan_counts(a.rowcnt_up) : -1;
    int64_t nnz = scan_counts(a.rowcnt);
    bool keep_words = false;
    int storage = choose_storage(nnz, nnz_up >= 0 ? nnz_up : nnz / 2, a.hitw ? hitw.cap : 0, &keep_words);
    if (incr && storage != 0) {
        // the grown matrix no longer fits in full: count again from scratch
        // (half storage is built whole)
        col_o_.release();
        val_o_.release();
        hi_o_.release();
        lo_o_.release();
        incr = false;
        a.first = 0;
        a.n_old = 0;
        a.extra = nullptr;
        a.oldcnt = nullptr;
        a.cursor = nullptr;
        a.rowcnt_up = rowcnt_up_.as<int64_t>();
        prepare_words(n);
        k_hbuild<NW, false><<<(unsigned)n, 64, 0, stream_>>>(a);
        HIPCHK(hipGetLastError());
        nnz_up = scan_counts(a.rowcnt_up);
        nnz = scan_counts(a.rowcnt);
        storage = choose_storage(nnz, nnz_up, a.hitw ? hitw.cap : 0, &keep_words);
    }
    if (!keep_words) {
        // only the full form's fill reads them, and only where they fit
        hitw.release();
        a.hitw = nullptr;
        a.woff = nullptr;
    }
    sym_ = storage == 1;
    if (sym_) (void)scan_counts(a.rowcnt_up);
    const int64_t nstore = sym_ ? nnz_up : nnz;
    nnz_ = nstore;
    matfree_ = storage == 2;
    if (matfree_) {
        // no stored entries: products regenerate them (k_hfree)
        col_.release();
        val_.release();
        hi_.release();
        lo_.release();
        col_o_.release();
        val_o_.release();
        hi_o_.release();
        lo_o_.release();
        D_.ensure(n * sizeof(double));
        k_diag<NW><<<grid_for(n, 128), 128, 0, stream_>>>(n, P_.M, a.keys, d_freq_.as<double>(),
                                                          d_diag_.as<uint32_t>(), (int)T_.diag_fcs.size(), a.elem,
                                                          a.ladder, 256, a.perm, D_.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream_));
        st.n = n;
        st.nnz_offdiag = nnz;
        st.probed_rows = incr ? n - n_old : n;
        return;
    }
    col_.ensure((nstore + kCsrPad) * sizeof(int32_t));
    if (form_ == kFP64) {
        val_.ensure((nstore + kCsrPad) * sizeof(double));
    } else {
        hi_.ensure((nstore + kCsrPad) * sizeof(float));
        lo_.ensure((nstore + kCsrPad) * sizeof(float));
    }
    a.rowptr = rowptr_.as<int64_t>();
    a.col = col_.as<int32_t>();
    a.val = val_.as<double>();
    a.hi = hi_.as<float>();
    a.lo = lo_.as<float>();
    a.half = sym_ ? 1 : 0;
    if (sym_) {
        rowabs_.ensure(n * sizeof(double));
        a.rowabs = rowabs_.as<double>();
    } else {
        rowabs_.release();
        hacc_.release();
        hacc_zero_ = 0;
    }
    if (incr) {
        k_copy_old_rows<<<grid_for(n_old, 4), 256, 0, stream_>>>(
            n_old, a.inv, old_perm_on ? inv_o_.as<int32_t>() : nullptr, old_perm_on ? perm_o_.as<int32_t>() : nullptr,
            rowptr_o_.as<int64_t>(), col_o_.as<int32_t>(), val_o_.as<double>(), hi_o_.as<float>(), lo_o_.as<float>(),
            a.rowptr, a.col, a.val, a.hi, a.lo, form_);
        HIPCHK(hipGetLastError());
        if (nnz > 0) {
                k_hbuild<NW, true><<<(unsigned)(n - n_old), 64, 0, stream_>>>(a);
            HIPCHK(hipGetLastError());
        }
    } else if (a.P > 0 && nnz > 0) {
            k_hbuild<NW, true><<<(unsigned)n, 64, 0, stream_>>>(a);
        HIPCHK(hipGetLastError());
    }
    if (sym_) {
        // the accumulator's bound: the largest sum of |H_ij| over a row
        hnorm_ = 0;
        if (nnz > 0) {
            DevBuf mx;
            mx.ensure(sizeof(double));
            size_t tmp = 0;
            HIPCHK(rocprim::reduce(nullptr, tmp, rowabs_.as<double>(), mx.as<double>(), 0.0, (size_t)n,
                                   rocprim::maximum<double>(), stream_));
            scan_tmp_.ensure(tmp);
            HIPCHK(rocprim::reduce(scan_tmp_.p, tmp, rowabs_.as<double>(), mx.as<double>(), 0.0, (size_t)n,
                                   rocprim::maximum<double>(), stream_));
            HIPCHK(hipMemcpyAsync(&hnorm_, mx.p, sizeof(double), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
        }
    }
    // columns ascending within each row.  rocPRIM's segmented sort counts
    // entries in 32 bits, so rows are sorted in slices of at most kSortSlice
    // entries, each with its own row offsets relative to the slice start.
    if (nstore > 0) {
        unsigned endbit = 1;
        while (endbit < 31 && ((uint64_t)(n - 1) >> endbit) != 0) endbit++;
        std::vector<int64_t> rph(n + 1);
        HIPCHK(hipMemcpyAsync(rph.data(), rowptr_.p, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        int64_t slice = kSortSlice;
        if (const char* e = std::getenv("VHCI_SORT_SLICE")) slice = std::max<int64_t>(1, std::atoll(e));
        const int64_t slice_cap = std::min<int64_t>(nstore, slice);
        DevBuf ctmp, vin, vout, rel;
        ctmp.ensure(slice_cap * sizeof(int32_t));
        if (form_ != kFP64) vin.ensure(slice_cap * sizeof(uint64_t));
        vout.ensure(slice_cap * sizeof(uint64_t));
        int64_t r0 = 0;
        while (r0 < n) {
            // the longest run of rows from r0 whose entries fit in a slice (a
            // row holds at most one entry per connection pattern, far below)
            int64_t r1 = std::upper_bound(rph.begin() + r0 + 1, rph.end(), rph[r0] + slice) - rph.begin() - 1;
            r1 = std::max(r1, r0 + 1);
            const int64_t k0 = rph[r0], cnt = rph[r1] - rph[r0];
            if (cnt > slice) throw std::runtime_error("H build: a row has more entries than a sort slice");
            const int64_t segs = r1 - r0;
            if (cnt > 0) {
                rel.ensure((segs + 1) * sizeof(uint32_t));
                k_rel_offsets<<<grid_for(segs + 1, 256), 256, 0, stream_>>>(rowptr_.as<int64_t>() + r0, segs + 1, k0,
                                                                           rel.as<uint32_t>());
                uint64_t* vals_in;
                if (form_ == kFP64) {
                    vals_in = val_.as<uint64_t>() + k0;
                } else {
                    k_pack_hilo<<<grid_for(cnt, 256), 256, 0, stream_>>>(cnt, hi_.as<float>() + k0,
                                                                         lo_.as<float>() + k0, vin.as<uint64_t>());
                    vals_in = vin.as<uint64_t>();
                }
                const uint32_t* ro = rel.as<uint32_t>();
                size_t tmp = 0;
                HIPCHK(rocprim::segmented_radix_sort_pairs(nullptr, tmp, col_.as<int32_t>() + k0, ctmp.as<int32_t>(),
                                                           vals_in, vout.as<uint64_t>(), (unsigned)cnt,
                                                           (unsigned)segs, ro, ro + 1, 0, endbit, stream_));
                scan_tmp_.ensure(tmp);
                HIPCHK(rocprim::segmented_radix_sort_pairs(scan_tmp_.p, tmp, col_.as<int32_t>() + k0,
                                                           ctmp.as<int32_t>(), vals_in, vout.as<uint64_t>(),
                                                           (unsigned)cnt, (unsigned)segs, ro, ro + 1, 0, endbit,
                                                           stream_));
                HIPCHK(hipMemcpyAsync(col_.as<int32_t>() + k0, ctmp.p, cnt * sizeof(int32_t),
                                      hipMemcpyDeviceToDevice, stream_));
                if (form_ == kFP64)
                    HIPCHK(hipMemcpyAsync(val_.as<double>() + k0, vout.p, cnt * sizeof(double),
                                          hipMemcpyDeviceToDevice, stream_));
                else
                    k_unpack_hilo<<<grid_for(cnt, 256), 256, 0, stream_>>>(cnt, vout.as<uint64_t>(),
                                                                           hi_.as<float>() + k0, lo_.as<float>() + k0);
                HIPCHK(hipGetLastError());
            }
            r0 = r1;
        }
        HIPCHK(hipStreamSynchronize(stream_));
    }
    D_.ensure(n * sizeof(double));
    k_diag<NW><<<grid_for(n, 128), 128, 0, stream_>>>(n, P_.M, a.keys, d_freq_.as<double>(), d_diag_.as<uint32_t>(),
                                                      (int)T_.diag_fcs.size(), a.elem, a.ladder, 256, a.perm,
                                                      D_.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    if (incr) {
        col_o_.release();
        val_o_.release();
        hi_o_.release();
        lo_o_.release();
    }
    built_n_ = n;
    st.n = n;
    st.nnz_offdiag = nnz;
    st.probed_rows = incr ? n - n_old : n;
    if (incr) {
        int64_t oe = 0;
        HIPCHK(hipMemcpy(&oe, rowptr_o_.as<int64_t>() + n_old, sizeof(int64_t), hipMemcpyDeviceToHost));
        st.carried = oe;
    }
}

HStats Engine::build_h() {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (n_ < 1) throw std::runtime_error("build_h: the basis is empty");
    auto t0 = clk::now();
    HStats st;
    switch (KW_) {
        case 1: build_h_t<4>(st); break;
        case 2: build_h_t<8>(st); break;
        case 3: build_h_t<12>(st); break;
        default: build_h_t<16>(st); break;
    }
    h_valid_ = true;
    st.build_ms = ms_since(t0);
    last_h_ = st;
    if (std::getenv("VHCI_DEBUG")) {
        size_t freeb = 0, totb = 0;
        HIPCHK(hipMemGetInfo(&freeb, &totb));
        static const char* const kind[3] = {"full CSR", "upper triangle", "matrix-free products"};
        std::fprintf(stderr,
                     "[h-build] n %lld off-diagonal entries %lld, %s (%lld stored) %.1f ms; context holds %.1f GB, "
                     "%.1f GB free\n",
                     (long long)n_, (long long)st.nnz_offdiag, kind[h_storage()], (long long)(matfree_ ? 0 : nnz_),
                     st.build_ms, pool_.held / 1e9, freeb / 1e9);
    }
    return st;
}

void Engine::get_h(std::vector<int64_t>& rowptr, std::vector<int32_t>& col, std::vector<double>& val,
                   std::vector<double>& diag) {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    if (matfree_) throw std::runtime_error("the Hamiltonian is matrix-free: no stored entries to return");
    rowptr.resize(n_ + 1);
    col.resize(nnz_);
    val.resize(nnz_);
    diag.resize(n_);
    HIPCHK(hipMemcpyAsync(rowptr.data(), rowptr_.p, (n_ + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    if (nnz_) HIPCHK(hipMemcpyAsync(col.data(), col_.p, nnz_ * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(diag.data(), D_.p, n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (nnz_) {
        if (form_ == kFP64) {
            HIPCHK(hipMemcpyAsync(val.data(), val_.p, nnz_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
        } else {
            std::vector<float> h(nnz_), l(nnz_);
            HIPCHK(hipMemcpyAsync(h.data(), hi_.p, nnz_ * sizeof(float), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipMemcpyAsync(l.data(), lo_.p, nnz_ * sizeof(float), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            for (int64_t k = 0; k < nnz_; k++) val[k] = (double)h[k] + (double)l[k];
        }
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (sym_) {
        // upper triangle -> both triangles (solver order, columns ascending:
        // row c takes its lower entries in increasing source row, then its
        // own upper ones)
        std::vector<int64_t> cnt(n_, 0);
        for (int64_t r = 0; r < n_; r++) {
            cnt[r] += rowptr[r + 1] - rowptr[r];
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) cnt[col[k]]++;
        }
        std::vector<int64_t> rp(n_ + 1, 0);
        for (int64_t r = 0; r < n_; r++) rp[r + 1] = rp[r] + cnt[r];
        std::vector<int32_t> c2(rp[n_]);
        std::vector<double> v2(rp[n_]);
        std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
        for (int64_t r = 0; r < n_; r++) {
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
                const int64_t c = col[k];
                c2[cur[c]] = (int32_t)r;
                v2[cur[c]++] = val[k];
            }
            // row r's lower entries all come from rows < r: placed already
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) {
                c2[cur[r]] = col[k];
                v2[cur[r]++] = val[k];
            }
        }
        rowptr.swap(rp);
        col.swap(c2);
        val.swap(v2);
    }
    const int64_t nnz_all = (int64_t)col.size();
    if (perm_on_) {
        // back to basis order: row perm[r] holds stored row r, columns mapped
        // through perm and sorted again
        std::vector<int32_t> perm(n_);
        HIPCHK(hipMemcpy(perm.data(), perm_.p, n_ * sizeof(int32_t), hipMemcpyDeviceToHost));
        std::vector<int32_t> inv(n_);
        for (int64_t r = 0; r < n_; r++) inv[perm[r]] = (int32_t)r;
        std::vector<int64_t> rp(n_ + 1, 0);
        for (int64_t i = 0; i < n_; i++) rp[i + 1] = rp[i] + (rowptr[inv[i] + 1] - rowptr[inv[i]]);
        std::vector<int32_t> c2(nnz_all);
        std::vector<double> v2(nnz_all), d2(n_);
        std::vector<std::pair<int32_t, double>> row;
        for (int64_t i = 0; i < n_; i++) {
            const int64_t r = inv[i];
            row.clear();
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) row.emplace_back(perm[col[k]], val[k]);
            std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t t = 0; t < row.size(); t++) {
                c2[rp[i] + t] = row[t].first;
                v2[rp[i] + t] = row[t].second;
            }
            d2[i] = diag[r];
        }
        rowptr.swap(rp);
        col.swap(c2);
        val.swap(v2);
        diag.swap(d2);
    }
}

void Engine::get_h_rows(const int64_t* rows, int64_t nr, std::vector<int64_t>& rowptr, std::vector<int32_t>& col,
                        std::vector<double>& val, std::vector<double>& diag) {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    if (matfree_) throw std::runtime_error("the Hamiltonian is matrix-free: no stored entries to return");
    std::vector<int32_t> perm, inv;
    if (perm_on_) {
        perm.resize(n_);
        HIPCHK(hipMemcpyAsync(perm.data(), perm_.p, n_ * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        inv.resize(n_);
        for (int64_t r = 0; r < n_; r++) inv[perm[r]] = (int32_t)r;
    }
    rowptr.assign(1, 0);
    col.clear();
    val.clear();
    diag.assign(nr, 0.0);
    for (int64_t k = 0; k < nr; k++)
        if (rows[k] < 0 || rows[k] >= n_) throw std::runtime_error("get_h_rows: row out of range");
    // half storage: the lower half of the requested rows (entries of the
    // stored triangle in their columns), as (solver row, value) lists
    std::vector<std::vector<std::pair<int32_t, double>>> lower;
    if (sym_ && nnz_ > 0) {
        lower.resize(nr);
        std::vector<int32_t> mk(n_, -1);
        for (int64_t k = 0; k < nr; k++) {
            const int64_t r = perm_on_ ? inv[rows[k]] : rows[k];
            if (mk[r] < 0) mk[r] = (int32_t)k;
        }
        DevBuf mark, ctr;
        mark.ensure(n_ * sizeof(int32_t));
        ctr.ensure(sizeof(unsigned long long));
        HIPCHK(hipMemcpyAsync(mark.p, mk.data(), n_ * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
        const bool f64 = form_ == kFP64;
        auto scan = [&](int32_t* oq, int32_t* orow, double* ov, int64_t cap) {
            k_fill<uint64_t><<<1, 64, 0, stream_>>>(ctr.as<uint64_t>(), 1, 0ull);
            k_lower_of<<<grid_for(n_, 4), 256, 0, stream_>>>(
                n_, rowptr_.as<int64_t>(), col_.as<int32_t>(), f64 ? val_.as<double>() : nullptr, hi_.as<float>(),
                lo_.as<float>(), mark.as<int32_t>(), ctr.as<unsigned long long>(), oq, orow, ov,
                (unsigned long long)cap);
            HIPCHK(hipGetLastError());
            unsigned long long c = 0;
            HIPCHK(hipMemcpyAsync(&c, ctr.p, sizeof(c), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            return (int64_t)c;
        };
        const int64_t m = scan(nullptr, nullptr, nullptr, 0);
        if (m > 0) {
            DevBuf oq, orow, ov;
            oq.ensure(m * sizeof(int32_t));
            orow.ensure(m * sizeof(int32_t));
            ov.ensure(m * sizeof(double));
            if (scan(oq.as<int32_t>(), orow.as<int32_t>(), ov.as<double>(), m) != m)
                throw std::runtime_error("get_h_rows: the lower-half scan changed between passes");
            std::vector<int32_t> hq(m), hr(m);
            std::vector<double> hv(m);
            HIPCHK(hipMemcpyAsync(hq.data(), oq.p, m * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipMemcpyAsync(hr.data(), orow.p, m * sizeof(int32_t), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipMemcpyAsync(hv.data(), ov.p, m * sizeof(double), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            for (int64_t t = 0; t < m; t++) lower[hq[t]].emplace_back(hr[t], hv[t]);
        }
        for (int64_t k = 0; k < nr; k++) {
            const int64_t r = perm_on_ ? inv[rows[k]] : rows[k];
            if (mk[r] != k) lower[k] = lower[mk[r]];
        }
    }
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t k = 0; k < nr; k++) {
        const int64_t b = rows[k];
        const int64_t r = perm_on_ ? inv[b] : b;
        int64_t se[2];
        HIPCHK(hipMemcpyAsync(se, rowptr_.as<int64_t>() + r, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(&diag[k], D_.as<double>() + r, sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        const int64_t cnt = se[1] - se[0];
        std::vector<int32_t> c(cnt);
        std::vector<double> v(cnt);
        if (cnt) {
            HIPCHK(hipMemcpyAsync(c.data(), col_.as<int32_t>() + se[0], cnt * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  stream_));
            if (form_ == kFP64) {
                HIPCHK(hipMemcpyAsync(v.data(), val_.as<double>() + se[0], cnt * sizeof(double), hipMemcpyDeviceToHost,
                                      stream_));
                HIPCHK(hipStreamSynchronize(stream_));
            } else {
                std::vector<float> h(cnt), l(cnt);
                HIPCHK(hipMemcpyAsync(h.data(), hi_.as<float>() + se[0], cnt * sizeof(float), hipMemcpyDeviceToHost,
                                      stream_));
                HIPCHK(hipMemcpyAsync(l.data(), lo_.as<float>() + se[0], cnt * sizeof(float), hipMemcpyDeviceToHost,
                                      stream_));
                HIPCHK(hipStreamSynchronize(stream_));
                for (int64_t t = 0; t < cnt; t++) v[t] = (double)h[t] + (double)l[t];
            }
        }
        row.clear();
        for (int64_t t = 0; t < cnt; t++) row.emplace_back(perm_on_ ? perm[c[t]] : c[t], v[t]);
        if (!lower.empty())
            for (const auto& e : lower[k]) row.emplace_back(perm_on_ ? perm[e.first] : e.first, e.second);
        std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (auto& e : row) {
            col.push_back(e.first);
            val.push_back(e.second);
        }
        rowptr.push_back((int64_t)col.size());
    }
}

// Structural nonzeros of the reference's triplet list (Ham.cpp:237-326):
// every pair whose quanta differ by an even L1 distance <= fcmax or an odd
// one <= fcmax-1 gets a triplet, zero or not; after H + H^T that is n plus
// twice the off-diagonal pairs.  Counted by probing the L1 ball.
int64_t Engine::structural_nnz() {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    if (T_.ball.empty() || n_ == 0) return -1;
    HIPCHK(hipSetDevice(device_));
    if (!d_ball_.p) {
        upload(d_ball_, T_.ball, stream_);
        upload(d_ball_walk_, walk_records(T_.ball), stream_);
        upload(d_boff_, T_.ball_off, stream_);
    }
    HBuildArgs a{};
    a.n = n_;
    a.M = P_.M;
    a.P = (int)T_.ball.size();
    a.keys = keys_.as<uint4>();
    a.slots = slots_.as<uint64_t>();
    a.mask = hmask_;
    a.bloom = bloom_.as<uint32_t>();
    a.bmask = bmask_;
    a.deltas = d_ball_.as<DeltaEnt>();
    a.walk = d_ball_walk_.as<uint64_t>();
    a.poff = d_boff_.as<uint32_t>();
    if (!a.walk) throw std::runtime_error("structural_nnz: walk records missing");
    rowcnt_.ensure((n_ + 1) * sizeof(int64_t));
    a.rowcnt = rowcnt_.as<int64_t>();
    switch (KW_) {
        case 1: k_hbuild<4, false><<<(unsigned)n_, 64, 0, stream_>>>(a); break;
        case 2: k_hbuild<8, false><<<(unsigned)n_, 64, 0, stream_>>>(a); break;
        case 3: k_hbuild<12, false><<<(unsigned)n_, 64, 0, stream_>>>(a); break;
        default: k_hbuild<16, false><<<(unsigned)n_, 64, 0, stream_>>>(a); break;
    }
    HIPCHK(hipGetLastError());
    std::vector<int64_t> cnt(n_);
    HIPCHK(hipMemcpyAsync(cnt.data(), rowcnt_.p, n_ * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    int64_t tot = n_;
    for (int64_t c : cnt) tot += c;
    return tot;
}

// ---------------------------------------------------------------- SpMM
int Engine::spmm_launches(int) const { return 1; }

// Full CSR when it fits beside the eigensolver's blocks, else the upper
// triangle, else matrix-free products.  VHCI_H_STORAGE=full|half|free forces
// one (VHCI_MATRIX_FREE=1 / 0 forces matrix-free / full).  The memory counted
// as available: free device memory, the previous matrix (released first) and
// the context pool's idle blocks (returned to the device when an
// allocation would fail).
int Engine::choose_storage(int64_t nnz, int64_t nnz_up, size_t words, bool* keep_words) {
    ks_mfac_ = 4.0;
    if (keep_words) *keep_words = false;
    if (const char* e = std::getenv("VHCI_H_STORAGE")) {
        const std::string v(e);
        if (v == "full") {
            if (keep_words) *keep_words = true;
            return 0;
        }
        if (v == "half") return 1;
        if (v == "free") return 2;
        throw std::runtime_error("VHCI_H_STORAGE must be full, half or free, not '" + v + "'");
    }
    if (const char* e = std::getenv("VHCI_MATRIX_FREE")) return std::string(e) == "1" ? 2 : 0;
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const double sv = 8.0;  // FP64 value, or FP32 hi + lo
    // the column sort's slice buffers
    const double sort = (double)kSortSlice * (4.0 + 8.0 + (form_ == kFP64 ? 0.0 : 8.0));
    const double n = (double)n_;
    // Krylov-Schur blocks: V, W (mfac nev + bs columns each), the kept Ritz
    // vectors and their products (1.6 nev each) and the resident Ritz block
    const double nev = std::max(1, P_.neig);
    auto eig = [&](double mfac) { return n * 8.0 * (2.0 * (mfac * nev + 8.0) + 2.0 * 1.6 * nev + nev); };
    const double held = (double)(col_.cap + val_.cap + hi_.cap + lo_.cap + col_o_.cap + val_o_.cap + hi_o_.cap +
                                 lo_o_.cap + hacc_.cap) +
                        (double)pool_.idle_bytes + (double)words;
    const double avail = 0.9 * ((double)free_b + held);
    // the sort's buffers are freed before the eigensolver starts
    const double full = (double)nnz * (sv + 4.0);
    // upper triangle, the row sums and the accumulator (n x 8)
    const double half = (double)nnz_up * (sv + 4.0) + n * 8.0 + n * 8.0 * 8.0;
    // full storage, with the Krylov-Schur basis down to 2.5 nev if need be
    // (more restarts cost less than the half form's slower products), then
    // half storage with the basis down to 2 nev
    for (double mf = 4.0; mf >= 2.5; mf -= 0.25)
        if (full + std::max(sort, eig(mf)) <= avail) {
            ks_mfac_ = mf;
            if (keep_words) *keep_words = full + sort + (double)words <= avail;
            return 0;
        }
    for (double mf = 4.0; mf >= 2.0; mf -= 0.25)
        if (half + std::max(sort, eig(mf)) <= avail) {
            ks_mfac_ = mf;
            return 1;
        }
    return 2;
}

template <int NW>
void Engine::spmm_free_t(const double* X, int ldx, double* Y, int ldy, int b) {
    HFreeArgs f{};
    HBuildArgs& a = f.h;
    a.n = n_;
    a.M = P_.M;
    a.P = P_deltas_;
    a.nq = 256;
    a.keys = keys_.as<uint4>();
    a.slots = slots_.as<uint64_t>();
    a.mask = hmask_;
    a.bloom = bloom_.as<uint32_t>();
    a.bmask = bmask_;
    a.deltas = d_deltas_.as<DeltaEnt>();
    a.walk = d_walk_.as<uint64_t>();
    if (!a.walk) throw std::runtime_error("pattern walk records missing");
    a.poff = d_doff_.as<uint32_t>();
    a.groups = d_groups_.as<GroupEnt>();
    a.fclist = d_fclist_.as<uint32_t>();
    a.elem = d_elem_.as<ElemFC>();
    a.ladder = d_ladder_.as<double>();
    a.perm = perm_on_ ? perm_.as<int32_t>() : nullptr;
    a.inv = perm_on_ ? inv_.as<int32_t>() : nullptr;
    f.D = D_.as<double>();
    for (int c0 = 0; c0 < b; c0 += 16) {
        f.X = X + c0;
        f.ldx = ldx;
        f.Y = Y + c0;
        f.ldy = ldy;
        f.bcols = std::min(16, b - c0);
        if (f.bcols <= 8)
            k_hfree<NW, 8><<<(unsigned)n_, 64, 0, stream_>>>(f);
        else
            k_hfree<NW, 16><<<(unsigned)n_, 64, 0, stream_>>>(f);
        HIPCHK(hipGetLastError());
    }
}

// Half storage: B = 4, 8 or 16 dense columns go straight to k_spmm_half;
// other widths and strided blocks through zero-padded dense chunks.
void Engine::spmm_half(const double* X, int ldx, double* Y, int ldy, int b, bool low, const float* X32,
                       int ldx32) {
    const int64_t n = n_;
    const int mode = form_ == kFP64 ? 0 : (low ? 1 : 2);
    const bool dense = ldx == b && ((uintptr_t)X & 15) == 0 && (b == 4 || b == 8) &&
                       (mode != 1 || !X32 || (ldx32 == b && ((uintptr_t)X32 & 15) == 0));
    if (!dense) {
        const int B = b <= 4 ? 4 : 8;
        hpack_x_.ensure((size_t)n * B * sizeof(double));
        hpack_y_.ensure((size_t)n * B * sizeof(double));
        const unsigned g = std::min<unsigned>(grid_for(n * B, 256), 8192);
        for (int c0 = 0; c0 < b; c0 += B) {
            const int bc = std::min(B, b - c0);
            k_pack_cols<<<g, 256, 0, stream_>>>(X + c0, ldx, n, bc, B, hpack_x_.as<double>());
            HIPCHK(hipGetLastError());
            spmm_half(hpack_x_.as<double>(), B, hpack_y_.as<double>(), B, B, low, nullptr, 0);
            k_unpack_cols<<<g, 256, 0, stream_>>>(hpack_y_.as<double>(), B, n, bc, Y + c0, ldy);
            HIPCHK(hipGetLastError());
        }
        return;
    }
    if (mode == 1 && !X32) {
        X32_.ensure((size_t)n * b * sizeof(float));
        k_to_f32<<<grid_for(n * b, 256), 256, 0, stream_>>>(X, ldx, X32_.as<float>(), b, n, b);
        X32 = X32_.as<float>();
    }
    // the accumulator (n x 8) starts at zero and every product leaves it so
    const int64_t acc_elems = n * 8;
    if (hacc_zero_ < (size_t)acc_elems * sizeof(unsigned long long)) {
        hacc_.ensure((size_t)acc_elems * sizeof(unsigned long long));
        k_fill<uint64_t><<<std::min<unsigned>(grid_for(acc_elems, 256), 8192), 256, 0, stream_>>>(
            hacc_.as<uint64_t>(), acc_elems, 0ull);
        hacc_zero_ = (size_t)acc_elems * sizeof(unsigned long long);
    }
    hcmax_.ensure(16 * sizeof(unsigned long long));
    k_fill<uint64_t><<<1, 64, 0, stream_>>>(hcmax_.as<uint64_t>(), b, 0ull);
    k_colmax<<<std::min<unsigned>(grid_for(n * b, 256), 2048), 256, 0, stream_>>>(
        n, b, X, ldx, hcmax_.as<unsigned long long>());
    HIPCHK(hipGetLastError());
    const int64_t* rp = rowptr_.as<int64_t>();
    const int32_t* cl = col_.as<int32_t>();
    const double* vl = val_.as<double>();
    const float* h = hi_.as<float>();
    const float* l = lo_.as<float>();
    const double* D = D_.as<double>();
    const unsigned long long* cm = hcmax_.as<unsigned long long>();
    unsigned long long* acc = hacc_.as<unsigned long long>();
    const unsigned grid = grid_for(n, 4);
    const unsigned gf = std::min<unsigned>(grid_for(n * b, 256), 8192);
    auto go = [&](auto tag) {
        constexpr int B = decltype(tag)::value;
        if (mode == 0)
            k_spmm_half<B, 0><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, hnorm_, cm, acc);
        else if (mode == 1)
            k_spmm_half<B, 1><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, hnorm_, cm, acc);
        else
            k_spmm_half<B, 2><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, hnorm_, cm, acc);
        HIPCHK(hipGetLastError());
        k_half_finish<B><<<gf, 256, 0, stream_>>>(n, acc, hnorm_, cm, Y, ldy);
        HIPCHK(hipGetLastError());
    };
    if (b == 4) go(std::integral_constant<int, 4>());
    else go(std::integral_constant<int, 8>());
}

void Engine::spmm(const double* X, int ldx, double* Y, int ldy, int b, bool low, const float* X32, int ldx32) {
    const int64_t n = n_;
    if (matfree_) {
        // FP64 throughout (there are no stored FP32 values to use)
        (void)low;
        (void)X32;
        (void)ldx32;
        switch (KW_) {
            case 1: spmm_free_t<4>(X, ldx, Y, ldy, b); break;
            case 2: spmm_free_t<8>(X, ldx, Y, ldy, b); break;
            case 3: spmm_free_t<12>(X, ldx, Y, ldy, b); break;
            default: spmm_free_t<16>(X, ldx, Y, ldy, b); break;
        }
        return;
    }
    if (sym_) {
        spmm_half(X, ldx, Y, ldy, b, low, X32, ldx32);
        return;
    }
    const int64_t* rp = rowptr_.as<int64_t>();
    const int32_t* cl = col_.as<int32_t>();
    const double* D = D_.as<double>();
    const int mode = form_ == kFP64 ? 0 : (low ? 1 : 2);
    if (mode == 1 && !X32) {
        X32_.ensure((size_t)n * b * sizeof(float));
        k_to_f32<<<grid_for(n * b, 256), 256, 0, stream_>>>(X, ldx, X32_.as<float>(), b, n, b);
        X32 = X32_.as<float>();
        ldx32 = b;
    }
    const double* vl = val_.as<double>();
    const float* h = hi_.as<float>();
    const float* l = lo_.as<float>();
    // rows per XCD run (xcd_block); VHCI_XCD_RUN_ROWS=0 keeps the plain order
    int run_rows = 1024;
    if (const char* e = std::getenv("VHCI_XCD_RUN_ROWS")) run_rows = std::max(0, std::atoi(e));
    // workgroups per run for rpb rows per workgroup, and the grid padded to whole rounds of runs
    auto xc_of = [&](int rpb) { return run_rows > 0 ? std::max(1, run_rows / rpb) : 0; };
    auto pad = [](unsigned g, int xc) {
        return xc > 0 ? (unsigned)((g + 8u * xc - 1) / (8u * xc) * (8u * xc)) : g;
    };
    if (b == 1) {
        const int xc = xc_of(16);
        const unsigned grid = pad(grid_for(n, 16), xc);
        if (mode == 0) k_spmv<0><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, xc);
        else if (mode == 1) k_spmv<1><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, xc);
        else k_spmv<2><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, xc);
        HIPCHK(hipGetLastError());
        return;
    }
    const unsigned grid = grid_for(n, 4);
    const bool rec = prof_.b > 0 && b == prof_.b;
    if (rec) {
        if (prof_.used + 2 > prof_.ev.size()) {
            for (int i = 0; i < 64; i++) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                prof_.ev.push_back(e);
            }
        }
        HIPCHK(hipEventRecord(prof_.ev[prof_.used], stream_));
        prof_.bytes += spmm_bytes(b, mode);
    }
    struct RecEnd {
        Engine* e;
        bool on;
        ~RecEnd() {
            if (on) {
                (void)hipEventRecord(e->prof_.ev[e->prof_.used + 1], e->stream_);
                e->prof_.used += 2;
            }
        }
    } rec_end{this, rec};
    const bool dense_x = ldx == b && ((uintptr_t)X & 15) == 0 &&
                         (mode != 1 || (ldx32 == b && ((uintptr_t)X32 & 15) == 0));
    if (dense_x && (b == 4 || b == 8 || b == 16)) {
        // rows of fewer than 128 entries on average go two to a wave
        bool half = nnz_ < 128 * n;
        if (const char* e = std::getenv("VHCI_SPMM_ROW_LANES")) half = std::atoi(e) == 32;
        const int xc = xc_of(half ? 8 : 4);
        auto go = [&](auto tag) {
            constexpr int B = decltype(tag)::value;
            if (half) {
                const unsigned g2 = pad(grid_for(n, 8), xc);
                if (mode == 0)
                    k_spmm_lane<B, 0, 32><<<g2, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, xc);
                else if (mode == 1)
                    k_spmm_lane<B, 1, 32><<<g2, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, xc);
                else
                    k_spmm_lane<B, 2, 32><<<g2, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, xc);
                return;
            }
            const unsigned g1 = pad(grid, xc);
            if (mode == 0) k_spmm_lane<B, 0><<<g1, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, xc);
            else if (mode == 1) k_spmm_lane<B, 1><<<g1, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, xc);
            else k_spmm_lane<B, 2><<<g1, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, X, Y, ldy, xc);
        };
        if (b == 4) go(std::integral_constant<int, 4>());
        else if (b == 8) go(std::integral_constant<int, 8>());
        else go(std::integral_constant<int, 16>());
        HIPCHK(hipGetLastError());
        return;
    }
    auto go = [&](auto tag) {
        constexpr int B = decltype(tag)::value;
        if (mode == 0)
            k_spmm<B, 0><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, ldx32, X, ldx, Y, ldy, b);
        else if (mode == 1)
            k_spmm<B, 1><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, ldx32, X, ldx, Y, ldy, b);
        else
            k_spmm<B, 2><<<grid, 256, 0, stream_>>>(n, rp, cl, vl, h, l, D, X32, ldx32, X, ldx, Y, ldy, b);
    };
    if (b <= 2) go(std::integral_constant<int, 2>());
    else if (b <= 4) go(std::integral_constant<int, 4>());
    else if (b <= 8) go(std::integral_constant<int, 8>());
    else if (b <= 16) go(std::integral_constant<int, 16>());
    else if (b <= 32) go(std::integral_constant<int, 32>());
    else go(std::integral_constant<int, 64>());
    HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------- H X
// without allocation.  The staging of apply_h_noalloc lives in `scratch`, a
// fixed array in the code object: each GPU the library runs on holds its own
// copy (112 MiB) until the process ends.  Host code never names the array
// in a kernel argument or copy; it uses the address hipGetSymbolAddress
// gives for the current device (noalloc_scratch).
constexpr int64_t kScratchElems = int64_t(4) << 20;  // n * b at most
constexpr int64_t kScratchDoubles = 3 * kScratchElems + kScratchElems / 2;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;                    // device -> address of its scratch
std::map<int, std::unique_ptr<std::mutex>> g_scratch_use;  // device -> lock held by a call
}  // namespace

// the device's scratch address (looked up once per device, with that device
// current) and the lock that serialises the calls using it
static double* noalloc_scratch(int device, std::mutex** lock) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("address of the no-allocation scratch array: ") +
                                     hipGetErrorString(e));
        it = g_scratch_ptr.emplace(device, static_cast<double*>(p)).first;
        g_scratch_use.emplace(device, std::make_unique<std::mutex>());
    }
    if (lock) *lock = g_scratch_use[device].get();
    return it->second;
}

// rows of a row-major n x b block moved through perm: gather (dst row r =
// src row perm[r]) or scatter (dst row perm[r] = src row r)
__global__ void k_perm_rows(const double* __restrict__ src, double* __restrict__ dst, int64_t n, int b,
                            const int32_t* __restrict__ perm, int gather) {
    const int64_t total = n * b;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / b;
        const int c = (int)(t - r * b);
        const int64_t pr = perm[r];
        if (gather)
            dst[t] = src[pr * b + c];
        else
            dst[pr * b + c] = src[t];
    }
}

void Engine::apply_h_noalloc(const double* X, double* Y, int b, bool fp32) {
    HIPCHK(hipSetDevice(device_));
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    if (b < 1 || b > 64) throw std::runtime_error("apply_h: 1..64 vectors");
    if (fp32 && form_ != kMixed) throw std::runtime_error("apply_h: the FP32 product exists in the mixed form only");
    if (fp32 && matfree_) throw std::runtime_error("apply_h: matrix-free products are FP64 only");
    const int64_t n = n_;
    const int64_t cnt = n * b;
    if (cnt > kScratchElems)
        throw std::length_error("apply_h_noalloc: n * b = " + std::to_string(cnt) + " is larger than the scratch limit of " +
                                std::to_string(kScratchElems) + " elements");
    std::mutex* lock = nullptr;
    double* S = noalloc_scratch(device_, &lock);
    std::lock_guard<std::mutex> lk(*lock);
    double* A = S;                 // X, then Y, in basis order
    double* B = S + cnt;           // X in solver order
    double* C = S + 2 * cnt;       // Y in solver order
    float* X32 = reinterpret_cast<float*>(S + 3 * cnt);  // FP32 copy of the solver-order X
    HIPCHK(hipMemcpyAsync(A, X, (size_t)cnt * sizeof(double), hipMemcpyHostToDevice, stream_));
    const unsigned g = std::min<unsigned>(grid_for(cnt, 256), 8192);
    const double* xs = A;
    if (perm_on_) {
        k_perm_rows<<<g, 256, 0, stream_>>>(A, B, n, b, perm_.as<int32_t>(), 1);
        HIPCHK(hipGetLastError());
        xs = B;
    }
    if (fp32) {
        k_to_f32<<<grid_for(cnt, 256), 256, 0, stream_>>>(xs, b, X32, b, n, b);
        HIPCHK(hipGetLastError());
    }
    spmm(xs, b, C, b, b, fp32, fp32 ? X32 : nullptr, b);
    const double* ys = C;
    if (perm_on_) {
        k_perm_rows<<<g, 256, 0, stream_>>>(C, A, n, b, perm_.as<int32_t>(), 0);
        HIPCHK(hipGetLastError());
        ys = A;
    }
    HIPCHK(hipMemcpyAsync(Y, ys, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
}

// plain device copy of `bytes` bytes (16-byte words), timed with HIP events:
// the calibration kernel for memory-counter readings
// four 16-byte loads in flight per thread (n a multiple of 4 * the grid's
// thread count is not required: the tail is copied one word at a time)
typedef unsigned cp_u4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const cp_u4* __restrict__ a, cp_u4* __restrict__ b, int64_t n) {
    const int64_t nt = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t i = t;
    for (; i + 3 * nt < n; i += 4 * nt) {
        const cp_u4 v0 = __builtin_nontemporal_load(a + i);
        const cp_u4 v1 = __builtin_nontemporal_load(a + i + nt);
        const cp_u4 v2 = __builtin_nontemporal_load(a + i + 2 * nt);
        const cp_u4 v3 = __builtin_nontemporal_load(a + i + 3 * nt);
        __builtin_nontemporal_store(v0, b + i);
        __builtin_nontemporal_store(v1, b + i + nt);
        __builtin_nontemporal_store(v2, b + i + 2 * nt);
        __builtin_nontemporal_store(v3, b + i + 3 * nt);
    }
    for (; i < n; i += nt) b[i] = a[i];
}

double Engine::copy_probe(size_t bytes) {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    const int64_t nw = (int64_t)(bytes / 16);
    DevBuf a, b;
    a.ensure(nw * 16);
    b.ensure(nw * 16);
    k_fill<uint64_t><<<grid_for(2 * nw, 256), 256, 0, stream_>>>(a.as<uint64_t>(), 2 * nw, 1ull);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const unsigned g = std::min<unsigned>(grid_for(nw, 256), 4096);
    // the first copy touches both buffers; the second is timed
    k_copy16<<<g, 256, 0, stream_>>>(a.as<cp_u4>(), b.as<cp_u4>(), nw);
    HIPCHK(hipEventRecord(e0, stream_));
    k_copy16<<<g, 256, 0, stream_>>>(a.as<cp_u4>(), b.as<cp_u4>(), nw);
    HIPCHK(hipEventRecord(e1, stream_));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

void Engine::apply_h(const double* X, double* Y, int b, bool fp32) {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    if (b < 1 || b > 64) throw std::runtime_error("apply_h: the block width must be 1 to 64");
    if (fp32 && form_ != kMixed) throw std::runtime_error("apply_h: the FP32 product exists in the mixed form only");
    if (fp32 && matfree_) throw std::runtime_error("apply_h: matrix-free products are FP64 only");
    const int64_t n = n_;
    const size_t cnt = (size_t)n * b;
    std::vector<int32_t> perm;
    if (perm_on_) {
        perm.resize(n);
        HIPCHK(hipMemcpy(perm.data(), perm_.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    // stored row r is basis state perm[r]
    std::vector<double> xs(cnt), ys(cnt);
    for (int64_t r = 0; r < n; r++) {
        const int64_t i = perm_on_ ? perm[r] : r;
        std::copy(X + i * b, X + (i + 1) * b, xs.begin() + r * b);
    }
    DevBuf xd, yd;
    xd.ensure(cnt * sizeof(double));
    yd.ensure(cnt * sizeof(double));
    HIPCHK(hipMemcpyAsync(xd.p, xs.data(), cnt * sizeof(double), hipMemcpyHostToDevice, stream_));
    spmm(xd.as<double>(), b, yd.as<double>(), b, b, fp32);
    HIPCHK(hipMemcpyAsync(ys.data(), yd.p, cnt * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int64_t r = 0; r < n; r++) {
        const int64_t i = perm_on_ ? perm[r] : r;
        std::copy(ys.begin() + r * b, ys.begin() + (r + 1) * b, Y + i * b);
    }
}

double Engine::time_spmm(int b, int reps, bool fp32, double* bytes_per_call) {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    const int64_t n = n_;
    DevBuf xb, yb;
    xb.ensure((size_t)n * b * sizeof(double));
    yb.ensure((size_t)n * b * sizeof(double));
    k_fill<double><<<grid_for(n * b, 256), 256, 0, stream_>>>(xb.as<double>(), n * b, 1.0 / std::sqrt((double)n));
    bool low = fp32 && form_ == kMixed;
    spmm(xb.as<double>(), b, yb.as<double>(), b, b, low);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, stream_));
    for (int r = 0; r < reps; r++) spmm(xb.as<double>(), b, yb.as<double>(), b, b, low);
    HIPCHK(hipEventRecord(e1, stream_));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bytes_per_call) *bytes_per_call = spmm_bytes(b, form_ == kFP64 ? 0 : (low ? 1 : 2));
    return ms / reps;
}

double Engine::spmm_bytes(int b, int mode) const {
    const double n = (double)n_;
    const double sv = mode == 1 ? 4.0 : 8.0;  // stored value bytes (hi + lo counts 8)
    const double sx = mode == 1 ? 4.0 : 8.0;  // gathered X element bytes
    // FP32 products also read the FP64 X row of the diagonal term
    const double xdiag = mode == 1 ? n * b * 8.0 : 0.0;
    // half storage: the upper triangle's entries, with the vector traffic
    // counted twice (the transposed products read and write rows again)
    const double vec = (sym_ ? 2.0 : 1.0) * (n * b * sx + n * b * 8.0);
    return (double)nnz_ * (sv + 4) + 8.0 * (n + 1) + 8.0 * n + vec + xdiag;
}

void Engine::spmm_profile(int b) {
    HIPCHK(hipStreamSynchronize(stream_));
    prof_.b = b;
    prof_.used = 0;
    prof_.bytes = 0;
    prof_.gemm_ms = 0;
    prof_.gemm_flops = 0;
    prof_.gemm_calls = 0;
}

void Engine::gemm_profile_read(double out[3]) {
    out[0] = prof_.gemm_ms;
    out[1] = prof_.gemm_flops;
    out[2] = (double)prof_.gemm_calls;
}

void Engine::phase_stats(double out[16]) const {
    out[0] = (double)last_h_.n;
    out[1] = (double)last_h_.nnz_offdiag;
    out[2] = (double)P_deltas_;
    out[3] = last_h_.build_ms;
    out[4] = (double)last_hb_.candidates;
    out[5] = (double)last_hb_.unique;
    out[6] = last_hb_.ms;
    out[7] = (double)last_pt2_.ns2;
    out[8] = (double)last_pt2_.pairs;
    out[9] = last_pt2_.sum_ms;
    out[10] = 16.0 * KW_;
    out[11] = (double)last_pt2_nopt_;
    out[12] = (double)last_h_.probed_rows;
    out[13] = (double)last_h_.carried;
    out[14] = (double)last_pt2_.chunks;
    out[15] = (double)h_storage();
}

void Engine::spmm_profile_read(double out[3]) {
    HIPCHK(hipStreamSynchronize(stream_));
    double ms = 0;
    for (size_t i = 0; i + 1 < prof_.used; i += 2) {
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, prof_.ev[i], prof_.ev[i + 1]));
        ms += t;
    }
    out[0] = ms;
    out[1] = prof_.bytes;
    out[2] = (double)(prof_.used / 2);
}

// ---------------------------------------------------------------- eigensolver
void Engine::dense_solve(int nev, std::vector<double>& E, EigStats& st) {
    const int64_t n = n_;
    DevBuf A, w, e;
    A.ensure((size_t)n * n * sizeof(double));
    w.ensure(n * sizeof(double));
    e.ensure(n * sizeof(double));
    info_.ensure(sizeof(int));
    if (matfree_) {
        // H times identity columns, 16 at a time: each entry is one matrix
        // element plus exact zeros, so A is the same as from stored entries
        DevBuf xi, yb;
        xi.ensure((size_t)n * 16 * sizeof(double));
        yb.ensure((size_t)n * 16 * sizeof(double));
        for (int64_t c0 = 0; c0 < n; c0 += 16) {
            const int bc = (int)std::min<int64_t>(16, n - c0);
            k_identity_block<<<grid_for(n * 16, 256), 256, 0, stream_>>>(xi.as<double>(), n, 16, c0, bc);
            spmm(xi.as<double>(), 16, yb.as<double>(), 16, 16, false);
            dim3 tg(grid_for(n, 32), 1);
            k_rm_to_cm<<<tg, 256, 0, stream_>>>(yb.as<double>(), 16, n, bc, A.as<double>() + c0 * n, n, nullptr);
            HIPCHK(hipGetLastError());
        }
    } else {
        k_fill<double><<<grid_for(n * n, 256), 256, 0, stream_>>>(A.as<double>(), n * n, 0.0);
        k_densify<<<(unsigned)n, 64, 0, stream_>>>(n, rowptr_.as<int64_t>(), col_.as<int32_t>(),
                                                   form_ == kFP64 ? val_.as<double>() : nullptr, hi_.as<float>(),
                                                   lo_.as<float>(), D_.as<double>(), A.as<double>(), sym_ ? 1 : 0);
        HIPCHK(hipGetLastError());
    }
    // The lower triangle is reduced from the first column, as Eigen's
    // SelfAdjointEigenSolver (src/Ham.cpp:372-379) and LAPACK's lower form
    // do.  For exactly degenerate eigenvalues (symmetry-equivalent modes)
    // the eigenvectors then come out in the same rotation as the
    // reference's, which fixes Cmax and so the first heat-bath selection;
    // reducing the upper triangle picks another rotation.
    BLASCHK(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, A.as<double>(),
                             (rocblas_int)n, w.as<double>(), e.as<double>(), info_.as<rocblas_int>()));
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    std::vector<double> wh(n);
    HIPCHK(hipMemcpyAsync(wh.data(), w.p, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (info != 0) throw std::runtime_error("Error: Eigenvalues did not converge.");
    E.assign(wh.begin(), wh.begin() + nev);
    X_.ensure((size_t)n * nev * sizeof(double));
    k_colmajor_to_rows<<<grid_for(n * nev, 256), 256, 0, stream_>>>(A.as<double>(), n, n, nev, X_.as<double>(),
                                                                     perm_on_ ? perm_.as<int32_t>() : nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    xrows_ = n;
    xcols_ = nev;
    st.dense = true;
}

EigStats Engine::dense_all(std::vector<double>& E) {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    auto t0 = clk::now();
    EigStats st;
    dense_solve((int)n_, E, st);
    st.ms = ms_since(t0);
    return st;
}

// workgroups of 256 threads of `kernel` that the current device holds at
// once (all CUs), kept per device and kernel
static int resident_groups(const void* kernel) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, int> memo;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = memo.find({dev, kernel});
    if (it != memo.end()) return it->second;
    int per_cu = 0, cus = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int r = std::max(1, per_cu) * std::max(1, cus);
    memo[{dev, kernel}] = r;
    return r;
}

// rows per chunk of tsm_tn's partial products for an n x p by n x q product
static int64_t tsm_rows_per_chunk(int64_t n, int p, int q) {
    // Row chunks (>= 256 rows each) times output tiles make the grid.  The
    // chunk count is set so that the grid is one wave of the workgroups the
    // device holds at once: every workgroup does the same work, so a second,
    // partly filled wave would cost as long as the first (384 chunks of the
    // 128-column tiles of a 296-column basis are 1,152 workgroups against
    // 1,024 resident).  At most 512 chunks, which k_tsm_reduce sums per
    // output.  VHCI_TSM_WAVES=k: k waves; 0: 384 chunks whatever the tiles.
    const void* kern;
    int tiles;
    if (q <= 8) {
        kern = reinterpret_cast<const void*>(&k_tsm_tn<128, 8>);
        tiles = (p + 127) / 128;
    } else if (q > 16 && p > 16) {
        kern = reinterpret_cast<const void*>(&k_tsm_tn_wide);
        tiles = ((p + 63) / 64) * ((q + 63) / 64);
    } else {
        kern = reinterpret_cast<const void*>(&k_tsm_tn<64, 16>);
        tiles = ((p + 63) / 64) * ((q + 15) / 16);
    }
    const char* we = std::getenv("VHCI_TSM_WAVES");
    const int waves = we ? std::max(0, std::atoi(we)) : 1;
    const int64_t want =
        waves == 0 ? 384 : std::clamp<int64_t>((int64_t)waves * resident_groups(kern) / tiles, 1, 512);
    return std::max<int64_t>(256, ((n + want - 1) / want + 31) / 32 * 32);
}

// bytes of tsm_tn's partial products for an n x p by n x q product
static size_t tsm_part_bytes(int64_t n, int p, int q) {
    const int64_t rpc = tsm_rows_per_chunk(n, p, q);
    return (size_t)((n + rpc - 1) / rpc) * p * q * sizeof(double);
}

// C (p x q, col-major, ldc) = A^T B for column-major A (n x p), B (n x q)
static void tsm_tn(hipStream_t s, DevBuf& part, int64_t n, int p, int q, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double* C, int ldc) {
    const int64_t rpc = tsm_rows_per_chunk(n, p, q);
    const int nchunk = (int)((n + rpc - 1) / rpc);
    part.ensure((size_t)nchunk * p * q * sizeof(double));
    if (q <= 8) {
        dim3 g(nchunk, (p + 127) / 128, 1);
        k_tsm_tn<128, 8><<<g, 256, 0, s>>>(n, p, q, A, lda, B, ldb, rpc, part.as<double>());
    } else if (q > 16 && p > 16) {
        dim3 g(nchunk, (p + 63) / 64, (q + 63) / 64);
        k_tsm_tn_wide<<<g, 256, 0, s>>>(n, p, q, A, lda, B, ldb, rpc, part.as<double>());
    } else {
        dim3 g(nchunk, (p + 63) / 64, (q + 15) / 16);
        k_tsm_tn<64, 16><<<g, 256, 0, s>>>(n, p, q, A, lda, B, ldb, rpc, part.as<double>());
    }
    k_tsm_reduce<<<grid_for((int64_t)p * q, 16), 256, 0, s>>>(part.as<double>(), nchunk, p, q, C, ldc);
    HIPCHK(hipGetLastError());
}

// Y = beta Y + alpha A S (A n x p, S p x q col-major with lds); gate (on the
// device, null: none): skipped when *gate is 0
static void tsm_nn(hipStream_t s, int64_t n, int p, int q, double alpha, const double* A, int64_t lda,
                   const double* S, int lds, double beta, double* Y, int64_t ldy, const double* gate = nullptr) {
    // the eigensolver's blocks (q = 8) take the 8-column form: half the
    // FMAs and LDS of the 16-column one, and 8 loads in flight per wave
    // (15-mode eigensolver 799 -> 780 ms, the same results bit for bit)
    for (int a0 = 0; a0 < p; a0 += 256) {
        const int pp = std::min(256, p - a0);
        const double b = a0 == 0 ? beta : 1.0;
        if (q <= 8) {
            dim3 g(std::min<unsigned>(grid_for(n, 256), 2048), 1);
            k_tsm_nn<8, 8><<<g, 256, 0, s>>>(n, pp, q, alpha, A + a0 * lda, lda, S + a0, lds, b, Y, ldy, gate);
        } else {
            dim3 g(std::min<unsigned>(grid_for(n, 256), 2048), (q + 15) / 16);
            k_tsm_nn<16, 4><<<g, 256, 0, s>>>(n, pp, q, alpha, A + a0 * lda, lda, S + a0, lds, b, Y, ldy, gate);
        }
    }
    HIPCHK(hipGetLastError());
}

// Symmetric eigendecomposition of a small k x k matrix (k <= 32) by cyclic
// Jacobi rotations; lam ascending, U column-major.
static void small_eigh(int k, std::vector<double> A, std::vector<double>& lam, std::vector<double>& U) {
    U.assign((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++) U[(size_t)i * k + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0, tot = 0;
        for (int j = 0; j < k; j++)
            for (int i = 0; i < k; i++) {
                const double v = A[(size_t)j * k + i] * A[(size_t)j * k + i];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < k - 1; p++)
            for (int q = p + 1; q < k; q++) {
                const double apq = A[(size_t)q * k + p];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * k + p], aqq = A[(size_t)q * k + q];
                const double tau = (aqq - app) / (2 * apq);
                const double tt = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1 + tau * tau));
                const double c = 1 / std::sqrt(1 + tt * tt), sn = tt * c;
                for (int r = 0; r < k; r++) {  // columns p, q
                    const double x = A[(size_t)p * k + r], y = A[(size_t)q * k + r];
                    A[(size_t)p * k + r] = c * x - sn * y;
                    A[(size_t)q * k + r] = sn * x + c * y;
                }
                for (int r = 0; r < k; r++) {  // rows p, q
                    const double x = A[(size_t)r * k + p], y = A[(size_t)r * k + q];
                    A[(size_t)r * k + p] = c * x - sn * y;
                    A[(size_t)r * k + q] = sn * x + c * y;
                }
                for (int r = 0; r < k; r++) {
                    const double x = U[(size_t)p * k + r], y = U[(size_t)q * k + r];
                    U[(size_t)p * k + r] = c * x - sn * y;
                    U[(size_t)q * k + r] = sn * x + c * y;
                }
            }
    }
    std::vector<int> idx(k);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return A[(size_t)x * k + x] < A[(size_t)y * k + y]; });
    lam.resize(k);
    std::vector<double> Us((size_t)k * k);
    for (int c = 0; c < k; c++) {
        lam[c] = A[(size_t)idx[c] * k + idx[c]];
        for (int r = 0; r < k; r++) Us[(size_t)c * k + r] = U[(size_t)idx[c] * k + r];
    }
    U.swap(Us);
}

// Y = H X for the k column-major vectors of X (ld), through row-major
// staging in chunks of at most 64 columns.  low selects the mixed form's
// FP32 product.
void Engine::apply_cm(const double* X, double* Y, int64_t ld, int k, bool low, EigStats& st, double* Y2) {
    const int64_t n = n_;
    const bool f32 = low && form_ == kMixed;
    for (int c = 0; c < k; c += 64) {
        const int kk = std::min(64, k - c);
        stage_in_.ensure((size_t)n * kk * sizeof(double));
        stage_out_.ensure((size_t)n * kk * sizeof(double));
        dim3 tg(grid_for(n, 32), (kk + 31) / 32);
        k_cm_to_rm<double><<<tg, 256, 0, stream_>>>(X + c * ld, ld, n, kk, stage_in_.as<double>(), kk, nullptr);
        if (f32) {
            stage_in32_.ensure((size_t)n * kk * sizeof(float));
            k_cm_to_rm<float><<<tg, 256, 0, stream_>>>(X + c * ld, ld, n, kk, stage_in32_.as<float>(), kk, nullptr);
            spmm(stage_in_.as<double>(), kk, stage_out_.as<double>(), kk, kk, true, stage_in32_.as<float>(), kk);
        } else {
            spmm(stage_in_.as<double>(), kk, stage_out_.as<double>(), kk, kk, false);
        }
        k_rm_to_cm<<<tg, 256, 0, stream_>>>(stage_out_.as<double>(), kk, n, kk, Y + c * ld, ld,
                                            Y2 ? Y2 + c * ld : nullptr);
        HIPCHK(hipGetLastError());
        st.spmm_calls++;
    }
    st.spmm_vectors += k;
    if (!f32) st.fp64_spmm_vectors += k;
}

// pseudo-random column-major block, reproducible from the seed
__global__ void k_random_cm(double* V, int64_t ld, int64_t n, int k, uint64_t seed) {
    const int c = blockIdx.y;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed ^ ((uint64_t)c * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)r * 0xD1B54A32D192ED03ULL);
        z = fmix64(z);
        V[r + c * ld] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}

void Engine::random_cm(double* V, int64_t ld, int k, uint64_t seed) {
    dim3 g(std::min<unsigned>(grid_for(n_, 256), 1024), k);
    k_random_cm<<<g, 256, 0, stream_>>>(V, ld, n_, k, seed);
    HIPCHK(hipGetLastError());
}

// Orthonormalise V[:, m0:m0+k] against V[:, :m0] (two passes of block
// classical Gram-Schmidt) and within itself (SVQB: eigendecomposition of the
// scaled Gram matrix, near-dependent directions dropped); returns the number
// of columns kept, packed at m0.
int Engine::ortho_cm(double* V, int64_t ld, int m0, int k, bool lanczos) {
    const int64_t n = n_;
    const int ni = (int)n;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    double* T = V + m0 * ld;
    S_.ensure((size_t)std::max(1, m0) * k * sizeof(double));
    gram_.ensure((size_t)k * k * sizeof(double));
    if (k > 32) otmp_.ensure((size_t)n * k * sizeof(double));
    work_.ensure((size_t)std::max<size_t>(work_.cap / sizeof(double), 2 * (size_t)k) * sizeof(double));
    info_.ensure(sizeof(int) * 4);
    const bool small = k <= 32;
    auto gram_host = [&](int kk, std::vector<double>& g) {
        if (small)
            tsm_tn(stream_, tsm_part_, n, kk, kk, T, ld, T, ld, gram_.as<double>(), kk);
        else
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, kk, kk, ni, &one, T, ld,
                                  T, ld, &zero, gram_.as<double>(), kk));
        g.resize((size_t)kk * kk);
        double* hg = static_cast<double*>(hr_.ensure(g.size() * sizeof(double)));
        HIPCHK(hipMemcpyAsync(hg, gram_.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        std::copy(hg, hg + g.size(), g.begin());
    };
    std::vector<double> g, before;
    auto cgs = [&]() {
        if (small) {
            tsm_tn(stream_, tsm_part_, n, m0, k, V, ld, T, ld, S_.as<double>(), m0);
            tsm_nn(stream_, n, m0, k, -1.0, V, ld, S_.as<double>(), m0, 1.0, T, ld);
        } else {
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, m0, k, ni, &one, V, ld,
                                  T, ld, &zero, S_.as<double>(), m0));
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, k, m0, &mone, V, ld,
                                  S_.as<double>(), m0, &one, T, ld));
        }
    };
    // squared column norms before projection (DGKS test)
    bool first_done = false;
    if (m0 > 0 && small && lanczos) {
        // A Lanczos block (its local components already removed): one pass
        // over [V, T] gives the projection coefficients V^T T (rows 0..m0)
        // and the Gram matrix of T (rows m0..m0+k).  The projection is
        // applied only when a coefficient exceeds kOrthoSkip times the
        // smallest column norm of T (on the 15-mode workload it is in about
        // 60 % of the steps; raising the threshold to 1e-10 stalls the
        // solver).
        const int p = m0 + k;
        S_.ensure(((size_t)p * k + 1) * sizeof(double));
        tsm_tn(stream_, tsm_part_, n, p, k, V, ld, T, ld, S_.as<double>(), p);
        // the gate flag follows R in S_, so one copy brings both
        k_gate_proj<<<1, 256, 0, stream_>>>(S_.as<double>(), m0, k, p, kOrthoSkip, S_.as<double>() + (size_t)p * k);
        const double* R = static_cast<double*>(hr_.ensure(((size_t)p * k + 1) * sizeof(double)));
        HIPCHK(hipMemcpyAsync(hr_.p, S_.p, ((size_t)p * k + 1) * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        const int flag = R[(size_t)p * k] != 0.0;
        g.assign((size_t)k * k, 0.0);
        for (int c = 0; c < k; c++)
            for (int i = 0; i < k; i++) g[(size_t)c * k + i] = R[(size_t)c * p + m0 + i];
        before.resize(k);
        for (int i = 0; i < k; i++) before[i] = g[(size_t)i * k + i];
        if (flag) {
            tsm_nn(stream_, n, m0, k, -1.0, V, ld, S_.as<double>(), p, 1.0, T, ld);
            // Gram matrix of the corrected block from the one in hand:
            // (T - V S)^T (T - V S) = T^T T - S^T S for orthonormal V.  A
            // column that loses more than half its squared norm gets a second
            // pass below, with its Gram matrix formed again.
            for (int c = 0; c < k; c++)
                for (int i = 0; i < k; i++) {
                    double d = 0.0;
                    for (int r = 0; r < m0; r++) d += R[(size_t)c * p + r] * R[(size_t)i * p + r];
                    g[(size_t)c * k + i] -= d;
                }
        }
        first_done = true;
    } else if (m0 > 0) {
        gram_host(k, g);
        before.resize(k);
        for (int i = 0; i < k; i++) before[i] = g[(size_t)i * k + i];
    }
    int passes = 0;
    for (int round = 0; round < 4; round++) {
        const bool given = first_done && round == 0;  // projection and Gram matrix from the fast path
        if (m0 > 0) {
            if (!given) cgs();
            passes++;
        }
        if (!given) gram_host(k, g);
        if (m0 > 0 && passes == 1 && round == 0) {
            // a column that lost more than half its squared norm to the
            // projection gets a second Gram-Schmidt pass (Daniel, Gragg,
            // Kaufman & Stewart, Math. Comp. 30 (1976))
            bool cancel = false;
            for (int i = 0; i < k; i++) cancel = cancel || g[(size_t)i * k + i] < 0.5 * before[i];
            if (cancel) {
                for (int i = 0; i < k; i++) before[i] = g[(size_t)i * k + i];
                continue;
            }
        }
        std::vector<double> dsc(k);
        for (int i = 0; i < k; i++) dsc[i] = g[(size_t)i * k + i] > 0 ? 1.0 / std::sqrt(g[(size_t)i * k + i]) : 0.0;
        for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++) g[(size_t)j * k + i] *= dsc[i] * dsc[j];
        std::vector<double> lam(k), U((size_t)k * k);
        if (small) {
            small_eigh(k, g, lam, U);
        } else {
            HIPCHK(hipMemcpyAsync(gram_.p, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
            BLASCHK(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_upper, k, gram_.as<double>(), k,
                                     work_.as<double>() + k, work_.as<double>(), info_.as<rocblas_int>()));
            HIPCHK(hipMemcpyAsync(lam.data(), work_.as<double>() + k, k * sizeof(double), hipMemcpyDeviceToHost,
                                  stream_));
            HIPCHK(hipMemcpyAsync(U.data(), gram_.p, U.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
        }
        const double lmax = lam[k - 1];
        std::vector<int> keep;
        for (int c = k - 1; c >= 0; c--)
            if (lam[c] > 1e-10 * std::max(1.0, lmax)) keep.push_back(c);
        const int kk = (int)keep.size();
        if (kk == 0) return 0;
        const double lmin = lam[keep.back()];
        std::vector<double> Z((size_t)k * kk);
        for (int c = 0; c < kk; c++) {
            const int src = keep[c];
            const double sc = 1.0 / std::sqrt(lam[src]);
            for (int i = 0; i < k; i++) Z[(size_t)c * k + i] = dsc[i] * U[(size_t)src * k + i] * sc;
        }
        S_.ensure((size_t)std::max(std::max(1, m0) * k, k * kk) * sizeof(double));
        // Z goes up from page-locked memory that stays valid until the copy
        // has run (it is rewritten only after a later synchronisation)
        double* hz = static_cast<double*>(hz_.ensure(Z.size() * sizeof(double)));
        std::copy(Z.begin(), Z.end(), hz);
        HIPCHK(hipMemcpyAsync(S_.p, hz, Z.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
        if (small) {
            const unsigned rg = std::min<unsigned>(grid_for(n, 256), 2048);
            if (k <= 8) k_rowmix<8><<<rg, 256, 0, stream_>>>(n, T, ld, k, kk, S_.as<double>(), nullptr);
            else if (k <= 16) k_rowmix<16><<<rg, 256, 0, stream_>>>(n, T, ld, k, kk, S_.as<double>(), nullptr);
            else k_rowmix<32><<<rg, 256, 0, stream_>>>(n, T, ld, k, kk, S_.as<double>(), nullptr);
            HIPCHK(hipGetLastError());
        } else {
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, kk, k, &one, T, ld,
                                  S_.as<double>(), k, &zero, otmp_.as<double>(), ni));
            HIPCHK(hipMemcpy2DAsync(T, ld * sizeof(double), otmp_.p, n * sizeof(double), n * sizeof(double), kk,
                                    hipMemcpyDeviceToDevice, stream_));
        }
        // no wait here: everything after this is ordered on the stream
        k = kk;
        // a well-conditioned block is done; otherwise repeat projection and
        // orthonormalisation once more on the new block
        if (lmin > 1e-4 * std::max(1.0, lmax) || round >= 2) return k;
        if (m0 > 0) {
            before.assign(k, 1.0);
            passes = 1;  // the next round's projection is a second pass
        }
    }
    return k;
}

// Block Davidson (Davidson-Liu): search space V and W = H V column-major
// n x m, projected matrix grown by new columns only, corrections = the
// diagonally preconditioned residuals of the unconverged wanted pairs; the
// b = nev + buffer Ritz block restarts the space at mmax.
void Engine::davidson(int nev, double tol, int maxit, std::vector<double>& E, EigStats& st) {
    const int64_t n = n_;
    const int ni = (int)n;
    const int b = (int)std::min<int64_t>(n, std::min(256, nev + std::max(4, nev / 5)));
    int mfac = 5;
    if (const char* e = std::getenv("VHCI_DAVIDSON_MFAC")) mfac = std::max(2, std::atoi(e));
    const int mmax = (int)std::min<int64_t>(n, (int64_t)mfac * b);
    const int64_t ld = n;
    V_.ensure((size_t)n * mmax * sizeof(double));
    W_.ensure((size_t)n * mmax * sizeof(double));
    HX_.ensure((size_t)n * b * sizeof(double));
    Xw_.ensure((size_t)n * b * sizeof(double));
    G_.ensure((size_t)mmax * mmax * sizeof(double));
    Gw_.ensure((size_t)mmax * mmax * sizeof(double));
    theta_.ensure((size_t)mmax * sizeof(double));
    work_.ensure((size_t)mmax * sizeof(double) * 2);
    info_.ensure(sizeof(int) * 4);
    double* V = V_.as<double>();
    double* W = W_.as<double>();
    double* HX = HX_.as<double>();
    double* X = Xw_.as<double>();
    double* G = G_.as<double>();
    double* Gw = Gw_.as<double>();
    double* th = theta_.as<double>();
    const double one = 1.0, zero = 0.0;
    std::vector<double> Dh(n);
    HIPCHK(hipMemcpyAsync(Dh.data(), D_.p, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    int k0 = 0;
    const bool resident = xcols_ > 0 && xrows_ > 0 && xrows_ <= n;
    if (resident) k0 = std::min(xcols_, b);
    {
        std::vector<int32_t> perm;
        if (perm_on_) {
            perm.resize(n);
            HIPCHK(hipMemcpy(perm.data(), perm_.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
        }
        // unit vectors go on the lowest diagonal entries of the states new
        // since the resident vectors were computed (of all states otherwise)
        std::vector<int64_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        const int64_t first = resident ? xrows_ : 0;
        std::vector<int64_t> cand;
        for (int64_t r = 0; r < n; r++)
            if ((perm_on_ ? perm[r] : r) >= first) cand.push_back(r);
        const int need = b - k0;
        if ((int64_t)cand.size() < need) cand = idx;
        auto cmp = [&](int64_t a, int64_t c) { return Dh[a] < Dh[c] || (Dh[a] == Dh[c] && a < c); };
        std::partial_sort(cand.begin(), cand.begin() + std::min<int64_t>(need, cand.size()), cand.end(), cmp);
        cand.resize(std::max(need, 1));
        DevBuf di;
        di.ensure(cand.size() * sizeof(int64_t));
        HIPCHK(hipMemcpyAsync(di.p, cand.data(), cand.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream_));
        dim3 g(grid_for(n, 256), b);
        k_start_block<<<g, 256, 0, stream_>>>(n, b, k0, resident ? X_.as<double>() : nullptr, xrows_, xcols_,
                                             di.as<int64_t>(), V, ld, perm_on_ ? perm_.as<int32_t>() : nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream_));
    }
    int m = ortho_cm(V, ld, 0, b);
    if (m == 0) throw std::runtime_error("Davidson: empty start block");
    bool low = form_ == kMixed;
    apply_cm(V, W, ld, m, low, st);
    BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, m, m, ni, &one, V, ld, W, ld,
                          &zero, G, mmax));
    const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
    const int rpb = 8192;
    const int nblk = (int)((n + rpb - 1) / rpb);
    part_.ensure((size_t)nblk * b * sizeof(double));
    std::vector<double> thh(mmax), rn(b);
    DevBuf Jd;
    Jd.ensure(b * sizeof(int));
    double prev_max = 1e300;
    int stall = 0;
    for (int it = 0; it < maxit; it++) {
        st.iterations = it + 1;
        HIPCHK(hipMemcpy2DAsync(Gw, m * sizeof(double), G, mmax * sizeof(double), m * sizeof(double), m,
                                hipMemcpyDeviceToDevice, stream_));
        BLASCHK(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_upper, m, Gw, m, th, work_.as<double>(),
                                 info_.as<rocblas_int>()));
        const int nb = std::min(b, m);
        BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, nb, m, &one, V, ld, Gw, m,
                              &zero, X, ld));
        BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, nb, m, &one, W, ld, Gw, m,
                              &zero, HX, ld));
        dim3 rg(nblk, nb);
        k_resid_cm<<<rg, 256, 0, stream_>>>(n, HX, X, ld, th, rpb, part_.as<double>());
        k_sum_parts<<<grid_for(nb, 64), 64, 0, stream_>>>(part_.as<double>(), nblk, nb, work_.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(thh.data(), th, m * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(rn.data(), work_.p, nb * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        std::vector<int> J;
        double maxrel = 0;
        bool all = true;
        for (int j = 0; j < std::min(nev, nb); j++) {
            const double r = std::sqrt(rn[j]);
            maxrel = std::max(maxrel, r / std::max(eps23, std::abs(thh[j])));
            if (r >= tol * std::max(eps23, std::abs(thh[j]))) {
                J.push_back(j);
                all = false;
            }
        }
        st.max_resid = maxrel;
        if (std::getenv("VHCI_DEBUG"))
            std::fprintf(stderr, "[davidson] it %d m %d maxrel %.3e unconverged %zu %s\n", it, m, maxrel, J.size(),
                         low ? "fp32" : "fp64");
        if (all && !low) {
            E.assign(thh.begin(), thh.begin() + nev);
            X_.ensure((size_t)n * nev * sizeof(double));
            dim3 tg(grid_for(n, 32), (nev + 31) / 32);
            k_cm_to_rm<double><<<tg, 256, 0, stream_>>>(X, ld, n, nev, X_.as<double>(), nev, perm_on_ ? perm_.as<int32_t>() : nullptr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream_));
            xrows_ = n;
            xcols_ = nev;
            return;
        }
        if (low) {
            const bool floor_hit = maxrel < 2e-7 || all || (maxrel > 0.5 * prev_max && ++stall >= 2);
            prev_max = std::min(prev_max, maxrel);
            if (floor_hit) {
                low = false;
                HIPCHK(hipMemcpyAsync(V, X, (size_t)n * nb * sizeof(double), hipMemcpyDeviceToDevice, stream_));
                apply_cm(V, W, ld, nb, false, st);
                BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, nb, nb, ni, &one, V,
                                      ld, W, ld, &zero, G, mmax));
                m = nb;
                continue;
            }
        }
        const int k = (int)J.size();
        if (m + k > mmax) {
            HIPCHK(hipMemcpyAsync(V, X, (size_t)n * nb * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            HIPCHK(hipMemcpyAsync(W, HX, (size_t)n * nb * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            dim3 dg(grid_for(nb, 64), nb);
            k_diag_block<<<dg, 64, 0, stream_>>>(G, nb, mmax, th);
            m = nb;
        }
        HIPCHK(hipMemcpyAsync(Jd.p, J.data(), k * sizeof(int), hipMemcpyHostToDevice, stream_));
        dim3 pg(std::min<unsigned>(grid_for(n, 256), 1024), k);
        k_precond_cm<<<pg, 256, 0, stream_>>>(n, k, Jd.as<int>(), HX, X, ld, th, D_.as<double>(), V + m * ld, ld);
        HIPCHK(hipGetLastError());
        const int kept = ortho_cm(V, ld, m, k);
        if (kept == 0) throw std::runtime_error("Error: Eigenvalues did not converge.");
        apply_cm(V + m * ld, W + m * ld, ld, kept, low, st);
        BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, m + kept, kept, ni, &one, V,
                              ld, W + m * ld, ld, &zero, G + (int64_t)m * mmax, mmax));
        m += kept;
    }
    throw std::runtime_error("Error: Eigenvalues did not converge.");
}

// Block Krylov-Schur (thick-restart block Lanczos with full
// re-orthogonalisation).  V holds m + bs column-major basis vectors and W =
// H V for the columns formed in the current cycle (FP32 cycles also carry
// W = H X for the kept Ritz vectors); after each cycle the Rayleigh-Ritz
// pairs of G = V^T W are tested (||H x - theta x|| < tol max(eps^(2/3),
// |theta|) for the nev lowest), and the space restarts from the keep lowest
// Ritz vectors plus the last residual block, which is orthogonal to all of
// them.  The start block mixes the resident Ritz vectors of the previous
// basis (zero-padded) with a small random component.  Mixed form: the cycles
// run with the FP32 product until the residuals stop improving; then the
// FP64 cycles start afresh from random combinations of the kept Ritz
// vectors.
void Engine::krylov_schur(int nev, double tol, int maxit, std::vector<double>& E, EigStats& st) {
    const int64_t n = n_;
    const int ni = (int)n;
    int bs = 8;
    if (const char* e = std::getenv("VHCI_KS_BLOCK")) bs = std::max(1, std::min(32, std::atoi(e)));
    // basis of m = 4 nev vectors (a multiple of bs, at least nev + 6
    // blocks), restarted from the keep = 1.6 nev lowest Ritz vectors: each
    // cycle adds about 2.4 nev new vectors per Rayleigh-Ritz step (on the
    // 15-mode, 100-state workload 7 % faster than m = 3.5 nev, and the
    // 8-36-mode request mix slightly faster too)
    const auto up = [&](double x) { return ((int)std::ceil(x) + bs - 1) / bs * bs; };
    double mfac = ks_mfac_;  // below 4 when half storage needed a smaller basis to fit
    if (const char* e = std::getenv("VHCI_KS_MFAC")) mfac = std::max(2.0, std::atof(e));
    int m = std::max(up(mfac * nev), up((double)nev + 6 * bs));
    if (const char* e = std::getenv("VHCI_KS_DIM")) m = std::max(nev + 2 * bs, (std::atoi(e) / bs) * bs);
    while (m + bs > n) m -= bs;
    if (m < nev + 2 * bs) throw std::runtime_error("Krylov-Schur: basis too small for the requested states");
    int keep = std::min(m - bs, std::max(up(1.6 * nev), up((double)nev + bs)));
    if (const char* e = std::getenv("VHCI_KS_KEEP")) keep = std::max(nev, std::min(m - bs, (std::atoi(e) / bs) * bs));
    const int64_t ld = n;
    V_.ensure((size_t)n * (m + bs) * sizeof(double));
    W_.ensure((size_t)n * m * sizeof(double));
    Xw_.ensure((size_t)n * keep * sizeof(double));
    HX_.ensure((size_t)n * keep * sizeof(double));
    G_.ensure((size_t)m * m * sizeof(double));
    Gw_.ensure((size_t)m * m * sizeof(double));
    theta_.ensure((size_t)m * sizeof(double));
    work_.ensure((size_t)std::max(m, 2 * bs) * 2 * sizeof(double));
    info_.ensure(sizeof(int) * 4);
    double* V = V_.as<double>();
    double* W = W_.as<double>();
    double* X = Xw_.as<double>();
    double* HX = HX_.as<double>();
    double* G = G_.as<double>();
    double* Gw = Gw_.as<double>();
    double* th = theta_.as<double>();
    const double one = 1.0, zero = 0.0;
    // start block
    uint64_t seed = 0x6B72796C6F76ull + (uint64_t)n;
    random_cm(V, ld, bs, seed);
    if (xcols_ > 0 && xrows_ > 0 && xrows_ <= n) {
        // V[:xrows, :bs] = 1e3 * Xprev R (R: xcols x bs, fixed pseudo-random) + the random block
        std::vector<double> R((size_t)xcols_ * bs);
        uint64_t z = seed;
        for (auto& r : R) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t y = z;
            y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
            y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
            y ^= y >> 31;
            r = (double)(y >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
        DevBuf Rd;
        Rd.ensure(R.size() * sizeof(double));
        HIPCHK(hipMemcpyAsync(Rd.p, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
        const double big = 1e3;
        if (perm_on_) {
            // previous vectors are in basis order: combine, then scatter into solver rows
            DevBuf T;
            T.ensure((size_t)xrows_ * bs * sizeof(double));
            const double zero0 = 0.0;
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, (int)xrows_, bs, xcols_,
                                  &big, X_.as<double>(), xcols_, Rd.as<double>(), xcols_, &zero0, T.as<double>(),
                                  (int)xrows_));
            dim3 g(std::min<unsigned>(grid_for(n, 256), 4096), bs);
            k_add_perm_rows<<<g, 256, 0, stream_>>>(V, ld, n, bs, T.as<double>(), xrows_, perm_.as<int32_t>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream_));
        } else {
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, (int)xrows_, bs, xcols_,
                                  &big, X_.as<double>(), xcols_, Rd.as<double>(), xcols_, &one, V, ld));
            HIPCHK(hipStreamSynchronize(stream_));
        }
    }
    int got = ortho_cm(V, ld, 0, bs);
    if (got < bs) throw std::runtime_error("Krylov-Schur: degenerate start block");
    bool low = form_ == kMixed;
    // Lanczos steps orthonormalised on the device (blocks of at most 8
    // columns; VHCI_HOST_ORTHO=1: every step on the host)
    const char* ho = std::getenv("VHCI_HOST_ORTHO");
    const bool dev_ortho = bs <= 8 && !(ho && std::string(ho) == "1");
    // VHCI_HOST_ORTHO=repeat: the device declines every step, so each one
    // is repeated on the host (a test of the repeat path)
    const bool decline_dev = ho && std::string(ho) == "repeat";
    // VHCI_GRAPH=1: a cycle's device steps are captured once in a HIP graph
    // and replayed in the cycles of the same shape (not while the bench
    // times the products, whose events the graph would replay)
    const char* ge = std::getenv("VHCI_GRAPH");
    const bool graph_ok = dev_ortho && ge && std::string(ge) == "1" && prof_.b == 0 && m + bs <= kOrthRows &&
                          !sym_ && !matfree_;  // full CSR products only
    struct CycleGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<const void*> key;
        int64_t vectors = 0, fp64 = 0;
        int calls = 0;
        void reset() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
        }
        ~CycleGraph() { reset(); }
    } cyc;
    if (dev_ortho) {
        // sized once, so that no buffer the queued steps use is replaced
        ostat_.ensure((size_t)(m / bs + 2) * sizeof(int));
        ozd_.ensure(64 * sizeof(double));
        S_.ensure(((size_t)(m + 2 * bs) * bs + 1) * sizeof(double));
        // the Gram's partial products for every width a step projects on
        size_t part_max = 0;
        for (int p = 1; p <= m + 2 * bs; p++) part_max = std::max(part_max, tsm_part_bytes(n, p, bs));
        tsm_part_.ensure(part_max);
    }
    const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
    const int rpb = 8192;
    const int nblk = (int)((n + rpb - 1) / rpb);
    part_.ensure((size_t)nblk * keep * sizeof(double));
    std::vector<double> thh(m), rn(nev);
    int k = 0;
    double prev_max = 1e300;
    int stall = 0;
    for (int it = 0; it < maxit; it++) {
        st.iterations = it + 1;
        bool refilled = false;
        // Steps after a cycle's first block (Lanczos steps) finish their
        // orthonormalisation on the device (k_orth_block), so they are
        // enqueued without waiting for the host; each leaves a status, read
        // once the cycle's steps are queued.  The first step whose status
        // asks for the general path is repeated on the host (ortho_cm), and
        // the steps after it are enqueued again, checked after 1, 2, 4, ...
        // steps so that repeated refusals cost no more than the host path.
        const int* dstat = dev_ortho ? ostat_.as<int>() : nullptr;
        // Lanczos step: H V_j has large components only along V_j and
        // V_(j-bs) (the block-tridiagonal part of the projected matrix; the
        // kept Ritz vectors couple to V_k alone).  Taking those out first
        // leaves the full pass a block that is already nearly orthogonal, so
        // one projection suffices.
        auto local_projection = [&](int j) {
            double* T = V + (int64_t)(j + bs) * ld;
            const int lo = std::max(k, j - bs);
            const int p = j + bs - lo;
            S_.ensure((size_t)std::max(p * bs, (j + bs) * bs) * sizeof(double));
            tsm_tn(stream_, tsm_part_, n, p, bs, V + (int64_t)lo * ld, ld, T, ld, S_.as<double>(), p);
            tsm_nn(stream_, n, p, bs, -1.0, V + (int64_t)lo * ld, ld, S_.as<double>(), p, 1.0, T, ld);
        };
        // one device step (status in slot): the product, the local
        // projection, then one pass over [V, T] for V^T T and T^T T, the
        // projection only when the gate is set (as in ortho_cm's Lanczos
        // path), and the block's Cholesky QR
        auto enqueue_step = [&](int j, int slot) {
            apply_cm(V + (int64_t)j * ld, W + (int64_t)j * ld, ld, bs, low, st, V + (int64_t)(j + bs) * ld);
            local_projection(j);
            double* T = V + (int64_t)(j + bs) * ld;
            const int m0 = j + bs, p = m0 + bs;
            double* R = S_.as<double>();
            tsm_tn(stream_, tsm_part_, n, p, bs, V, ld, T, ld, R, p);
            k_gate_proj<<<1, 256, 0, stream_>>>(R, m0, bs, p, kOrthoSkip, R + (size_t)p * bs);
            tsm_nn(stream_, n, m0, bs, -1.0, V, ld, R, p, 1.0, T, ld, R + (size_t)p * bs);
            k_orth_block<<<1, 256, 0, stream_>>>(R, m0, bs, p, ozd_.as<double>(), ostat_.as<int>() + slot,
                                                  decline_dev ? 1 : 0);
            const unsigned rg = std::min<unsigned>(grid_for(n, 256), 2048);
            k_rowmix<8><<<rg, 256, 0, stream_>>>(n, T, ld, bs, bs, ozd_.as<double>(), ostat_.as<int>() + slot);
            HIPCHK(hipGetLastError());
        };
        // the device steps j0, j0 + bs, ... (nsteps) from a captured graph,
        // captured again when the cycle's shape or a buffer it uses changed
        auto launch_graph = [&](int j0, int nsteps) {
            const std::vector<const void*> key = {V, W, S_.p, tsm_part_.p, ozd_.p, ostat_.p, stage_in_.p,
                                                  stage_out_.p, stage_in32_.p, X32_.p,
                                                  reinterpret_cast<const void*>((intptr_t)j0),
                                                  reinterpret_cast<const void*>((intptr_t)nsteps),
                                                  reinterpret_cast<const void*>((intptr_t)(low ? 1 : 0))};
            if (!cyc.exec || key != cyc.key) {
                cyc.reset();
                const int64_t spmv0 = st.spmm_vectors, calls0 = st.spmm_calls, fp640 = st.fp64_spmm_vectors;
                HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
                for (int q = 0; q < nsteps; q++) enqueue_step(j0 + q * bs, q);
                HIPCHK(hipStreamEndCapture(stream_, &cyc.graph));
                HIPCHK(hipGraphInstantiate(&cyc.exec, cyc.graph, nullptr, nullptr, 0));
                cyc.key = key;
                cyc.vectors = st.spmm_vectors - spmv0;
                cyc.calls = (int)(st.spmm_calls - calls0);
                cyc.fp64 = st.fp64_spmm_vectors - fp640;
                st.spmm_vectors = spmv0;
                st.spmm_calls = (int)calls0;
                st.fp64_spmm_vectors = (int)fp640;
                st.graphs_captured++;
            }
            HIPCHK(hipGraphLaunch(cyc.exec, stream_));
            st.spmm_vectors += cyc.vectors;
            st.spmm_calls += cyc.calls;
            st.fp64_spmm_vectors += (int)cyc.fp64;
            st.graph_launches++;
        };
        auto lanczos_steps = [&](int j0, bool host_first) -> int {
            int queued = 0;  // device steps enqueued since the last status check
            int first_dev = -1;
            int batch = host_first ? 1 : m;  // steps between checks
            auto check = [&]() -> int {
                if (queued == 0) return -1;
                std::vector<int> h(queued);
                HIPCHK(hipMemcpyAsync(h.data(), dstat, queued * sizeof(int), hipMemcpyDeviceToHost, stream_));
                HIPCHK(hipStreamSynchronize(stream_));
                for (int q = 0; q < queued; q++)
                    if (h[q] != 0) {
                        st.device_ortho_redone++;
                        return first_dev + q * bs;
                    }
                st.device_ortho_steps += queued;
                queued = 0;
                return -1;
            };
            for (int j = j0; j < m; j += bs) {
                const bool dev = dev_ortho && j > k && j + bs <= kOrthRows && !(host_first && j == j0);
                if (dev && graph_ok && j == k + bs && j0 == k) {
                    // every device step of the cycle from one captured graph
                    const int nsteps = (m - j) / bs;
                    launch_graph(j, nsteps);
                    first_dev = j;
                    queued = nsteps;
                    break;
                }
                if (dev) {
                    if (queued == 0) first_dev = j;
                    enqueue_step(j, queued);
                    queued++;
                    if (queued >= batch) {
                        const int bad = check();
                        if (bad >= 0) return bad;
                        batch *= 2;
                    }
                    continue;
                }
                const int bad = check();
                if (bad >= 0) return bad;
                // W_j = H V_j, also written to V_(j+bs) as the next block's start
                apply_cm(V + (int64_t)j * ld, W + (int64_t)j * ld, ld, bs, low, st, V + (int64_t)(j + bs) * ld);
                if (j > k) local_projection(j);
                int kept = ortho_cm(V, ld, j + bs, bs, j > k);
                if (kept < bs) refilled = true;
                for (int tries = 0; kept < bs && tries < 4; tries++) {
                    // invariant subspace reached for some directions: refill
                    random_cm(V + (int64_t)(j + bs + kept) * ld, ld, bs - kept, seed + 977 * (it * m + j + tries + 1));
                    int more = ortho_cm(V, ld, j + bs + kept, bs - kept);
                    kept += more;
                }
                if (kept < bs) throw std::runtime_error("Krylov-Schur: could not extend the basis");
            }
            return check();
        };
        for (int jf = lanczos_steps(k, false); jf >= 0;) jf = lanczos_steps(jf, true);
        // G[:, k:m] = V[:, :m]^T W[:, k:m] (upper triangle of the new columns).
        // One rocBLAS product over all n rows runs at 2.3 TFLOP/s for this
        // shape; split into row chunks of one batched product (MFMA tiles)
        // with the partial products summed in a fixed order, it is a little
        // faster than the vector-ALU tall-skinny kernel.  Chunks of 8192
        // rows, longer for bases beyond 2^21 states so that at most 256
        // partial products are held.
        // GEMM timing (bench profiling): events around the projection and
        // the rotation, read after the rotation's
        const bool gprof = prof_.b > 0;
        if (gprof) {
            if (prof_.gev.empty()) {
                prof_.gev.resize(4);
                for (auto& e : prof_.gev) HIPCHK(hipEventCreate(&e));
            }
            HIPCHK(hipEventRecord(prof_.gev[2], stream_));
        }
        {
            const int64_t chunk = std::max<int64_t>(8192, (n + 255) / 256);
            const int nfull = (int)(n / chunk);
            const int64_t tail = n - (int64_t)nfull * chunk;
            const int q = m - k;
            const int nb = nfull + (tail > 0 ? 1 : 0);
            rr_part_.ensure((size_t)std::max(1, nb) * m * q * sizeof(double));
            double* pp = rr_part_.as<double>();
            const double* Wk = W + (int64_t)k * ld;
            if (nfull > 0)
                BLASCHK(rocblas_dgemm_strided_batched(blas_, rocblas_operation_transpose, rocblas_operation_none, m,
                                                      q, (int)chunk, &one, V, ld, chunk, Wk, ld, chunk, &zero, pp, m,
                                                      (int64_t)m * q, nfull));
            if (tail > 0)
                BLASCHK(rocblas_dgemm(blas_, rocblas_operation_transpose, rocblas_operation_none, m, q, (int)tail,
                                      &one, V + (int64_t)nfull * chunk, ld, Wk + (int64_t)nfull * chunk, ld, &zero,
                                      pp + (int64_t)nfull * m * q, m));
            k_sum_batches<<<grid_for((int64_t)m * q, 256), 256, 0, stream_>>>(pp, nb, m, q, G + (int64_t)k * m, m);
            HIPCHK(hipGetLastError());
        }
        if (gprof) HIPCHK(hipEventRecord(prof_.gev[3], stream_));
        HIPCHK(hipMemcpyAsync(Gw, G, (size_t)m * m * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        BLASCHK(rocsolver_dsyevd(blas_, rocblas_evect_original, rocblas_fill_upper, m, Gw, m, th, work_.as<double>(),
                                 info_.as<rocblas_int>()));
        if (gprof) HIPCHK(hipEventRecord(prof_.gev[0], stream_));
        // FP64 cycles: W = H V holds column for column, and every block but
        // the last maps into the basis, so H V = V G + V_m B E^T with
        // B = V_m^T W_last (V_m: the block after the last, columns m..m+bs).
        // The Ritz residual H x_i - theta_i x_i is then V_m B s_i(last bs
        // rows): its norm comes from bs x bs and bs x nev numbers, and H X
        // is not formed.  The FP32 cycles' products carry FP32 rounding in
        // every column, so they form H X and measure the residuals directly;
        // so does a cycle in which a block lost directions and was refilled
        // at random (the relation then misses what was dropped), with H X
        // from the product since W's kept columns are not carried in FP64.
        const bool relation = !low && !refilled;
        // the rotation X = V S (and H X = W S for FP32 cycles) on the
        // matrix cores; VHCI_ROT=blas takes rocBLAS dgemm instead
        const char* rot_env = std::getenv("VHCI_ROT");
        const bool rot_blas = rot_env && std::string(rot_env) == "blas";
        if (rot_blas) {
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, keep, m, &one, V, ld, Gw,
                                  m, &zero, X, ld));
            if (low)
                BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, keep, m, &one, W, ld,
                                      Gw, m, &zero, HX, ld));
        } else {
            rot_mfma(stream_, n, m, keep, V, ld, Gw, m, X, ld, rot_s_);
            if (low) rot_mfma(stream_, n, m, keep, W, ld, Gw, m, HX, ld, rot_s_);
        }
        if (gprof) {
            HIPCHK(hipEventRecord(prof_.gev[1], stream_));
            HIPCHK(hipEventSynchronize(prof_.gev[1]));
            float t = 0, tp = 0;
            HIPCHK(hipEventElapsedTime(&t, prof_.gev[0], prof_.gev[1]));
            HIPCHK(hipEventElapsedTime(&tp, prof_.gev[2], prof_.gev[3]));
            prof_.gemm_ms += t + tp;
            prof_.gemm_flops += (low ? 2.0 : 1.0) * 2.0 * (double)n * m * keep + 2.0 * (double)n * m * (m - k);
            prof_.gemm_calls += (low ? 2 : 1) + 1;
        }
        if (!low && !relation) apply_cm(X, HX, ld, nev, false, st);
        if (relation) {
            S_.ensure((size_t)std::max<size_t>(S_.cap / sizeof(double), (size_t)bs * bs) * sizeof(double));
            tsm_tn(stream_, tsm_part_, n, bs, bs, V + (int64_t)m * ld, ld, W + (int64_t)(m - bs) * ld, ld,
                   S_.as<double>(), bs);
            k_relation_resid<<<grid_for(nev, 64), 64, 0, stream_>>>(S_.as<double>(), bs, Gw, m, nev, work_.as<double>());
        } else {
            dim3 rg(nblk, nev);
            k_resid_cm<<<rg, 256, 0, stream_>>>(n, HX, X, ld, th, rpb, part_.as<double>());
            k_sum_parts<<<grid_for(nev, 64), 64, 0, stream_>>>(part_.as<double>(), nblk, nev, work_.as<double>());
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(rn.data(), work_.p, nev * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(thh.data(), th, m * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        double maxrel = 0;
        int unconv = 0;
        for (int j = 0; j < nev; j++) {
            const double r = std::sqrt(rn[j]);
            maxrel = std::max(maxrel, r / std::max(eps23, std::abs(thh[j])));
            if (r >= tol * std::max(eps23, std::abs(thh[j]))) unconv++;
        }
        st.max_resid = maxrel;
        if (std::getenv("VHCI_DEBUG"))
            std::fprintf(stderr,
                         "[krylov-schur] cycle %d m %d keep %d bs %d maxrel %.3e unconverged %d %s (device steps %d, "
                         "repeated %d, graphs %d captured %d launched)\n",
                         it, m, keep, bs, maxrel, unconv, low ? "fp32" : "fp64", st.device_ortho_steps,
                         st.device_ortho_redone, st.graphs_captured, st.graph_launches);
        if (unconv == 0 && !low) {
            E.assign(thh.begin(), thh.begin() + nev);
            X_.ensure((size_t)n * nev * sizeof(double));
            dim3 tg(grid_for(n, 32), (nev + 31) / 32);
            k_cm_to_rm<double><<<tg, 256, 0, stream_>>>(X, ld, n, nev, X_.as<double>(), nev, perm_on_ ? perm_.as<int32_t>() : nullptr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream_));
            xrows_ = n;
            xcols_ = nev;
            return;
        }
        // thick restart
        HIPCHK(hipMemcpyAsync(V, X, (size_t)n * keep * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(V + (int64_t)keep * ld, V + (int64_t)m * ld, (size_t)n * bs * sizeof(double),
                              hipMemcpyDeviceToDevice, stream_));
        bool switched = false;
        if (low) {
            const bool floor_hit = maxrel < 1e-7 || unconv == 0 || (maxrel > 0.5 * prev_max && ++stall >= 2);
            prev_max = std::min(prev_max, maxrel);
            if (floor_hit) {
                low = false;
                switched = true;
            }
        }
        if (switched) {
            // The FP32 cycles' Krylov relation does not hold for the FP64
            // product, so the FP64 cycles start afresh from a block of random
            // combinations of the kept Ritz vectors.
            std::vector<double> R((size_t)keep * bs);
            uint64_t z = seed ^ 0x5A5A5A5Aull;
            for (auto& r : R) {
                z += 0x9E3779B97F4A7C15ull;
                uint64_t y = z;
                y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
                y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
                y ^= y >> 31;
                r = (double)(y >> 11) * (1.0 / 9007199254740992.0) - 0.5;
            }
            DevBuf Rd;
            Rd.ensure(R.size() * sizeof(double));
            HIPCHK(hipMemcpyAsync(Rd.p, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
            BLASCHK(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, ni, bs, keep, &one, X, ld,
                                  Rd.as<double>(), keep, &zero, V, ld));
            HIPCHK(hipStreamSynchronize(stream_));
            if (ortho_cm(V, ld, 0, bs) < bs) throw std::runtime_error("Krylov-Schur: degenerate restart block");
            k = 0;
            continue;
        } else {
            // W[:, :keep] = H X is read only by the next cycle's H X product,
            // which FP64 cycles do not form
            if (low)
                HIPCHK(hipMemcpyAsync(W, HX, (size_t)n * keep * sizeof(double), hipMemcpyDeviceToDevice, stream_));
            dim3 dg(grid_for(keep, 64), keep);
            k_diag_block<<<dg, 64, 0, stream_>>>(G, keep, m, th);
        }
        HIPCHK(hipGetLastError());
        k = keep;
    }
    throw std::runtime_error("Error: Eigenvalues did not converge.");
}

EigStats Engine::eigsolve(int nev, double tol, int maxit, std::vector<double>& E) {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (!h_valid_) throw std::runtime_error("Hamiltonian not built");
    auto t0 = clk::now();
    EigStats st;
    const int64_t n = n_;
    if (nev < 1) throw std::runtime_error("nev must be positive");
    if (nev > n) nev = (int)n;
    if (nev > 256) throw std::runtime_error("at most 256 eigenpairs are supported");
    if (n <= std::max<int64_t>(4 * (nev + std::max(4, nev / 5)), 1024)) {
        dense_solve(nev, E, st);
    } else {
        const char* alg = std::getenv("VHCI_EIGSOLVER");
        if (alg && std::string(alg) == "davidson")
            davidson(nev, tol, maxit, E, st);
        else
            krylov_schur(nev, tol, maxit, E, st);
    }
    st.ms = ms_since(t0);
    return st;
}

void Engine::get_vectors(double* C, int ncol) const {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    if (ncol > xcols_) throw std::runtime_error("requested more eigenvector columns than resident");
    std::vector<double> xo((size_t)xrows_ * xcols_);
    HIPCHK(hipMemcpyAsync(xo.data(), X_.p, xo.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int64_t i = 0; i < xrows_; i++)
        for (int c = 0; c < ncol; c++) C[i * ncol + c] = xo[i * xcols_ + c];
}

void Engine::freq_assignments(int ncol, std::vector<FreqColumn>& out) const {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (ncol < 1 || ncol > xcols_ || xrows_ < 1) throw std::runtime_error("no resident eigenvectors to assign");
    const int64_t n = xrows_;
    const int64_t rpc = std::max<int64_t>(64, (n + 1023) / 1024);
    const int nchunk = (int)((n + rpc - 1) / rpc);
    DevBuf part, half, cnt, rows, vals;
    part.ensure((size_t)nchunk * ncol * sizeof(double));
    half.ensure((size_t)ncol * sizeof(double));
    cnt.ensure((size_t)ncol * sizeof(uint32_t));
    const double* X = X_.as<double>();
    k_col_absmax<<<nchunk, 256, 0, stream_>>>(n, X, xcols_, ncol, rpc, part.as<double>());
    k_col_half<<<grid_for(ncol, 64), 64, 0, stream_>>>(part.as<double>(), nchunk, ncol, half.as<double>());
    HIPCHK(hipGetLastError());
    std::vector<uint32_t> c(ncol);
    uint32_t cap = 64;
    for (int pass = 0; pass < 2; pass++) {
        rows.ensure((size_t)ncol * cap * sizeof(int64_t));
        vals.ensure((size_t)ncol * cap * sizeof(double));
        k_fill<uint32_t><<<grid_for(ncol, 256), 256, 0, stream_>>>(cnt.as<uint32_t>(), ncol, 0u);
        k_col_select<<<nchunk, 256, 0, stream_>>>(n, X, xcols_, ncol, rpc, half.as<double>(), cnt.as<uint32_t>(), cap,
                                                  rows.as<int64_t>(), vals.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c.data(), cnt.p, ncol * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        const uint32_t most = *std::max_element(c.begin(), c.end());
        if (most <= cap) break;
        cap = most;  // a wider second pass holds every selected row
    }
    std::vector<int64_t> hr((size_t)ncol * cap);
    std::vector<double> hv((size_t)ncol * cap);
    HIPCHK(hipMemcpyAsync(hr.data(), rows.p, hr.size() * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(hv.data(), vals.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    out.assign(ncol, FreqColumn{});
    std::vector<uint32_t> want;
    for (int k = 0; k < ncol; k++) {
        std::vector<std::pair<int64_t, double>> sel;
        for (uint32_t t = 0; t < c[k]; t++) sel.emplace_back(hr[(size_t)k * cap + t], hv[(size_t)k * cap + t]);
        std::sort(sel.begin(), sel.end());  // ascending basis row, as the reference's scan
        for (auto& e : sel) {
            out[k].row.push_back(e.first);
            out[k].absval.push_back(e.second);
            want.push_back((uint32_t)e.first);
        }
    }
    // quanta of the selected rows
    std::vector<uint8_t> keys((size_t)std::max<size_t>(want.size(), 1) * KW_ * 16);
    if (!want.empty()) {
        DevBuf idx, kout;
        idx.ensure(want.size() * sizeof(uint32_t));
        kout.ensure(keys.size());
        HIPCHK(hipMemcpyAsync(idx.p, want.data(), want.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
        k_gather_keys<<<grid_for((int64_t)want.size() * KW_, 256), 256, 0, stream_>>>(
            keys_.as<uint4>(), idx.as<uint32_t>(), (int64_t)want.size(), KW_, kout.as<uint4>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(keys.data(), kout.p, keys.size(), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    size_t w = 0;
    const int M = P_.M, K = KW_ * 16;
    for (int k = 0; k < ncol; k++)
        for (size_t t = 0; t < out[k].row.size(); t++, w++) {
            std::vector<int32_t> q(M);
            for (int m = 0; m < M; m++) q[m] = keys[w * K + m];
            out[k].quanta.push_back(std::move(q));
        }
}

void Engine::set_vectors(const double* C, int ncol) {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    X_.ensure((size_t)n_ * ncol * sizeof(double));
    HIPCHK(hipMemcpyAsync(X_.p, C, (size_t)n_ * ncol * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    xrows_ = n_;
    xcols_ = ncol;
}

// ---------------------------------------------------------------- heat bath
// Candidate counts of every basis state for threshold eps (from cmax_):
// candcnt_/candoff_ hold the counts and their exclusive prefix; returns the
// total (HB.cpp:24-26, the loop over the |f|-sorted constants).
int64_t Engine::hb_count(double eps) {
    const int64_t n = n_;
    const int nhb = (int)T_.hb_fc.size();
    candcnt_.ensure((n + 1) * sizeof(int64_t));
    candoff_.ensure((n + 1) * sizeof(int64_t));
    k_hb_count<<<grid_for(n, 256), 256, 0, stream_>>>(n, cmax_.as<double>(), d_hbfc_.as<double>(), nhb,
                                                      d_hbvaroff_.as<int64_t>(), eps, candcnt_.as<int64_t>());
    k_fill<int64_t><<<1, 1, 0, stream_>>>(candcnt_.as<int64_t>() + n, 1, 0);
    size_t tmp = 0;
    HIPCHK(rocprim::exclusive_scan(nullptr, tmp, candcnt_.as<int64_t>(), candoff_.as<int64_t>(), (int64_t)0,
                                   (size_t)(n + 1), rocprim::plus<int64_t>(), stream_));
    scan_tmp_.ensure(tmp);
    HIPCHK(rocprim::exclusive_scan(scan_tmp_.p, tmp, candcnt_.as<int64_t>(), candoff_.as<int64_t>(), (int64_t)0,
                                   (size_t)(n + 1), rocprim::plus<int64_t>(), stream_));
    int64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, candoff_.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return total;
}

// The distinct new states among the counted candidates whose key hash tag
// (the hash's high 32 bits) lies in [tlo, thi), in first-discovery order:
// newkeys_, nnew_, and the hash (nslots_, pool_keys_, pool_pos_) that finds
// them.  The dedup pool starts at an estimate and grows fourfold while it
// overflows; false (nothing kept) when it would need more than max_pool
// entries.
template <int NW>
bool Engine::hb_gen_t(int64_t total, uint64_t tlo, uint64_t thi, uint64_t max_pool, HbStats& st) {
    const int64_t n = n_;
    nnew_ = 0;
    const double frac = (double)(thi - tlo) / 4294967296.0;
    uint64_t ucap = (uint64_t)std::min<double>((double)total,
                                               std::max(1048576.0, frac * (double)std::min<int64_t>(total, 32 * n)));
    counters_.ensure(4 * sizeof(uint32_t));
    for (int attempt = 0; attempt < 12; attempt++) {
        uint64_t cap = next_pow2(std::max<uint64_t>(2048, 2 * ucap));
        if (cap > (1ull << 32)) return false;
        // the pool holds at most max_pool states; an overflow at that size
        // ends the attempt (the caller splits the range)
        const bool capped = std::min<uint64_t>(cap / 2, 0xffffffffull - 1) >= max_pool;
        const uint32_t pool_cap = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(cap / 2, 0xffffffffull - 1), max_pool);
        nslots_.ensure(cap * sizeof(uint64_t));
        pool_keys_.ensure((size_t)pool_cap * KW_ * 16);
        pool_rank_.ensure((size_t)pool_cap * sizeof(uint64_t));
        k_fill<uint64_t><<<grid_for(cap, 256), 256, 0, stream_>>>(nslots_.as<uint64_t>(), (int64_t)cap, kEmpty);
        k_fill<uint32_t><<<1, 4, 0, stream_>>>(counters_.as<uint32_t>(), 4, 0u);
        // ranks start above every candidate, so a rank read while another
        // thread inserts is never a value left by an earlier selection
        k_fill<uint64_t><<<grid_for(pool_cap, 256), 256, 0, stream_>>>(pool_rank_.as<uint64_t>(), (int64_t)pool_cap,
                                                                       ~0ull);
        HbGenArgs a{};
        a.n = n;
        a.total = total;
        a.candoff = candoff_.as<int64_t>();
        a.keys = keys_.as<uint4>();
        a.bslots = slots_.as<uint64_t>();
        a.bmask = hmask_;
        a.bbloom = bloom_.as<uint32_t>();
        a.bbmask = bmask_;
        a.vars = d_hbvars_.as<HbVar>();
        a.nslots = nslots_.as<uint64_t>();
        a.nmask = cap - 1;
        a.pool_keys = pool_keys_.as<uint4>();
        a.pool_rank = pool_rank_.as<uint64_t>();
        a.counters = counters_.as<uint32_t>();
        a.pool_cap = pool_cap;
        a.tlo = tlo;
        a.thi = thi;
        k_hb_gen<NW><<<std::min<unsigned>(grid_for(n * 64, 256), 16384), 256, 0, stream_>>>(a);
        HIPCHK(hipGetLastError());
        uint32_t cnt[4];
        HIPCHK(hipMemcpyAsync(cnt, counters_.p, sizeof(cnt), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        if (cnt[1] & 1u) throw std::runtime_error("Error: a heat-bath state exceeds 255 quanta in one mode.");
        if (cnt[1] & 2u) {
            if (capped) return false;
            ucap *= 4;
            continue;
        }
        // collect, sort by first-discovery rank, gather keys
        col_rank_.ensure((size_t)pool_cap * sizeof(uint64_t));
        col_idx_.ensure((size_t)pool_cap * sizeof(uint32_t));
        k_fill<uint32_t><<<1, 1, 0, stream_>>>(counters_.as<uint32_t>() + 2, 1, 0u);
        k_collect<<<grid_for(cap, 256), 256, 0, stream_>>>(nslots_.as<uint64_t>(), cap, pool_rank_.as<uint64_t>(),
                                                          counters_.as<uint32_t>() + 2, col_rank_.as<uint64_t>(),
                                                          col_idx_.as<uint32_t>());
        uint32_t u = 0;
        HIPCHK(hipMemcpyAsync(&u, counters_.as<uint32_t>() + 2, sizeof(u), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        srt_rank_.ensure((size_t)std::max<uint32_t>(u, 1) * sizeof(uint64_t));
        srt_idx_.ensure((size_t)std::max<uint32_t>(u, 1) * sizeof(uint32_t));
        unsigned endbit = 1;
        while (endbit < 64 && ((uint64_t)total >> endbit) != 0) endbit++;
        size_t stmp = 0;
        HIPCHK(rocprim::radix_sort_pairs(nullptr, stmp, col_rank_.as<uint64_t>(), srt_rank_.as<uint64_t>(),
                                         col_idx_.as<uint32_t>(), srt_idx_.as<uint32_t>(), (size_t)u, 0, endbit,
                                         stream_));
        scan_tmp_.ensure(stmp);
        HIPCHK(rocprim::radix_sort_pairs(scan_tmp_.p, stmp, col_rank_.as<uint64_t>(), srt_rank_.as<uint64_t>(),
                                         col_idx_.as<uint32_t>(), srt_idx_.as<uint32_t>(), (size_t)u, 0, endbit,
                                         stream_));
        nmask_ = cap - 1;
        pool_pos_.ensure((size_t)pool_cap * sizeof(uint32_t));
        k_pool_pos<<<grid_for((int64_t)u, 256), 256, 0, stream_>>>(srt_idx_.as<uint32_t>(), u, pool_pos_.as<uint32_t>());
        newkeys_.ensure((size_t)std::max<uint32_t>(u, 1) * KW_ * 16);
        k_gather_keys<<<grid_for((int64_t)u * KW_, 256), 256, 0, stream_>>>(
            pool_keys_.as<uint4>(), srt_idx_.as<uint32_t>(), u, KW_, newkeys_.as<uint4>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream_));
        nnew_ = u;
        st.unique += u;
        return true;
    }
    return false;
}

template <int NW>
void Engine::hb_select_t(double eps, HbStats& st) {
    const int64_t total = hb_count(eps);
    st.candidates = total;
    nnew_ = 0;
    if (total == 0) return;
    if (!hb_gen_t<NW>(total, 0, uint64_t(1) << 32, uint64_t(1) << 32, st))
        throw std::runtime_error("heat-bath selection: the new states do not fit one hash of 2^32 slots");
}

HbStats Engine::hb_select(const double* cmax_host, int nopt, double eps) {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    auto t0 = clk::now();
    HbStats st;
    const int64_t n = n_;
    cmax_.ensure(n * sizeof(double));
    if (cmax_host) {
        HIPCHK(hipMemcpyAsync(cmax_.p, cmax_host, n * sizeof(double), hipMemcpyHostToDevice, stream_));
    } else {
        if (xrows_ != n || xcols_ < 1) throw std::runtime_error("no resident eigenvectors for this basis");
        if (nopt < 1 || nopt > xcols_) nopt = std::min(xcols_, std::max(1, nopt));
        k_cmax<<<grid_for(n, 256), 256, 0, stream_>>>(n, X_.as<double>(), xcols_, nopt, cmax_.as<double>());
        HIPCHK(hipGetLastError());
    }
    switch (KW_) {
        case 1: hb_select_t<4>(eps, st); break;
        case 2: hb_select_t<8>(eps, st); break;
        case 3: hb_select_t<12>(eps, st); break;
        default: hb_select_t<16>(eps, st); break;
    }
    st.ms = ms_since(t0);
    last_hb_ = st;
    if (std::getenv("VHCI_DEBUG"))
        std::fprintf(stderr, "[heat-bath] n %lld eps %.3g candidates %lld unique %lld %.1f ms\n", (long long)n, eps,
                     (long long)st.candidates, (long long)st.unique, st.ms);
    return st;
}

void Engine::get_new_states(int32_t* q) const {
    HIPCHK(hipSetDevice(device_));
    PoolScope ps(&pool_);
    const int M = P_.M, K = KW_ * 16;
    std::vector<uint8_t> packed((size_t)nnew_ * K);
    if (nnew_)
        HIPCHK(hipMemcpyAsync(packed.data(), newkeys_.p, packed.size(), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int64_t i = 0; i < nnew_; i++)
        for (int m = 0; m < M; m++) q[i * M + m] = packed[(size_t)i * K + m];
}

void Engine::append_new_states() {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (nnew_ == 0) return;
    const int64_t n1 = n_ + nnew_;
    if (n1 >= (int64_t)INT32_MAX) throw std::runtime_error("basis larger than 2^31-1 states");
    const size_t K = (size_t)KW_ * 16;
    if (keys_.cap < n1 * K) {
        DevBuf nk;
        nk.ensure(std::max<size_t>(n1 * K, keys_.cap * 3 / 2));
        HIPCHK(hipMemcpyAsync(nk.p, keys_.p, n_ * K, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        nk.swap(keys_);
    }
    HIPCHK(hipMemcpyAsync(keys_.as<uint8_t>() + n_ * K, newkeys_.p, nnew_ * K, hipMemcpyDeviceToDevice, stream_));
    const int64_t n0 = n_;
    n_ = n1;
    nnew_ = 0;
    h_valid_ = false;
    if ((uint64_t)n1 * 2 > hmask_ + 1) {
        rebuild_hash(n1);
    } else {
        const uint4* k = keys_.as<uint4>();
        switch (KW_) {
            case 1: k_hash_insert<4><<<grid_for(n1 - n0, 256), 256, 0, stream_>>>(k, n0, n1, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
            case 2: k_hash_insert<8><<<grid_for(n1 - n0, 256), 256, 0, stream_>>>(k, n0, n1, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
            case 3: k_hash_insert<12><<<grid_for(n1 - n0, 256), 256, 0, stream_>>>(k, n0, n1, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
            default: k_hash_insert<16><<<grid_for(n1 - n0, 256), 256, 0, stream_>>>(k, n0, n1, slots_.as<uint64_t>(), hmask_, bloom_.as<uint32_t>(), bmask_); break;
        }
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(stream_));
    // a heat-bath iteration boundary: the blocks the last iteration's phases
    // gave back are sized for the smaller basis and go back to the device
    pool_.trim();
}

void Engine::pool_stats(int64_t out[5]) const {
    out[0] = pool_.mallocs;
    out[1] = pool_.frees;
    out[2] = pool_.reuses;
    out[3] = pool_.held;
    out[4] = pool_.idle_bytes;
}

// ---------------------------------------------------------------- PT2
// Push form; returns false (nothing launched) when the pair list would not
// fit in a quarter of the free device memory.
// One chunk of the push form: every variational state i applies every
// pattern it can use; a result a in the chunk's external states (hash tag in
// [tlo, thi)) gives the pair (a, i).  The pairs, sorted by a (stably, so i
// ascends within a), feed one wave per external state (k_pt2_num).  Returns
// false, with nothing summed, when the chunk has more than max_pairs pairs
// or their buffers cannot be allocated.
template <int NW>
bool Engine::pt2_push_t(const double* Ed, int nopt, int nwaves, int64_t max_pairs, bool carry, uint64_t tlo,
                        uint64_t thi, Pt2Stats& st) {
    const int64_t n = n_, ns2 = nnew_;
    Pt2PushArgs a{};
    a.n = n;
    a.M = P_.M;
    a.P = P_deltas_;
    a.vkeys = keys_.as<uint4>();
    // the chunk's states in a hash sized for them (load <= 1/2; the dedup
    // hash of the selection is sized from an estimate and is often far
    // larger), found at their canonical positions
    const uint64_t cap = next_pow2(std::max<uint64_t>(1024, 2 * (uint64_t)ns2));
    s2slots_.ensure(cap * sizeof(uint64_t));
    k_fill<uint64_t><<<grid_for(cap, 256), 256, 0, stream_>>>(s2slots_.as<uint64_t>(), (int64_t)cap, kEmpty);
    const uint64_t fbits = bloom_bits(cap);
    s2bloom_.ensure(fbits / 8);
    k_fill<uint32_t><<<grid_for((int64_t)(fbits / 32), 256), 256, 0, stream_>>>(s2bloom_.as<uint32_t>(),
                                                                                (int64_t)(fbits / 32), 0u);
    k_hash_insert<NW><<<grid_for(ns2, 256), 256, 0, stream_>>>(newkeys_.as<uint4>(), 0, ns2, s2slots_.as<uint64_t>(),
                                                               cap - 1, s2bloom_.as<uint32_t>(), fbits - 1);
    HIPCHK(hipGetLastError());
    a.aslots = s2slots_.as<uint64_t>();
    a.amask = cap - 1;
    a.abloom = s2bloom_.as<uint32_t>();
    a.abmask = fbits - 1;
    a.apool = newkeys_.as<uint4>();
    a.apos = nullptr;
    a.deltas = d_deltas_.as<DeltaEnt>();
    a.walk = d_walk_.as<uint64_t>();
    if (!a.walk) throw std::runtime_error("pattern walk records missing");
    a.poff = d_doff_.as<uint32_t>();
    a.tlo = tlo;
    a.thi = thi;
    rowcnt_.ensure((n + 1) * sizeof(int64_t));
    pairoff_.ensure((n + 1) * sizeof(int64_t));
    a.cnt = rowcnt_.as<int64_t>();
    // pairs estimated from every 64th state first: a chunk far over the
    // budget is split without a full count
    constexpr int64_t kSample = 64;
    double est = -1.0;
    if (n >= 64 * kSample) {
        const int64_t ns = (n + kSample - 1) / kSample;
        a.rstride = kSample;
        k_pt2_push<NW, false><<<(unsigned)ns, 64, 0, stream_>>>(a);
        HIPCHK(hipGetLastError());
        std::vector<int64_t> sc(ns);
        HIPCHK(hipMemcpyAsync(sc.data(), a.cnt, ns * sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        est = (double)std::accumulate(sc.begin(), sc.end(), int64_t(0)) * ((double)n / (double)ns);
        if (est > 1.25 * (double)max_pairs) {
            last_pairs_ = (int64_t)(1.25 * est);
            return false;
        }
    }
    unsigned endbit = 1;
    while (endbit < 32 && ((uint64_t)ns2 >> endbit) != 0) endbit++;
    DevBuf pk, pv, sk, sv, seg, cur;
    int64_t npairs = -1;
    const uint32_t* skp = nullptr;  // sorted: external state of each pair
    const uint64_t* svp = nullptr;  // sorted: (variational state << 32 | pattern)
    // One pass (k_pt2_push1) into buffers sized from the estimate; the
    // counted form (count, scan, fill) when there is no estimate, when the
    // pairs overflow the buffers, or with VHCI_PT2_COUNTED=1.  Both give the
    // same sorted pairs.
    const char* ce = std::getenv("VHCI_PT2_COUNTED");
    if (est >= 0.0 && !(ce && std::string(ce) == "1")) {
        const int64_t cap = std::min<int64_t>(max_pairs, (int64_t)(1.25 * est) + (int64_t(1) << 20));
        bool ok = true;
        try {
            pk.ensure((size_t)cap * sizeof(uint64_t));  // keys (a << 32 | i); after the sort, the values
            pv.ensure((size_t)cap * sizeof(uint32_t));  // patterns; after the sort, the external states
            sk.ensure((size_t)cap * sizeof(uint64_t));
            sv.ensure((size_t)cap * sizeof(uint32_t));
            seg.ensure((ns2 + 1) * sizeof(int64_t));
            cur.ensure(sizeof(unsigned long long));
        } catch (const std::runtime_error&) {
            ok = false;
        }
        if (ok) {
            k_fill<uint64_t><<<1, 64, 0, stream_>>>(cur.as<uint64_t>(), 1, 0ull);
            k_pt2_push1<NW><<<(unsigned)n, 64, 0, stream_>>>(a, cur.as<unsigned long long>(), pk.as<uint64_t>(),
                                                             pv.as<uint32_t>(), (unsigned long long)cap);
            HIPCHK(hipGetLastError());
            unsigned long long got = 0;
            HIPCHK(hipMemcpyAsync(&got, cur.p, sizeof(got), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            if ((int64_t)got <= cap) {
                npairs = (int64_t)got;
                if (npairs > 0) {
                    size_t stmp = 0;
                    HIPCHK(rocprim::radix_sort_pairs(nullptr, stmp, pk.as<uint64_t>(), sk.as<uint64_t>(),
                                                     pv.as<uint32_t>(), sv.as<uint32_t>(), (size_t)npairs, 0,
                                                     32 + endbit, stream_));
                    scan_tmp_.ensure(stmp);
                    HIPCHK(rocprim::radix_sort_pairs(scan_tmp_.p, stmp, pk.as<uint64_t>(), sk.as<uint64_t>(),
                                                     pv.as<uint32_t>(), sv.as<uint32_t>(), (size_t)npairs, 0,
                                                     32 + endbit, stream_));
                    k_pair_split<<<std::min<unsigned>(grid_for(npairs, 256), 65536), 256, 0, stream_>>>(
                        sk.as<uint64_t>(), sv.as<uint32_t>(), npairs, pv.as<uint32_t>(), pk.as<uint64_t>());
                    HIPCHK(hipGetLastError());
                }
                skp = pv.as<uint32_t>();
                svp = pk.as<uint64_t>();
            }
        }
    }
    if (npairs < 0) {
        a.rstride = 1;
        k_pt2_push<NW, false><<<(unsigned)n, 64, 0, stream_>>>(a);
        k_fill<int64_t><<<1, 1, 0, stream_>>>(a.cnt + n, 1, 0);
        size_t tmp = 0;
        HIPCHK(rocprim::exclusive_scan(nullptr, tmp, a.cnt, pairoff_.as<int64_t>(), (int64_t)0, (size_t)(n + 1),
                                       rocprim::plus<int64_t>(), stream_));
        scan_tmp_.ensure(tmp);
        HIPCHK(rocprim::exclusive_scan(scan_tmp_.p, tmp, a.cnt, pairoff_.as<int64_t>(), (int64_t)0, (size_t)(n + 1),
                                       rocprim::plus<int64_t>(), stream_));
        HIPCHK(hipMemcpyAsync(&npairs, pairoff_.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        if (npairs > max_pairs) {
            last_pairs_ = npairs;
            return false;
        }
        try {
            pk.ensure(std::max<int64_t>(1, npairs) * sizeof(uint32_t));
            pv.ensure(std::max<int64_t>(1, npairs) * sizeof(uint64_t));
            sk.ensure(std::max<int64_t>(1, npairs) * sizeof(uint32_t));
            sv.ensure(std::max<int64_t>(1, npairs) * sizeof(uint64_t));
            seg.ensure((ns2 + 1) * sizeof(int64_t));
        } catch (const std::runtime_error&) {
            return false;  // the chunk is split instead
        }
        a.off = pairoff_.as<int64_t>();
        a.pkey = pk.as<uint32_t>();
        a.pval = pv.as<uint64_t>();
        if (npairs > 0) {
            k_pt2_push<NW, true><<<(unsigned)n, 64, 0, stream_>>>(a);
            HIPCHK(hipGetLastError());
            size_t stmp = 0;
            HIPCHK(rocprim::radix_sort_pairs(nullptr, stmp, pk.as<uint32_t>(), sk.as<uint32_t>(), pv.as<uint64_t>(),
                                             sv.as<uint64_t>(), (size_t)npairs, 0, endbit, stream_));
            scan_tmp_.ensure(stmp);
            HIPCHK(rocprim::radix_sort_pairs(scan_tmp_.p, stmp, pk.as<uint32_t>(), sk.as<uint32_t>(),
                                             pv.as<uint64_t>(), sv.as<uint64_t>(), (size_t)npairs, 0, endbit,
                                             stream_));
        }
        skp = sk.as<uint32_t>();
        svp = sv.as<uint64_t>();
    }
    if (std::getenv("VHCI_DEBUG"))
        std::fprintf(stderr, "[pt2] chunk [%#llx, %#llx): %lld external states, %lld pairs (%s)\n",
                     (unsigned long long)tlo, (unsigned long long)thi, (long long)ns2, (long long)npairs,
                     skp == sk.as<uint32_t>() ? "counted" : "one pass");
    k_seg_start<<<grid_for(ns2 + 1, 256), 256, 0, stream_>>>(skp, npairs, ns2, seg.as<int64_t>());
    HIPCHK(hipGetLastError());
    const bool mixed = form_ == kMixed;
    Pt2NumArgs b{};
    b.ns2 = ns2;
    b.M = P_.M;
    b.nq = 256;
    b.nopt = nopt;
    b.akeys = newkeys_.as<uint4>();
    b.seg = seg.as<int64_t>();
    b.sval = svp;
    b.deltas = d_deltas_.as<DeltaEnt>();
    b.groups = d_groups_.as<GroupEnt>();
    b.fclist = d_fclist_.as<uint32_t>();
    b.elem = d_elem_.as<ElemFC>();
    b.ladder = d_ladder_.as<double>();
    b.dfc = d_diag_.as<uint32_t>();
    b.ndfc = (int)T_.diag_fcs.size();
    b.freq = d_freq_.as<double>();
    b.C = X_.as<double>();
    b.C32 = C32_.as<float>();
    b.ldc = mixed ? nopt : xcols_;
    b.E = Ed;
    b.part = pt2_part_.as<double>();
    b.carry = carry ? 1 : 0;
    if (mixed && nopt <= 128)
        k_pt2_num<NW, true, 2><<<nwaves, 64, 0, stream_>>>(b);
    else if (mixed)
        k_pt2_num<NW, true, 4><<<nwaves, 64, 0, stream_>>>(b);
    else if (nopt <= 128)
        k_pt2_num<NW, false, 2><<<nwaves, 64, 0, stream_>>>(b);
    else
        k_pt2_num<NW, false, 4><<<nwaves, 64, 0, stream_>>>(b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    st.pairs += npairs;
    return true;
}

// One chunk of the pull form (VHCI_PT2_PULL=1, or a table without
// patterns): one wave per external state probes its patterns against the
// variational hash.
template <int NW>
void Engine::pt2_pull_t(const double* Ed, int nopt, int nwaves, bool carry) {
    const bool mixed = form_ == kMixed;
    Pt2Args a{};
    a.ns2 = nnew_;
    a.M = P_.M;
    a.P = P_deltas_;
    a.nq = 256;
    a.nopt = nopt;
    a.akeys = newkeys_.as<uint4>();
    a.vkeys = keys_.as<uint4>();
    a.vslots = slots_.as<uint64_t>();
    a.vmask = hmask_;
    a.vbloom = bloom_.as<uint32_t>();
    a.vbmask = bmask_;
    a.deltas = d_deltas_.as<DeltaEnt>();
    a.poff = d_doff_.as<uint32_t>();
    a.groups = d_groups_.as<GroupEnt>();
    a.fclist = d_fclist_.as<uint32_t>();
    a.elem = d_elem_.as<ElemFC>();
    a.ladder = d_ladder_.as<double>();
    a.dfc = d_diag_.as<uint32_t>();
    a.ndfc = (int)T_.diag_fcs.size();
    a.freq = d_freq_.as<double>();
    a.C = X_.as<double>();
    a.C32 = C32_.as<float>();
    a.ldc = mixed ? nopt : xcols_;
    a.E = Ed;
    a.part = pt2_part_.as<double>();
    a.carry = carry ? 1 : 0;
    if (mixed)
        k_pt2<NW, true><<<nwaves, 64, 0, stream_>>>(a);
    else
        k_pt2<NW, false><<<nwaves, 64, 0, stream_>>>(a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
}

// DoPT2 over the external space in chunks.  The space is split by the hash
// tag of its states (high 32 bits of the key hash): a chunk is a tag range,
// and its states, their dedup hash and their (a, i) pairs are all that is
// held at once.  The ranges start as one; a range whose states need more
// than the state budget, or whose pairs exceed the pair budget, is halved
// (the lower half first), so the chunks are the same on every run with the
// same budgets.  Every chunk continues the same per-wave compensated sums,
// which are reduced once at the end.
template <int NW>
void Engine::pt2_t(const std::vector<double>& E, int nopt, double eps2, Pt2Stats& st, std::vector<double>& dE) {
    dE.assign(nopt, 0.0);
    auto t0 = clk::now();
    HbStats hs;
    const int64_t total = hb_count(eps2);
    hs.candidates = total;
    st.candidates = total;
    nnew_ = 0;
    int64_t max_pairs = kPt2ChunkPairs, max_states = kPt2ChunkStates;
    if (const char* e = std::getenv("VHCI_PT2_CHUNK_PAIRS")) max_pairs = std::max<int64_t>(1, std::atoll(e));
    if (const char* e = std::getenv("VHCI_PT2_CHUNK_STATES")) max_states = std::max<int64_t>(1, std::atoll(e));
    const bool pull = std::getenv("VHCI_PT2_PULL") != nullptr || P_deltas_ == 0;
    DevBuf Ed, out;
    Ed.ensure(nopt * sizeof(double));
    out.ensure(nopt * sizeof(double));
    HIPCHK(hipMemcpyAsync(Ed.p, E.data(), nopt * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (form_ == kMixed) {
        C32_.ensure((size_t)n_ * nopt * sizeof(float));
        k_to_f32<<<grid_for(n_ * nopt, 256), 256, 0, stream_>>>(X_.as<double>(), xcols_, C32_.as<float>(), nopt, n_,
                                                               nopt);
    }
    double sel_ms = ms_since(t0), sum_ms = 0;
    int nwaves = 0;
    bool summed = false;
    std::vector<std::pair<uint64_t, uint64_t>> work{{0, uint64_t(1) << 32}};
    // a range is cut into `parts` equal ranges (a power of two), pushed so
    // that the lowest comes next
    auto split = [&](uint64_t lo, uint64_t hi, uint64_t parts) {
        if (hi - lo < 2) throw std::runtime_error("PT2: one hash tag holds more external states or pairs than a chunk");
        parts = std::min<uint64_t>(std::max<uint64_t>(2, next_pow2(parts)), hi - lo);
        const uint64_t w = (hi - lo) / parts;
        for (uint64_t k = parts; k-- > 0;) work.push_back({lo + k * w, k + 1 == parts ? hi : lo + (k + 1) * w});
    };
    while (total > 0 && !work.empty()) {
        const auto [lo, hi] = work.back();
        work.pop_back();
        auto t1 = clk::now();
        const int64_t u0 = hs.unique;
        if (!hb_gen_t<NW>(total, lo, hi, (uint64_t)max_states, hs)) {
            sel_ms += ms_since(t1);
            split(lo, hi, 2);
            continue;
        }
        sel_ms += ms_since(t1);
        const int64_t ns2 = nnew_;
        if (ns2 == 0) continue;
        if (nwaves == 0) {
            // one chunk: as many waves as states, up to 8192 (the grid of
            // the unchunked sum); several: 8192
            nwaves = work.empty() ? (int)std::min<int64_t>(ns2, 256 * 32) : 256 * 32;
            pt2_part_.ensure((size_t)nwaves * nopt * 2 * sizeof(double));
        }
        auto t2 = clk::now();
        if (pull) {
            pt2_pull_t<NW>(Ed.as<double>(), nopt, nwaves, summed);
        } else if (!pt2_push_t<NW>(Ed.as<double>(), nopt, nwaves, max_pairs, summed, lo, hi, st)) {
            sum_ms += ms_since(t2);
            hs.unique = u0;
            // as many parts as the counted pairs need (two when the buffers
            // could not be allocated)
            split(lo, hi, last_pairs_ > max_pairs ? (uint64_t)((last_pairs_ + max_pairs - 1) / max_pairs) : 2);
            last_pairs_ = 0;
            continue;
        }
        sum_ms += ms_since(t2);
        summed = true;
        st.ns2 += ns2;
        st.chunks++;
    }
    if (summed) {
        auto t3 = clk::now();
        k_pt2_reduce<<<grid_for(nopt, 64), 64, 0, stream_>>>(pt2_part_.as<double>(), nwaves, nopt, out.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dE.data(), out.p, nopt * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        sum_ms += ms_since(t3);
    }
    hs.ms = sel_ms;
    last_hb_ = hs;
    st.select_ms = sel_ms;
    st.sum_ms = sum_ms;
}

// the eigensolver's work blocks (the Krylov basis, its products, the
// Rayleigh-Ritz buffers) go back to the device before PT2; the next
// eigensolve allocates them again
void Engine::release_solver_work() {
    for (DevBuf* b : {&V_, &W_, &HX_, &Xw_, &T_b_, &G_, &S_, &part_, &X32_, &stage_in_, &stage_in32_, &stage_out_,
                      &otmp_, &Gw_, &tsm_part_, &rr_part_, &rot_s_})
        b->release();
    pool_.trim();
}

Pt2Stats Engine::pt2(const std::vector<double>& E, int nopt, double eps2, std::vector<double>& dE) {
    PoolScope ps(&pool_);
    HIPCHK(hipSetDevice(device_));
    if (nopt < 1 || nopt > xcols_ || (int)E.size() < nopt) throw std::runtime_error("PT2: bad number of states");
    if (nopt > 256) throw std::runtime_error("PT2: at most 256 states");
    const int64_t n = n_;
    if (xrows_ != n) throw std::runtime_error("no resident eigenvectors for this basis");
    release_solver_work();
    cmax_.ensure(n * sizeof(double));
    k_cmax<<<grid_for(n, 256), 256, 0, stream_>>>(n, X_.as<double>(), xcols_, nopt, cmax_.as<double>());
    HIPCHK(hipGetLastError());
    Pt2Stats st;
    switch (KW_) {
        case 1: pt2_t<4>(E, nopt, eps2, st, dE); break;
        case 2: pt2_t<8>(E, nopt, eps2, st, dE); break;
        case 3: pt2_t<12>(E, nopt, eps2, st, dE); break;
        default: pt2_t<16>(E, nopt, eps2, st, dE); break;
    }
    last_pt2_ = st;
    last_pt2_n_ = n_;
    last_pt2_nopt_ = nopt;
    if (std::getenv("VHCI_DEBUG"))
        std::fprintf(stderr,
                     "[pt2] n %lld candidates %lld external %lld pairs %lld chunks %d select %.1f ms sum %.1f ms\n",
                     (long long)n_, (long long)st.candidates, (long long)st.ns2, (long long)st.pairs, st.chunks,
                     st.select_ms, st.sum_ms);
    return st;
}

}  // namespace vhci
