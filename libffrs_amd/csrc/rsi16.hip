// rsi16.hip -- RSi16 (GF(65537), NTT-based) systematic encode and repair on gfx950.
//
// Reference semantics (paths relative to the reference repository root):
//   field ops     libffrs/galois.hpp:93-146, 168-216, 258-395 (scalar and SIMD-lane branches)
//   NTT tables    libffrs/ntt.hpp:55-123; butterflies :225-326; nttr/inttr :569-591
//   encode        libffrs/rsi16v_impl.hpp:93-119 via NTT::pntt_message_residue (ntt.hpp:180-199)
//   repair        libffrs/rsi16v_impl.hpp:122-168 (_repair), :271-440 (_sugiyama),
//                 :443-512 (_find_roots_ntt, _forney), ntt.hpp:133-152 (pntt),
//                 ntt.hpp:343-564 (reversal, Newton inverse, NTT-domain division)
//   known pos.    libffrs/rsi16v_impl.hpp:171-231, 532-563
//
// Layout: a wave is split into lane groups that each own one codeword. The default
// kernels (ecc_len <= 256) keep every polynomial in registers, ecc/G coefficients per
// lane for groups of G = min(64, ecc/2) lanes, and exchange butterfly partners across
// lanes with DPP / ds_swizzle / ds_bpermute; encode for ecc_len <= 32 runs one codeword
// per lane. The LDS kernels (other group sizes, larger ecc_len, interleaved encode) keep
// the polynomials in LDS with stages ordered by wavefront-scope fences. A group never
// waits on another, so the groups of one wave may follow different control flow.
//
// Exactness: the encode path reproduces the reference's lazy residue arithmetic
// word for word (the u32 values, not just their residues), because the parity depends
// on them in corner cases; VectorOps / ScalarOps select the SIMD-lane or the scalar
// variant of mod_p / mul / mul_residue / sub. The repair path uses the scalar
// full-reduction operations, whose outputs are canonical field elements; its Sugiyama
// solver computes the reference's values in coefficient form (see pr_sugiyama).
#include "common.hpp"

#include <algorithm>
#include <type_traits>
#include <cstring>

namespace ffrs {

constexpr uint32_t P16 = 0x10001u;
constexpr uint32_t RSI16_MAX_ECC = 4096;

// ---------------------------------------------------------------- field operations

// Every product operand is at most p = 65537 (canonical values, mod_p results, table
// entries, u16 symbols): telling the compiler so lets it use the full-rate 24-bit multiply,
// whose low 32 bits are the 32-bit product for such operands.
__host__ __device__ inline void ffrs_u24(uint32_t a, uint32_t b)
{
    __builtin_assume(a <= P16);
    __builtin_assume(b <= P16);
}

struct ScalarOps {  // integral branches
    __host__ __device__ static inline uint32_t add(uint32_t a, uint32_t b)
    {
        uint32_t r = a + b;
        return r >= P16 ? r - P16 : r;
    }
    __host__ __device__ static inline uint32_t sub(uint32_t a, uint32_t b)
    {
        uint32_t r = a - b;
        return r >= 0x80000000u ? r + P16 : r;
    }
    __host__ __device__ static inline uint32_t mul(uint32_t a, uint32_t b)
    {
        ffrs_u24(a, b);
        uint32_t r = a * b;
        if (r == 0 && a && b)
            return 1;
        r = (r & 0xffff) - (r >> 16);
        return r >= 0x80000000u ? r + P16 : r;
    }
    __host__ __device__ static inline uint32_t mul_residue(uint32_t a, uint32_t b)
    {
        ffrs_u24(a, b);
        uint32_t r = a * b;
        if (r == 0 && a && b)
            return 1;
        return (r & 0xffff) - (r >> 16);
    }
    __host__ __device__ static inline uint32_t mod_p(uint32_t x)
    {
        uint32_t r = (x & 0xffff) - (x >> 16) - (x >> 31);
        return r >= 0x80000000u ? r + P16 : r;
    }
};

struct VectorOps {  // SIMD-lane branches
    __host__ __device__ static inline uint32_t add(uint32_t a, uint32_t b)
    {
        uint32_t r = a + b;
        return r - (r >= P16 ? P16 : 0u);
    }
    __host__ __device__ static inline uint32_t sub(uint32_t a, uint32_t b) { return a - b + (b > a ? P16 : 0u); }
    __host__ __device__ static inline uint32_t mul(uint32_t a, uint32_t b)
    {
        ffrs_u24(a, b);
        uint32_t r = a * b;
        uint32_t ovf = (r == 0 && (a & b) != 0) ? 1u : 0u;
        uint32_t lo = r & 0xffff, hi = r >> 16;
        return lo - hi + ovf + (hi > lo ? P16 : 0u);
    }
    __host__ __device__ static inline uint32_t mul_residue(uint32_t a, uint32_t b)
    {
        ffrs_u24(a, b);
        uint32_t r = a * b;
        uint32_t ovf = (r == 0 && (a & b) != 0) ? 1u : 0u;
        return (r & 0xffff) - (r >> 16) + ovf;
    }
    __host__ __device__ static inline uint32_t mod_p(uint32_t x)
    {
        uint32_t lo = x & 0xffff, hi = x >> 16;
        return lo - hi - (x >> 31) + (hi >= lo ? P16 : 0u);
    }
};

using S = ScalarOps;

__host__ __device__ inline uint32_t s_neg(uint32_t a) { return a == 0 ? 0u : P16 - a; }

// a^e with canonical scalar products
__host__ __device__ inline uint32_t s_pow(uint32_t a, uint32_t e)
{
    uint32_t r = 1;
    while (e) {
        if (e & 1)
            r = S::mul(r, a);
        a = S::mul(a, a);
        e >>= 1;
    }
    return r;
}

// Table conventions of galois.hpp:98-114 over GF(65537): inv(0) = 1, div(0, b) = 0,
// div(a, 0) = a; otherwise the field inverse / quotient (a^65535 = a^-1).
__host__ __device__ inline uint32_t g_inv(uint32_t a) { return a == 0 ? 1u : s_pow(a, 65535u); }
__host__ __device__ inline uint32_t g_div(uint32_t a, uint32_t b)
{
    if (a == 0)
        return 0;
    if (b == 0)
        return a;
    return S::mul(a, g_inv(b));
}

__host__ __device__ inline uint32_t rbo_bits(uint32_t b, unsigned bits)
{
    return bits ? (__builtin_bitreverse32(b) >> (32 - bits)) : 0u;
}

// ---------------------------------------------------------------- device tables

struct Rsi16Dev {
    uint32_t n, k, E, N, nbits;
    uint32_t ecc_inv, ecc2_inv;
    const uint32_t *roots_ecc, *roots_i_ecc;     // E
    const uint32_t *roots_ecc2, *roots_i_ecc2;   // 2E
    const uint32_t *pntt_shift;                  // n
    const uint32_t *pintt_shift;                 // N
    const uint32_t *roots_ntt;                   // N
    const uint32_t *mix;                         // E
    const uint32_t *rbo_ecc;                     // E
    const uint32_t *inv_tab;                     // 65537: inv_tab[a] = g_inv(a)
};

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

struct Grp {
    uint32_t gl;  // lane within the group
    uint32_t g;   // group size
};

// ---- butterflies over x[0..len), spread over the group's lanes

// Cooperative CT butterflies (RBO in, natural out), ntt.hpp:225-249, full reduction.
__device__ inline void ct_full(uint32_t *x, uint32_t len, const uint32_t *roots, Grp G)
{
    for (uint32_t s = 1, ef = len >> 1; s < len; s <<= 1, ef >>= 1) {
        for (uint32_t b = G.gl; b < len / 2; b += G.g) {
            uint32_t j = b & (s - 1);
            uint32_t i = ((b - j) << 1) + j;
            uint32_t a = x[i], c = x[i + s];
            if (j) {
                c = S::mul(c, roots[ef * j]);
            }
            x[i] = S::add(a, c);
            x[i + s] = S::sub(a, c);
        }
        wave_sync();
    }
}

// Cooperative CT butterflies with the lazy residue arithmetic, ntt.hpp:252-275.
template <class O>
__device__ inline void ct_residue(uint32_t *x, uint32_t len, const uint32_t *roots, Grp G)
{
    for (uint32_t s = 1, ef = len >> 1; s < len; s <<= 1, ef >>= 1) {
        for (uint32_t b = G.gl; b < len / 2; b += G.g) {
            uint32_t j = b & (s - 1);
            uint32_t i = ((b - j) << 1) + j;
            uint32_t a = x[i], c = x[i + s];
            if (j) {
                uint32_t m = O::mul_residue(O::mod_p(c), roots[ef * j]);
                x[i] = a + m;
                x[i + s] = a - m;
            } else {
                x[i] = a + c;
                x[i + s] = a - c;
            }
        }
        wave_sync();
    }
}

// Cooperative GS butterflies (natural in, RBO out), ntt.hpp:280-302.
template <class O>
__device__ inline void gs_full(uint32_t *x, uint32_t len, const uint32_t *roots, Grp G)
{
    for (uint32_t s = len >> 1, ef = 0; s > 0; s >>= 1, ++ef) {
        for (uint32_t b = G.gl; b < len / 2; b += G.g) {
            uint32_t j = b & (s - 1);
            uint32_t i = ((b - j) << 1) + j;
            uint32_t a = x[i], c = x[i + s];
            x[i] = O::add(a, c);
            uint32_t d = O::sub(a, c);
            x[i + s] = j ? O::mul(d, roots[j << ef]) : d;
        }
        wave_sync();
    }
}

__device__ inline void nttr(const Rsi16Dev &T, uint32_t *x, Grp G) { gs_full<S>(x, T.E, T.roots_ecc, G); }

__device__ inline void inttr(const Rsi16Dev &T, uint32_t *x, Grp G)
{
    ct_full(x, T.E, T.roots_i_ecc, G);
    for (uint32_t j = G.gl; j < T.E; j += G.g)
        x[j] = S::mul(x[j], T.ecc_inv);
    wave_sync();
}

// degree + 1 of r[0..max) (simd.hpp:234-242); every lane of the group scans.
__device__ inline uint32_t poly_len(const uint32_t *r, uint32_t max)
{
    while (max > 0 && r[max - 1] == 0)
        --max;
    return max;
}

// The same length computed by the whole group at once: each lane scans its strided
// elements, then a maximum over the group's lanes (xor shuffles by less than g stay inside
// the group, whose lanes are an aligned block of g). Every lane of the group must call it.
__device__ inline uint32_t group_len(const uint32_t *r, uint32_t max, Grp G)
{
    uint32_t hi = 0;
    for (uint32_t j = G.gl; j < max; j += G.g)
        if (r[j] != 0)
            hi = j + 1;
    for (uint32_t s = 1; s < G.g; s <<= 1) {
        uint32_t o = __shfl_xor(hi, s);
        hi = o > hi ? o : hi;
    }
    return hi;
}

// ---------------------------------------------------------------- encode

// One codeword per lane group: pntt_message_residue + _mix_ecc_residue
// (rsi16v_impl.hpp:93-119). buf holds 2 E words. Symbols are u16, or u32 field elements (the
// reference encodes GFT sources as they are); u16 outputs are written as & 0xFFFF, u32 outputs
// keep the word.
template <class O, class TI = uint16_t, class TO = uint16_t>
__device__ inline void encode_group(const Rsi16Dev &T, const TI *m, TO *o, uint32_t *buf, Grp G,
                                    uint32_t stride = 1)
{
    const uint32_t E = T.E;
    uint32_t *acc = buf + E;
    uint32_t nb = T.n / E;
    for (uint32_t c = 0; c + 1 < nb; ++c) {
        for (uint32_t j = G.gl; j < E; j += G.g)
            buf[j] = m[(size_t)(c * E + j) * stride];
        wave_sync();
        ct_residue<O>(buf, E, T.roots_ecc, G);
        for (uint32_t j = G.gl; j < E; j += G.g) {
            uint32_t v = O::mul_residue(O::mod_p(buf[j]), T.pntt_shift[c * E + j]);
            acc[j] = c ? acc[j] + v : v;
        }
        wave_sync();
    }
    for (uint32_t j = G.gl; j < E; j += G.g)
        acc[j] = O::mul(O::mod_p(acc[j]), T.mix[j]);
    wave_sync();
    gs_full<O>(acc, E, T.roots_i_ecc, G);
    for (uint32_t j = G.gl; j < E; j += G.g) {
        if constexpr (sizeof(TO) == 2)
            o[(size_t)j * stride] = (uint16_t)(acc[j] & 0xffff);
        else
            o[(size_t)j * stride] = acc[j];
    }
}

// Codeword cw = b * I + c is column c of interleaved block b: symbol i of it sits at
// i * I + c of the block (pyrsi16.hpp:222-232, 324-361); I = 1 is the plain layout.
template <class O, class TI = uint16_t, class TO = uint16_t>
__global__ __launch_bounds__(64) void rsi16_encode_kernel(const TI *__restrict__ msgs, uint32_t count,
                                                           TO *__restrict__ out, Rsi16Dev T, uint32_t g,
                                                           uint32_t I)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E;
    Grp G{threadIdx.x % g, g};
    uint32_t grp = threadIdx.x / g;
    uint32_t cpw = 64 / g;
    uint32_t cw = blockIdx.x * cpw + grp;
    if (cw >= count)
        return;
    const size_t b = cw / I, c = cw % I;
    encode_group<O, TI, TO>(T, msgs + b * T.k * I + c, out + b * E * I + c, s_mem + grp * 2 * E, G, I);
}

// ---- one codeword per lane, transforms in registers (ecc_len <= 32)

// ct_residue with the E values in registers; the butterflies of a stage are independent,
// so running them in sequence gives the same words as the cooperative version.
template <class O, int E>
__device__ inline void ct_residue_reg(uint32_t (&x)[E], const uint32_t *__restrict__ roots)
{
#pragma unroll
    for (int s = 1; s < E; s <<= 1) {
        const int ef = E / (2 * s);
#pragma unroll
        for (int b = 0; b < E / 2; ++b) {
            const int j = b & (s - 1);
            const int i = ((b - j) << 1) + j;
            const uint32_t a = x[i], c = x[i + s];
            if (j) {
                const uint32_t m = O::mul_residue(O::mod_p(c), roots[ef * j]);
                x[i] = a + m;
                x[i + s] = a - m;
            } else {
                x[i] = a + c;
                x[i + s] = a - c;
            }
        }
    }
}

// gs_full with the E values in registers.
template <class O, int E>
__device__ inline void gs_reg(uint32_t (&x)[E], const uint32_t *__restrict__ roots)
{
    int ef = 0;
#pragma unroll
    for (int s = E / 2; s > 0; s >>= 1, ++ef) {
#pragma unroll
        for (int b = 0; b < E / 2; ++b) {
            const int j = b & (s - 1);
            const int i = ((b - j) << 1) + j;
            const uint32_t a = x[i], c = x[i + s];
            x[i] = O::add(a, c);
            const uint32_t d = O::sub(a, c);
            x[i + s] = j ? O::mul(d, roots[j << ef]) : d;
        }
    }
}

// encode_group for one codeword in one lane: the twiddles and shifts are wave-uniform
// scalar loads; with ALIGNED (16-byte aligned rows, ecc_len >= 8) the message is read in
// 16-byte pieces (every chunk of such a row starts 16-byte aligned) and the parity written
// the same way.
template <class O, int E, bool ALIGNED>
__device__ inline void encode_lane(const Rsi16Dev &T, const uint16_t *__restrict__ m, uint16_t *__restrict__ o)
{
    const uint32_t nb = T.n / E;
    // message chunks loaded PF at a time before any is transformed (for short ecc_len each
    // chunk is little work, and long rows would otherwise wait on one load per chunk)
    constexpr uint32_t PF = E <= 8 ? 8 : (E <= 16 ? 4 : 2);
    uint32_t acc[E];
    for (uint32_t c0 = 0; c0 + 1 < nb; c0 += PF) {
        uint32_t xs[PF][E];
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u) {
            if (c0 + u + 1 < nb) {
                const uint16_t *mc = m + (c0 + u) * E;
                if constexpr (ALIGNED) {
#pragma unroll
                    for (int q = 0; q < E / 8; ++q) {
                        const uint4 v = reinterpret_cast<const uint4 *>(mc)[q];
                        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int h = 0; h < 4; ++h) {
                            xs[u][8 * q + 2 * h] = w[h] & 0xffffu;
                            xs[u][8 * q + 2 * h + 1] = w[h] >> 16;
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < E; ++j)
                        xs[u][j] = mc[j];
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u) {
            const uint32_t c = c0 + u;
            if (c + 1 < nb) {
                ct_residue_reg<O, E>(xs[u], T.roots_ecc);
                const uint32_t *sh = T.pntt_shift + c * E;
#pragma unroll
                for (int j = 0; j < E; ++j) {
                    const uint32_t v = O::mul_residue(O::mod_p(xs[u][j]), sh[j]);
                    acc[j] = c ? acc[j] + v : v;
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < E; ++j)
        acc[j] = O::mul(O::mod_p(acc[j]), T.mix[j]);
    gs_reg<O, E>(acc, T.roots_i_ecc);
    if constexpr (ALIGNED) {
#pragma unroll
        for (int q = 0; q < E / 8; ++q) {
            uint4 v;
            v.x = (acc[8 * q + 0] & 0xffffu) | (acc[8 * q + 1] << 16);
            v.y = (acc[8 * q + 2] & 0xffffu) | (acc[8 * q + 3] << 16);
            v.z = (acc[8 * q + 4] & 0xffffu) | (acc[8 * q + 5] << 16);
            v.w = (acc[8 * q + 6] & 0xffffu) | (acc[8 * q + 7] << 16);
            reinterpret_cast<uint4 *>(o)[q] = v;
        }
    } else {
#pragma unroll
        for (int j = 0; j < E; ++j)
            o[j] = (uint16_t)(acc[j] & 0xffff);
    }
}

// encode_lane with the message chunks of one codeword spread over L lanes (gl = lane in the
// group): each lane sums its chunks' shifted residues, the L partial sums are added across
// the group (integer sums mod 2^32, so the order does not change them), and every lane
// finishes the same words and stores its share. For long rows with short ecc_len.
template <class O, int E, bool ALIGNED>
__device__ inline void encode_split(const Rsi16Dev &T, const uint16_t *__restrict__ m, uint16_t *__restrict__ o,
                                    uint32_t gl, uint32_t L)
{
    const uint32_t nb = T.n / E;
    uint32_t acc[E];
#pragma unroll
    for (int j = 0; j < E; ++j)
        acc[j] = 0;
    for (uint32_t c = gl; c + 1 < nb; c += L) {
        uint32_t x[E];
        const uint16_t *mc = m + c * E;
        if constexpr (ALIGNED) {
#pragma unroll
            for (int q = 0; q < E / 8; ++q) {
                const uint4 v = reinterpret_cast<const uint4 *>(mc)[q];
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    x[8 * q + 2 * h] = w[h] & 0xffffu;
                    x[8 * q + 2 * h + 1] = w[h] >> 16;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < E; ++j)
                x[j] = mc[j];
        }
        ct_residue_reg<O, E>(x, T.roots_ecc);
        const uint32_t *sh = T.pntt_shift + c * E;
#pragma unroll
        for (int j = 0; j < E; ++j)
            acc[j] += O::mul_residue(O::mod_p(x[j]), sh[j]);
    }
    for (uint32_t s = 1; s < L; s <<= 1)
#pragma unroll
        for (int j = 0; j < E; ++j)
            acc[j] += (uint32_t)__shfl_xor((int)acc[j], (int)s);
#pragma unroll
    for (int j = 0; j < E; ++j)
        acc[j] = O::mul(O::mod_p(acc[j]), T.mix[j]);
    gs_reg<O, E>(acc, T.roots_i_ecc);
#pragma unroll
    for (int j = 0; j < E; ++j)
        if ((uint32_t)j % L == gl)
            o[j] = (uint16_t)(acc[j] & 0xffff);
}

// One codeword per lane (ecc_len <= 32; rows 16-byte aligned for ecc_len >= 8).
template <class O, int E>
__global__ __launch_bounds__(256) void rsi16_encode_lane_kernel(const uint16_t *__restrict__ msgs, uint32_t count,
                                                                uint16_t *__restrict__ out, Rsi16Dev T)
{
    const uint32_t cw = blockIdx.x * 256 + threadIdx.x;
    if (cw >= count)
        return;
    encode_lane<O, E, (E >= 8)>(T, msgs + (size_t)cw * T.k, out + (size_t)cw * E);
}

// ---------------------------------------------------------------- syndromes

// NTT::pntt (ntt.hpp:133-152) of the codeword msg || ecc into syn[0..E); buf is E scratch.
template <class TM = uint16_t, class TE = uint16_t>
__device__ inline void pntt_codeword(const Rsi16Dev &T, const TM *m, const TE *e, uint32_t *buf, uint32_t *syn, Grp G,
                                     size_t ms = 1, size_t es = 1)
{
    const uint32_t E = T.E;
    uint32_t nb = T.n / E;
    for (uint32_t c = 0; c < nb; ++c) {
        for (uint32_t j = G.gl; j < E; j += G.g) {
            uint32_t idx = c * E + j;
            buf[j] = idx < T.k ? (uint32_t)m[idx * ms] : (uint32_t)e[(idx - T.k) * es];
        }
        wave_sync();
        ct_full(buf, E, T.roots_ecc, G);
        for (uint32_t j = G.gl; j < E; j += G.g) {
            uint32_t v = S::mul(buf[j], T.pntt_shift[c * E + j]);
            syn[j] = c ? S::add(syn[j], v) : v;
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------- Sugiyama

// NTT::_reverse_ntt_vec (ntt.hpp:344-354) in gather form: the sequential swaps form an
// involution k(i) = rbo_ecc[E - rbo_ecc[i]], so v'[i] = v[k(i)] * roots_i_ecc[j] * roots_ecc[shift j].
__device__ inline void reverse_ntt(const Rsi16Dev &T, uint32_t *v, uint32_t shift, uint32_t *tmp, Grp G)
{
    const uint32_t E = T.E, mask = E - 1;
    for (uint32_t i = G.gl; i < E; i += G.g)
        tmp[i] = v[i];
    wave_sync();
    for (uint32_t i = G.gl; i < E; i += G.g) {
        if (i == 0)
            continue;
        uint32_t j = T.rbo_ecc[i];
        uint32_t k = T.rbo_ecc[E - j];
        uint32_t x = tmp[k];
        x = S::mul(x, T.roots_i_ecc[j]);
        x = S::mul(x, T.roots_ecc[(shift * j) & mask]);
        v[i] = x;
    }
    wave_sync();
}

// NTT::poly_inv_mod_xn_rev (ntt.hpp:363-469), scalar instantiation. Returns false where
// the reference's py_assert fires.
__device__ inline bool poly_inv_mod_xn_rev(const Rsi16Dev &T, uint32_t *g, uint32_t g_len, uint32_t g0,
                                           uint32_t a_len, uint32_t *a, Grp G)
{
    const uint32_t E = T.E;
    uint32_t a0 = g_inv(g0);
    for (uint32_t i = G.gl; i < E; i += G.g)
        a[i] = a0;
    wave_sync();
    uint32_t l = 2;
    uint32_t a_max = a_len;
    while (l < a_len) {
        if (g_len + l - 2 > E)
            return false;
        for (uint32_t i = G.gl; i < E; i += G.g) {
            uint32_t ai = a[i];
            uint32_t ag = S::mul(ai, g[i]);
            a[i] = S::mul(ai, S::sub(2, ag));
        }
        wave_sync();
        inttr(T, a, G);
        for (uint32_t i = G.gl; i < E; i += G.g)
            if (i >= l)
                a[i] = 0;
        wave_sync();
        nttr(T, a, G);
        l *= 2;
    }
    for (uint32_t i = G.gl; i < E; i += G.g) {
        uint32_t ag = S::mul(a[i], g[i]);
        g[i] = S::mul(a[i], S::sub(2, ag));
    }
    wave_sync();
    inttr(T, g, G);
    // a = g % x^l, then reverse a[0..a_max), then zero a[a_len..E)
    for (uint32_t i = G.gl; i < E; i += G.g) {
        uint32_t src = i < a_max ? a_max - 1 - i : i;
        uint32_t v = src < l ? g[src] : 0u;
        a[i] = i < a_len ? v : 0u;
    }
    wave_sync();
    nttr(T, a, G);
    return true;
}

// NTT::poly_div_ntt (ntt.hpp:471-564), scalar instantiation. q_len via *q_len_out;
// returns false where the reference raises.
__device__ inline bool poly_div_ntt(const Rsi16Dev &T, const uint32_t *f, uint32_t f_len, const uint32_t *g,
                                    uint32_t g_len, uint32_t f0, uint32_t g0, uint32_t *t, uint32_t *q,
                                    uint32_t *tmp, uint32_t *q_len_out, Grp G)
{
    const uint32_t E = T.E;
    uint32_t q_len = f_len - g_len + 1;
    if (g_len == 0)
        return false;
    if (f_len < g_len) {
        for (uint32_t i = G.gl; i < E; i += G.g)
            q[i] = 0;
        wave_sync();
        *q_len_out = 0;
        return true;
    }
    for (uint32_t i = G.gl; i < E; i += G.g)
        t[i] = g[i];
    wave_sync();
    reverse_ntt(T, t, g_len, tmp, G);
    if (!poly_inv_mod_xn_rev(T, t, g_len, g0, q_len, q, G))
        return false;
    for (uint32_t j = G.gl; j < E; j += G.g)
        t[j] = S::mul(S::sub(f[j], f0), q[j]);
    wave_sync();
    inttr(T, t, G);
    uint32_t right_n = q_len < E - g_len ? q_len : E - g_len;
    for (uint32_t i = G.gl; i < E; i += G.g) {
        uint32_t v;
        if (i < right_n)
            v = t[g_len + i];
        else if (i < q_len)
            v = t[i - right_n];
        else
            v = 0;
        q[i] = v;
    }
    wave_sync();
    nttr(T, q, G);
    *q_len_out = q_len;
    return true;
}

// poly_div_ntt's outcome (true unless it raises) from the lengths alone: g_len == 0 raises,
// f_len < g_len returns, and poly_inv_mod_xn_rev raises where its Newton loop meets
// g_len + l - 2 > E (ntt.hpp:363-564).
__device__ inline bool div_ok(uint32_t E, uint32_t f_len, uint32_t g_len)
{
    if (g_len == 0)
        return false;
    if (f_len < g_len)
        return true;
    const uint32_t q_len = f_len - g_len + 1;
    for (uint32_t l = 2; l < q_len; l *= 2)
        if (g_len + l - 2 > E)
            return false;
    return true;
}

// Scratch words (times E) of sugiyama.
constexpr uint32_t SUGI_WORDS = 5;

// RSi16v<1>::_sugiyama (rsi16v_impl.hpp:271-440). r1 holds the syndromes on entry and
// the normalised evaluator on exit; w holds SUGI_WORDS * E scratch words. Returns the
// locator length, or 0xFFFFFFFF where the reference raises.
//
// The reference keeps a1, a2 and r2 as E-point transforms and multiplies pointwise, which
// is multiplication modulo x^E - 1 of the coefficient vectors. Here every polynomial stays
// in coefficient form and the same products are formed as cyclic convolutions, so every
// value is the same field element. In the usual step the quotient has two terms
// (q_len == 2): poly_inv_mod_xn_rev then reduces to a = (-a0^2 g[g_len - 2],
// 2 a0 - a0^2 g0) with a0 = 1 / g0 (its Newton loop does not run), two coefficients of
// (f - f0) * a give the quotient, and the updates are two-tap convolutions. Any other step
// goes through the transforms exactly as the reference does. The last division, whose
// result the reference discards, only decides whether it raises (div_ok). Lengths are
// scanned below E.
__device__ inline uint32_t sugiyama(const Rsi16Dev &T, uint32_t *a1_out, uint32_t *r1_out, uint32_t *w, Grp G)
{
    const uint32_t E = T.E, mask = E - 1;
    uint32_t *a1 = a1_out, *r1 = r1_out, *a2 = w, *r2 = w + E;
    uint32_t *t = w + 2 * E, *q = w + 3 * E, *tmp = w + 4 * E;
    for (uint32_t i = G.gl; i < E; i += G.g) {
        a1[i] = 0;
        a2[i] = i == 0 ? 1u : 0u;  // the transform's all-ones vector
        r2[i] = r1[i];
    }
    wave_sync();
    uint32_t q_len = 0, a2_len = 1, a1_len;
    uint32_t r1_len = group_len(r1, E, G);
    uint32_t r2_len = r1_len;
    uint32_t r2l = r2[0];
    const uint32_t lanes = r1_len != 0 ? 0xffffffffu : 0u;
    {
        const uint32_t r1h = r1[(r1_len - 1) & mask];
        const uint32_t r1h2 = r1[(r1_len - 2) & mask];
        const uint32_t a11 = s_neg(T.inv_tab[r1h]);
        const uint32_t a10 = S::mul(r1h2, S::mul(a11, a11));
        wave_sync();
        if (G.gl == 0) {
            a1[0] = a10;
            a1[1] = a11;
        }
        a1_len = 2 & lanes;
        for (uint32_t i = G.gl; i < E; i += G.g) {
            uint32_t v = S::mul(a10, r2[i]);  // r2 still holds the original r1
            if (i >= 1)
                v = S::add(v, S::mul(a11, r2[i - 1]));
            tmp[i] = v;
        }
        wave_sync();
        r1_len = group_len(tmp, E - 1, G) & lanes;
        for (uint32_t i = G.gl; i < E; i += G.g)
            r1[i] = i < r1_len ? tmp[i] : 0u;
        wave_sync();
    }
    if (r1_len >= E / 2) {
        for (uint32_t it = 1; it < E / 2; ++it) {
            if (r1_len < E / 2) {  // the reference divides once more and stops
                if (!div_ok(E, r2_len, r1_len))
                    return 0xffffffffu;
                break;
            }
            const uint32_t r1h = r1[(r1_len - 1) & mask];
            const uint32_t r1l = r1[0];
            const uint32_t g_len = r1_len, f_len = r2_len;
            if (f_len == g_len + 1) {  // q_len == 2 (g_len >= E / 2 >= 1 here)
                q_len = 2;
                // the reference's quotient factors c0 = -a0^2 r1s and c1 = 2 a0 - a0^2 r1h = a0
                // (a0 r1h = 1 exactly), so q = c0 v + c1 v' = a0 (v' - w v) with w = a0 r1s
                const uint32_t a0 = T.inv_tab[r1h];
                const uint32_t w = S::mul(a0, r1[(g_len - 2) & mask]);
                // quotient coefficients t[idx] of t = (r2 - f0) * (c0 + c1 x) mod x^E - 1
                const uint32_t right_n = E - g_len < 2 ? E - g_len : 2u;
                const uint32_t idx0 = right_n >= 1 ? g_len : 0u;
                const uint32_t idx1 = right_n >= 2 ? g_len + 1 : 1u - right_n;
                const uint32_t at[4] = {idx0, (idx0 - 1) & mask, idx1, (idx1 - 1) & mask};
                uint32_t v[4];
                for (int u = 0; u < 4; ++u) {
                    const uint32_t x = r2[at[u]];
                    v[u] = at[u] == 0 ? S::sub(x, r2l) : x;
                }
                const uint32_t q0 = S::mul(a0, S::sub(v[1], S::mul(w, v[0])));
                const uint32_t q1 = S::mul(a0, S::sub(v[3], S::mul(w, v[2])));
                wave_sync();
                // a2 - q a1 and r2 - q r1 in place of a2 and r2
                for (uint32_t i = G.gl; i < E; i += G.g) {
                    const uint32_t j = (i - 1) & mask;
                    a2[i] = S::sub(a2[i], S::add(S::mul(q0, a1[i]), S::mul(q1, a1[j])));
                    r2[i] = S::sub(r2[i], S::add(S::mul(q0, r1[i]), S::mul(q1, r1[j])));
                }
                wave_sync();
            } else {  // through the transforms; a1 and r1 are transformed back afterwards
                nttr(T, r1, G);
                nttr(T, a1, G);
                nttr(T, a2, G);
                nttr(T, r2, G);
                if (!poly_div_ntt(T, r2, r2_len, r1, r1_len, r2l, r1h, t, q, tmp, &q_len, G))
                    return 0xffffffffu;
                for (uint32_t i = G.gl; i < E; i += G.g) {
                    const uint32_t qi = q[i];
                    a2[i] = S::sub(a2[i], S::mul(qi, a1[i]));
                    r2[i] = S::sub(r2[i], S::mul(qi, r1[i]));
                }
                wave_sync();
                inttr(T, a2, G);
                inttr(T, r2, G);
                inttr(T, a1, G);
                inttr(T, r1, G);
            }
            uint32_t t_len = q_len + a1_len - 1;
            t_len = a2_len > t_len ? a2_len : t_len;
            uint32_t nq_len = q_len + r1_len - 1;
            nq_len = r2_len > nq_len ? r2_len : nq_len;
            uint32_t *sw = a1;  // a2 <- a1, a1 <- the new remainder's cofactor
            a1 = a2;
            a2 = sw;
            sw = r1;
            r1 = r2;
            r2 = sw;
            a2_len = a1_len;
            r2_len = r1_len;
            r2l = r1l;
            a1_len = t_len;
            r1_len = group_len(r1, nq_len < E ? nq_len : E, G);
        }
    }
    const uint32_t a1_max = a1_len & lanes;
    a1_len = group_len(a1, a1_max < E ? a1_max : E, G);
    const uint32_t inv0 = T.inv_tab[a1[0]];
    wave_sync();
    for (uint32_t i = G.gl; i < E; i += G.g) {
        a1_out[i] = S::mul(a1[i], inv0);
        r1_out[i] = S::mul(r1[i], inv0);
    }
    wave_sync();
    return a1_len & lanes;
}

// RSi16v::_eval (rsi16v_impl.hpp:628-636), Horner over r[len-1 .. 0]
__device__ inline uint32_t eval_poly(const uint32_t *r, uint32_t len, uint32_t x)
{
    uint32_t s = 0;
    for (uint32_t i = len; i-- > 0;)
        s = S::add(S::mul(s, x), r[i]);
    return s;
}

// ---------------------------------------------------------------- repair

// Per group LDS: syn/evaluator E, locator E, derivative E, roots chunk E, Sugiyama scratch.
constexpr uint32_t REPAIR_WORDS = 4 + SUGI_WORDS;

// Repair of one codeword by one lane group (rsi16v_impl.hpp:122-168); lane 0 of the
// group writes the status to *st. base holds REPAIR_WORDS * E words.
__device__ inline void repair_group(const Rsi16Dev &T, uint16_t *m, uint16_t *e, uint32_t *base, Grp G, uint8_t *st)
{
    const uint32_t E = T.E;
    uint32_t *syn = base, *loc = base + E, *der = base + 2 * E, *chunk = base + 3 * E, *w = base + 4 * E;

    pntt_codeword(T, m, e, chunk, syn, G);
    if (group_len(syn, E, G) == 0) {
        if (G.gl == 0)
            *st = FFRS_RSI16_NO_ERRORS;
        return;
    }
    uint32_t L = sugiyama(T, loc, syn, w, G);
    if (L == 0xffffffffu) {
        if (G.gl == 0)
            *st = FFRS_RSI16_RAISES;
        return;
    }
    uint32_t eval_len = E - L + 1;
    // _deriv (rsi16v_impl.hpp:608-616)
    for (uint32_t i = G.gl; i < E; i += G.g)
        der[i] = i + 1 < E ? S::mul(loc[i + 1], i + 1) : 0u;
    wave_sync();
    uint32_t der_len = group_len(der, E, G);
    if (L > E) {  // py_assert at rsi16v_impl.hpp:454
        if (G.gl == 0)
            *st = FFRS_RSI16_RAISES;
        return;
    }
    uint32_t lanes = L != 0 ? 0xffffffffu : 0u;
    // Horner from the highest non-zero coefficient below emax: the zeros above it add nothing
    uint32_t dmax = der_len & lanes, emax = group_len(syn, (eval_len & lanes) < E ? (eval_len & lanes) : E, G);
    uint32_t nb = T.n / E;
    // NTT::pintt_ecc (ntt.hpp:206-220), one chunk of E positions at a time. The zeros are
    // listed first (the Sugiyama scratch is free now) and corrected afterwards with one
    // root per lane: corrections touch only the block, never the values the scan reads,
    // so the order does not change the result. A non-zero locator of degree < E has
    // fewer than E roots.
    uint32_t *list = w, *nroots = w + E;
    if (G.gl == 0)
        *nroots = 0;
    wave_sync();
    for (uint32_t c = 0; c < nb; ++c) {
        for (uint32_t j = G.gl; j < E; j += G.g)
            chunk[j] = S::mul(loc[j], T.pintt_shift[c * E + j]);
        wave_sync();
        gs_full<S>(chunk, E, T.roots_i_ecc, G);
        if (lanes) {
            for (uint32_t j = G.gl; j < E; j += G.g)
                if (chunk[j] == 0) {
                    uint32_t slot = atomicAdd(nroots, 1u);
                    if (slot < E)
                        list[slot] = c * E + j;
                }
        }
        wave_sync();
    }
    const uint32_t total = *nroots;
    const uint32_t listed = total < E ? total : E;
    for (uint32_t r = G.gl; r < listed; r += G.g) {
        uint32_t pos = list[r];
        uint32_t ri = rbo_bits(pos, T.nbits);
        uint32_t x = T.roots_ntt[ri];
        uint32_t xi = T.roots_ntt[(T.N - ri) & (T.N - 1)];
        uint32_t num = eval_poly(syn, emax, xi);
        uint32_t den = eval_poly(der, dmax, xi);
        uint32_t err = S::mul(g_div(num, den), x);  // _forney (rsi16v_impl.hpp:500-512)
        uint16_t *at = pos < T.k ? &m[pos] : &e[pos - T.k];
        *at = (uint16_t)(S::add((uint32_t)*at, err) & 0xffff);
    }
    if (G.gl == 0)
        *st = total > 0 ? FFRS_RSI16_REPAIR_OK : (lanes ? FFRS_RSI16_ERROR_LOCATION_FAIL : FFRS_RSI16_NO_ERRORS);
}

// ---- repair with the polynomials in registers (ecc_len <= 256)
//
// The same operations as repair_group in the same order. A group of G = min(E / 2, 64)
// lanes serves one codeword and each lane holds C = E / G coefficients of every polynomial:
// lane gl holds coefficients gl + h G, h < C. Butterfly partners at distance s < G sit in
// lane gl ^ s (same h) and are exchanged with one cross-lane move (DPP for s <= 8,
// ds_swizzle for 16, ds_bpermute for 32); partners at distance s >= G are two of the lane's
// own coefficients. Each lane's twiddles are loaded once per codeword. Permutations gather
// with ds_bpermute. Lanes of a group take the same branches (every decision depends on
// group-wide values only), and no move reads a lane outside the group (an aligned block of
// G lanes).

// Field operations for operands in [0, 65536] (every value the register repair handles:
// u16 symbols, table entries and results of these operations). The results equal
// ScalarOps': a product of two such operands is below 2^33, so its 16-bit halves come from
// the full-rate 24-bit multiplies, and min(t, t + p) replaces the sign test (for t < 2^31
// the sum is larger; for a wrapped t the sum is the reduced value).
struct F {
    __device__ static inline uint32_t add(uint32_t a, uint32_t b)
    {
        const uint32_t r = a + b, t = r - P16;
        return t < r ? t : r;
    }
    __device__ static inline uint32_t sub(uint32_t a, uint32_t b)
    {
        const uint32_t r = a - b, t = r + P16;
        return t < r ? t : r;
    }
    __device__ static inline uint32_t mul(uint32_t a, uint32_t b)
    {
        __builtin_assume(a <= 0x10000u);
        __builtin_assume(b <= 0x10000u);
        const uint64_t x = (uint64_t)a * (uint64_t)b;
        const uint32_t t = ((uint32_t)x & 0xffffu) - (uint32_t)(x >> 16), u = t + P16;
        return t < u ? t : u;
    }
    // a b (a <= p, b <= 65536) folded once: lo16 - hi of the product, in [-65537, 65535]
    __device__ static inline int32_t part(uint32_t a, uint32_t b)
    {
        __builtin_assume(a <= P16);
        __builtin_assume(b <= 0x10000u);
        const uint64_t x = (uint64_t)a * (uint64_t)b;
        return (int32_t)((uint32_t)x & 0xffffu) - (int32_t)(uint32_t)(x >> 16);
    }
    // c + r0 + r1 reduced (c <= 65536, r0 and r1 from part): the sum plus 2p lies in
    // [0, 327680], and one more fold leaves [-5, 65535]
    __device__ static inline uint32_t fold3(uint32_t c, int32_t r0, int32_t r1)
    {
        const uint32_t y = (uint32_t)((int32_t)c + r0 + r1 + 2 * (int32_t)P16);
        const uint32_t t = (y & 0xffffu) - (y >> 16), u = t + P16;
        return t < u ? t : u;
    }
};

// v of lane (this lane ^ X)
template <uint32_t X>
__device__ inline uint32_t xchg(uint32_t v, uint32_t lane)
{
    if constexpr (X == 1)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm 1,0,3,2
    else if constexpr (X == 2)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm 2,3,0,1
    else if constexpr (X == 4) {
        // row_half_mirror reads lane ^ 7, then quad_perm 3,2,1,0 reads lane ^ 3
        const int h = __builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true);
        return (uint32_t)__builtin_amdgcn_mov_dpp(h, 0x1B, 0xF, 0xF, true);
    } else if constexpr (X == 8)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true);  // row_ror 8
    else if constexpr (X == 16)
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);  // bit mode, xor 16
    else
        return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane ^ X) << 2), (int)v);
}

__device__ inline uint32_t lane_read(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(lane << 2), (int)v);
}

// xchg with s a constant after unrolling
__device__ __attribute__((always_inline)) inline uint32_t xchg_s(uint32_t v, uint32_t s, uint32_t lane)
{
    switch (s) {
    case 1:
        return xchg<1>(v, lane);
    case 2:
        return xchg<2>(v, lane);
    case 4:
        return xchg<4>(v, lane);
    case 8:
        return xchg<8>(v, lane);
    case 16:
        return xchg<16>(v, lane);
    default:
        return xchg<32>(v, lane);
    }
}

// Lanes per codeword of the register repair: each lane holds E / G coefficients of every
// polynomial (two up to E = 128; four coefficients per lane with E / 4 lanes measured
// slower there: 0.74 vs 0.65 ms for 262,144 RSi16(256,32) codewords).
__host__ __device__ constexpr uint32_t reg_group(uint32_t E)
{
    return E / 2 < 64 ? E / 2 : 64;
}

template <int E>
struct RegShape {
    static constexpr uint32_t G = reg_group(E);  // lanes per codeword
    static constexpr int C = E / (int)G;                    // coefficients per lane
    static constexpr int LE = __builtin_ctz(E);             // butterfly stages
    static constexpr int SX = __builtin_ctz(G);             // stages with partners in other lanes
    static constexpr int LC = __builtin_ctz(C);             // stages within the lane
};

template <int E>
struct PR {
    uint32_t v[RegShape<E>::C];  // coefficient gl + h G in v[h]
};

template <int E>
struct LaneR {
    using Sh = RegShape<E>;
    uint32_t gl, base, lane;  // lane in the group, first lane of the group, lane in the wave
    // Stage t (s = 2^t) of both butterfly orders uses roots[j << (LE - 1 - t)], j = i & (s - 1).
    // Stages with partners in other lanes (t < SX): j = gl & (s - 1), kept where this lane is
    // the upper partner and 1 elsewhere. Stages within the lane (s = G 2^u): the pair
    // (h, h + 2^u) with bit u of h clear uses entry 2^u - 1 + (h & (2^u - 1)).
    uint32_t x_e[Sh::SX > 0 ? Sh::SX : 1], x_i[Sh::SX > 0 ? Sh::SX : 1];
    uint32_t w_e[Sh::C - 1], w_i[Sh::C - 1];

    __device__ void init(const Rsi16Dev &T, uint32_t lane_in_wave)
    {
        lane = lane_in_wave;
        gl = lane % Sh::G;
        base = lane - gl;
#pragma unroll
        for (int t = 0; t < Sh::SX; ++t) {
            const uint32_t s = 1u << t, j = gl & (s - 1);
            const bool upper = (gl & s) != 0;
            const uint32_t idx = j << (Sh::LE - 1 - t);
            x_e[t] = upper ? T.roots_ecc[idx] : 1u;
            x_i[t] = upper ? T.roots_i_ecc[idx] : 1u;
        }
#pragma unroll
        for (int u = 0; u < Sh::LC; ++u)
#pragma unroll
            for (int hl = 0; hl < (1 << u); ++hl) {
                const uint32_t j = gl + (uint32_t)hl * Sh::G;
                const uint32_t idx = j << (Sh::LE - 1 - (Sh::SX + u));
                w_e[(1 << u) - 1 + hl] = T.roots_ecc[idx];
                w_i[(1 << u) - 1 + hl] = T.roots_i_ecc[idx];
            }
    }
};

// coefficient idx (the same for the whole group)
template <int E>
__device__ inline uint32_t pr_get(const PR<E> &p, uint32_t idx, const LaneR<E> &L)
{
    constexpr uint32_t G = RegShape<E>::G;
    const uint32_t slot = idx / G;
    // select by the bits of slot (a lookup into p.v would go through scratch memory)
    uint32_t w[RegShape<E>::C];
#pragma unroll
    for (int h = 0; h < RegShape<E>::C; ++h)
        w[h] = p.v[h];
#pragma unroll
    for (int b = 0; (1 << b) < RegShape<E>::C; ++b)
#pragma unroll
        for (int h = 0; h + (1 << b) < RegShape<E>::C; h += 2 << b)
            w[h] = ((slot >> b) & 1) ? w[h + (1 << b)] : w[h];
    return lane_read(w[0], L.base + idx % G);
}

// coefficient k of p, k < E chosen per lane
template <int E>
__device__ inline uint32_t pr_gather(const PR<E> &p, uint32_t k, const LaneR<E> &L)
{
    constexpr uint32_t G = RegShape<E>::G;
    const uint32_t src = L.base + k % G, slot = k / G;
    uint32_t w[RegShape<E>::C];
#pragma unroll
    for (int h = 0; h < RegShape<E>::C; ++h)
        w[h] = lane_read(p.v[h], src);
#pragma unroll
    for (int b = 0; (1 << b) < RegShape<E>::C; ++b)
#pragma unroll
        for (int h = 0; h + (1 << b) < RegShape<E>::C; h += 2 << b)
            w[h] = ((slot >> b) & 1) ? w[h + (1 << b)] : w[h];
    return w[0];
}

// p[(i - 1) mod E] for each of the lane's coefficients
template <int E>
__device__ inline PR<E> pr_prev(const PR<E> &p, const LaneR<E> &L)
{
    constexpr uint32_t G = RegShape<E>::G;
    constexpr int C = RegShape<E>::C;
    const uint32_t src = L.base + (L.gl + G - 1) % G;
    uint32_t r[C];
#pragma unroll
    for (int h = 0; h < C; ++h)
        r[h] = lane_read(p.v[h], src);
    PR<E> o;
#pragma unroll
    for (int h = 0; h < C; ++h)
        o.v[h] = L.gl ? r[h] : r[(h + C - 1) % C];
    return o;
}

// degree + 1 of p restricted to [0, max) (group_len)
template <int E>
__device__ inline uint32_t pr_len(const PR<E> &p, uint32_t max, const LaneR<E> &L)
{
    constexpr uint32_t G = RegShape<E>::G;
    uint32_t hi = 0;
#pragma unroll
    for (int h = 0; h < RegShape<E>::C; ++h) {
        const uint32_t i = L.gl + (uint32_t)h * G;
        if (i < max && p.v[h] != 0)
            hi = i + 1;
    }
#pragma unroll
    for (int t = 0; t < RegShape<E>::SX; ++t) {
        const uint32_t o = xchg_s(hi, 1u << t, L.lane);
        hi = o > hi ? o : hi;
    }
    return hi;
}

// ct_full (RBO in, natural out, full reduction), ntt.hpp:225-249
template <int E>
__device__ inline void pr_ct(PR<E> &x, const uint32_t *xw, const uint32_t *ww, const LaneR<E> &L)
{
    constexpr int C = RegShape<E>::C;
#pragma unroll
    for (int t = 0; t < RegShape<E>::SX; ++t) {  // s = 2^t < G
        // the upper lane sends c * w, the lower lane sends a (its w is 1)
        const bool upper = (L.gl & (1u << t)) != 0;
#pragma unroll
        for (int h = 0; h < C; ++h) {
            const uint32_t sa = F::mul(x.v[h], xw[t]);
            const uint32_t ra = xchg_s(sa, 1u << t, L.lane);
            x.v[h] = upper ? F::sub(ra, sa) : F::add(sa, ra);
        }
    }
#pragma unroll
    for (int u = 0; u < RegShape<E>::LC; ++u)  // s = G 2^u
#pragma unroll
        for (int h = 0; h < C; ++h)
            if (!(h & (1 << u))) {
                const uint32_t a = x.v[h];
                const uint32_t c = F::mul(x.v[h + (1 << u)], ww[(1 << u) - 1 + (h & ((1 << u) - 1))]);
                x.v[h] = F::add(a, c);
                x.v[h + (1 << u)] = F::sub(a, c);
            }
}

// gs_full (natural in, RBO out) with full reduction, ntt.hpp:280-302
template <int E>
__device__ inline void pr_gs(PR<E> &x, const uint32_t *xw, const uint32_t *ww, const LaneR<E> &L)
{
    constexpr int C = RegShape<E>::C;
#pragma unroll
    for (int u = RegShape<E>::LC - 1; u >= 0; --u)  // s = G 2^u
#pragma unroll
        for (int h = 0; h < C; ++h)
            if (!(h & (1 << u))) {
                const uint32_t a = x.v[h], c = x.v[h + (1 << u)];
                x.v[h] = F::add(a, c);
                x.v[h + (1 << u)] = F::mul(F::sub(a, c), ww[(1 << u) - 1 + (h & ((1 << u) - 1))]);
            }
#pragma unroll
    for (int t = RegShape<E>::SX - 1; t >= 0; --t) {  // s = 2^t < G
        const bool upper = (L.gl & (1u << t)) != 0;
#pragma unroll
        for (int h = 0; h < C; ++h) {
            const uint32_t r = xchg_s(x.v[h], 1u << t, L.lane);
            const uint32_t d = upper ? F::sub(r, x.v[h]) : F::add(x.v[h], r);
            x.v[h] = F::mul(d, xw[t]);
        }
    }
}

template <int E>
__device__ inline void pr_nttr(PR<E> &x, const LaneR<E> &L)
{
    pr_gs<E>(x, L.x_e, L.w_e, L);
}

template <int E>
__device__ inline void pr_inttr(const Rsi16Dev &T, PR<E> &x, const LaneR<E> &L)
{
    pr_ct<E>(x, L.x_i, L.w_i, L);
#pragma unroll
    for (int h = 0; h < RegShape<E>::C; ++h)
        x.v[h] = F::mul(x.v[h], T.ecc_inv);
}

// NTT::_reverse_ntt_vec (ntt.hpp:344-354) in gather form, as reverse_ntt
template <int E>
__device__ inline void pr_reverse_ntt(const Rsi16Dev &T, PR<E> &v, uint32_t shift, const LaneR<E> &L)
{
    PR<E> out;
#pragma unroll
    for (int h = 0; h < RegShape<E>::C; ++h) {
        const uint32_t i = L.gl + (uint32_t)h * RegShape<E>::G;
        const uint32_t j = T.rbo_ecc[i];
        const uint32_t k = T.rbo_ecc[(E - j) & (E - 1)];
        uint32_t x = pr_gather<E>(v, k, L);
        x = F::mul(x, T.roots_i_ecc[j]);
        x = F::mul(x, T.roots_ecc[(shift * j) & (E - 1)]);
        out.v[h] = i == 0 ? v.v[h] : x;
    }
    v = out;
}

// NTT::poly_inv_mod_xn_rev (ntt.hpp:363-469), as poly_inv_mod_xn_rev
template <int E>
__device__ inline bool pr_poly_inv(const Rsi16Dev &T, PR<E> &g, uint32_t g_len, uint32_t g0, uint32_t a_len, PR<E> &a,
                                   const LaneR<E> &L)
{
    constexpr int C = RegShape<E>::C;
    constexpr uint32_t G = RegShape<E>::G;
    const uint32_t a0 = T.inv_tab[g0];
#pragma unroll
    for (int h = 0; h < C; ++h)
        a.v[h] = a0;
    uint32_t l = 2;
    const uint32_t a_max = a_len;
    while (l < a_len) {
        if (g_len + l - 2 > (uint32_t)E)
            return false;
#pragma unroll
        for (int h = 0; h < C; ++h)
            a.v[h] = F::mul(a.v[h], F::sub(2, F::mul(a.v[h], g.v[h])));
        pr_inttr<E>(T, a, L);
#pragma unroll
        for (int h = 0; h < C; ++h)
            if (L.gl + (uint32_t)h * G >= l)
                a.v[h] = 0;
        pr_nttr<E>(a, L);
        l *= 2;
    }
#pragma unroll
    for (int h = 0; h < C; ++h)
        g.v[h] = F::mul(a.v[h], F::sub(2, F::mul(a.v[h], g.v[h])));
    pr_inttr<E>(T, g, L);
#pragma unroll
    for (int h = 0; h < C; ++h) {
        const uint32_t i = L.gl + (uint32_t)h * G;
        const uint32_t src = i < a_max ? a_max - 1 - i : i;
        const uint32_t v = pr_gather<E>(g, src & (E - 1), L);
        a.v[h] = (i < a_len && src < l) ? v : 0u;
    }
    pr_nttr<E>(a, L);
    return true;
}

// NTT::poly_div_ntt (ntt.hpp:471-564), as poly_div_ntt
template <int E>
__device__ inline bool pr_poly_div(const Rsi16Dev &T, const PR<E> &f, uint32_t f_len, const PR<E> &g, uint32_t g_len,
                                   uint32_t f0, uint32_t g0, PR<E> &t, PR<E> &q, uint32_t *q_len_out,
                                   const LaneR<E> &L)
{
    constexpr int C = RegShape<E>::C;
    constexpr uint32_t G = RegShape<E>::G;
    const uint32_t q_len = f_len - g_len + 1;
    if (g_len == 0)
        return false;
    if (f_len < g_len) {
#pragma unroll
        for (int h = 0; h < C; ++h)
            q.v[h] = 0;
        *q_len_out = 0;
        return true;
    }
    t = g;
    pr_reverse_ntt<E>(T, t, g_len, L);
    if (!pr_poly_inv<E>(T, t, g_len, g0, q_len, q, L))
        return false;
#pragma unroll
    for (int h = 0; h < C; ++h)
        t.v[h] = F::mul(F::sub(f.v[h], f0), q.v[h]);
    pr_inttr<E>(T, t, L);
    const uint32_t right_n = q_len < E - g_len ? q_len : E - g_len;
#pragma unroll
    for (int h = 0; h < C; ++h) {
        const uint32_t i = L.gl + (uint32_t)h * G;
        const uint32_t src = i < right_n ? g_len + i : (i < q_len ? i - right_n : 0u);
        const uint32_t v = pr_gather<E>(t, src & (E - 1), L);
        q.v[h] = i < q_len ? v : 0u;
    }
    pr_nttr<E>(q, L);
    *q_len_out = q_len;
    return true;
}

// RSi16v<1>::_sugiyama (rsi16v_impl.hpp:271-440) in coefficient form, as sugiyama: r1 holds
// the syndromes on entry and the normalised evaluator on exit. Returns the locator length,
// or 0xFFFFFFFF where the reference raises.
template <int E>
__device__ inline uint32_t pr_sugiyama(const Rsi16Dev &T, PR<E> &a1, PR<E> &r1, const LaneR<E> &L)
{
    constexpr uint32_t G = RegShape<E>::G, mask = E - 1;
    constexpr int C = RegShape<E>::C;
    PR<E> a2, r2 = r1;  // a2: the transform's all-ones vector
#pragma unroll
    for (int h = 0; h < C; ++h) {
        a2.v[h] = (h == 0 && L.gl == 0) ? 1u : 0u;
        a1.v[h] = 0;
    }
    uint32_t q_len = 0, a2_len = 1, a1_len;
    uint32_t r1_len = pr_len<E>(r1, E, L);
    uint32_t r2_len = r1_len;
    uint32_t r2l = pr_get<E>(r2, 0, L);
    const uint32_t lanes = r1_len != 0 ? 0xffffffffu : 0u;
    // r2[r2_len - 1] and r2[r2_len - 2], carried from the step that read them while r2 was r1
    uint32_t h1, h2;
    bool carried = true;
    {
        const uint32_t r1h = pr_get<E>(r1, (r1_len - 1) & mask, L);
        const uint32_t r1h2 = pr_get<E>(r1, (r1_len - 2) & mask, L);
        h1 = r1h;  // r2 is this polynomial from here on
        h2 = r1h2;
        const uint32_t a11 = s_neg(T.inv_tab[r1h]);
        const uint32_t a10 = F::mul(r1h2, F::mul(a11, a11));
        if constexpr (G > 1) {  // coefficients 0 and 1 live in lanes 0 and 1
            if (L.gl == 0)
                a1.v[0] = a10;
            else if (L.gl == 1)
                a1.v[0] = a11;
        } else {  // G == 1: they are the lane's two coefficients
            a1.v[0] = a10;
            a1.v[1] = a11;
        }
        a1_len = 2 & lanes;
        // tmp[i] = a10 r2[i] + a11 r2[i - 1]  (r2 still holds the original r1)
        const PR<E> prev = pr_prev<E>(r2, L);
        PR<E> tmp;
#pragma unroll
        for (int h = 0; h < C; ++h) {
            tmp.v[h] = F::mul(a10, r2.v[h]);
            if (h > 0 || L.gl >= 1)
                tmp.v[h] = F::add(tmp.v[h], F::mul(a11, prev.v[h]));
        }
        r1_len = pr_len<E>(tmp, E - 1, L) & lanes;
#pragma unroll
        for (int h = 0; h < C; ++h)
            r1.v[h] = L.gl + (uint32_t)h * G < r1_len ? tmp.v[h] : 0u;
    }
    if (r1_len >= E / 2) {
        for (uint32_t it = 1; it < E / 2; ++it) {
            if (r1_len < E / 2) {  // the reference divides once more and stops
                if (!div_ok(E, r2_len, r1_len))
                    return 0xffffffffu;
                break;
            }
            const uint32_t r1h = pr_get<E>(r1, (r1_len - 1) & mask, L);
            const uint32_t r1l = pr_get<E>(r1, 0, L);
            const uint32_t g_len = r1_len, f_len = r2_len;
            PR<E> t, rn;
            if (f_len == g_len + 1) {  // q_len == 2 (g_len >= E / 2 >= 1 here)
                q_len = 2;
                // the reference's quotient factors c0 = -a0^2 r1s and c1 = 2 a0 - a0^2 r1h = a0
                // (a0 r1h = 1 exactly), so q = c0 v + c1 v' = a0 (v' - w v) with w = a0 r1s
                const uint32_t a0 = T.inv_tab[r1h];
                const uint32_t r1s = pr_get<E>(r1, (g_len - 2) & mask, L);
                const uint32_t w = F::mul(a0, r1s);
                // quotient coefficients t[idx] of t = (r2 - f0) * (c0 + c1 x) mod x^E - 1
                const uint32_t right_n = (uint32_t)E - g_len < 2 ? (uint32_t)E - g_len : 2u;
                const uint32_t idx0 = right_n >= 1 ? g_len : 0u;
                const uint32_t idx1 = right_n >= 2 ? g_len + 1 : 1u - right_n;
                uint32_t v[4];
                const uint32_t at[4] = {idx0, (idx0 - 1) & mask, idx1, (idx1 - 1) & mask};
                // the two leading coefficients of r2 are known when idx0 = r2_len - 1 (then
                // idx0 >= E / 2 >= 2, so neither is coefficient 0)
                const bool known = carried && right_n >= 1;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    uint32_t x;
                    if (u < 2 && known)
                        x = u == 0 ? h1 : h2;
                    else if (u == 3 && right_n >= 2)
                        x = known ? h1 : pr_get<E>(r2, at[u], L);  // idx1 - 1 = idx0
                    else
                        x = pr_get<E>(r2, at[u], L);
                    v[u] = at[u] == 0 ? F::sub(x, r2l) : x;
                }
                h1 = r1h;
                h2 = r1s;
                carried = true;
                const uint32_t q0 = F::mul(a0, F::sub(v[1], F::mul(w, v[0])));
                const uint32_t q1 = F::mul(a0, F::sub(v[3], F::mul(w, v[2])));
                const PR<E> pa = pr_prev<E>(a1, L), pr = pr_prev<E>(r1, L);
                const uint32_t n0 = P16 - q0, n1 = P16 - q1;  // a2 - q0 a1 - q1 pa = a2 + n0 a1 + n1 pa
#pragma unroll
                for (int h = 0; h < C; ++h) {
                    t.v[h] = F::fold3(a2.v[h], F::part(n0, a1.v[h]), F::part(n1, pa.v[h]));
                    rn.v[h] = F::fold3(r2.v[h], F::part(n0, r1.v[h]), F::part(n1, pr.v[h]));
                }
            } else {  // through the transforms
                h1 = r1h;
                carried = false;  // r1[g_len - 2] is not read on this path
                PR<E> R1 = r1, R2 = r2, A1 = a1, A2 = a2, scratch, Q;
                pr_nttr<E>(R1, L);
                pr_nttr<E>(R2, L);
                pr_nttr<E>(A1, L);
                pr_nttr<E>(A2, L);
                if (!pr_poly_div<E>(T, R2, r2_len, R1, r1_len, r2l, r1h, scratch, Q, &q_len, L))
                    return 0xffffffffu;
#pragma unroll
                for (int h = 0; h < C; ++h) {
                    t.v[h] = F::sub(A2.v[h], F::mul(Q.v[h], A1.v[h]));
                    rn.v[h] = F::sub(R2.v[h], F::mul(Q.v[h], R1.v[h]));
                }
                pr_inttr<E>(T, t, L);
                pr_inttr<E>(T, rn, L);
            }
            uint32_t t_len = q_len + a1_len - 1;
            t_len = a2_len > t_len ? a2_len : t_len;
            uint32_t nq_len = q_len + r1_len - 1;
            nq_len = r2_len > nq_len ? r2_len : nq_len;
            a2 = a1;
            r2 = r1;
            a2_len = a1_len;
            r2_len = r1_len;
            r2l = r1l;
            a1 = t;
            r1 = rn;
            a1_len = t_len;
            r1_len = pr_len<E>(rn, nq_len, L);
        }
    }
    a1_len = pr_len<E>(a1, a1_len & lanes, L);
    const uint32_t inv0 = T.inv_tab[pr_get<E>(a1, 0, L)];
#pragma unroll
    for (int h = 0; h < C; ++h) {
        a1.v[h] = F::mul(a1.v[h], inv0);
        r1.v[h] = F::mul(r1.v[h], inv0);
    }
    return a1_len & lanes;
}

// repair_group with the polynomials in registers; base holds REG_WORDS * E words (evaluator
// and derivative copies for Forney, the root list and its count).
constexpr uint32_t REG_WORDS = 4;

template <int E, bool SIN = false>
__device__ inline void repair_group_reg(const Rsi16Dev &T, uint16_t *m, uint16_t *e, uint32_t *base, const LaneR<E> &L,
                                        uint8_t *st, const uint32_t *syn_in = nullptr)
{
    constexpr uint32_t G = RegShape<E>::G;
    constexpr int C = RegShape<E>::C;
    // NTT::pntt (ntt.hpp:133-152) of m || e, one chunk of E symbols at a time; the symbols
    // and shifts of PF chunks are loaded before any is transformed, so their latencies overlap
    constexpr uint32_t PF = C >= 4 ? 1 : 4;
    PR<E> syn;
    const uint32_t nb = T.n / E;
    if constexpr (SIN) {  // syndromes computed ahead (rsi16_synd_mfma_kernel)
#pragma unroll
        for (int h = 0; h < C; ++h)
            syn.v[h] = syn_in[L.gl + (uint32_t)h * G];
    }
    for (uint32_t c0 = 0; !SIN && c0 < nb; c0 += PF) {
        PR<E> x[PF], sh[PF];
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u)
            if (c0 + u < nb) {
#pragma unroll
                for (int h = 0; h < C; ++h) {
                    const uint32_t i = (c0 + u) * E + L.gl + (uint32_t)h * G;
                    x[u].v[h] = i < T.k ? (uint32_t)m[i] : (uint32_t)e[i - T.k];
                    sh[u].v[h] = T.pntt_shift[i];
                }
            }
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u)
            if (c0 + u < nb) {
                pr_ct<E>(x[u], L.x_e, L.w_e, L);
#pragma unroll
                for (int h = 0; h < C; ++h) {
                    const uint32_t v = F::mul(x[u].v[h], sh[u].v[h]);
                    syn.v[h] = c0 + u ? F::add(syn.v[h], v) : v;
                }
            }
    }
    if (pr_len<E>(syn, E, L) == 0) {
        if (L.gl == 0)
            *st = FFRS_RSI16_NO_ERRORS;
        return;
    }
    PR<E> loc;
    const uint32_t Ln = pr_sugiyama<E>(T, loc, syn, L);
    if (Ln == 0xffffffffu) {
        if (L.gl == 0)
            *st = FFRS_RSI16_RAISES;
        return;
    }
    const uint32_t eval_len = E - Ln + 1;
    // _deriv (rsi16v_impl.hpp:608-616): der[i] = (i + 1) loc[i + 1]
    PR<E> der;
#pragma unroll
    for (int h = 0; h < C; ++h) {
        const uint32_t i = L.gl + (uint32_t)h * G;
        const uint32_t nx = pr_gather<E>(loc, (i + 1) & (E - 1), L);
        der.v[h] = i + 1 < (uint32_t)E ? F::mul(nx, i + 1) : 0u;
    }
    const uint32_t der_len = pr_len<E>(der, E, L);
    if (Ln > (uint32_t)E) {  // py_assert at rsi16v_impl.hpp:454
        if (L.gl == 0)
            *st = FFRS_RSI16_RAISES;
        return;
    }
    const uint32_t lanes = Ln != 0 ? 0xffffffffu : 0u;
    // Horner from the highest non-zero coefficient below emax: the zeros above it add nothing
    const uint32_t dmax = der_len & lanes, emax = pr_len<E>(syn, eval_len & lanes, L);
    uint32_t *ev = base, *dv = base + E, *list = base + 2 * E, *nroots = base + 3 * E;
#pragma unroll
    for (int h = 0; h < C; ++h) {
        ev[L.gl + (uint32_t)h * G] = syn.v[h];
        dv[L.gl + (uint32_t)h * G] = der.v[h];
    }
    if (L.gl == 0)
        *nroots = 0;
    wave_sync();
    // NTT::pintt_ecc (ntt.hpp:206-220) chunk by chunk, shifts of PF chunks loaded ahead;
    // zeros listed, corrected afterwards
    for (uint32_t c0 = 0; c0 < nb; c0 += PF) {
        PR<E> sh[PF];
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u)
            if (c0 + u < nb) {
#pragma unroll
                for (int h = 0; h < C; ++h)
                    sh[u].v[h] = T.pintt_shift[(c0 + u) * E + L.gl + (uint32_t)h * G];
            }
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u)
            if (c0 + u < nb) {
                const uint32_t c = c0 + u;
                PR<E> ch;
#pragma unroll
                for (int h = 0; h < C; ++h)
                    ch.v[h] = F::mul(loc.v[h], sh[u].v[h]);
                pr_gs<E>(ch, L.x_i, L.w_i, L);
                if (lanes) {
#pragma unroll
                    for (int h = 0; h < C; ++h)
                        if (ch.v[h] == 0) {
                            const uint32_t slot = atomicAdd(nroots, 1u);
                            if (slot < (uint32_t)E)
                                list[slot] = c * E + L.gl + (uint32_t)h * G;
                        }
                }
            }
    }
    wave_sync();
    const uint32_t total = *nroots;
    const uint32_t listed = total < (uint32_t)E ? total : (uint32_t)E;
    for (uint32_t r = L.gl; r < listed; r += G) {
        const uint32_t pos = list[r];
        const uint32_t ri = rbo_bits(pos, T.nbits);
        const uint32_t x = T.roots_ntt[ri];
        const uint32_t xi = T.roots_ntt[(T.N - ri) & (T.N - 1)];
        const uint32_t num = eval_poly(ev, emax, xi);
        const uint32_t den = eval_poly(dv, dmax, xi);
        // _forney (rsi16v_impl.hpp:500-512) with the table conventions of g_div
        const uint32_t q = num == 0 ? 0u : (den == 0 ? num : F::mul(num, T.inv_tab[den]));
        const uint32_t err = F::mul(q, x);
        uint16_t *at = pos < T.k ? &m[pos] : &e[pos - T.k];
        *at = (uint16_t)(F::add((uint32_t)*at, err) & 0xffff);
    }
    if (L.gl == 0)
        *st = total > 0 ? FFRS_RSI16_REPAIR_OK : (lanes ? FFRS_RSI16_ERROR_LOCATION_FAIL : FFRS_RSI16_NO_ERRORS);
}

// encode_group with the values in registers (lane layout of LaneR): the residue butterflies
// of ct_residue (only lanes with a non-zero twiddle index multiply, so every lazy value is
// the same integer as there), the shifted sums, the mix and gs_full, all for one codeword
// per lane group.
template <class O, int E>
__device__ inline void encode_group_reg(const Rsi16Dev &T, const uint16_t *__restrict__ m, uint16_t *__restrict__ o,
                                        const LaneR<E> &L)
{
    constexpr uint32_t G = RegShape<E>::G;
    constexpr int C = RegShape<E>::C;
    constexpr uint32_t PF = C >= 4 ? 1 : 4;
    const uint32_t nb = T.n / E;
    PR<E> acc;
    for (uint32_t c0 = 0; c0 + 1 < nb; c0 += PF) {
        PR<E> x[PF], sh[PF];
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u)
            if (c0 + u + 1 < nb) {
#pragma unroll
                for (int h = 0; h < C; ++h) {
                    const uint32_t i = (c0 + u) * E + L.gl + (uint32_t)h * G;
                    x[u].v[h] = m[i];
                    sh[u].v[h] = T.pntt_shift[i];
                }
            }
#pragma unroll
        for (uint32_t u = 0; u < PF; ++u) {
            if (c0 + u + 1 >= nb)
                continue;
            PR<E> &y = x[u];
            // ct_residue (ntt.hpp:252-275): partners in other lanes
#pragma unroll
            for (int t = 0; t < RegShape<E>::SX; ++t) {
                const uint32_t sbit = 1u << t;
                const bool upper = (L.gl & sbit) != 0;
                const bool mult = upper && (L.gl & (sbit - 1)) != 0;
#pragma unroll
                for (int h = 0; h < C; ++h) {
                    const uint32_t sa = mult ? O::mul_residue(O::mod_p(y.v[h]), L.x_e[t]) : y.v[h];
                    const uint32_t ra = xchg_s(sa, sbit, L.lane);
                    y.v[h] = upper ? ra - sa : sa + ra;
                }
            }
            // partners within the lane
#pragma unroll
            for (int uu = 0; uu < RegShape<E>::LC; ++uu)
#pragma unroll
                for (int h = 0; h < C; ++h)
                    if (!(h & (1 << uu))) {
                        const int hl = h & ((1 << uu) - 1);
                        const uint32_t a = y.v[h], cc = y.v[h + (1 << uu)];
                        const bool mult = L.gl != 0 || hl != 0;
                        const uint32_t mm = mult ? O::mul_residue(O::mod_p(cc), L.w_e[(1 << uu) - 1 + hl]) : cc;
                        y.v[h] = a + mm;
                        y.v[h + (1 << uu)] = a - mm;
                    }
#pragma unroll
            for (int h = 0; h < C; ++h) {
                const uint32_t v = O::mul_residue(O::mod_p(y.v[h]), sh[u].v[h]);
                acc.v[h] = c0 + u ? acc.v[h] + v : v;
            }
        }
    }
#pragma unroll
    for (int h = 0; h < C; ++h)
        acc.v[h] = O::mul(O::mod_p(acc.v[h]), T.mix[L.gl + (uint32_t)h * G]);
    // gs_full on canonical values: O's add, sub and mul equal F's there
    pr_gs<E>(acc, L.x_i, L.w_i, L);
#pragma unroll
    for (int h = 0; h < C; ++h)
        o[L.gl + (uint32_t)h * G] = (uint16_t)(acc.v[h] & 0xffff);
}

template <class O, int E>
__global__ __launch_bounds__(64) void rsi16_encode_reg_kernel(const uint16_t *__restrict__ msgs, uint32_t count,
                                                               uint16_t *__restrict__ out, Rsi16Dev T)
{
    constexpr uint32_t G = RegShape<E>::G, cpw = 64 / G;
    const uint32_t grp = threadIdx.x / G;
    const uint32_t cw = blockIdx.x * cpw + grp;
    if (cw >= count)
        return;
    LaneR<E> L;
    L.init(T, threadIdx.x);
    encode_group_reg<O, E>(T, msgs + (size_t)cw * T.k, out + (size_t)cw * E, L);
}

template <int E, bool SIN = false>
__global__ __launch_bounds__(64) void rsi16_repair_reg_kernel(uint16_t *__restrict__ msgs,
                                                               uint16_t *__restrict__ eccs, uint32_t count,
                                                               uint8_t *__restrict__ status, Rsi16Dev T,
                                                               const uint32_t *__restrict__ syn = nullptr)
{
    extern __shared__ uint32_t s_mem[];
    constexpr uint32_t G = RegShape<E>::G, cpw = 64 / G;
    const uint32_t grp = threadIdx.x / G;
    const uint32_t cw = blockIdx.x * cpw + grp;
    if (cw >= count)
        return;
    LaneR<E> L;
    L.init(T, threadIdx.x);
    repair_group_reg<E, SIN>(T, msgs + (size_t)cw * T.k, eccs + (size_t)cw * E, s_mem + grp * REG_WORDS * E, L,
                             status + cw, SIN ? syn + (size_t)cw * E : nullptr);
}

__global__ __launch_bounds__(64) void rsi16_repair_kernel(uint16_t *__restrict__ msgs, uint16_t *__restrict__ eccs,
                                                           uint32_t count, uint8_t *__restrict__ status, Rsi16Dev T,
                                                           uint32_t g)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E;
    Grp G{threadIdx.x % g, g};
    uint32_t grp = threadIdx.x / g;
    uint32_t cpw = 64 / g;
    uint32_t cw = blockIdx.x * cpw + grp;
    if (cw >= count)
        return;
    repair_group(T, msgs + (size_t)cw * T.k, eccs + (size_t)cw * E, s_mem + grp * REPAIR_WORDS * E, G, status + cw);
}

// Syndromes only (py_synd).
__global__ __launch_bounds__(64) void rsi16_synds_kernel(const uint16_t *__restrict__ msgs,
                                                          const uint16_t *__restrict__ eccs, uint32_t count,
                                                          uint32_t *__restrict__ out, Rsi16Dev T, uint32_t g)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E;
    Grp G{threadIdx.x % g, g};
    uint32_t grp = threadIdx.x / g;
    uint32_t cw = blockIdx.x * (64 / g) + grp;
    if (cw >= count)
        return;
    uint32_t *syn = s_mem + grp * 2 * E, *chunk = syn + E;
    pntt_codeword(T, msgs + (size_t)cw * T.k, eccs + (size_t)cw * E, chunk, syn, G);
    for (uint32_t j = G.gl; j < E; j += G.g)
        out[(size_t)cw * E + j] = syn[j];
}

// Sugiyama on given syndromes (py_sugiyama<1>).
__global__ __launch_bounds__(64) void rsi16_sugiyama_kernel(const uint32_t *__restrict__ synds, uint32_t count,
                                                             uint32_t *__restrict__ loc_out,
                                                             uint32_t *__restrict__ ev_out,
                                                             uint32_t *__restrict__ len_out, Rsi16Dev T, uint32_t g)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E;
    Grp G{threadIdx.x % g, g};
    uint32_t grp = threadIdx.x / g;
    uint32_t cw = blockIdx.x * (64 / g) + grp;
    if (cw >= count)
        return;
    uint32_t *r1 = s_mem + grp * (2 + SUGI_WORDS) * E, *a1 = r1 + E, *w = r1 + 2 * E;
    for (uint32_t j = G.gl; j < E; j += G.g)
        r1[j] = synds[(size_t)cw * E + j];
    wave_sync();
    uint32_t L = sugiyama(T, a1, r1, w, G);
    for (uint32_t j = G.gl; j < E; j += G.g) {
        loc_out[(size_t)cw * E + j] = a1[j];
        ev_out[(size_t)cw * E + j] = r1[j];
    }
    if (G.gl == 0)
        len_out[cw] = L;
}

// py_roots: zero positions of the partial inverse NTT of E coefficients.
__global__ __launch_bounds__(64) void rsi16_roots_kernel(const uint32_t *__restrict__ poly, uint8_t *__restrict__ hit,
                                                          Rsi16Dev T, uint32_t g)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E;
    if (threadIdx.x >= g)
        return;
    Grp G{threadIdx.x, g};
    uint32_t *chunk = s_mem;
    for (uint32_t c = 0; c < T.n / E; ++c) {
        for (uint32_t j = G.gl; j < E; j += G.g)
            chunk[j] = S::mul(poly[j], T.pintt_shift[c * E + j]);
        wave_sync();
        gs_full<S>(chunk, E, T.roots_i_ecc, G);
        for (uint32_t j = G.gl; j < E; j += G.g)
            hit[c * E + j] = chunk[j] == 0;
        wave_sync();
    }
}

// _repair with error positions (rsi16v_impl.hpp:171-231) past its all-zero syndrome check: from
// the syndromes in ev[0..E), apply(r, pos[r], e_r) for every position with its error value e_r.
// LDS words: ev 2E, loc E, tmp 2E, der E.
template <class F>
__device__ inline void known_errors(const Rsi16Dev &T, const uint32_t *pos, uint32_t npos, uint32_t *ev,
                                    uint32_t *loc, uint32_t *tmp, uint32_t *der, Grp G, F apply)
{
    const uint32_t E = T.E;
    // _error_locator (rsi16v_impl.hpp:532-542): coefficients c_1..c_npos of prod (1 - x_j z)
    if (G.gl == 0) {
        for (uint32_t i = 0; i < E; ++i)
            loc[i] = 0;
        for (uint32_t j = 0; j < npos; ++j) {
            uint32_t x = T.roots_ntt[rbo_bits(pos[j], T.nbits)];
            for (uint32_t i = npos - 1; i > 0; --i)
                loc[i] = S::sub(loc[i], S::mul(loc[i - 1], x));
            loc[0] = S::sub(loc[0], x);
        }
    }
    wave_sync();
    // _error_evaluator (rsi16v_impl.hpp:546-563): Omega = S(z) * (1 + sum c_i z^i) mod z^E via length-2E NTTs
    for (uint32_t i = G.gl; i < 2 * E; i += G.g) {
        if (i >= E)
            ev[i] = 0;
        tmp[i] = i == 0 ? 1u : (i <= npos ? loc[i - 1] : 0u);
    }
    wave_sync();
    gs_full<S>(ev, 2 * E, T.roots_ecc2, G);
    gs_full<S>(tmp, 2 * E, T.roots_ecc2, G);
    for (uint32_t i = G.gl; i < 2 * E; i += G.g)
        ev[i] = S::mul(ev[i], tmp[i]);
    wave_sync();
    ct_full(ev, 2 * E, T.roots_i_ecc2, G);
    for (uint32_t i = G.gl; i < 2 * E; i += G.g)
        ev[i] = i < E ? S::mul(ev[i], T.ecc2_inv) : 0u;
    // _deriv_shifted (rsi16v_impl.hpp:619-626)
    for (uint32_t i = G.gl; i < E; i += G.g)
        der[i] = S::mul(loc[i], i + 1);
    wave_sync();
    uint32_t dlen = poly_len(der, E);
    for (uint32_t r = G.gl; r < npos; r += G.g) {
        uint32_t p = pos[r];
        uint32_t ri = rbo_bits(p, T.nbits);
        uint32_t x = T.roots_ntt[ri];
        uint32_t xi = T.roots_ntt[(T.N - ri) & (T.N - 1)];
        uint32_t num = eval_poly(ev, E, xi);
        uint32_t den = eval_poly(der, dlen, xi);
        apply(r, p, S::mul(g_div(num, den), x));
    }
}

// Known-position repair, rsi16v_impl.hpp:171-231 (one codeword, one group).
__global__ __launch_bounds__(64) void rsi16_repair_known_kernel(uint16_t *__restrict__ m, uint16_t *__restrict__ e,
                                                                 const uint32_t *__restrict__ pos, uint32_t npos,
                                                                 uint8_t *__restrict__ status, Rsi16Dev T, uint32_t g)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E;
    if (threadIdx.x >= g)
        return;
    Grp G{threadIdx.x, g};
    uint32_t *ev = s_mem, *loc = s_mem + 2 * E, *tmp = s_mem + 3 * E, *der = s_mem + 5 * E, *chunk = s_mem + 6 * E;
    pntt_codeword(T, m, e, chunk, ev, G);
    bool any = false;
    for (uint32_t j = 0; j < E; ++j)
        any |= ev[j] != 0;
    if (!any) {
        if (G.gl == 0)
            *status = FFRS_RSI16_NO_ERRORS;
        return;
    }
    known_errors(T, pos, npos, ev, loc, tmp, der, G, [&](uint32_t, uint32_t p, uint32_t err) {
        uint16_t *at = p < T.k ? &m[p] : &e[p - T.k];
        *at = (uint16_t)(S::add((uint32_t)*at, err) & 0xffff);
    });
    if (G.gl == 0)
        *status = FFRS_RSI16_REPAIR_OK;
}

// ---------------------------------------------------------------- encode as an int8 MFMA contraction

// The parity of a message is a fixed linear map of it mod p: parity_j = sum_m v_m G[m][j] mod p,
// then & 0xFFFF (pyrsi16.hpp:128-140), with G[m] the encode of the unit message at m (the residue
// NTT, mix and GS butterfly of rsi16v_impl.hpp:93-119 are linear). That holds for the scalar
// words always, and for the SIMD words wherever no mix constant is 65536 (there a zero residue
// reduces to p, and p * 65536 wraps to 65536 in 32 bits). This kernel evaluates the map exactly
// on the 16x16x64 int8 matrix cores:
//  - symbol v = v0 + 256 v1; the limbs enter as int8 v_a - 128 (the byte XOR 0x80);
//  - G entry g = g0 + 128 g1 + 16384 g2 (g0, g1 < 128, g2 <= 4), three non-negative int8 limbs;
//  - P_ab = sum_m (v_a[m] - 128) g_b[m][j] accumulates in int32 (|P| <= 128 * 127 * k);
//  - value_j = sum_ab 256^a 128^b P_ab + K_j mod p, where K_j = 128 * 257 * sum_b 128^b C_b[j]
//    (C_b[j] = sum_m g_b[m][j]) puts the offsets back; with 2^16 = -1 mod p every partial sum
//    is folded as lo16 - hi16 before it can overflow.
// One wave takes 16 codewords: B operand lane l = codeword l % 16, symbols 16 (l / 16) .. +16 of a
// 64-symbol chunk; A operand = the matching G limbs of 16 parity rows; the C tile holds parity
// rows 4 (l / 16) .. +4 of codeword l % 16. Symbols past k meet zero rows of G.
typedef int v4i_t __attribute__((ext_vector_type(4)));

// x - 65536 floor(x / 65536) - floor(x / 65536) = x mod p (up to a multiple of p), small enough
// for the next step
__device__ inline int32_t fold16(int32_t x) { return (x & 0xffff) - (x >> 16); }

//
// SYN: the same contraction gives the syndromes of NTT::pntt (ntt.hpp:133-152), S_j = sum over the
// n symbols of c_m w^(j rbo(m)) mod p, with the chunks of the message (msgs, k per codeword)
// followed by those of the parity (eccs, E per codeword, ecc_bytes in the buffer) against the
// matching rows of that matrix (nq_m message chunks); the canonical values go to syn (u32).
template <int CT, bool SYN>
__device__ inline void mfma_contract(const uint16_t *__restrict__ msgs, uint32_t msg_bytes, uint32_t count,
                                     uint16_t *__restrict__ out, const uint4 *__restrict__ gpack,
                                     const uint32_t *__restrict__ kconst, uint32_t k, uint32_t E, uint32_t nq,
                                     const uint16_t *__restrict__ eccs, uint32_t ecc_bytes, uint32_t nq_m,
                                     uint32_t *__restrict__ syn)
{
    const uint32_t lane = threadIdx.x & 63, c = lane & 15, h = lane >> 4;
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6), waves = gridDim.x * 4;
    const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(msgs, msg_bytes);
    const __amdgpu_buffer_rsrc_t rsrc_e = make_rsrc(SYN ? eccs : msgs, SYN ? ecc_bytes : msg_bytes);
    const uint32_t T = E / 16;
    for (uint32_t grp = wave; grp * 16 < count; grp += waves) {
        const uint32_t cw = grp * 16 + c;
        const bool valid = cw < count;
        const uint32_t row = cw * k * 2;  // byte offset of the codeword's message (< 2^32: launches are split)
        const uint32_t erow = cw * E * 2;
        for (uint32_t t0 = 0; t0 < T; t0 += CT) {
            v4i_t acc[CT][2][3];
#pragma unroll
            for (int tt = 0; tt < CT; ++tt)
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b)
                        acc[tt][a][b] = v4i_t{0, 0, 0, 0};
            for (uint32_t q = 0; q < nq; ++q) {
                uint4 x0 = make_uint4(0, 0, 0, 0), x1 = x0;
                if (valid) {  // 32-byte aligned reads, wholly inside or wholly past the buffer
                    if (!SYN || q < nq_m) {
                        const uint32_t off = row + (64 * q + 16 * h) * 2;
                        x0 = buf_load_u32x4(rsrc, off);
                        x1 = buf_load_u32x4(rsrc, off + 16);
                    } else {
                        const uint32_t off = erow + (64 * (q - nq_m) + 16 * h) * 2;
                        x0 = buf_load_u32x4(rsrc_e, off);
                        x1 = buf_load_u32x4(rsrc_e, off + 16);
                    }
                }
                const v4i_t lo = {(int)(__builtin_amdgcn_perm(x0.y, x0.x, 0x06040200u) ^ 0x80808080u),
                                  (int)(__builtin_amdgcn_perm(x0.w, x0.z, 0x06040200u) ^ 0x80808080u),
                                  (int)(__builtin_amdgcn_perm(x1.y, x1.x, 0x06040200u) ^ 0x80808080u),
                                  (int)(__builtin_amdgcn_perm(x1.w, x1.z, 0x06040200u) ^ 0x80808080u)};
                const v4i_t hi = {(int)(__builtin_amdgcn_perm(x0.y, x0.x, 0x07050301u) ^ 0x80808080u),
                                  (int)(__builtin_amdgcn_perm(x0.w, x0.z, 0x07050301u) ^ 0x80808080u),
                                  (int)(__builtin_amdgcn_perm(x1.y, x1.x, 0x07050301u) ^ 0x80808080u),
                                  (int)(__builtin_amdgcn_perm(x1.w, x1.z, 0x07050301u) ^ 0x80808080u)};
#pragma unroll
                for (int tt = 0; tt < CT; ++tt) {
                    const uint32_t t = t0 + tt;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const uint4 g = gpack[((size_t)(b * T + t) * nq + q) * 64 + lane];
                        const v4i_t A = {(int)g.x, (int)g.y, (int)g.z, (int)g.w};
                        acc[tt][0][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, lo, acc[tt][0][b], 0, 0, 0);
                        acc[tt][1][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, hi, acc[tt][1][b], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int tt = 0; tt < CT; ++tt) {
                const uint32_t t = t0 + tt;
                uint32_t v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t j = 16 * t + 4 * h + i;
                    int32_t z[3];
#pragma unroll
                    for (int b = 0; b < 3; ++b)
                        z[b] = fold16(fold16(acc[tt][0][b][i]) + 256 * fold16(acc[tt][1][b][i]));
                    int32_t f = fold16(z[0] + 128 * z[1] + 16384 * z[2] + (int32_t)kconst[j]);
                    f = f < 0 ? f + (int32_t)P16 : f;
                    f = f >= (int32_t)P16 ? f - (int32_t)P16 : f;
                    v[i] = SYN ? (uint32_t)f : (uint32_t)f & 0xffffu;
                }
                if (valid) {
                    if constexpr (SYN)
                        *reinterpret_cast<uint4 *>(syn + (size_t)cw * E + 16 * t + 4 * h) =
                            make_uint4(v[0], v[1], v[2], v[3]);
                    else
                        *reinterpret_cast<uint2 *>(out + (size_t)cw * E + 16 * t + 4 * h) =
                            make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16));
                }
            }
        }
    }
}

template <int CT>
__global__ __launch_bounds__(256) void rsi16_encode_mfma_kernel(const uint16_t *__restrict__ msgs, uint32_t msg_bytes,
                                                                uint32_t count, uint16_t *__restrict__ out,
                                                                const uint4 *__restrict__ gpack,
                                                                const uint32_t *__restrict__ kconst, uint32_t k,
                                                                uint32_t E, uint32_t nq)
{
    mfma_contract<CT, false>(msgs, msg_bytes, count, out, gpack, kconst, k, E, nq, nullptr, 0, 0, nullptr);
}

// The syndromes of NTT::pntt (ntt.hpp:133-152) as the same contraction (mfma_contract, SYN).
template <int CT>
__global__ __launch_bounds__(256) void rsi16_synd_mfma_kernel(const uint16_t *__restrict__ msgs, uint32_t msg_bytes,
                                                              const uint16_t *__restrict__ eccs, uint32_t ecc_bytes,
                                                              uint32_t count, const uint4 *__restrict__ vpack,
                                                              const uint32_t *__restrict__ vconst, uint32_t k,
                                                              uint32_t E, uint32_t nq, uint32_t nq_m,
                                                              uint32_t *__restrict__ syn)
{
    mfma_contract<CT, true>(msgs, msg_bytes, count, nullptr, vpack, vconst, k, E, nq, eccs, ecc_bytes, nq_m, syn);
}

// ---------------------------------------------------------------- interleaved repair, SIMD words

// The reference's vector instantiations repair the columns of an interleaved block in groups of
// W = 4, 8 or 16 run in lockstep (pyrsi16.hpp:389-453 -> RSi16v<W>::_repair, _sugiyama,
// _find_roots_ntt: rsi16v_impl.hpp:122-168, 271-497; NTT::poly_div_ntt, poly_inv_mod_xn_rev:
// ntt.hpp:363-564; vector helpers simd.hpp:60-254), and the columns of a group change one
// another's results: the Sugiyama loop runs while any lane is active, lanes that finished keep
// their polynomials but have their lengths overwritten and measured again against bounds taken
// over the group, and Forney evaluates every lane with the group's longest derivative and
// evaluator. With W = 16 the Newton step of the inverse ignores its lane mask
// (rsi16v_avx512.cpp, assign_masked). Here one GPU lane plays one SIMD lane: the W lanes of a
// cluster take the same branches, decided on maxima and any() over the cluster, and each lane
// runs the per-lane arithmetic (on canonical values the vector and scalar words agree) over
// its own polynomials, kept in a workspace with the lanes' elements side by side.

struct LaneVec {  // element i of one lane's vector
    uint32_t *p;
    uint32_t stride;
    __device__ uint32_t &operator[](uint32_t i) const { return p[(size_t)i * stride]; }
};

template <int W>
__device__ inline uint32_t cl_max(uint32_t v)
{
#pragma unroll
    for (int m = 1; m < W; m <<= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)v, m, W);
        v = o > v ? o : v;
    }
    return v;
}

template <int W>
__device__ inline bool cl_any(bool b) { return cl_max<W>(b ? 1u : 0u) != 0; }

// ntt.hpp:280-302 (natural in, RBO out) and ntt.hpp:225-249 (RBO in, natural out), one lane
__device__ inline void lv_gs(LaneVec x, const uint32_t *roots, uint32_t len)
{
    for (uint32_t st = len / 2, ef = 0; st > 0; st /= 2, ++ef)
        for (uint32_t b = 0; b < len; b += 2 * st) {
            for (uint32_t i = b; i < b + st; ++i) {
                const uint32_t a = x[i], c = x[i + st];
                x[i] = S::add(a, c);
                const uint32_t d = S::sub(a, c);
                x[i + st] = i == b ? d : S::mul(d, roots[(i - b) << ef]);
            }
        }
}

__device__ inline void lv_ct(LaneVec x, const uint32_t *roots, uint32_t len)
{
    for (uint32_t st = 1, ef = len >> 1; st < len; st *= 2, ef >>= 1)
        for (uint32_t b = 0; b < len; b += 2 * st) {
            for (uint32_t i = b; i < b + st; ++i) {
                const uint32_t a = x[i];
                const uint32_t m = i == b ? x[i + st] : S::mul(x[i + st], roots[ef * (i - b)]);
                x[i] = S::add(a, m);
                x[i + st] = S::sub(a, m);
            }
        }
}

__device__ inline void lv_nttr(const Rsi16Dev &T, LaneVec x) { lv_gs(x, T.roots_ecc, T.E); }
__device__ inline void lv_inttr(const Rsi16Dev &T, LaneVec x)
{
    lv_ct(x, T.roots_i_ecc, T.E);
    for (uint32_t i = 0; i < T.E; ++i)
        x[i] = S::mul(x[i], T.ecc_inv);
}

// simd.hpp:228-254: degree + 1 within the first `bound` coefficients (a bound past the
// polynomial is clamped to it)
__device__ inline uint32_t lv_len(LaneVec x, uint32_t bound, uint32_t E)
{
    uint32_t m = bound < E ? bound : E;
    while (m > 0 && x[m - 1] == 0)
        --m;
    return m;
}

// NTT::_reverse_ntt_vec, ntt.hpp:344-354
__device__ inline void lv_reverse_ntt(const Rsi16Dev &T, LaneVec v, uint32_t shift)
{
    const uint32_t E = T.E, mask = E - 1;
    for (uint32_t i = 1; i < E; ++i) {
        const uint32_t j = T.rbo_ecc[i], kk = T.rbo_ecc[E - j];
        if (i < kk) {
            const uint32_t t = v[i];
            v[i] = v[kk];
            v[kk] = t;
        }
        v[i] = S::mul(S::mul(v[i], T.roots_i_ecc[j]), T.roots_ecc[(shift * j) & mask]);
    }
}

// The per-lane state of the coupled repair, in the workspace.
struct SimdLane {
    LaneVec blk, ev, loc, der, q, t, a2, r2, tmp;
};

// NTT::poly_inv_mod_xn_rev, ntt.hpp:363-469, vector instantiation; false where it raises
template <int W>
__device__ inline bool lv_poly_inv(const Rsi16Dev &T, LaneVec g, uint32_t g_len, uint32_t g0, uint32_t a_len, LaneVec a)
{
    const uint32_t E = T.E;
    const uint32_t a0 = g_inv(g0);
    for (uint32_t i = 0; i < E; ++i)
        a[i] = a0;
    uint32_t l = 2;
    const uint32_t a_max = cl_max<W>(a_len);
    bool cond = l < a_len;
    while (cl_any<W>(cond)) {
        if (cl_max<W>(cond ? g_len : 0u) + l - 2 > E)  // py_assert at ntt.hpp:388-391
            return false;
        if (cond || W == 16)  // rsi16v_avx512.cpp: the W = 16 masked assignment updates every lane
            for (uint32_t i = 0; i < E; ++i)
                a[i] = S::mul(a[i], S::sub(2, S::mul(a[i], g[i])));
        lv_inttr(T, a);
        for (uint32_t i = l; i < E; ++i)
            a[i] = 0;
        lv_nttr(T, a);
        l *= 2;
        cond = l < a_len;
    }
    for (uint32_t i = 0; i < E; ++i)
        g[i] = S::mul(a[i], S::sub(2, S::mul(a[i], g[i])));
    lv_inttr(T, g);
    for (uint32_t i = 0; i < E; ++i)
        a[i] = i < l ? g[i] : 0u;
    for (uint32_t i = 0, m = a_max - 1; i < a_max / 2; ++i, --m) {  // std::reverse(&a[0], &a[a_max])
        const uint32_t t = a[i];
        a[i] = a[m];
        a[m] = t;
    }
    for (uint32_t i = 0; i < a_len; ++i)  // vec::copy_n(&a[0], a_max - a_len, a_len, &a[0], 0)
        a[i] = a[a_max - a_len + i];
    for (uint32_t i = a_len; i < E; ++i)
        a[i] = 0;
    lv_nttr(T, a);
    return true;
}

// NTT::poly_div_ntt, ntt.hpp:471-564: q <- f // g (transforms); false where it raises. W = 1
// is the reference's integral instantiation (a zero divisor always raises, f_len < g_len
// gives q = 0), W = 4, 8, 16 its vector one.
template <int W>
__device__ inline bool lv_poly_div(const Rsi16Dev &T, LaneVec f, uint32_t f_len, LaneVec g, uint32_t g_len, uint32_t f0,
                                   uint32_t g0, LaneVec t, LaneVec q, uint32_t &q_len)
{
    const uint32_t E = T.E;
    if constexpr (W == 1) {
        if (g_len == 0)  // py_assert(g_len > 0)
            return false;
        if (f_len < g_len) {
            for (uint32_t i = 0; i < E; ++i)
                q[i] = 0;
            q_len = 0;
            return true;
        }
    } else if (cl_any<W>(g_len == 0 && f_len > 0)) {  // py_assert(!any((g_len == 0) && (f_len > 0)))
        return false;
    }
    q_len = (g_len != 0 && f_len >= g_len) ? f_len - g_len + 1 : 0u;
    for (uint32_t i = 0; i < E; ++i)
        t[i] = g[i];
    lv_reverse_ntt(T, t, g_len);
    if (!lv_poly_inv<W>(T, t, g_len, g0, q_len, q))
        return false;
    for (uint32_t i = 0; i < E; ++i)
        t[i] = S::mul(S::sub(f[i], f0), q[i]);
    lv_inttr(T, t);
    const uint32_t room = E - g_len, right_n = q_len < room ? q_len : room;
    for (uint32_t i = 0; i < right_n; ++i)
        q[i] = t[g_len + i];
    for (uint32_t i = 0; i < q_len - right_n; ++i)
        q[right_n + i] = t[i];
    for (uint32_t i = q_len; i < E; ++i)
        q[i] = 0;
    lv_nttr(T, q);
    return true;
}

// RSi16v<W>::_sugiyama, rsi16v_impl.hpp:271-440: L.ev holds the syndromes on entry and the
// evaluator on exit, L.loc the locator; returns the locator length, or 0xFFFFFFFF where the
// reference raises.
template <int W>
__device__ inline uint32_t lv_sugiyama(const Rsi16Dev &T, const SimdLane &L)
{
    const uint32_t E = T.E, mask = E - 1;
    LaneVec a1 = L.loc, r1 = L.ev, q = L.q, t = L.t, a2 = L.a2, r2 = L.r2;
    for (uint32_t i = 0; i < E; ++i) {
        a1[i] = 0;
        a2[i] = 1;
        q[i] = 0;
        t[i] = 0;
    }
    uint32_t q_len = 0, t_len = 0, a2_len = 1;
    uint32_t r1_len = lv_len(r1, E, E);
    for (uint32_t i = 0; i < E; ++i)
        r2[i] = r1[i];
    uint32_t r2_len = r1_len, r2l = r2[0];
    const uint32_t lanes = r1_len != 0 ? 0xffffffffu : 0u;
    uint32_t a1_len;
    {  // the degree-one first step (:308-335)
        const uint32_t h = r1[(r1_len - 1) & mask], h2 = r1[(r1_len - 2) & mask];
        const uint32_t c1 = s_neg(g_inv(h));
        const uint32_t c0 = S::mul(h2, S::mul(c1, c1));
        a1[0] = c0;
        a1[1] = c1;
        a1_len = 2 & lanes;
        for (uint32_t i = 0; i < E; ++i)
            r1[i] = S::mul(c0, r1[i]);
        for (uint32_t i = 1; i < E; ++i)
            r1[i] = S::add(r1[i], S::mul(c1, r2[i - 1]));
        r1_len = lv_len(r1, E - 1, E) & lanes;
        for (uint32_t i = r1_len; i < E; ++i)
            r1[i] = 0;
        lv_nttr(T, a1);
        lv_nttr(T, r2);
    }
    bool cond = r1_len >= E / 2;
    if (cl_any<W>(cond)) {
        for (uint32_t it = 1; it < E / 2; ++it) {
            const uint32_t r1h = r1[(r1_len - 1) & mask], r1l = r1[0];
            lv_nttr(T, r1);
            if (!lv_poly_div<W>(T, r2, cond ? r2_len : 0u, r1, r1_len, r2l, r1h, t, q, q_len))
                return 0xffffffffu;
            for (uint32_t i = 0; i < E; ++i) {
                const uint32_t qi = q[i];
                t[i] = S::sub(a2[i], S::mul(qi, a1[i]));  // a2 - q a1 (poly_mul_ntt, _vec_sub)
                q[i] = S::sub(r2[i], S::mul(qi, r1[i]));  // r2 - q r1
            }
            t_len = q_len + a1_len - 1;
            t_len = a2_len > t_len ? a2_len : t_len;
            q_len = q_len + r1_len - 1;
            q_len = r2_len > q_len ? r2_len : q_len;
            for (uint32_t i = 0; i < E; ++i) {
                a2[i] = a1[i];
                r2[i] = r1[i];
            }
            a2_len = a1_len;
            r2_len = r1_len;
            r2l = r1l;
            cond = r1_len >= E / 2;
            if (!cl_any<W>(cond)) {
                lv_inttr(T, r1);
                break;
            }
            if (cond)  // vec::copy_n_masked
                for (uint32_t i = 0; i < E; ++i) {
                    a1[i] = t[i];
                    r1[i] = q[i];
                }
            a1_len = t_len;  // assigned in every lane
            r1_len = q_len;
            lv_inttr(T, r1);
            r1_len = lv_len(r1, cl_max<W>(cond ? r1_len : 0u), E);
        }
    }
    lv_inttr(T, a1);
    a1_len = lv_len(a1, cl_max<W>(a1_len & lanes), E);
    const uint32_t inv0 = g_inv(a1[0]);
    for (uint32_t i = 0; i < E; ++i) {
        a1[i] = S::mul(a1[i], inv0);
        r1[i] = S::mul(r1[i], inv0);
    }
    return a1_len & lanes;
}

__device__ inline uint32_t lv_eval(LaneVec r, uint32_t len, uint32_t x)
{
    uint32_t s = 0;
    for (uint32_t i = len - 1; i < len; --i)
        s = S::add(S::mul(s, x), r[i]);
    return s;
}

// The reference's NTT methods and RSi16._sugiyama<W> (pyntt.hpp:28-84, rsi16v_impl.hpp:670-795,
// pyrsi16.hpp:653-674) on W vectors, one lane per vector as one SIMD lane of the reference's
// vector words (W = 1: its scalar words): vector l's elements sit at ws[i * W + l] in regions of
// `len` elements. Division, inverse and Sugiyama take the lanes' group decisions as the
// reference's vector code does (the same lane functions as the coupled interleaved repair).
// result[l]: 1 where the operation raises in the reference (division, inverse), or the locator
// length (Sugiyama; 0xFFFFFFFF where it raises).
enum : int { NTTOP_NTT, NTTOP_INTT, NTTOP_NTTR, NTTOP_INTTR, NTTOP_MUL, NTTOP_DIV, NTTOP_INV, NTTOP_SUGIYAMA };

template <int W>
__global__ __launch_bounds__(64) void rsi16_ntt_util_kernel(Rsi16Dev T, int op, uint32_t *__restrict__ ws, uint32_t len,
                                                            const uint32_t *__restrict__ roots_i_ntt, uint32_t ntt_len_i,
                                                            uint32_t n_arg, uint32_t *__restrict__ result)
{
    const uint32_t lane = threadIdx.x, E = T.E, N = T.N;
    auto vec = [&](int r) { return LaneVec{ws + (size_t)r * len * W + lane, (uint32_t)W}; };
    LaneVec a = vec(0), b = vec(1);
    uint32_t res = 0;
    switch (op) {
    case NTTOP_NTT:  // ntt.hpp:596-599
        lv_ct(a, T.roots_ntt, N);
        break;
    case NTTOP_INTT:  // ntt.hpp:601-606
        lv_gs(a, roots_i_ntt, N);
        for (uint32_t i = 0; i < N; ++i)
            a[i] = S::mul(a[i], ntt_len_i);
        break;
    case NTTOP_NTTR:
        lv_nttr(T, a);
        break;
    case NTTOP_INTTR:
        lv_inttr(T, a);
        break;
    case NTTOP_MUL:  // py_poly_mul: nttr both, poly_mul_ntt (pointwise), inttr
        lv_nttr(T, a);
        lv_nttr(T, b);
        for (uint32_t i = 0; i < E; ++i)
            a[i] = S::mul(a[i], b[i]);
        lv_inttr(T, a);
        break;
    case NTTOP_DIV: {  // py_poly_div: lengths and end coefficients first, then poly_div_ntt
        LaneVec t = vec(2), q = vec(3);
        const uint32_t a_len = lv_len(a, E, E), b_len = lv_len(b, E, E);
        const uint32_t al = a[0], bh = b[(b_len - 1) & (E - 1)];
        lv_nttr(T, a);
        lv_nttr(T, b);
        uint32_t q_len = 0;
        if (!lv_poly_div<W>(T, a, a_len, b, b_len, al, bh, t, q, q_len)) {
            res = 1;
            break;
        }
        lv_inttr(T, q);
        for (uint32_t i = 0; i < E; ++i)
            a[i] = q[i];
        break;
    }
    case NTTOP_INV: {  // py_poly_inv
        LaneVec r = vec(2);
        const uint32_t g0 = a[0], g_len = lv_len(a, E, E);
        lv_nttr(T, a);
        if (!lv_poly_inv<W>(T, a, g_len, g0, n_arg, r)) {
            res = 1;
            break;
        }
        lv_inttr(T, r);
        for (uint32_t i = 0, m = n_arg - 1; i < n_arg / 2; ++i, --m) {  // std::reverse(&res[0], &res[n])
            const uint32_t x = r[i];
            r[i] = r[m];
            r[m] = x;
        }
        for (uint32_t i = 0; i < E; ++i)
            a[i] = r[i];
        break;
    }
    default: {  // NTTOP_SUGIYAMA: syndromes in a, locator out in b, evaluator out in a
        SimdLane L{vec(8), a, b, vec(2), vec(3), vec(4), vec(5), vec(6), vec(7)};
        res = lv_sugiyama<W>(T, L);
        break;
    }
    }
    result[lane] = res;
}

// One lane per SIMD lane: group `g0 + lane / W` of the ncols columns from col0 of an interleaved
// block (columns col0 + (g0 + lane / W) * W + lane % W, those past ncols zero), its codeword
// loaded into the workspace, repaired there; gstatus[group] = the group's status (the maximum
// over its lanes) or FFRS_RSI16_RAISES. The ecc rows hold u16 symbols or u32 field elements.
template <int W, class TE = uint16_t>
__global__ __launch_bounds__(64) void rsi16_repair_simd_kernel(Rsi16Dev T, const uint16_t *__restrict__ msg,
                                                               const TE *__restrict__ ecc, uint32_t il, uint32_t col0,
                                                               uint32_t ncols, uint32_t g0, uint32_t lanes_total,
                                                               uint32_t *__restrict__ work, uint8_t *__restrict__ gstatus)
{
    const uint32_t lane = blockIdx.x * 64 + threadIdx.x;
    if (lane >= lanes_total)  // lanes_total is a multiple of W: clusters are whole
        return;
    const uint32_t n = T.n, k = T.k, E = T.E;
    const uint32_t grp = g0 + lane / W, cc = grp * W + lane % W, col = col0 + cc;
    const bool valid = cc < ncols;
    auto row = [&](uint32_t r) { return LaneVec{work + (size_t)r * lanes_total + lane, lanes_total}; };
    SimdLane L{row(0), row(n), row(n + E), row(n + 2 * E), row(n + 3 * E), row(n + 4 * E), row(n + 5 * E),
               row(n + 6 * E), row(n + 7 * E)};
    for (uint32_t i = 0; i < n; ++i)
        L.blk[i] = !valid ? 0u : (i < k ? msg[(size_t)i * il + col] : ecc[(size_t)(i - k) * il + col]);
    // NTT::pntt, ntt.hpp:133-152: the syndromes
    for (uint32_t c = 0; c < n / E; ++c) {
        for (uint32_t i = 0; i < E; ++i)
            L.tmp[i] = L.blk[c * E + i];
        lv_ct(L.tmp, T.roots_ecc, E);
        for (uint32_t i = 0; i < E; ++i) {
            const uint32_t v = S::mul(L.tmp[i], T.pntt_shift[c * E + i]);
            L.ev[i] = c == 0 ? v : S::add(L.ev[i], v);
        }
    }
    bool nz = false;
    for (uint32_t i = 0; i < E; ++i)
        nz |= L.ev[i] != 0;
    uint8_t st = FFRS_RSI16_NO_ERRORS;
    if (cl_any<W>(nz)) {
        const uint32_t Lloc = lv_sugiyama<W>(T, L);
        if (cl_any<W>(Lloc == 0xffffffffu) || cl_max<W>(Lloc) > E) {  // rsi16v_impl.hpp:454
            st = FFRS_RSI16_RAISES;
        } else {
            const bool active = Lloc != 0;
            const uint32_t elen = active ? E - Lloc + 1 : 0u;
            for (uint32_t i = 0; i + 1 < E; ++i)  // _deriv, rsi16v_impl.hpp:608-616
                L.der[i] = S::mul(L.loc[i + 1], i + 1);
            L.der[E - 1] = 0;
            const uint32_t dlen = active ? lv_len(L.der, E, E) : 0u;
            const uint32_t dmax = cl_max<W>(dlen), emax = cl_max<W>(elen);
            // _find_roots_ntt with NTT::pintt_ecc chunk by chunk (ntt.hpp:206-220)
            uint32_t count = 0;
            for (uint32_t c = 0; c < n / E; ++c) {
                for (uint32_t i = 0; i < E; ++i)
                    L.tmp[i] = S::mul(L.loc[i], T.pintt_shift[c * E + i]);
                lv_gs(L.tmp, T.roots_i_ecc, E);
                for (uint32_t i = 0; i < E; ++i) {
                    if (active && L.tmp[i] == 0) {
                        const uint32_t m = c * E + i;
                        const uint32_t x = T.roots_ntt[rbo_bits(m, T.nbits)];
                        const uint32_t xi = g_inv(x);
                        const uint32_t num = lv_eval(L.ev, emax, xi), den = lv_eval(L.der, dmax, xi);
                        L.blk[m] = S::add(L.blk[m], S::mul(g_div(num, den), x));  // _forney, :500-512
                        ++count;
                    }
                }
            }
            const uint8_t lst = count > 0 ? FFRS_RSI16_REPAIR_OK
                                          : (active ? FFRS_RSI16_ERROR_LOCATION_FAIL : FFRS_RSI16_NO_ERRORS);
            st = (uint8_t)cl_max<W>(lst);
        }
    }
    if (lane % W == 0)
        gstatus[grp] = st;
}

// Writes the repaired columns of groups [g0, g_end) back (message values & 0xFFFF, ecc values
// likewise or whole for u32 rows), from the workspace of the launch that repaired groups g0..
// (lanes_total lanes).
template <int W, class TE = uint16_t>
__global__ __launch_bounds__(256) void rsi16_simd_writeback_kernel(const Rsi16Dev T, uint16_t *__restrict__ msg,
                                                                   TE *__restrict__ ecc, uint32_t il, uint32_t col0,
                                                                   uint32_t ncols, uint32_t g0, uint32_t g_end,
                                                                   uint32_t lanes_total,
                                                                   const uint32_t *__restrict__ work)
{
    const uint32_t lane = blockIdx.x * 64 + (threadIdx.x & 63);
    if (lane >= lanes_total)
        return;
    const uint32_t grp = g0 + lane / W, cc = grp * W + lane % W, col = col0 + cc;
    if (grp >= g_end || cc >= ncols)
        return;
    const uint32_t n = T.n, k = T.k;
    for (uint32_t i = threadIdx.x >> 6; i < n; i += 4) {
        const uint32_t v = work[(size_t)i * lanes_total + lane];
        if (i < k)
            msg[(size_t)i * il + col] = (uint16_t)v;
        else
            ecc[(size_t)(i - k) * il + col] = (TE)(sizeof(TE) == 2 ? (v & 0xffffu) : v);
    }
}

// ---------------------------------------------------------------- CIRC16 (pycirc16.hpp)

// One CIRC16 block (pycirc16.hpp:45-64, 423-441): message row r (r < outer k) holds the inner
// messages of the I interleave blocks side by side, inner k symbols each, so inner codeword
// q = r I + kk has its message at q * inner_k; the outer parity has outer E rows of the same
// shape; the inner parity of codeword q sits at q * inner_E, message rows first, then the
// rows of the outer parity. The repair keeps the outer parity and the inner parity as u32
// field elements (CircRepair's GFT buffers), so values of 65536 survive between the steps.

// load_ecc (pycirc16.hpp:158-162): the u16 ecc into the u32 state
__global__ __launch_bounds__(256) void circ16_expand_kernel(const uint16_t *__restrict__ ecc, uint32_t *__restrict__ rso32,
                                                             uint32_t rso_len, uint32_t *__restrict__ rsi32,
                                                             uint32_t rsi_len)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < rso_len + rsi_len; i += gridDim.x * 256) {
        if (i < rso_len)
            rso32[i] = ecc[i];
        else
            rsi32[i - rso_len] = ecc[i];
    }
}

// compute_rsi_synd (pycirc16.hpp:164-180): NTT::pntt syndromes of `rows` inner codewords, one
// lane group each; message q at m + q k (u16 message rows or u32 outer-parity rows), parity at
// e + q E.
template <class TM>
__global__ __launch_bounds__(64) void circ16_synd_kernel(const Rsi16Dev T, const TM *__restrict__ m,
                                                          const uint32_t *__restrict__ e, uint32_t rows,
                                                          uint32_t *__restrict__ syn, uint32_t g)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t E = T.E, cpw = 64 / g, grp = threadIdx.x / g;
    const uint32_t q = blockIdx.x * cpw + grp;
    if (q >= rows)
        return;
    Grp G{threadIdx.x % g, g};
    uint32_t *chunk = s_mem + grp * 2 * E, *ev = chunk + E;
    pntt_codeword<TM, uint32_t>(T, m + (size_t)q * T.k, e + (size_t)q * E, chunk, ev, G);
    for (uint32_t j = G.gl; j < E; j += G.g)
        syn[(size_t)q * E + j] = ev[j];
}

// repair_outer_zeros (pycirc16.hpp:182-244), one wave per inner codeword q of the outer parity
// rows (message at rso32 + q k, parity at rsio32 + q E): where its syndromes are not all zero
// and it holds 1 .. E zero symbols, the known-position repair at those positions; where every
// repaired value is 0 mod 0x10000 the values are written back and the syndromes recomputed.
__global__ __launch_bounds__(64) void circ16_outer_zeros_kernel(const Rsi16Dev T, uint32_t *__restrict__ rso32,
                                                                 uint32_t *__restrict__ rsio32, uint32_t rows,
                                                                 uint32_t *__restrict__ syn)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t q = blockIdx.x, lane = threadIdx.x, E = T.E, k = T.k;
    if (q >= rows)
        return;
    uint32_t *m = rso32 + (size_t)q * k, *e = rsio32 + (size_t)q * E, *sy = syn + (size_t)q * E;
    uint32_t *ev = s_mem, *loc = s_mem + 2 * E, *tmp = s_mem + 3 * E, *der = s_mem + 5 * E, *chunk = s_mem + 6 * E;
    uint32_t *pos = s_mem + 7 * E, *err = s_mem + 8 * E;
    bool nz = false;
    for (uint32_t j = lane; j < E; j += 64)
        nz |= sy[j] != 0;
    if (!__ballot(nz))
        return;
    uint32_t cnt = 0;
    for (uint32_t j = lane; j < k + E; j += 64)
        cnt += (j < k ? m[j] : e[j - k]) == 0 ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1)
        cnt += (uint32_t)__shfl_xor((int)cnt, o);
    if (cnt == 0 || cnt > E)  // none, or too many zeros: left as they are
        return;
    // the zero positions in ascending order (message part, then parity part)
    uint32_t base = 0;
    for (uint32_t j0 = 0; j0 < k + E; j0 += 64) {
        const uint32_t j = j0 + lane;
        const bool z = j < k + E && (j < k ? m[j] : e[j - k]) == 0;
        const uint64_t mk = __ballot(z);
        if (z)
            pos[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u))] = j;
        base += (uint32_t)__builtin_popcountll(mk);
    }
    for (uint32_t j = lane; j < E; j += 64)
        ev[j] = sy[j];
    wave_sync();
    Grp G{lane, 64};
    known_errors(T, pos, cnt, ev, loc, tmp, der, G, [&](uint32_t r, uint32_t, uint32_t v) { err[r] = v; });
    wave_sync();
    bool bad = false;
    for (uint32_t r = lane; r < cnt; r += 64)
        bad |= (err[r] & 0xffffu) != 0;  // the symbol was 0: the repaired value is the error
    if (__ballot(bad))
        return;
    for (uint32_t r = lane; r < cnt; r += 64) {
        const uint32_t p = pos[r];
        if (p < k)
            m[p] = err[r];
        else
            e[p - k] = err[r];
    }
    __syncthreads();  // the stores above are read back by other lanes below
    pntt_codeword<uint32_t, uint32_t>(T, m, e, chunk, ev, G);  // synd_block (pyrsi16.hpp:186-189)
    for (uint32_t j = lane; j < E; j += 64)
        sy[j] = ev[j];
}

// repair_inner_zeros (pycirc16.hpp:246-289) and find_error_locations (:291-295), one wave per
// inner codeword q: where the syndromes are not all zero and its parity holds a zero, mix_ecc of
// the syndromes (pyrsi16.hpp:218-220); where that is 0 mod 0x10000 everywhere it becomes the
// parity and the syndromes are cleared. flag[q] = the syndromes are not all zero afterwards.
__global__ __launch_bounds__(64) void circ16_inner_zeros_kernel(const Rsi16Dev T, uint32_t *__restrict__ rsi32,
                                                                 uint32_t rows, uint32_t *__restrict__ syn,
                                                                 uint8_t *__restrict__ flag)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t q = blockIdx.x, lane = threadIdx.x, E = T.E;
    if (q >= rows)
        return;
    uint32_t *e = rsi32 + (size_t)q * E, *sy = syn + (size_t)q * E;
    bool nz = false, zero = false;
    for (uint32_t j = lane; j < E; j += 64) {
        nz |= sy[j] != 0;
        zero |= e[j] == 0;
    }
    bool any = __ballot(nz) != 0;
    if (any && __ballot(zero)) {
        for (uint32_t j = lane; j < E; j += 64)
            s_mem[j] = S::mul(sy[j], T.mix[j]);
        wave_sync();
        gs_full<S>(s_mem, E, T.roots_i_ecc, Grp{lane, 64});
        bool bad = false;
        for (uint32_t j = lane; j < E; j += 64)
            bad |= (s_mem[j] & 0xffffu) != 0;
        if (!__ballot(bad)) {
            for (uint32_t j = lane; j < E; j += 64) {
                e[j] = s_mem[j];
                sy[j] = 0;
            }
            any = false;
        }
    }
    if (lane == 0)
        flag[q] = any ? 1 : 0;
}

// repair_outer_errors with error locations (pycirc16.hpp:297-313 -> pyrsi16.hpp:458-522): the
// known-position repair of outer codeword `col` (symbols at msg + i il, parity rows at
// rso32 + (i - k) il) at the rows pos[kk * E ..] of its interleave block kk = col / inner_k,
// where that block has 1 .. E error rows (cnt[kk]). With positions the SIMD words give each
// column the scalar result. One wave per column.
__global__ __launch_bounds__(64) void circ16_outer_known_kernel(const Rsi16Dev T, uint16_t *__restrict__ msg,
                                                                 uint32_t *__restrict__ rso32, uint32_t il, uint32_t ik,
                                                                 const uint32_t *__restrict__ pos,
                                                                 const uint32_t *__restrict__ cnt)
{
    extern __shared__ uint32_t s_mem[];
    const uint32_t col = blockIdx.x, lane = threadIdx.x, E = T.E, k = T.k;
    if (col >= il)
        return;
    const uint32_t kk = col / ik, npos = cnt[kk];
    if (npos == 0 || npos > E)
        return;
    uint32_t *ev = s_mem, *loc = s_mem + 2 * E, *tmp = s_mem + 3 * E, *der = s_mem + 5 * E, *chunk = s_mem + 6 * E;
    Grp G{lane, 64};
    uint16_t *m = msg + col;
    uint32_t *e = rso32 + col;
    pntt_codeword<uint16_t, uint32_t>(T, m, e, chunk, ev, G, il, il);
    bool nz = false;
    for (uint32_t j = lane; j < E; j += 64)
        nz |= ev[j] != 0;
    if (!__ballot(nz))
        return;
    known_errors(T, pos + (size_t)kk * E, npos, ev, loc, tmp, der, G, [&](uint32_t, uint32_t p, uint32_t v) {
        if (p < k)
            m[(size_t)p * il] = (uint16_t)(S::add((uint32_t)m[(size_t)p * il], v) & 0xffffu);
        else
            e[(size_t)(p - k) * il] = S::add(e[(size_t)(p - k) * il], v);
    });
}

// dump_ecc (pycirc16.hpp:323-327) of `blocks` blocks: out block b = [outer parity (u32 state)
// & 0xFFFF | inner parity of the message rows | inner parity of the outer parity rows]
__global__ __launch_bounds__(256) void circ16_assemble_kernel(const uint32_t *__restrict__ rso32,
                                                               const uint16_t *__restrict__ rsi16,
                                                               const uint16_t *__restrict__ rsio16, uint32_t rso_len,
                                                               uint32_t rsi_len, uint32_t rsio_len, uint32_t blocks,
                                                               uint16_t *__restrict__ out)
{
    const uint32_t el = rso_len + rsi_len + rsio_len;
    const size_t total = (size_t)el * blocks;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
        const size_t b = t / el;
        const uint32_t x = (uint32_t)(t - b * el);
        uint16_t v;
        if (x < rso_len)
            v = (uint16_t)(rso32[b * rso_len + x] & 0xffffu);
        else if (x < rso_len + rsi_len)
            v = rsi16[b * rsi_len + (x - rso_len)];
        else
            v = rsio16[b * rsio_len + (x - rso_len - rsi_len)];
        out[t] = v;
    }
}

// ---------------------------------------------------------------- mixed-size plans

// A plan's work list is sorted by (block_len, ecc_len) class and cut into chunks of the
// class's codewords-per-wave; single-wave workgroups take chunks from a device counter.
struct Rsi16PlanItem {
    uint64_t msg_off;  // element (u16) offsets
    uint64_t ecc_off;
    uint32_t idx;
    uint32_t cls;
};

struct Rsi16PlanChunk {
    uint32_t first;
    uint16_t count;
    uint16_t cls;
};

struct Rsi16PlanClass {
    Rsi16Dev T;
    uint32_t g;      // lanes per codeword of the LDS code
    uint32_t l_enc;  // lanes per codeword of the per-lane encode code (encode_split when > 1)
};

// zeroes the work counters of a plan's launches (one per chunk range)
__global__ void rsi16_plan_reset_kernel(uint32_t *queue, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        queue[i] = 0;
}

// One persistent launch per chunk range: every wave takes the next chunk of its range from
// the range's counter. E = 0: any ecc_len, one codeword per lane group with the polynomials
// in LDS (repair_group / encode_group). E > 0, repair: the register code of
// rsi16_repair_reg_kernel<E>. E > 0, encode: one codeword per lane (encode_lane).
template <int MODE, int E>  // MODE 0: encode (vector words), 1: encode (scalar words), 2: repair
__global__ __launch_bounds__(64) void rsi16_plan_kernel(uint16_t *__restrict__ msgs, uint16_t *__restrict__ eccs,
                                                         uint8_t *__restrict__ status,
                                                         const Rsi16PlanItem *__restrict__ items,
                                                         const Rsi16PlanChunk *__restrict__ chunks, uint32_t nchunks,
                                                         const Rsi16PlanClass *__restrict__ classes, uint32_t *queue)
{
    extern __shared__ uint32_t s_mem[];
    __shared__ uint32_t s_chunk;
    for (;;) {
        if (threadIdx.x == 0)
            s_chunk = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t c = __builtin_amdgcn_readfirstlane(s_chunk);  // one wave: the class fields go to SGPRs
        if (c >= nchunks)
            break;
        const Rsi16PlanChunk ch = chunks[c];
        const Rsi16PlanClass &cl = classes[ch.cls];
        if constexpr (E == 0) {
            const uint32_t g = cl.g, Ec = cl.T.E;
            Grp G{threadIdx.x % g, g};
            const uint32_t grp = threadIdx.x / g;
            if (grp < ch.count) {
                const Rsi16PlanItem it = items[ch.first + grp];
                if constexpr (MODE == 2)
                    repair_group(cl.T, msgs + it.msg_off, eccs + it.ecc_off, s_mem + grp * REPAIR_WORDS * Ec, G,
                                 status + it.idx);
                else if constexpr (MODE == 0)
                    encode_group<VectorOps>(cl.T, msgs + it.msg_off, eccs + it.ecc_off, s_mem + grp * 2 * Ec, G);
                else
                    encode_group<ScalarOps>(cl.T, msgs + it.msg_off, eccs + it.ecc_off, s_mem + grp * 2 * Ec, G);
            }
        } else if constexpr (MODE == 2) {
            constexpr uint32_t G = RegShape<E>::G;
            const uint32_t grp = threadIdx.x / G;
            if (grp < ch.count) {
                const Rsi16PlanItem it = items[ch.first + grp];
                LaneR<E> L;
                L.init(cl.T, threadIdx.x);
                repair_group_reg<E>(cl.T, msgs + it.msg_off, eccs + it.ecc_off, s_mem + grp * REG_WORDS * E, L,
                                    status + it.idx);
            }
        } else if constexpr (E > 32) {  // encode, the register code of rsi16_encode_reg_kernel<E>
            using O = typename std::conditional<MODE == 0, VectorOps, ScalarOps>::type;
            constexpr uint32_t G = RegShape<E>::G;
            const uint32_t grp = threadIdx.x / G;
            if (grp < ch.count) {
                const Rsi16PlanItem it = items[ch.first + grp];
                LaneR<E> L;
                L.init(cl.T, threadIdx.x);
                encode_group_reg<O, E>(cl.T, msgs + it.msg_off, eccs + it.ecc_off, L);
            }
        } else {
            using O = typename std::conditional<MODE == 0, VectorOps, ScalarOps>::type;
            const uint32_t L = cl.l_enc, grp = threadIdx.x / L, gl = threadIdx.x % L;
            if (grp < ch.count) {
                const Rsi16PlanItem it = items[ch.first + grp];
                const uint16_t *m = msgs + it.msg_off;
                uint16_t *o = eccs + it.ecc_off;
                bool aligned = false;
                if constexpr (E >= 8)
                    aligned = (((uintptr_t)m | (uintptr_t)o) & 15) == 0;
                if (L == 1) {
                    if constexpr (E >= 8) {
                        if (aligned)
                            encode_lane<O, E, true>(cl.T, m, o);
                        else
                            encode_lane<O, E, false>(cl.T, m, o);
                    } else {
                        encode_lane<O, E, false>(cl.T, m, o);
                    }
                } else {
                    if constexpr (E >= 8) {
                        if (aligned)
                            encode_split<O, E, true>(cl.T, m, o, gl, L);
                        else
                            encode_split<O, E, false>(cl.T, m, o, gl, L);
                    } else {
                        encode_split<O, E, false>(cl.T, m, o, gl, L);
                    }
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace ffrs


// ---------------------------------------------------------------- host side + C ABI

using namespace ffrs;

struct ffrs_rsi16 {
    ffrs_ctx *ctx = nullptr;
    uint32_t n = 0, k = 0, E = 0, N = 0, nbits = 0, primitive = 0, root = 0;
    uint32_t *d_tables = nullptr;
    Rsi16Dev dev{};
    uint32_t g_enc = 0, g_rep = 0;  // lanes per codeword of the encode and repair kernels
    void *d_work = nullptr;          // workspace of the SIMD-word interleaved repair
    size_t work_size = 0;
    // the encode as an int8 MFMA contraction (rsi16_encode_mfma_kernel), where it applies
    uint4 *d_gpack = nullptr;        // G limbs in the A-operand lane layout
    uint32_t *d_kconst = nullptr;    // K_j
    uint32_t mfma_nq = 0;            // 64-symbol chunks of the message (0: no MFMA encode)
    bool mix_minus_one = false;      // a mix constant is 65536: the SIMD words are not the linear map
    int encode_mode = 0;             // 0 automatic, 1 butterflies, 2 MFMA wherever it applies
    // the syndromes as the same contraction (rsi16_synd_mfma_kernel) ahead of the repair
    uint4 *d_vpack = nullptr;        // limbs of w^(j rbo(m)), m < n, in the A-operand lane layout
    uint32_t *d_vconst = nullptr;    // their offset constants
    uint32_t synd_nq = 0, synd_nqm = 0;  // chunks in all, message chunks (0: not built)
    uint32_t *d_syn = nullptr;       // syndromes of the last repair launch
    size_t syn_size = 0;
    int repair_synd = 0;             // 0 automatic, 1 in the repair kernel, 2 MFMA ahead of it
};

namespace {

// Field tables as the reference builds them (galois.hpp:133-144 with gf_mul_mod):
// exp over i = 0..65536 inclusive, so log[1] = 65536 and log[0] = 0.
struct Gfi16Host {
    std::vector<uint32_t> exp_t, log_t;
    explicit Gfi16Host(uint32_t primitive) : exp_t(65537), log_t(65537, 0)
    {
        uint32_t x = 1;
        for (uint32_t i = 0; i < 65537; ++i) {
            exp_t[i] = x;
            log_t[x] = i;
            uint32_t r;
            if (__builtin_mul_overflow(x, primitive, &r))
                r = 1;
            x = r % P16;
        }
    }
    uint32_t inv(uint32_t a) const { return exp_t[65536 - log_t[a]]; }
    uint32_t div(uint32_t a, uint32_t b) const
    {
        if (a == 0)
            return 0;
        uint32_t r = log_t[a] + 65536 - log_t[b];
        if (r >= 65536)
            r -= 65536;
        return exp_t[r];
    }
    uint32_t pow(uint32_t a, uint32_t e) const { return exp_t[(uint32_t)(log_t[a] * e) % 65536u]; }
};

uint32_t group_size(uint32_t E) { return std::min<uint32_t>(64, std::max<uint32_t>(1, E / 2)); }

int check_count(size_t count)
{
    if (count > 0x7fffffffu)
        return fail(FFRS_ERR_INVALID, "too many codewords in one call");
    return FFRS_OK;
}

template <class O>
int launch_encode_lane(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t count, uint16_t *d_ecc)
{
    dim3 grid((unsigned)((count + 255) / 256));
    auto *k = rs->E == 2 ? rsi16_encode_lane_kernel<O, 2> : rs->E == 4 ? rsi16_encode_lane_kernel<O, 4>
              : rs->E == 8 ? rsi16_encode_lane_kernel<O, 8> : rs->E == 16 ? rsi16_encode_lane_kernel<O, 16>
              : rsi16_encode_lane_kernel<O, 32>;
    hipLaunchKernelGGL(k, grid, dim3(256), 0, rs->ctx->stream, d_msg, (uint32_t)count, d_ecc, rs->dev);
    FFRS_HIP(hipGetLastError());
    mark_kernel(rs->ctx);
    return FFRS_OK;
}

template <int E>
int launch_encode_reg(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t count, uint16_t *d_ecc, int vec)
{
    constexpr uint32_t cpw = 64 / RegShape<E>::G;
    const dim3 grid((unsigned)((count + cpw - 1) / cpw));
    if (vec)
        hipLaunchKernelGGL((rsi16_encode_reg_kernel<VectorOps, E>), grid, dim3(64), 0, rs->ctx->stream, d_msg,
                           (uint32_t)count, d_ecc, rs->dev);
    else
        hipLaunchKernelGGL((rsi16_encode_reg_kernel<ScalarOps, E>), grid, dim3(64), 0, rs->ctx->stream, d_msg,
                           (uint32_t)count, d_ecc, rs->dev);
    FFRS_HIP(hipGetLastError());
    mark_kernel(rs->ctx);
    return FFRS_OK;
}

// The MFMA contraction over `count` codewords, in launches of < 2^31 message bytes.
int launch_encode_mfma(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t count, uint16_t *d_ecc)
{
    const size_t k = rs->k, E = rs->E;
    const size_t step = std::max<size_t>(16, (((size_t)1 << 31) / (k * 2)) & ~(size_t)15);
    auto *kern = E == 16 ? rsi16_encode_mfma_kernel<1> : E == 32 ? rsi16_encode_mfma_kernel<2>
                                                                 : rsi16_encode_mfma_kernel<4>;
    for (size_t b = 0; b < count; b += step) {
        const size_t c = std::min(step, count - b);
        const size_t groups = (c + 15) / 16;
        const int grid = persistent_grid(rs->ctx, 8, (groups + 3) / 4);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, rs->ctx->stream, d_msg + b * k, (uint32_t)(c * k * 2),
                           (uint32_t)c, d_ecc + b * E, (const uint4 *)rs->d_gpack, (const uint32_t *)rs->d_kconst,
                           (uint32_t)k, (uint32_t)E, rs->mfma_nq);
        FFRS_HIP(hipGetLastError());
    }
    mark_kernel(rs->ctx);
    return FFRS_OK;
}

int launch_encode(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t count, uint16_t *d_ecc, int vec, uint32_t I = 1)
{
    marks_begin(rs->ctx);
    if (count == 0)
        return FFRS_OK;
    // the linear map on the matrix cores where it is the encode of the words asked for (and,
    // automatically, where it is measured faster than the butterflies: ecc_len <= 32)
    if (rs->mfma_nq && I == 1 && rs->encode_mode != 1 && (rs->encode_mode == 2 || rs->E <= 32) &&
        !(vec && rs->mix_minus_one) && (((uintptr_t)d_msg | (uintptr_t)d_ecc) & 15) == 0)
        return launch_encode_mfma(rs, d_msg, count, d_ecc);
    // one lane per codeword (g = 1): the register kernel for ecc_len <= 32, plain layout
    const bool aligned = rs->E < 8 || (((uintptr_t)d_msg | (uintptr_t)d_ecc) & 15) == 0;
    if (rs->g_enc == 1 && rs->E <= 32 && I == 1 && aligned)
        return vec ? launch_encode_lane<VectorOps>(rs, d_msg, count, d_ecc)
                   : launch_encode_lane<ScalarOps>(rs, d_msg, count, d_ecc);
    uint32_t g = rs->g_enc == 1 && rs->E <= 32 ? group_size(rs->E) : rs->g_enc, cpw = 64 / g;
    if (I == 1 && rs->E >= 2 && rs->E <= 256 && g == group_size(rs->E)) {  // the default shape: registers
        switch (rs->E) {
        case 2:
            return launch_encode_reg<2>(rs, d_msg, count, d_ecc, vec);
        case 4:
            return launch_encode_reg<4>(rs, d_msg, count, d_ecc, vec);
        case 8:
            return launch_encode_reg<8>(rs, d_msg, count, d_ecc, vec);
        case 16:
            return launch_encode_reg<16>(rs, d_msg, count, d_ecc, vec);
        case 32:
            return launch_encode_reg<32>(rs, d_msg, count, d_ecc, vec);
        case 64:
            return launch_encode_reg<64>(rs, d_msg, count, d_ecc, vec);
        case 128:
            return launch_encode_reg<128>(rs, d_msg, count, d_ecc, vec);
        default:
            return launch_encode_reg<256>(rs, d_msg, count, d_ecc, vec);
        }
    }
    size_t lds = (size_t)cpw * 2 * rs->E * 4;
    dim3 grid((unsigned)((count + cpw - 1) / cpw));
    if (int r = allow_lds(rsi16_encode_kernel<VectorOps>, lds))
        return r;
    if (int r = allow_lds(rsi16_encode_kernel<ScalarOps>, lds))
        return r;
    if (vec)
        hipLaunchKernelGGL(rsi16_encode_kernel<VectorOps>, grid, dim3(64), lds, rs->ctx->stream, d_msg,
                           (uint32_t)count, d_ecc, rs->dev, g, I);
    else
        hipLaunchKernelGGL(rsi16_encode_kernel<ScalarOps>, grid, dim3(64), lds, rs->ctx->stream, d_msg,
                           (uint32_t)count, d_ecc, rs->dev, g, I);
    FFRS_HIP(hipGetLastError());
    mark_kernel(rs->ctx);
    return FFRS_OK;
}

template <int E>
int launch_repair_reg(ffrs_rsi16 *rs, uint16_t *d_msg, uint16_t *d_ecc, size_t count, uint8_t *d_status)
{
    constexpr uint32_t cpw = 64 / RegShape<E>::G;
    const size_t lds = (size_t)cpw * REG_WORDS * E * 4;
    // the syndromes on the matrix cores first, where their tables exist (ecc 16, 32, 64)
    const bool msyn = rs->synd_nq && rs->repair_synd != 1 && (((uintptr_t)d_msg | (uintptr_t)d_ecc) & 15) == 0;
    if (msyn) {
        const size_t step = std::max<size_t>(16, (((size_t)1 << 31) / (rs->k * 2)) & ~(size_t)15);
        const size_t need = std::min(count, step) * E * 4;
        if (rs->syn_size < need) {
            if (rs->d_syn)
                FFRS_HIP(hipFree(rs->d_syn));
            rs->d_syn = nullptr;
            rs->syn_size = 0;
            hipError_t e = hipMalloc(&rs->d_syn, need);
            if (e != hipSuccess)
                return fail(FFRS_ERR_NOMEM, "syndrome workspace of %zu bytes: %s", need, hipGetErrorString(e));
            rs->syn_size = need;
        }
        constexpr int CT = E >= 16 && E <= 64 ? E / 16 : 1;  // every tile in one pass (tables: ecc 16 .. 64)
        auto *ks = rsi16_synd_mfma_kernel<CT>;
        if (int r = allow_lds(rsi16_repair_reg_kernel<E, true>, lds))
            return r;
        for (size_t b = 0; b < count; b += step) {
            const size_t c = std::min(step, count - b);
            const size_t groups = (c + 15) / 16;
            hipLaunchKernelGGL(ks, dim3(persistent_grid(rs->ctx, 8, (groups + 3) / 4)), dim3(256), 0, rs->ctx->stream,
                               (const uint16_t *)(d_msg + b * rs->k), (uint32_t)(c * rs->k * 2),
                               (const uint16_t *)(d_ecc + b * E), (uint32_t)(c * E * 2), (uint32_t)c,
                               (const uint4 *)rs->d_vpack, (const uint32_t *)rs->d_vconst, rs->k, (uint32_t)E,
                               rs->synd_nq, rs->synd_nqm, rs->d_syn);
            FFRS_HIP(hipGetLastError());
            mark_kernel(rs->ctx);
            hipLaunchKernelGGL((rsi16_repair_reg_kernel<E, true>), dim3((unsigned)((c + cpw - 1) / cpw)), dim3(64), lds,
                               rs->ctx->stream, d_msg + b * rs->k, d_ecc + b * E, (uint32_t)c, d_status + b, rs->dev,
                               (const uint32_t *)rs->d_syn);
            FFRS_HIP(hipGetLastError());
            mark_kernel(rs->ctx);
        }
        return FFRS_OK;
    }
    if (int r = allow_lds(rsi16_repair_reg_kernel<E>, lds))
        return r;
    hipLaunchKernelGGL(rsi16_repair_reg_kernel<E>, dim3((unsigned)((count + cpw - 1) / cpw)), dim3(64), lds,
                       rs->ctx->stream, d_msg, d_ecc, (uint32_t)count, d_status, rs->dev);
    FFRS_HIP(hipGetLastError());
    mark_kernel(rs->ctx);
    return FFRS_OK;
}

int launch_repair(ffrs_rsi16 *rs, uint16_t *d_msg, uint16_t *d_ecc, size_t count, uint8_t *d_status)
{
    marks_begin(rs->ctx);
    if (count == 0)
        return FFRS_OK;
    if (rs->E >= 2 && rs->E <= 256 && rs->g_rep == group_size(rs->E)) {  // the default shape: registers
        switch (rs->E) {
        case 2:
            return launch_repair_reg<2>(rs, d_msg, d_ecc, count, d_status);
        case 4:
            return launch_repair_reg<4>(rs, d_msg, d_ecc, count, d_status);
        case 8:
            return launch_repair_reg<8>(rs, d_msg, d_ecc, count, d_status);
        case 16:
            return launch_repair_reg<16>(rs, d_msg, d_ecc, count, d_status);
        case 32:
            return launch_repair_reg<32>(rs, d_msg, d_ecc, count, d_status);
        case 64:
            return launch_repair_reg<64>(rs, d_msg, d_ecc, count, d_status);
        case 128:
            return launch_repair_reg<128>(rs, d_msg, d_ecc, count, d_status);
        default:
            return launch_repair_reg<256>(rs, d_msg, d_ecc, count, d_status);
        }
    }
    uint32_t g = rs->g_rep, cpw = 64 / g;
    size_t lds = (size_t)cpw * REPAIR_WORDS * rs->E * 4;
    dim3 grid((unsigned)((count + cpw - 1) / cpw));
    if (int r = allow_lds(rsi16_repair_kernel, lds))
        return r;
    hipLaunchKernelGGL(rsi16_repair_kernel, grid, dim3(64), lds, rs->ctx->stream, d_msg, d_ecc, (uint32_t)count,
                       d_status, rs->dev, g);
    FFRS_HIP(hipGetLastError());
    mark_kernel(rs->ctx);
    return FFRS_OK;
}

}  // namespace

namespace {

// NTT::gs_butterfly, ntt.hpp:280-302 (natural in, RBO out), scalar words
void gs_host(const std::vector<uint32_t> &roots, uint32_t *x, uint32_t len)
{
    for (uint32_t st = len / 2, ef = 0; st > 0; st /= 2, ++ef)
        for (uint32_t b = 0; b < len; b += 2 * st)
            for (uint32_t i = b; i < b + st; ++i) {
                const uint32_t a = x[i], c = x[i + st];
                x[i] = S::add(a, c);
                x[i + st] = i == b ? S::sub(a, c) : S::mul(S::sub(a, c), roots[(i - b) << ef]);
            }
}

// The tables of rsi16_encode_mfma_kernel: G[m] = the parity of the unit message at m, i.e. the GS
// butterfly of mix_i X_m^i with X_m = w^rbo(m) (the message residue of rsi16v_impl.hpp:93-97 is
// sum_m v_m X_m^i, then :108-119), split into three 7-bit limbs and packed in the A-operand
// lane order, and K_j. Shapes with ecc_len not a multiple of 16, above 256, or message_len above
// 16384 keep the butterfly kernels.
int build_mfma_encode(ffrs_rsi16 *rs, const std::vector<uint32_t> &mix, const std::vector<uint32_t> &roots_i_ecc,
                      const std::vector<uint32_t> &roots_ntt)
{
    const uint32_t k = rs->k, E = rs->E;
    if (E % 16 || E > 256 || k > 16384)
        return FFRS_OK;
    const uint32_t nq = (k + 63) / 64, T = E / 16, kp = nq * 64;
    std::vector<uint8_t> limb((size_t)3 * kp * E, 0);  // [b][m][j]
    std::vector<uint64_t> C((size_t)3 * E, 0);
    std::vector<uint32_t> row(E);
    for (uint32_t m = 0; m < k; ++m) {
        const uint32_t x = roots_ntt[rbo_bits(m, rs->nbits)];
        uint32_t pw = 1;
        for (uint32_t i = 0; i < E; ++i) {
            row[i] = S::mul(mix[i], pw);
            pw = S::mul(pw, x);
        }
        gs_host(roots_i_ecc, row.data(), E);
        for (uint32_t j = 0; j < E; ++j) {
            const uint32_t g = row[j], gb[3] = {g & 127u, (g >> 7) & 127u, g >> 14};
            for (int b = 0; b < 3; ++b) {
                limb[((size_t)b * kp + m) * E + j] = (uint8_t)gb[b];
                C[(size_t)b * E + j] += gb[b];
            }
        }
    }
    std::vector<uint8_t> pack((size_t)3 * T * nq * 64 * 16);
    for (uint32_t b = 0; b < 3; ++b)
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t q = 0; q < nq; ++q)
                for (uint32_t l = 0; l < 64; ++l)
                    for (uint32_t bt = 0; bt < 16; ++bt) {
                        const uint32_t m = 64 * q + 16 * (l / 16) + bt, j = 16 * t + l % 16;
                        pack[(((size_t)(b * T + t) * nq + q) * 64 + l) * 16 + bt] = limb[((size_t)b * kp + m) * E + j];
                    }
    std::vector<uint32_t> K(E);
    for (uint32_t j = 0; j < E; ++j) {
        const uint64_t sum = (C[j] + 128 * C[E + j] + 16384 * C[2 * (size_t)E + j]) % P16;
        K[j] = (uint32_t)(sum * 128 % P16 * 257 % P16);
    }
    hipError_t e = hipMalloc(&rs->d_gpack, pack.size());
    if (e == hipSuccess)
        e = hipMalloc(&rs->d_kconst, K.size() * 4);
    if (e == hipSuccess)
        e = hipMemcpy(rs->d_gpack, pack.data(), pack.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(rs->d_kconst, K.data(), K.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(FFRS_ERR_DEVICE, "MFMA encode tables: %s", hipGetErrorString(e));
    rs->mfma_nq = nq;
    return FFRS_OK;
}

// The syndrome matrix V[m][j] = w^(j rbo(m)) (NTT::pntt, ntt.hpp:133-152) for the MFMA
// contraction: rows m < k in ceil(k / 64) message chunks, rows k + i (i < E) in the parity
// chunks after them, packed and offset as build_mfma_encode does.
int build_mfma_synd(ffrs_rsi16 *rs, const std::vector<uint32_t> &roots_ntt)
{
    const uint32_t k = rs->k, E = rs->E;
    if (E % 16 || E > 64 || k > 16384)
        return FFRS_OK;
    const uint32_t nqm = (k + 63) / 64, nqe = (E + 63) / 64, nq = nqm + nqe, T = E / 16, kp = nq * 64;
    std::vector<uint8_t> limb((size_t)3 * kp * E, 0);  // [b][row][j], rows in chunk order
    std::vector<uint64_t> C((size_t)3 * E, 0);
    for (uint32_t m = 0; m < k + E; ++m) {
        const uint32_t r = m < k ? m : nqm * 64 + (m - k);  // row of the packed matrix
        const uint32_t x = roots_ntt[rbo_bits(m, rs->nbits)];
        uint32_t pw = 1;
        for (uint32_t j = 0; j < E; ++j) {
            const uint32_t g = pw, gb[3] = {g & 127u, (g >> 7) & 127u, g >> 14};
            for (int b = 0; b < 3; ++b) {
                limb[((size_t)b * kp + r) * E + j] = (uint8_t)gb[b];
                C[(size_t)b * E + j] += gb[b];
            }
            pw = S::mul(pw, x);
        }
    }
    std::vector<uint8_t> pack((size_t)3 * T * nq * 64 * 16);
    for (uint32_t b = 0; b < 3; ++b)
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t q = 0; q < nq; ++q)
                for (uint32_t l = 0; l < 64; ++l)
                    for (uint32_t bt = 0; bt < 16; ++bt) {
                        const uint32_t m = 64 * q + 16 * (l / 16) + bt, j = 16 * t + l % 16;
                        pack[(((size_t)(b * T + t) * nq + q) * 64 + l) * 16 + bt] = limb[((size_t)b * kp + m) * E + j];
                    }
    std::vector<uint32_t> K(E);
    for (uint32_t j = 0; j < E; ++j) {
        const uint64_t sum = (C[j] + 128 * C[E + j] + 16384 * C[2 * (size_t)E + j]) % P16;
        K[j] = (uint32_t)(sum * 128 % P16 * 257 % P16);
    }
    hipError_t e = hipMalloc(&rs->d_vpack, pack.size());
    if (e == hipSuccess)
        e = hipMalloc(&rs->d_vconst, K.size() * 4);
    if (e == hipSuccess)
        e = hipMemcpy(rs->d_vpack, pack.data(), pack.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(rs->d_vconst, K.data(), K.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(FFRS_ERR_DEVICE, "MFMA syndrome tables: %s", hipGetErrorString(e));
    rs->synd_nq = nq;
    rs->synd_nqm = nqm;
    return FFRS_OK;
}

}  // namespace

extern "C" {

int ffrs_rsi16_create(ffrs_ctx *ctx, size_t block_len, size_t ecc_len, uint32_t primitive, ffrs_rsi16 **out)
{
    if (!ctx || !out)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (ecc_len < 2 || (ecc_len & (ecc_len - 1)))
        return fail(FFRS_ERR_PARAM, "ecc_len must be a power of 2: %zu", ecc_len);
    if (block_len % ecc_len || block_len <= ecc_len)
        return fail(FFRS_ERR_PARAM, "block_len must be a multiple of ecc_len larger than it (%zu, %zu)", block_len,
                    ecc_len);
    if (block_len > 65536)
        return fail(FFRS_ERR_PARAM, "ntt root of unity sanity check failed (block_len > 65536)");
    if (ecc_len > RSI16_MAX_ECC)
        return fail(FFRS_ERR_PARAM, "ecc_len %zu above this build's limit of %u", ecc_len, RSI16_MAX_ECC);
    if (primitive < 2 || primitive > 65536)
        return fail(FFRS_ERR_PARAM, "%u is not a primitive element of GF(65537)", primitive);
    Gfi16Host gf(primitive);
    if (!(gf.pow(primitive, 32768) == 65536 && gf.pow(primitive, 65536) == 1))
        return fail(FFRS_ERR_PARAM, "%u is not a primitive element of GF(65537)", primitive);

    // ntt.hpp:55-123
    uint32_t n = (uint32_t)block_len, E = (uint32_t)ecc_len;
    uint32_t N = 2;
    while (N < n)
        N <<= 1;
    uint32_t nbits = (uint32_t)__builtin_ctz(N);
    uint32_t root = gf.pow(primitive, 65536 / N);
    if (gf.pow(root, N) != 1 || gf.pow(root, N / 2) != 65536)
        return fail(FFRS_ERR_PARAM, "ntt root of unity sanity check failed");
    uint32_t root_i = gf.inv(root);
    uint32_t er = gf.pow(primitive, 65536 / E), er_i = gf.inv(er);
    uint32_t e2 = gf.pow(primitive, 65536 / (2 * E)), e2_i = gf.inv(e2);
    if (gf.pow(er, E) != 1 || gf.pow(er, E / 2) != 65536 || gf.pow(e2, 2 * E) != 1 || gf.pow(e2, E) != 65536)
        return fail(FFRS_ERR_PARAM, "ntt root of unity sanity check failed");

    std::vector<uint32_t> roots_ecc(E), roots_i_ecc(E), roots_ecc2(2 * E), roots_i_ecc2(2 * E), pntt_shift(n),
        pintt_shift(N), roots_ntt(N), mix(E), rbo_ecc(E);
    for (uint32_t i = 0; i < N; ++i)
        roots_ntt[i] = gf.pow(root, i);
    for (uint32_t i = 0; i < E; ++i) {
        roots_ecc[i] = gf.pow(er, i);
        roots_i_ecc[i] = gf.pow(er_i, i);
        rbo_ecc[i] = rbo_bits(i * (N / E), nbits);
    }
    for (uint32_t i = 0; i < 2 * E; ++i) {
        roots_ecc2[i] = gf.pow(e2, i);
        roots_i_ecc2[i] = gf.pow(e2_i, i);
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t blk = (i / E) * E, j = i - blk;
        pntt_shift[i] = gf.pow(root, j * rbo_bits(blk, nbits));
    }
    for (uint32_t i = 0; i < N; ++i) {
        uint32_t blk = (i / E) * E, j = i - blk;
        pintt_shift[i] = gf.div(gf.pow(root_i, j * rbo_bits(blk, nbits)), N);
    }
    // rsi16v_impl.hpp:68-86
    uint32_t mw = roots_ntt[rbo_bits(n - E, nbits)];
    uint32_t mw_i = gf.inv(mw);
    for (uint32_t i = 0; i < E; ++i)
        mix[i] = s_neg(gf.div(gf.pow(mw_i, i), E));

    auto *rs = new ffrs_rsi16();
    rs->ctx = ctx;
    rs->n = n;
    rs->k = n - E;
    rs->E = E;
    rs->N = N;
    rs->nbits = nbits;
    rs->primitive = primitive;
    rs->root = root;
    // field inverses with the table convention inv(0) = 1, from powers of the primitive root 3
    std::vector<uint32_t> inv_tab(P16);
    {
        std::vector<uint32_t> pw(P16 - 1);
        uint32_t x = 1;
        for (uint32_t i = 0; i < P16 - 1; ++i) {
            pw[i] = x;
            x = S::mul(x, 3);
        }
        inv_tab[0] = 1;
        for (uint32_t i = 0; i < P16 - 1; ++i)
            inv_tab[pw[i]] = pw[(P16 - 1 - i) & 0xffff];
    }
    size_t words = 6 * (size_t)E + 2 * (size_t)E * 2 + n + 2 * (size_t)N + P16;
    if (int r = ctx_set_device(ctx)) {
        delete rs;
        return r;
    }
    hipError_t herr = hipMalloc(&rs->d_tables, words * 4);
    if (herr != hipSuccess) {
        delete rs;
        return fail(FFRS_ERR_NOMEM, "table allocation failed: %s", hipGetErrorString(herr));
    }
    std::vector<uint32_t> host;
    host.reserve(words);
    auto put = [&](const std::vector<uint32_t> &v) {
        size_t off = host.size();
        host.insert(host.end(), v.begin(), v.end());
        return rs->d_tables + off;
    };
    Rsi16Dev &d = rs->dev;
    d.n = n;
    d.k = n - E;
    d.E = E;
    d.N = N;
    d.nbits = nbits;
    d.ecc_inv = gf.inv(E);
    d.ecc2_inv = gf.inv(2 * E);
    d.roots_ecc = put(roots_ecc);
    d.roots_i_ecc = put(roots_i_ecc);
    d.roots_ecc2 = put(roots_ecc2);
    d.roots_i_ecc2 = put(roots_i_ecc2);
    d.pntt_shift = put(pntt_shift);
    d.pintt_shift = put(pintt_shift);
    d.roots_ntt = put(roots_ntt);
    d.mix = put(mix);
    d.rbo_ecc = put(rbo_ecc);
    d.inv_tab = put(inv_tab);
    herr = hipMemcpy(rs->d_tables, host.data(), host.size() * 4, hipMemcpyHostToDevice);
    if (herr != hipSuccess) {
        ffrs_rsi16_destroy(rs);
        return fail(FFRS_ERR_DEVICE, "table upload failed: %s", hipGetErrorString(herr));
    }
    rs->g_rep = group_size(rs->E);
    rs->g_enc = rs->E <= 32 ? 1u : group_size(rs->E);  // register kernel, one codeword per lane
    rs->mix_minus_one = std::any_of(mix.begin(), mix.end(), [](uint32_t v) { return v == 65536; });
    if (int r = build_mfma_encode(rs, mix, roots_i_ecc, roots_ntt)) {
        ffrs_rsi16_destroy(rs);
        return r;
    }
    if (int r = build_mfma_synd(rs, roots_ntt)) {
        ffrs_rsi16_destroy(rs);
        return r;
    }
    *out = rs;
    return FFRS_OK;
}

int ffrs_rsi16_set_repair_syndromes(ffrs_rsi16 *rs, int mode)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (mode < 0 || mode > 2)
        return fail(FFRS_ERR_INVALID, "syndrome mode %d (0 automatic, 1 in the repair kernel, 2 MFMA where it applies)",
                    mode);
    rs->repair_synd = mode;
    return FFRS_OK;
}

int ffrs_rsi16_set_encode_kernel(ffrs_rsi16 *rs, int kernel)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (kernel < 0 || kernel > 2)
        return fail(FFRS_ERR_INVALID, "encode kernel %d (0 automatic, 1 butterflies, 2 MFMA where it applies)",
                    kernel);
    rs->encode_mode = kernel;
    return FFRS_OK;
}

void ffrs_rsi16_destroy(ffrs_rsi16 *rs)
{
    if (!rs)
        return;
    (void)hipSetDevice(rs->ctx->device);
    if (rs->d_tables)
        (void)hipFree(rs->d_tables);
    if (rs->d_work)
        (void)hipFree(rs->d_work);
    if (rs->d_gpack)
        (void)hipFree(rs->d_gpack);
    if (rs->d_kconst)
        (void)hipFree(rs->d_kconst);
    if (rs->d_vpack)
        (void)hipFree(rs->d_vpack);
    if (rs->d_vconst)
        (void)hipFree(rs->d_vconst);
    if (rs->d_syn)
        (void)hipFree(rs->d_syn);
    delete rs;
}

uint32_t ffrs_rsi16_root(const ffrs_rsi16 *rs) { return rs ? rs->root : 0; }

int ffrs_rsi16_set_groups(ffrs_rsi16 *rs, unsigned g_encode, unsigned g_repair)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    unsigned cap = std::min<unsigned>(64, rs->E);
    for (unsigned g : {g_encode, g_repair})
        if (g && ((g & (g - 1)) || g > cap))
            return fail(FFRS_ERR_INVALID, "lanes per codeword must be a power of two <= %u, got %u", cap, g);
    if (g_repair && (size_t)(64 / g_repair) * REPAIR_WORDS * rs->E * 4 > 160 * 1024)
        return fail(FFRS_ERR_INVALID, "%u lanes per codeword need more LDS than a CU has", g_repair);
    if (g_encode)
        rs->g_enc = g_encode;
    if (g_repair)
        rs->g_rep = g_repair;
    return FFRS_OK;
}

int ffrs_rsi16_get_groups(const ffrs_rsi16 *rs, unsigned *g_encode, unsigned *g_repair)
{
    if (!rs || !g_encode || !g_repair)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    *g_encode = rs->g_enc;
    *g_repair = rs->g_rep;
    return FFRS_OK;
}
size_t ffrs_rsi16_ntt_len(const ffrs_rsi16 *rs) { return rs ? rs->N : 0; }

int ffrs_rsi16_encode_device(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t count, uint16_t *d_ecc, int vector_lanes)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (int r = check_count(count))
        return r;
    if (count && (!d_msg || !d_ecc))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = ctx_set_device(rs->ctx))
        return r;
    return launch_encode(rs, d_msg, count, d_ecc, vector_lanes);
}

int ffrs_rsi16_repair_device(ffrs_rsi16 *rs, uint16_t *d_msg, uint16_t *d_ecc, size_t count, uint8_t *d_status)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (int r = check_count(count))
        return r;
    if (count && (!d_msg || !d_ecc || !d_status))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = ctx_set_device(rs->ctx))
        return r;
    return launch_repair(rs, d_msg, d_ecc, count, d_status);
}

int ffrs_rsi16_encode(ffrs_rsi16 *rs, const uint16_t *msg, size_t count, uint16_t *ecc, int vector_lanes)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (int r = check_count(count))
        return r;
    if (count == 0)
        return FFRS_OK;
    if (!msg || !ecc)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = count * rs->k * 2, eb = count * rs->E * 2;
    void *d_m, *d_e;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    if (int r = launch_encode(rs, (const uint16_t *)d_m, count, (uint16_t *)d_e, vector_lanes))
        return r;
    FFRS_HIP(hipMemcpyAsync(ecc, d_e, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rsi16_encode_interleaved_device(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t blocks, size_t interleave,
                                         uint16_t *d_ecc, int vector_lanes)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (interleave == 0 || interleave > 0xffffu)
        return fail(FFRS_ERR_INVALID, "interleave must be in [1, 65535], got %zu", interleave);
    if (int r = check_count(blocks * interleave))
        return r;
    if (blocks && (!d_msg || !d_ecc))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = ctx_set_device(rs->ctx))
        return r;
    return launch_encode(rs, d_msg, blocks * interleave, d_ecc, vector_lanes, (uint32_t)interleave);
}

int ffrs_rsi16_encode_interleaved(ffrs_rsi16 *rs, const uint16_t *msg, size_t blocks, size_t interleave, uint16_t *ecc,
                                  int vector_lanes)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (interleave == 0 || interleave > 0xffffu)
        return fail(FFRS_ERR_INVALID, "interleave must be in [1, 65535], got %zu", interleave);
    if (int r = check_count(blocks * interleave))
        return r;
    if (blocks == 0)
        return FFRS_OK;
    if (!msg || !ecc)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = blocks * interleave * rs->k * 2, eb = blocks * interleave * rs->E * 2;
    void *d_m, *d_e;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    if (int r = launch_encode(rs, (const uint16_t *)d_m, blocks * interleave, (uint16_t *)d_e, vector_lanes,
                              (uint32_t)interleave))
        return r;
    FFRS_HIP(hipMemcpyAsync(ecc, d_e, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rsi16_repair_interleaved_simd(ffrs_rsi16 *rs, uint16_t *msg, uint16_t *ecc, size_t interleave,
                                       unsigned width, uint8_t *status)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (width != 4 && width != 8 && width != 16)
        return fail(FFRS_ERR_INVALID, "SIMD width must be 4, 8 or 16, got %u", width);
    if (interleave == 0 || interleave > 0xffffu)
        return fail(FFRS_ERR_INVALID, "interleave must be in [1, 65535], got %zu", interleave);
    if (!msg || !ecc || !status)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    const uint32_t W = width, il = (uint32_t)interleave;
    const size_t mb = (size_t)rs->k * il * 2, eb = (size_t)rs->E * il * 2;
    const uint32_t groups = (il + W - 1) / W;
    const size_t rows = (size_t)rs->n + 8 * (size_t)rs->E;  // per-lane workspace, u32 each
    // lanes per launch: every group of the call, or as many as 256 MB of workspace hold
    size_t lanes_cap = std::max<size_t>(64, ((size_t)256 << 20) / (rows * 4) / 64 * 64);
    const uint32_t gpb = (uint32_t)std::min<size_t>(groups, lanes_cap / W);  // groups per launch
    void *d_m, *d_e, *d_s;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    if (int r = stage_reserve(ctx, 2, groups, &d_s))
        return r;
    const size_t wbytes = (size_t)gpb * W * rows * 4;
    if (rs->work_size < wbytes) {
        if (rs->d_work)
            FFRS_HIP(hipFree(rs->d_work));
        rs->d_work = nullptr;
        rs->work_size = 0;
        hipError_t e = hipMalloc(&rs->d_work, wbytes);
        if (e != hipSuccess)
            return fail(FFRS_ERR_NOMEM, "repair workspace of %zu bytes: %s", wbytes, hipGetErrorString(e));
        rs->work_size = wbytes;
    }
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_e, ecc, eb, hipMemcpyHostToDevice, ctx->stream));
    std::vector<uint8_t> gst(groups);
    uint8_t res = FFRS_RSI16_NO_ERRORS;
    bool raised = false;
    for (uint32_t g0 = 0; g0 < groups && !raised; g0 += gpb) {
        const uint32_t ng = std::min(gpb, groups - g0), lanes = ng * W;
        const dim3 grid((lanes + 63) / 64);
        auto *rk = W == 4 ? rsi16_repair_simd_kernel<4> : W == 8 ? rsi16_repair_simd_kernel<8>
                                                                  : rsi16_repair_simd_kernel<16>;
        hipLaunchKernelGGL(rk, grid, dim3(64), 0, ctx->stream, rs->dev, (const uint16_t *)d_m,
                           (const uint16_t *)d_e, il, 0u, il, g0, lanes, (uint32_t *)rs->d_work, (uint8_t *)d_s);
        FFRS_HIP(hipGetLastError());
        FFRS_HIP(hipMemcpyAsync(gst.data() + g0, (uint8_t *)d_s + g0, ng, hipMemcpyDeviceToHost, ctx->stream));
        FFRS_HIP(hipStreamSynchronize(ctx->stream));
        // groups are written back in order up to the first that raises (pyrsi16.hpp:410-430)
        uint32_t end = g0;
        for (; end < g0 + ng && gst[end] != FFRS_RSI16_RAISES; ++end)
            res = std::max(res, gst[end]);
        raised = end < g0 + ng;
        if (end > g0) {
            auto *wk = W == 4 ? rsi16_simd_writeback_kernel<4> : W == 8 ? rsi16_simd_writeback_kernel<8>
                                                                        : rsi16_simd_writeback_kernel<16>;
            hipLaunchKernelGGL(wk, grid, dim3(256), 0, ctx->stream, rs->dev, (uint16_t *)d_m, (uint16_t *)d_e, il, 0u,
                               il, g0, end, lanes, (const uint32_t *)rs->d_work);
            FFRS_HIP(hipGetLastError());
        }
    }
    FFRS_HIP(hipMemcpyAsync(msg, d_m, mb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(ecc, d_e, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    *status = raised ? (uint8_t)FFRS_RSI16_RAISES : res;
    return FFRS_OK;
}

int ffrs_rsi16_repair(ffrs_rsi16 *rs, uint16_t *msg, uint16_t *ecc, size_t count, uint8_t *status)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (int r = check_count(count))
        return r;
    if (count == 0)
        return FFRS_OK;
    if (!msg || !ecc || !status)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = count * rs->k * 2, eb = count * rs->E * 2;
    void *d_m, *d_e, *d_s;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    if (int r = stage_reserve(ctx, 2, count, &d_s))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_e, ecc, eb, hipMemcpyHostToDevice, ctx->stream));
    if (int r = launch_repair(rs, (uint16_t *)d_m, (uint16_t *)d_e, count, (uint8_t *)d_s))
        return r;
    FFRS_HIP(hipMemcpyAsync(msg, d_m, mb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(ecc, d_e, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(status, d_s, count, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rsi16_repair_known(ffrs_rsi16 *rs, uint16_t *msg, uint16_t *ecc, const uint32_t *positions, size_t npos,
                            uint8_t *status)
{
    if (!rs || !status || (npos && (!msg || !ecc || !positions)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (npos == 0) {
        *status = FFRS_RSI16_NO_ERRORS;
        return FFRS_OK;
    }
    if (npos > rs->E)
        return fail(FFRS_ERR_SIZE, "%zu error positions exceed ecc_len %u", npos, rs->E);
    for (size_t i = 0; i < npos; ++i)
        if (positions[i] >= rs->n)
            return fail(FFRS_ERR_INVALID, "error position %u >= block_len %u", positions[i], rs->n);
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = rs->k * 2, eb = rs->E * 2;
    void *d_m, *d_e, *d_p;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb + 64, &d_e))
        return r;
    if (int r = stage_reserve(ctx, 2, npos * 4 + 64, &d_p))
        return r;
    uint8_t *d_s = (uint8_t *)d_e + eb;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_e, ecc, eb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_p, positions, npos * 4, hipMemcpyHostToDevice, ctx->stream));
    uint32_t g = group_size(rs->E);
    size_t lds = (size_t)7 * rs->E * 4;
    if (int r = allow_lds(rsi16_repair_known_kernel, lds))
        return r;
    hipLaunchKernelGGL(rsi16_repair_known_kernel, dim3(1), dim3(64), lds, ctx->stream, (uint16_t *)d_m,
                       (uint16_t *)d_e, (const uint32_t *)d_p, (uint32_t)npos, d_s, rs->dev, g);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(msg, d_m, mb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(ecc, d_e, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(status, d_s, 1, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rsi16_synds(ffrs_rsi16 *rs, const uint16_t *msg, const uint16_t *ecc, size_t count, uint32_t *synds)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (int r = check_count(count))
        return r;
    if (count == 0)
        return FFRS_OK;
    if (!msg || !ecc || !synds)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = count * rs->k * 2, eb = count * rs->E * 2, ob = count * rs->E * 4;
    void *d_m, *d_e, *d_o;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    if (int r = stage_reserve(ctx, 2, ob, &d_o))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_e, ecc, eb, hipMemcpyHostToDevice, ctx->stream));
    uint32_t g = group_size(rs->E), cpw = 64 / g;
    size_t lds = (size_t)cpw * 2 * rs->E * 4;
    if (int r = allow_lds(rsi16_synds_kernel, lds))
        return r;
    hipLaunchKernelGGL(rsi16_synds_kernel, dim3((unsigned)((count + cpw - 1) / cpw)), dim3(64), lds, ctx->stream,
                       (const uint16_t *)d_m, (const uint16_t *)d_e, (uint32_t)count, (uint32_t *)d_o, rs->dev, g);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(synds, d_o, ob, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rsi16_sugiyama(ffrs_rsi16 *rs, const uint32_t *synds, size_t count, uint32_t *locator, uint32_t *evaluator,
                        uint32_t *length)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (int r = check_count(count))
        return r;
    if (count == 0)
        return FFRS_OK;
    if (!synds || !locator || !evaluator || !length)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t sb = count * rs->E * 4;
    void *d_s, *d_o, *d_l;
    if (int r = stage_reserve(ctx, 0, sb, &d_s))
        return r;
    if (int r = stage_reserve(ctx, 1, 2 * sb, &d_o))
        return r;
    if (int r = stage_reserve(ctx, 2, count * 4, &d_l))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_s, synds, sb, hipMemcpyHostToDevice, ctx->stream));
    uint32_t g = group_size(rs->E), cpw = 64 / g;
    size_t lds = (size_t)cpw * (2 + SUGI_WORDS) * rs->E * 4;
    uint32_t *lo = (uint32_t *)d_o, *eo = lo + count * rs->E;
    if (int r = allow_lds(rsi16_sugiyama_kernel, lds))
        return r;
    hipLaunchKernelGGL(rsi16_sugiyama_kernel, dim3((unsigned)((count + cpw - 1) / cpw)), dim3(64), lds, ctx->stream,
                       (const uint32_t *)d_s, (uint32_t)count, lo, eo, (uint32_t *)d_l, rs->dev, g);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(locator, lo, sb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(evaluator, eo, sb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(length, d_l, count * 4, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rsi16_ntt_op(ffrs_rsi16 *rs, int op, unsigned width, const uint32_t *a, const uint32_t *b, uint32_t n,
                      uint32_t *out, uint32_t *out2, uint32_t *result)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (width != 1 && width != 4 && width != 8 && width != 16)
        return fail(FFRS_ERR_INVALID, "width %u (1, 4, 8 or 16)", width);
    if (op < NTTOP_NTT || op > NTTOP_SUGIYAMA)
        return fail(FFRS_ERR_INVALID, "operation %d", op);
    const bool two = op == NTTOP_MUL || op == NTTOP_DIV;
    if (!a || !out || !result || (two && !b) || (op == NTTOP_SUGIYAMA && !out2))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    const uint32_t W = width, E = rs->E, N = rs->N;
    if (op == NTTOP_INV && n > E)
        return fail(FFRS_ERR_PARAM, "n = %u (0 .. ecc_len)", n);
    const uint32_t len = op <= NTTOP_INTTR ? N : E;
    const size_t region = (size_t)len * W * 4, wbytes = 9 * region + (size_t)N * 4 + 64;
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    if (rs->work_size < wbytes) {
        if (rs->d_work)
            FFRS_HIP(hipFree(rs->d_work));
        rs->d_work = nullptr;
        rs->work_size = 0;
        hipError_t e = hipMalloc(&rs->d_work, wbytes);
        if (e != hipSuccess)
            return fail(FFRS_ERR_NOMEM, "workspace of %zu bytes: %s", wbytes, hipGetErrorString(e));
        rs->work_size = wbytes;
    }
    // inputs vector-major [W][len] -> element-major [len][W]
    std::vector<uint32_t> h(2 * (size_t)len * W, 0);
    for (uint32_t l = 0; l < W; ++l)
        for (uint32_t i = 0; i < len; ++i) {
            h[(size_t)i * W + l] = a[(size_t)l * len + i];
            if (two)
                h[(size_t)len * W + (size_t)i * W + l] = b[(size_t)l * len + i];
        }
    Gfi16Host gfh(rs->primitive);
    std::vector<uint32_t> ri(N);  // powers of the inverse root (ntt.hpp:74-79)
    const uint64_t rinv = gfh.inv(rs->root);
    ri[0] = 1;
    for (uint32_t i = 1; i < N; ++i)
        ri[i] = (uint32_t)(ri[i - 1] * rinv % P16);
    const uint32_t ntt_len_i = gfh.inv(N % P16);
    uint32_t *ws = (uint32_t *)rs->d_work, *d_ri = ws + 9 * (size_t)len * W, *d_res = d_ri + N;
    FFRS_HIP(hipMemcpyAsync(ws, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_ri, ri.data(), ri.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    auto *kern = W == 1 ? rsi16_ntt_util_kernel<1> : W == 4 ? rsi16_ntt_util_kernel<4>
                 : W == 8 ? rsi16_ntt_util_kernel<8> : rsi16_ntt_util_kernel<16>;
    hipLaunchKernelGGL(kern, dim3(1), dim3(W), 0, ctx->stream, rs->dev, op, ws, len, (const uint32_t *)d_ri, ntt_len_i,
                       n, d_res);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(h.data(), ws, 2 * region, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(result, d_res, W * 4, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t l = 0; l < W; ++l)
        for (uint32_t i = 0; i < len; ++i) {
            out[(size_t)l * len + i] = h[(size_t)i * W + l];
            if (op == NTTOP_SUGIYAMA)
                out2[(size_t)l * len + i] = h[(size_t)len * W + (size_t)i * W + l];
        }
    return FFRS_OK;
}

int ffrs_rsi16_roots(ffrs_rsi16 *rs, const uint32_t *poly, uint8_t *is_root)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (!poly || !is_root)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    void *d_p, *d_h;
    if (int r = stage_reserve(ctx, 0, rs->E * 4, &d_p))
        return r;
    if (int r = stage_reserve(ctx, 1, rs->n, &d_h))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_p, poly, rs->E * 4, hipMemcpyHostToDevice, ctx->stream));
    uint32_t g = group_size(rs->E);
    hipLaunchKernelGGL(rsi16_roots_kernel, dim3(1), dim3(64), (size_t)rs->E * 4, ctx->stream, (const uint32_t *)d_p,
                       (uint8_t *)d_h, rs->dev, g);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(is_root, d_h, rs->n, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- mixed-size plans (host)

struct ffrs_rsi16_plan {
    ffrs_ctx *ctx = nullptr;
    std::vector<ffrs_rsi16 *> codecs;
    Rsi16PlanItem *d_items = nullptr;
    Rsi16PlanChunk *d_chunks = nullptr;      // repair chunks, then encode chunks
    Rsi16PlanClass *d_classes = nullptr;
    uint32_t *d_queue = nullptr;             // one counter per range
    struct Range {
        uint32_t key;    // kernel specialisation: ecc_len, or 0 for the LDS code
        uint32_t first;  // first chunk in d_chunks
        uint32_t n;      // chunks
        uint32_t lds;    // bytes of dynamic LDS
    };
    std::vector<Range> rep_ranges, enc_ranges;
    uint32_t count = 0, nchunks = 0;
    uint64_t msg_end = 0, ecc_end = 0;
    // the caller's list (in its order, class per item) and the classes, kept to lay the plan
    // out again when a class's shape changes
    std::vector<Rsi16PlanItem> list;
    std::vector<Rsi16PlanClass> classes;
    std::vector<std::pair<uint32_t, uint32_t>> keys;  // (block_len, ecc_len) per class
    bool use_graph = false;
    struct Graph {
        int mode;
        const void *m, *e, *st;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
};

namespace {

void rsi16_plan_free(ffrs_rsi16_plan *p)
{
    (void)hipSetDevice(p->ctx->device);
    for (auto &g : p->graphs)
        (void)hipGraphExecDestroy(g.exec);
    for (void *d : {(void *)p->d_items, (void *)p->d_chunks, (void *)p->d_classes, (void *)p->d_queue})
        if (d)
            (void)hipFree(d);
    for (auto *c : p->codecs)
        ffrs_rsi16_destroy(c);
    delete p;
}

// Sorts the work list by class (costliest first: largest ecc_len, then largest block_len),
// cuts it into chunks of one class holding a wave's worth for the kernel that runs them
// (consecutive chunks with the same kernel form one range, one launch) and uploads the list,
// the chunks and the classes. Drops the captured graphs, which hold the old layout.
int rsi16_plan_layout(ffrs_rsi16_plan *p)
{
    const uint32_t count = p->count;
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const uint32_t ca = p->list[a].cls, cb = p->list[b].cls;
        const auto &ka = p->keys[ca], &kb = p->keys[cb];
        if (ka.second != kb.second)
            return ka.second > kb.second;
        if (ka.first != kb.first)
            return ka.first > kb.first;
        return ca < cb;
    });
    std::vector<Rsi16PlanItem> items(count);
    for (uint32_t j = 0; j < count; ++j)
        items[j] = p->list[order[j]];
    std::vector<Rsi16PlanChunk> chunks;
    p->rep_ranges.clear();
    p->enc_ranges.clear();
    auto build = [&](bool repair, std::vector<ffrs_rsi16_plan::Range> &ranges) {
        for (uint32_t j = 0; j < count; ++j) {
            const uint32_t c = items[j].cls;
            const ffrs_rsi16 *codec = p->codecs[c];
            const uint32_t E = codec->E, g = p->classes[c].g;
            const uint32_t G = reg_group(E);
            const bool dflt = E >= 2 && E <= 256 && g == group_size(E);
            const bool reg = repair ? dflt : (E >= 2 && E <= 32) || dflt;
            const uint32_t key = reg ? E : 0u;
            const uint32_t cpw = !reg ? 64 / g : repair ? 64 / G : (E <= 32 ? 64 / p->classes[c].l_enc : 64 / G);
            const uint32_t lds = repair ? (reg ? (64 / G) * REG_WORDS * E * 4 : (64 / g) * REPAIR_WORDS * E * 4)
                                        : (reg ? 0u : (64 / g) * 2 * E * 4);
            const bool new_range = ranges.empty() || ranges.back().key != key;
            if (new_range)
                ranges.push_back({key, (uint32_t)chunks.size(), 0, 0});
            if (new_range || chunks.back().cls != c || chunks.back().count == cpw) {
                chunks.push_back({j, 0, (uint16_t)c});
                ranges.back().n++;
            }
            chunks.back().count++;
            ranges.back().lds = std::max(ranges.back().lds, lds);
        }
    };
    build(true, p->rep_ranges);
    build(false, p->enc_ranges);
    p->nchunks = (uint32_t)chunks.size();
    for (auto &g : p->graphs)
        (void)hipGraphExecDestroy(g.exec);
    p->graphs.clear();
    if (count) {
        FFRS_HIP(hipMemcpy(p->d_items, items.data(), items.size() * sizeof(items[0]), hipMemcpyHostToDevice));
        FFRS_HIP(hipMemcpy(p->d_chunks, chunks.data(), chunks.size() * sizeof(chunks[0]), hipMemcpyHostToDevice));
    }
    FFRS_HIP(hipMemcpy(p->d_classes, p->classes.data(), p->classes.size() * sizeof(p->classes[0]),
                       hipMemcpyHostToDevice));
    return FFRS_OK;
}

template <int MODE, int E>
int rsi16_plan_launch_range(ffrs_rsi16_plan *p, const ffrs_rsi16_plan::Range &r, uint32_t *queue, uint16_t *msgs,
                            uint16_t *eccs, uint8_t *status)
{
    ffrs_ctx *ctx = p->ctx;
    auto kern = rsi16_plan_kernel<MODE, E>;
    if (int rc = allow_lds(kern, r.lds))
        return rc;
    int per_cu = (int)std::max<size_t>(1, std::min<size_t>(32, (160 * 1024) / std::max<size_t>(r.lds, 1)));
    int grid = persistent_grid(ctx, per_cu, r.n);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), r.lds, ctx->stream, msgs, eccs, status, p->d_items,
                       p->d_chunks + r.first, r.n, p->d_classes, queue);
    FFRS_HIP(hipGetLastError());
    return FFRS_OK;
}

template <int MODE>
int rsi16_plan_launch_key(ffrs_rsi16_plan *p, const ffrs_rsi16_plan::Range &r, uint32_t *queue, uint16_t *msgs,
                          uint16_t *eccs, uint8_t *status)
{
    switch (r.key) {
    case 2:
        return rsi16_plan_launch_range<MODE, 2>(p, r, queue, msgs, eccs, status);
    case 4:
        return rsi16_plan_launch_range<MODE, 4>(p, r, queue, msgs, eccs, status);
    case 8:
        return rsi16_plan_launch_range<MODE, 8>(p, r, queue, msgs, eccs, status);
    case 16:
        return rsi16_plan_launch_range<MODE, 16>(p, r, queue, msgs, eccs, status);
    case 32:
        return rsi16_plan_launch_range<MODE, 32>(p, r, queue, msgs, eccs, status);
    default:
        break;
    }
    switch (r.key) {
    case 64:
        return rsi16_plan_launch_range<MODE, 64>(p, r, queue, msgs, eccs, status);
    case 128:
        return rsi16_plan_launch_range<MODE, 128>(p, r, queue, msgs, eccs, status);
    case 256:
        return rsi16_plan_launch_range<MODE, 256>(p, r, queue, msgs, eccs, status);
    default:
        break;
    }
    return rsi16_plan_launch_range<MODE, 0>(p, r, queue, msgs, eccs, status);
}

int rsi16_plan_enqueue(ffrs_rsi16_plan *p, int mode, uint16_t *msgs, uint16_t *eccs, uint8_t *status)
{
    ffrs_ctx *ctx = p->ctx;
    const auto &ranges = mode == 2 ? p->rep_ranges : p->enc_ranges;
    uint32_t *queue = p->d_queue + (mode == 2 ? 0 : p->rep_ranges.size());
    hipLaunchKernelGGL(rsi16_plan_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, queue, (uint32_t)ranges.size());
    FFRS_HIP(hipGetLastError());
    for (size_t i = 0; i < ranges.size(); ++i) {
        int r = mode == 2 ? rsi16_plan_launch_key<2>(p, ranges[i], queue + i, msgs, eccs, status)
                : mode == 0 ? rsi16_plan_launch_key<0>(p, ranges[i], queue + i, msgs, eccs, status)
                            : rsi16_plan_launch_key<1>(p, ranges[i], queue + i, msgs, eccs, status);
        if (r)
            return r;
    }
    return FFRS_OK;
}

int rsi16_plan_run(ffrs_rsi16_plan *p, int mode, uint16_t *msgs, size_t msg_elems, uint16_t *eccs, size_t ecc_elems,
                   uint8_t *status)
{
    ffrs_ctx *ctx = p->ctx;
    if (int r = ctx_set_device(ctx))
        return r;
    if (p->count == 0)
        return FFRS_OK;
    if (msg_elems < p->msg_end || ecc_elems < p->ecc_end)
        return fail(FFRS_ERR_SIZE, "buffers of %zu / %zu symbols: the plan needs %llu / %llu", msg_elems, ecc_elems,
                    (unsigned long long)p->msg_end, (unsigned long long)p->ecc_end);
    marks_begin(ctx);
    if (!p->use_graph) {
        if (int r = rsi16_plan_enqueue(p, mode, msgs, eccs, status))
            return r;
        mark_kernel(ctx);
        return FFRS_OK;
    }
    for (auto &g : p->graphs)
        if (g.mode == mode && g.m == msgs && g.e == eccs && g.st == status) {
            FFRS_HIP(hipGraphLaunch(g.exec, ctx->stream));
            mark_kernel(ctx);
            return FFRS_OK;
        }
    hipGraph_t graph = nullptr;
    FFRS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int r = rsi16_plan_enqueue(p, mode, msgs, eccs, status);
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (r)
        return r;
    if (e != hipSuccess)
        return fail(FFRS_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
        return fail(FFRS_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
    p->graphs.push_back({mode, msgs, eccs, status, exec});
    FFRS_HIP(hipGraphLaunch(exec, ctx->stream));
    mark_kernel(ctx);
    return FFRS_OK;
}

}  // namespace

extern "C" {

int ffrs_rsi16_plan_create(ffrs_ctx *ctx, size_t count, const uint32_t *block_len, const uint32_t *ecc_len,
                           const uint64_t *msg_off, const uint64_t *ecc_off, uint32_t primitive, ffrs_rsi16_plan **out)
{
    if (!ctx || !out || (count && (!block_len || !ecc_len || !msg_off || !ecc_off)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (primitive == 0)
        primitive = 3;
    if (count > 0xffffffffu)
        return fail(FFRS_ERR_INVALID, "too many codewords in one plan");
    auto *p = new ffrs_rsi16_plan();
    p->ctx = ctx;
    p->count = (uint32_t)count;
    std::vector<std::pair<uint32_t, uint32_t>> keys;  // (block_len, ecc_len) per class
    std::vector<uint32_t> cls(count);
    std::vector<Rsi16PlanClass> classes;
    for (size_t i = 0; i < count; ++i) {
        std::pair<uint32_t, uint32_t> key{block_len[i], ecc_len[i]};
        size_t c = std::find(keys.begin(), keys.end(), key) - keys.begin();
        if (c == keys.size()) {
            ffrs_rsi16 *codec = nullptr;
            if (int r = ffrs_rsi16_create(ctx, block_len[i], ecc_len[i], primitive, &codec)) {
                rsi16_plan_free(p);
                return r;
            }
            keys.push_back(key);
            p->codecs.push_back(codec);
            uint32_t g = codec->g_rep;
            // encode lanes per codeword: at least four message chunks per lane
            uint32_t l_enc = 1;
            const uint32_t chunks_per_cw = codec->n / codec->E - 1;
            while (l_enc < 64 && 8 * l_enc <= chunks_per_cw)
                l_enc *= 2;
            classes.push_back({codec->dev, g, l_enc});
        }
        cls[i] = (uint32_t)c;
    }
    p->list.resize(count);
    for (size_t i = 0; i < count; ++i) {
        const ffrs_rsi16 *codec = p->codecs[cls[i]];
        p->list[i] = {msg_off[i], ecc_off[i], (uint32_t)i, cls[i]};
        p->msg_end = std::max<uint64_t>(p->msg_end, msg_off[i] + codec->k);
        p->ecc_end = std::max<uint64_t>(p->ecc_end, ecc_off[i] + codec->E);
    }
    p->classes = classes;
    p->keys = keys;
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    auto alloc = [&](auto **dst, size_t bytes) {
        if (e == hipSuccess)
            e = hipMalloc((void **)dst, std::max<size_t>(1, bytes));
    };
    alloc(&p->d_items, count * sizeof(Rsi16PlanItem));
    alloc(&p->d_chunks, 2 * count * sizeof(Rsi16PlanChunk));  // repair and encode chunks, at most one per item each
    alloc(&p->d_classes, classes.size() * sizeof(Rsi16PlanClass));
    // one counter per range: at most one repair and one encode range per class
    alloc(&p->d_queue, 4 * (2 * classes.size() + 2));
    if (e != hipSuccess) {
        rsi16_plan_free(p);
        return fail(FFRS_ERR_NOMEM, "plan upload failed: %s", hipGetErrorString(e));
    }
    if (int r = rsi16_plan_layout(p)) {
        rsi16_plan_free(p);
        return r;
    }
    *out = p;
    return FFRS_OK;
}

int ffrs_rsi16_plan_set_class_groups(ffrs_rsi16_plan *plan, uint32_t block_len, uint32_t ecc_len, unsigned l_encode,
                                     unsigned g_repair)
{
    if (!plan)
        return fail(FFRS_ERR_INVALID, "plan is NULL");
    size_t c = std::find(plan->keys.begin(), plan->keys.end(), std::make_pair(block_len, ecc_len)) - plan->keys.begin();
    if (c == plan->keys.size())
        return fail(FFRS_ERR_INVALID, "the plan has no codewords of RSi16(%u, %u)", block_len, ecc_len);
    const ffrs_rsi16 *codec = plan->codecs[c];
    const uint32_t E = codec->E;
    auto pow2 = [](unsigned x) { return x && (x & (x - 1)) == 0; };
    if (l_encode && (!pow2(l_encode) || l_encode > 64))
        return fail(FFRS_ERR_INVALID, "encode lanes per codeword %u (a power of two up to 64)", l_encode);
    if (g_repair && (!pow2(g_repair) || g_repair > std::min<uint32_t>(64, E) ||
                     (size_t)(64 / g_repair) * REPAIR_WORDS * E * 4 > 160 * 1024))
        return fail(FFRS_ERR_INVALID, "repair lanes per codeword %u for ecc_len %u", g_repair, E);
    std::lock_guard<std::mutex> lk(plan->ctx->mu);
    if (int r = ctx_set_device(plan->ctx))
        return r;
    if (l_encode)
        plan->classes[c].l_enc = l_encode;
    plan->classes[c].g = g_repair ? g_repair : group_size(E);
    return rsi16_plan_layout(plan);
}

void ffrs_rsi16_plan_destroy(ffrs_rsi16_plan *plan)
{
    if (plan)
        rsi16_plan_free(plan);
}

int ffrs_rsi16_plan_use_graph(ffrs_rsi16_plan *plan, int enable)
{
    if (!plan)
        return fail(FFRS_ERR_INVALID, "plan is NULL");
    plan->use_graph = enable != 0;
    return FFRS_OK;
}

int ffrs_rsi16_plan_encode(ffrs_rsi16_plan *plan, const uint16_t *d_msgs, size_t msg_elems, uint16_t *d_eccs,
                           size_t ecc_elems, int vector_lanes)
{
    if (!plan || (plan->count && (!d_msgs || !d_eccs)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    return rsi16_plan_run(plan, vector_lanes ? 0 : 1, const_cast<uint16_t *>(d_msgs), msg_elems, d_eccs, ecc_elems,
                          nullptr);
}

int ffrs_rsi16_plan_repair(ffrs_rsi16_plan *plan, uint16_t *d_msgs, size_t msg_elems, uint16_t *d_eccs,
                           size_t ecc_elems, uint8_t *d_status)
{
    if (!plan || (plan->count && (!d_msgs || !d_eccs || !d_status)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    return rsi16_plan_run(plan, 2, d_msgs, msg_elems, d_eccs, ecc_elems, d_status);
}

}  // extern "C"

// ---------------------------------------------------------------- CIRC16 (host)

struct ffrs_circ16 {
    ffrs_rsi16 *ri = nullptr, *ro = nullptr;  // inner and outer codecs (not owned)
    uint32_t I = 1, W = 1;                     // interleave blocks; lanes of the SIMD words (1: scalar)
    uint32_t ik = 0, iE = 0, ok = 0, oE = 0, on = 0, rI = 0;
    size_t msg_len = 0, rso_len = 0, rsi_len = 0, rsio_len = 0, ecc_len = 0;
    void *d_buf = nullptr;
    size_t buf_size = 0;
};

namespace {

int circ_reserve(ffrs_circ16 *c, size_t bytes)
{
    if (c->buf_size >= bytes)
        return FFRS_OK;
    if (c->d_buf)
        FFRS_HIP(hipFree(c->d_buf));
    c->d_buf = nullptr;
    c->buf_size = 0;
    hipError_t e = hipMalloc(&c->d_buf, bytes);
    if (e != hipSuccess)
        return fail(FFRS_ERR_NOMEM, "CIRC16 workspace of %zu bytes: %s", bytes, hipGetErrorString(e));
    c->buf_size = bytes;
    return FFRS_OK;
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// The generic lane-group encode of `count` codewords (interleave stride I), u16 or u32 symbols in
// and out, in the SIMD or scalar words.
template <class TI, class TO>
int circ_encode(ffrs_rsi16 *rs, const TI *d_in, size_t count, TO *d_out, bool vec, uint32_t I)
{
    if (count == 0)
        return FFRS_OK;
    const uint32_t g = rs->g_enc == 1 && rs->E <= 32 ? group_size(rs->E) : rs->g_enc, cpw = 64 / g;
    const size_t lds = (size_t)cpw * 2 * rs->E * 4;
    auto *kv = rsi16_encode_kernel<VectorOps, TI, TO>;
    auto *ks = rsi16_encode_kernel<ScalarOps, TI, TO>;
    auto *kern = vec ? kv : ks;
    if (int r = allow_lds(kern, lds))
        return r;
    hipLaunchKernelGGL(kern, dim3((unsigned)((count + cpw - 1) / cpw)), dim3(64), lds, rs->ctx->stream, d_in,
                       (uint32_t)count, d_out, rs->dev, g, I);
    FFRS_HIP(hipGetLastError());
    return FFRS_OK;
}

// Repair workspace of one block: u16 message and ecc, u32 outer parity, u32 inner parity,
// u32 syndromes, row flags, error rows per interleave block and their counts, group statuses.
struct CircWs {
    uint16_t *msg, *ecc;
    uint32_t *rso32, *rsi32, *syn, *pos, *cnt;
    uint8_t *flag, *gst;
    uint32_t *work;
    size_t work_lanes;
};

size_t circ_repair_layout(const ffrs_circ16 *c, CircWs *w, char *base)
{
    const size_t rows = (size_t)c->on * c->I, ngr = (c->ik + c->W - 1) / c->W;
    const size_t wrows = (size_t)c->ro->n + 8 * (size_t)c->oE;  // per-lane words of the SIMD repair
    size_t lanes = ngr * c->W;
    const size_t cap = std::max<size_t>(c->W, ((size_t)256 << 20) / (wrows * 4) / c->W * c->W);
    lanes = std::min(lanes, cap);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + o : nullptr;
        o += al256(bytes);
        return p;
    };
    w->msg = (uint16_t *)take(c->msg_len * 2);
    w->ecc = (uint16_t *)take(c->ecc_len * 2);
    w->rso32 = (uint32_t *)take(c->rso_len * 4);
    w->rsi32 = (uint32_t *)take(rows * c->iE * 4);
    w->syn = (uint32_t *)take(rows * c->iE * 4);
    w->flag = (uint8_t *)take(rows);
    w->pos = (uint32_t *)take((size_t)c->I * c->oE * 4);
    w->cnt = (uint32_t *)take((size_t)c->I * 4);
    w->gst = (uint8_t *)take(ngr);
    w->work = (uint32_t *)take(lanes * wrows * 4);
    w->work_lanes = lanes;
    return o;
}

// load_ecc, compute_rsi_synd, repair_outer_zeros, repair_inner_zeros and the row flags of
// find_error_locations, for every interleave block at once (they touch disjoint rows).
int circ_prepare(ffrs_circ16 *c, const CircWs &w)
{
    hipStream_t st = c->ri->ctx->stream;
    const uint32_t rows = c->on * c->I, mrows = c->ok * c->I, erows = c->oE * c->I, iE = c->iE;
    const size_t rsi_all = (size_t)rows * iE;
    hipLaunchKernelGGL(circ16_expand_kernel, dim3(256), dim3(256), 0, st, (const uint16_t *)w.ecc, w.rso32,
                       (uint32_t)c->rso_len, w.rsi32, (uint32_t)rsi_all);
    FFRS_HIP(hipGetLastError());
    const uint32_t g = group_size(iE), cpw = 64 / g;
    const size_t lds = (size_t)cpw * 2 * iE * 4;
    if (int r = allow_lds(circ16_synd_kernel<uint16_t>, lds))
        return r;
    if (int r = allow_lds(circ16_synd_kernel<uint32_t>, lds))
        return r;
    hipLaunchKernelGGL(circ16_synd_kernel<uint16_t>, dim3((mrows + cpw - 1) / cpw), dim3(64), lds, st, c->ri->dev,
                       (const uint16_t *)w.msg, (const uint32_t *)w.rsi32, mrows, w.syn, g);
    FFRS_HIP(hipGetLastError());
    hipLaunchKernelGGL(circ16_synd_kernel<uint32_t>, dim3((erows + cpw - 1) / cpw), dim3(64), lds, st, c->ri->dev,
                       (const uint32_t *)w.rso32, (const uint32_t *)(w.rsi32 + (size_t)mrows * iE), erows,
                       w.syn + (size_t)mrows * iE, g);
    FFRS_HIP(hipGetLastError());
    const size_t lz = (size_t)9 * iE * 4;
    if (int r = allow_lds(circ16_outer_zeros_kernel, lz))
        return r;
    hipLaunchKernelGGL(circ16_outer_zeros_kernel, dim3(erows), dim3(64), lz, st, c->ri->dev, w.rso32,
                       w.rsi32 + (size_t)mrows * iE, erows, w.syn + (size_t)mrows * iE);
    FFRS_HIP(hipGetLastError());
    if (int r = allow_lds(circ16_inner_zeros_kernel, (size_t)iE * 4))
        return r;
    hipLaunchKernelGGL(circ16_inner_zeros_kernel, dim3(rows), dim3(64), (size_t)iE * 4, st, c->ri->dev, w.rsi32,
                       rows, w.syn, w.flag);
    FFRS_HIP(hipGetLastError());
    return FFRS_OK;
}

// The error rows of interleave block kk from the row flags (find_error_locations: rows i with
// non-zero inner syndromes of codeword i I + kk, ascending).
std::vector<uint32_t> circ_rows(const ffrs_circ16 *c, const std::vector<uint8_t> &flag, uint32_t kk)
{
    std::vector<uint32_t> rows;
    for (uint32_t i = 0; i < c->on; ++i)
        if (flag[(size_t)i * c->I + kk])
            rows.push_back(i);
    return rows;
}

}  // namespace

extern "C" {

int ffrs_circ16_create(ffrs_rsi16 *inner, ffrs_rsi16 *outer, size_t interleave, unsigned simd_width, ffrs_circ16 **out)
{
    if (!inner || !outer || !out)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (inner->ctx != outer->ctx)
        return fail(FFRS_ERR_INVALID, "inner and outer codecs belong to different contexts");
    if (inner->primitive != outer->primitive)
        return fail(FFRS_ERR_INVALID, "inner and outer codecs use different primitives");
    if (simd_width != 1 && simd_width != 4 && simd_width != 8 && simd_width != 16)
        return fail(FFRS_ERR_INVALID, "SIMD width must be 1, 4, 8 or 16, got %u", simd_width);
    if (interleave == 0 || (size_t)inner->k * interleave > 0xffffu)
        return fail(FFRS_ERR_INVALID, "inner message_len * interleave must be in [1, 65535], got %zu",
                    (size_t)inner->k * interleave);
    auto *c = new ffrs_circ16;
    c->ri = inner;
    c->ro = outer;
    c->I = (uint32_t)interleave;
    c->W = simd_width;
    c->ik = inner->k;
    c->iE = inner->E;
    c->ok = outer->k;
    c->oE = outer->E;
    c->on = outer->n;
    c->rI = c->ik * c->I;
    c->msg_len = (size_t)c->ok * c->rI;
    c->rso_len = (size_t)c->oE * c->rI;
    c->rsi_len = (size_t)c->iE * c->ok * c->I;
    c->rsio_len = (size_t)c->iE * c->oE * c->I;
    c->ecc_len = c->rso_len + c->rsi_len + c->rsio_len;
    *out = c;
    return FFRS_OK;
}

void ffrs_circ16_destroy(ffrs_circ16 *c)
{
    if (!c)
        return;
    if (c->d_buf) {
        (void)hipSetDevice(c->ri->ctx->device);
        (void)hipFree(c->d_buf);
    }
    delete c;
}

size_t ffrs_circ16_message_len(const ffrs_circ16 *c) { return c ? c->msg_len : 0; }
size_t ffrs_circ16_ecc_len(const ffrs_circ16 *c) { return c ? c->ecc_len : 0; }

int ffrs_circ16_encode(ffrs_circ16 *c, const uint16_t *msg, size_t blocks, uint16_t *ecc)
{
    if (!c)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (blocks == 0)
        return FFRS_OK;
    if (!msg || !ecc)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = check_count(blocks * c->rI))
        return r;
    ffrs_ctx *ctx = c->ri->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    const size_t mb = al256(blocks * c->msg_len * 2), ob = al256(blocks * c->rso_len * 4),
                 ib = al256(blocks * c->rsi_len * 2), iob = al256(blocks * c->rsio_len * 2),
                 eb = al256(blocks * c->ecc_len * 2);
    if (int r = circ_reserve(c, mb + ob + ib + iob + eb))
        return r;
    char *base = (char *)c->d_buf;
    uint16_t *d_msg = (uint16_t *)base, *d_rsi = (uint16_t *)(base + mb + ob), *d_rsio = (uint16_t *)(base + mb + ob + ib),
             *d_out = (uint16_t *)(base + mb + ob + ib + iob);
    uint32_t *d_rso = (uint32_t *)(base + mb);
    hipStream_t st = ctx->stream;
    const bool vec = c->W > 1;
    FFRS_HIP(hipMemcpyAsync(d_msg, msg, blocks * c->msg_len * 2, hipMemcpyHostToDevice, st));
    // encode_block (pycirc16.hpp:371-381): the outer parity as field elements, then the inner
    // parity of the message rows and of the outer parity rows
    if (int r = circ_encode<uint16_t, uint32_t>(c->ro, d_msg, blocks * c->rI, d_rso, vec, c->rI))
        return r;
    if (int r = launch_encode(c->ri, d_msg, blocks * c->ok * c->I, d_rsi, vec ? 1 : 0))
        return r;
    if (int r = circ_encode<uint32_t, uint16_t>(c->ri, d_rso, blocks * c->oE * c->I, d_rsio, vec, 1))
        return r;
    hipLaunchKernelGGL(circ16_assemble_kernel, dim3(1024), dim3(256), 0, st, (const uint32_t *)d_rso,
                       (const uint16_t *)d_rsi, (const uint16_t *)d_rsio, (uint32_t)c->rso_len, (uint32_t)c->rsi_len,
                       (uint32_t)c->rsio_len, (uint32_t)blocks, d_out);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(ecc, d_out, blocks * c->ecc_len * 2, hipMemcpyDeviceToHost, st));
    FFRS_HIP(hipStreamSynchronize(st));
    return FFRS_OK;
}

int ffrs_circ16_find_outer_error_locations(ffrs_circ16 *c, const uint16_t *msg, const uint16_t *ecc,
                                           size_t interleave, uint32_t *rows, size_t *count)
{
    if (!c || !msg || !ecc || !rows || !count)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (interleave >= c->I)
        return fail(FFRS_ERR_PARAM, "interleave %zu >= %u", interleave, c->I);
    ffrs_ctx *ctx = c->ri->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    CircWs w;
    if (int r = circ_reserve(c, circ_repair_layout(c, &w, nullptr)))
        return r;
    circ_repair_layout(c, &w, (char *)c->d_buf);
    hipStream_t st = ctx->stream;
    FFRS_HIP(hipMemcpyAsync(w.msg, msg, c->msg_len * 2, hipMemcpyHostToDevice, st));
    FFRS_HIP(hipMemcpyAsync(w.ecc, ecc, c->ecc_len * 2, hipMemcpyHostToDevice, st));
    if (int r = circ_prepare(c, w))
        return r;
    std::vector<uint8_t> flag((size_t)c->on * c->I);
    FFRS_HIP(hipMemcpyAsync(flag.data(), w.flag, flag.size(), hipMemcpyDeviceToHost, st));
    FFRS_HIP(hipStreamSynchronize(st));
    auto v = circ_rows(c, flag, (uint32_t)interleave);
    std::copy(v.begin(), v.end(), rows);
    *count = v.size();
    return FFRS_OK;
}

int ffrs_circ16_repair(ffrs_circ16 *c, uint16_t *msg, uint16_t *ecc, int *raised)
{
    if (!c || !msg || !ecc || !raised)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    *raised = 0;
    ffrs_ctx *ctx = c->ri->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    CircWs w;
    if (int r = circ_reserve(c, circ_repair_layout(c, &w, nullptr)))
        return r;
    circ_repair_layout(c, &w, (char *)c->d_buf);
    hipStream_t st = ctx->stream;
    FFRS_HIP(hipMemcpyAsync(w.msg, msg, c->msg_len * 2, hipMemcpyHostToDevice, st));
    FFRS_HIP(hipMemcpyAsync(w.ecc, ecc, c->ecc_len * 2, hipMemcpyHostToDevice, st));
    if (int r = circ_prepare(c, w))
        return r;
    std::vector<uint8_t> flag((size_t)c->on * c->I);
    FFRS_HIP(hipMemcpyAsync(flag.data(), w.flag, flag.size(), hipMemcpyDeviceToHost, st));
    FFRS_HIP(hipStreamSynchronize(st));
    // repair_outer_errors (pycirc16.hpp:297-313) per interleave block: with the error rows as
    // positions where there are at most outer E of them, else the repair without positions
    std::vector<uint32_t> pos((size_t)c->I * c->oE, 0), cnt(c->I, 0);
    std::vector<uint32_t> blind;  // blocks repaired without positions, ascending
    for (uint32_t kk = 0; kk < c->I; ++kk) {
        auto rv = circ_rows(c, flag, kk);
        if (rv.empty())
            continue;
        if (rv.size() <= c->oE) {
            cnt[kk] = (uint32_t)rv.size();
            std::copy(rv.begin(), rv.end(), pos.begin() + (size_t)kk * c->oE);
        } else {
            blind.push_back(kk);
        }
    }
    if (std::any_of(cnt.begin(), cnt.end(), [](uint32_t v) { return v != 0; })) {
        FFRS_HIP(hipMemcpyAsync(w.pos, pos.data(), pos.size() * 4, hipMemcpyHostToDevice, st));
        FFRS_HIP(hipMemcpyAsync(w.cnt, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice, st));
        const size_t lds = (size_t)7 * c->oE * 4;
        if (int r = allow_lds(circ16_outer_known_kernel, lds))
            return r;
        hipLaunchKernelGGL(circ16_outer_known_kernel, dim3(c->rI), dim3(64), lds, st, c->ro->dev, w.msg, w.rso32,
                           c->rI, c->ik, (const uint32_t *)w.pos, (const uint32_t *)w.cnt);
        FFRS_HIP(hipGetLastError());
    }
    // without positions: the reference's SIMD groups of W columns from the block's first column
    // (pyrsi16.hpp:391-456), in block order; a group that raises stops the repair there
    const uint32_t W = c->W, ngr = (c->ik + W - 1) / W;
    const uint32_t gpb = (uint32_t)std::max<size_t>(1, w.work_lanes / W);
    uint32_t raise_kk = c->I;
    std::vector<uint8_t> gst(ngr);
    for (uint32_t kk : blind) {
        bool stop = false;
        for (uint32_t g0 = 0; g0 < ngr && !stop; g0 += gpb) {
            const uint32_t ng = std::min(gpb, ngr - g0), lanes = ng * W;
            const dim3 grid((lanes + 63) / 64);
            auto *rk = W == 1 ? rsi16_repair_simd_kernel<1, uint32_t> : W == 4 ? rsi16_repair_simd_kernel<4, uint32_t>
                       : W == 8 ? rsi16_repair_simd_kernel<8, uint32_t> : rsi16_repair_simd_kernel<16, uint32_t>;
            hipLaunchKernelGGL(rk, grid, dim3(64), 0, st, c->ro->dev, (const uint16_t *)w.msg,
                               (const uint32_t *)w.rso32, c->rI, kk * c->ik, c->ik, g0, lanes, w.work, w.gst);
            FFRS_HIP(hipGetLastError());
            FFRS_HIP(hipMemcpyAsync(gst.data() + g0, w.gst + g0, ng, hipMemcpyDeviceToHost, st));
            FFRS_HIP(hipStreamSynchronize(st));
            uint32_t end = g0;
            while (end < g0 + ng && gst[end] != FFRS_RSI16_RAISES)
                ++end;
            stop = end < g0 + ng;
            if (end > g0) {
                auto *wk = W == 1 ? rsi16_simd_writeback_kernel<1, uint32_t>
                           : W == 4 ? rsi16_simd_writeback_kernel<4, uint32_t>
                           : W == 8 ? rsi16_simd_writeback_kernel<8, uint32_t> : rsi16_simd_writeback_kernel<16, uint32_t>;
                hipLaunchKernelGGL(wk, grid, dim3(256), 0, st, c->ro->dev, w.msg, w.rso32, c->rI, kk * c->ik, c->ik, g0,
                                   end, lanes, (const uint32_t *)w.work);
                FFRS_HIP(hipGetLastError());
            }
        }
        if (stop) {
            raise_kk = kk;
            break;
        }
    }
    if (raise_kk < c->I) {
        // the reference raises inside repair_all: the blocks before raise_kk are repaired, block
        // raise_kk up to its raising group, the others untouched; ecc is not written back
        std::vector<uint16_t> rep(c->msg_len);
        FFRS_HIP(hipMemcpyAsync(rep.data(), w.msg, c->msg_len * 2, hipMemcpyDeviceToHost, st));
        FFRS_HIP(hipStreamSynchronize(st));
        const size_t lim = (size_t)(raise_kk + 1) * c->ik;
        for (size_t r = 0; r < c->ok; ++r)
            std::copy(rep.begin() + r * c->rI, rep.begin() + r * c->rI + lim, msg + r * c->rI);
        *raised = 1;
        return FFRS_OK;
    }
    // recompute_inner_ecc (pycirc16.hpp:317-321) and dump_ecc
    const bool vec = W > 1;
    uint16_t *d_rsi = w.ecc + c->rso_len, *d_rsio = d_rsi + c->rsi_len;
    if (int r = launch_encode(c->ri, w.msg, (size_t)c->ok * c->I, d_rsi, vec ? 1 : 0))
        return r;
    if (int r = circ_encode<uint32_t, uint16_t>(c->ri, w.rso32, (size_t)c->oE * c->I, d_rsio, vec, 1))
        return r;
    hipLaunchKernelGGL(circ16_assemble_kernel, dim3(64), dim3(256), 0, st, (const uint32_t *)w.rso32,
                       (const uint16_t *)d_rsi, (const uint16_t *)d_rsio, (uint32_t)c->rso_len, 0u, 0u, 1u, w.ecc);
    FFRS_HIP(hipGetLastError());
    FFRS_HIP(hipMemcpyAsync(msg, w.msg, c->msg_len * 2, hipMemcpyDeviceToHost, st));
    FFRS_HIP(hipMemcpyAsync(ecc, w.ecc, c->ecc_len * 2, hipMemcpyDeviceToHost, st));
    FFRS_HIP(hipStreamSynchronize(st));
    return FFRS_OK;
}

}  // extern "C"
