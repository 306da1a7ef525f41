// rs256.hip -- RS256 (GF(2^8)) systematic encode and repair on gfx950.
//
// Reference semantics (paths relative to the reference repository root):
//   field       libffrs/galois.hpp:93-146, 624-667 (alpha = primitive, poly 0x11d default)
//   generator   libffrs/reed_solomon.hpp:55-88
//   encode      libffrs/reed_solomon.hpp:169-389, libffrs/pyrs256.hpp:99-127
//   syndromes   libffrs/reed_solomon.hpp:412-453
//   repair      libffrs/reed_solomon.hpp:555-744 (BM :655-702, roots :511-552, Forney :704-743)
//
// Kernels
//   rs256_encode_kernel  persistent workgroups, one codeword per lane. The parity is the
//                        remainder of x^ecc M(x) by g(x), advanced S bytes per step with
//                        S slice tables T_s[b] = b x^(ecc+s) mod g held in LDS.
//   rs256_repair_kernel  one codeword per lane: the same remainder over the message,
//                        XOR the received parity -> R(x) = C(x) mod g(x); S_i = R(alpha^i);
//                        Berlekamp-Massey with the reference's truncation and status rules;
//                        Chien search over positions [0, n); Forney; in-place corrections.
#include "common.hpp"

#include <algorithm>
#include <type_traits>
#include <cstring>

namespace ffrs {

// ---------------------------------------------------------------- host tables

struct Gf256Host {
    uint8_t exp_t[256];  // reference convention: exp[255] = 1
    uint8_t log_t[256];  // reference convention: log[1] = 255, log[0] = 0
    uint8_t mul(uint8_t a, uint8_t b) const
    {
        if (!a || !b)
            return 0;
        unsigned r = (unsigned)log_t[a] + log_t[b];
        if (r >= 255)
            r -= 255;
        return exp_t[r];
    }
};

// Shift-and-add product modulo the field polynomial (galois.hpp:624-651).
static uint8_t gf256_mul_serial(uint8_t a, uint8_t b, uint8_t poly1)
{
    uint8_t r = 0;
    for (int i = 7; i >= 0; --i) {
        r = (r & 0x80) ? (uint8_t)((r << 1) ^ poly1) : (uint8_t)(r << 1);
        if (a & (1u << i))
            r ^= b;
    }
    return r;
}

static bool gf256_build(unsigned primitive, unsigned polynomial, Gf256Host &gf)
{
    uint8_t poly1 = (uint8_t)(polynomial & 0xff);
    if (poly1 & 0x80)
        return false;  // the reference's wide multiply asserts this (galois.hpp:715)
    memset(gf.log_t, 0, sizeof gf.log_t);
    uint8_t x = 1;
    for (int i = 0; i < 256; ++i) {
        gf.exp_t[i] = x;
        gf.log_t[x] = (uint8_t)i;
        x = gf256_mul_serial(x, (uint8_t)primitive, poly1);
    }
    bool seen[256] = {false};
    for (int i = 0; i < 255; ++i) {
        if (gf.exp_t[i] == 0 || seen[gf.exp_t[i]])
            return false;  // primitive does not generate the multiplicative group
        seen[gf.exp_t[i]] = true;
    }
    return true;
}

// Device-side field tables: EXP2[x] = alpha^(x mod 255) for x <= 509 (the largest index
// is log a + 255 - log b in a quotient) and 0 above; LOG2[a] = log_alpha(a) in [0, 254]
// with LOG2[0] = ZLOG (511). A product of logs that involves a zero lands in the zero
// region of EXP2.
constexpr int ZLOG = 511;
constexpr int EXP2_LEN = 1024;

// ---------------------------------------------------------------- device code

struct Rs256Tables {
    const uint32_t *lut;     // S * 256 rows of lut_row(EW) dwords, slice-major
    const uint8_t *exp2;     // EXP2_LEN
    const uint16_t *log2;    // 256
};

// LUT row stride in dwords: rows of 32 or 64 bytes are padded by 16 bytes so that random
// row reads start at every 16-byte bank group of the LDS instead of every other one.
__host__ __device__ constexpr int lut_row(int EW) { return (EW % 8 == 0) ? EW + 4 : EW; }

// Replicated two-slice layouts (S == LUT_REP), 64 KB per slice, slice s at byte s << 16.
// Rows of 4, 8 or 16 bytes (EW = 1, 2, 4): row b of slice s is stored C = 64 / EW times, copy c
// at byte offset (s << 16) | (b << 8) | (c * 4 EW), so that each row fills the 64 banks once.
// The LDS serves 64 / EW lanes of a 4 EW-byte read together, and lane l reads copy l % C: the
// lanes served together always sit in distinct banks, whatever rows the data selects.
// Rows of 32 bytes (EW == 8): chunk h (16
// bytes) of row b of slice s is stored LUT_COPIES times, copy c at byte offset
// (s << 16) | (b << 8) | (h << 7) | (c << 4), i.e. in 16-byte bank group 8h + c of the 64
// banks. The LDS serves a 16-byte read sixteen lanes at a time (one lane per bank group when
// there are no conflicts). Lane l reads copy l % 8, and lanes whose bit 3 is set read chunk 1
// with their first read and chunk 0 with their second, the others the reverse: the sixteen
// lanes served together then always sit in sixteen different bank groups, whatever rows the
// data selects. (Unreplicated rows collide about three ways per group of sixteen.)
// Rows and the remainder are held halfword-paired: with h_0..h_15 the remainder's 16-bit
// symbol pairs (h_0 = bytes 0 and 1, the leading ones), dword j holds h_j in its low half and
// h_(j+8) in its high half. A two-byte LFSR shift then moves whole dwords (register renaming)
// and one 16-bit shift, instead of a byte-align per dword. Lanes that read chunk 1 first keep
// the two halves of their remainder swapped (registers 0..3 hold dwords 4..7), so that every
// lane XORs its first read into registers 0..3.
constexpr int LUT_REP = 0;
constexpr int LUT_COPIES = 8;
__host__ __device__ constexpr int lut_dwords(int EW, int S)
{
    return S == LUT_REP ? 2 * 256 * 2 * LUT_COPIES * 4 : S * 256 * lut_row(EW);
}

// Lanes of the replicated layout that read chunk 1 first (bit 3 of the lane index).
__device__ inline bool lut_swapped() { return (threadIdx.x & 8u) != 0; }

// a ^ b ^ c in one instruction (truth table 0x96)
__device__ inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }

// LDS home of an LFSR table of layout (EW, S): the replicated layout lives in a static array
// (a known LDS address, so row offsets fold into the read instructions), every other layout at
// the start of the kernel's dynamic LDS `dyn`.
template <int EW, int S>
__device__ inline uint32_t *lfsr_table(uint32_t *dyn)
{
    if constexpr (S == LUT_REP) {
        __shared__ uint4 s_rep[lut_dwords(8, LUT_REP) / 4];
        return reinterpret_cast<uint32_t *>(s_rep);
    } else {
        return dyn;
    }
}
// Dynamic LDS bytes that the table of layout (EW, S) takes.
__host__ __device__ constexpr int lfsr_dyn_dwords(int EW, int S) { return S == LUT_REP ? 0 : lut_dwords(EW, S); }

// One row of EW dwords of an LDS table, in one read.
template <int EW>
__device__ inline void lds_row(const char *p, uint32_t (&r)[EW])
{
    if constexpr (EW == 4) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p);
        r[0] = v.x; r[1] = v.y; r[2] = v.z; r[3] = v.w;
    } else if constexpr (EW == 2) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p);
        r[0] = v.x; r[1] = v.y;
    } else {
        static_assert(EW == 1, "rows of 4, 8 or 16 bytes");
        r[0] = *reinterpret_cast<const uint32_t *>(p);
    }
}

template <int EW, int S>
__device__ inline void lut_step(uint32_t (&rem)[EW], uint32_t w, const uint32_t *__restrict__ s_lut)
{
    if constexpr (S == LUT_REP && EW < 8) {
        // two bytes per half-step: slice 1 row x.byte0, slice 0 row x.byte1 at this lane's copy
        // (offsets built by byte permutes: copy in byte 0, row in byte 1, slice in byte 2)
        constexpr uint32_t C = 64 / EW;
        const uint32_t lane_off1 = ((threadIdx.x % C) * 4 * EW) | 0x10000u;
        const char *lds0 = reinterpret_cast<const char *>(s_lut);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t x = rem[0] ^ (h ? w >> 16 : w);
            uint32_t a1[EW], a0[EW];
            lds_row<EW>(lds0 + __builtin_amdgcn_perm(x, lane_off1, 0x0c020400u), a1);
            lds_row<EW>(lds0 + __builtin_amdgcn_perm(x, lane_off1, 0x0c0c0500u), a0);
#pragma unroll
            for (int d = 0; d < EW; ++d) {
                const uint32_t nx = d + 1 < EW ? alignbyte(rem[(d + 1) % EW], rem[d], 2) : rem[d] >> 16;
                rem[d] = xor3(nx, a1[d], a0[d]);
            }
        }
    } else if constexpr (S == LUT_REP) {
        static_assert(EW == 8, "replicated layouts hold rows of 4 to 32 bytes");
        // byte permutes build the offsets: this lane's copy and first chunk in byte 0, the row
        // in byte 1, the slice in byte 2; the second chunk is 128 bytes away
        const bool swapped = lut_swapped();
        const uint32_t lane_off1 = ((threadIdx.x % LUT_COPIES) << 4) | (swapped ? 0x80u : 0u) | 0x10000u;
        const uint32_t sh3 = swapped ? 16u : 0u, sh7 = 16u - sh3;
        const char *lds0 = reinterpret_cast<const char *>(s_lut);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t x = (swapped ? rem[4] : rem[0]) ^ (h ? w >> 16 : w);
            // slice 1 row x.byte0 (the first byte needs one more shift), slice 0 row x.byte1
            const uint32_t o1 = __builtin_amdgcn_perm(x, lane_off1, 0x0c020400u);
            const uint32_t o0 = __builtin_amdgcn_perm(x, lane_off1, 0x0c0c0500u);
            const uint4 a1 = *reinterpret_cast<const uint4 *>(lds0 + o1);
            const uint4 b1 = *reinterpret_cast<const uint4 *>(lds0 + (o1 ^ 0x80u));
            const uint4 a0 = *reinterpret_cast<const uint4 *>(lds0 + o0);
            const uint4 b0 = *reinterpret_cast<const uint4 *>(lds0 + (o0 ^ 0x80u));
            // h_8 moves to the low half of dword 7 (register 7, or 3 when swapped); dword 3
            // takes dword 4 unshifted (register 3, or 7 when swapped)
            const uint32_t t3 = rem[4] >> sh3, t7 = rem[0] >> sh7;
            rem[0] = xor3(rem[1], a1.x, a0.x);
            rem[1] = xor3(rem[2], a1.y, a0.y);
            rem[2] = xor3(rem[3], a1.z, a0.z);
            rem[3] = xor3(t3, a1.w, a0.w);
            rem[4] = xor3(rem[5], b1.x, b0.x);
            rem[5] = xor3(rem[6], b1.y, b0.y);
            rem[6] = xor3(rem[7], b1.z, b0.z);
            rem[7] = xor3(t7, b1.w, b0.w);
        }
    } else if constexpr (S == 4) {
        uint32_t x = rem[0] ^ w;
        const uint32_t *t3 = s_lut + (3 * 256 + (x & 0xff)) * lut_row(EW);
        const uint32_t *t2 = s_lut + (2 * 256 + ((x >> 8) & 0xff)) * lut_row(EW);
        const uint32_t *t1 = s_lut + (1 * 256 + ((x >> 16) & 0xff)) * lut_row(EW);
        const uint32_t *t0 = s_lut + (0 * 256 + (x >> 24)) * lut_row(EW);
        if constexpr (EW % 4 == 0) {
#pragma unroll
            for (int d = 0; d < EW; d += 4) {
                uint4 a = *reinterpret_cast<const uint4 *>(t3 + d);
                uint4 b = *reinterpret_cast<const uint4 *>(t2 + d);
                uint4 c = *reinterpret_cast<const uint4 *>(t1 + d);
                uint4 e = *reinterpret_cast<const uint4 *>(t0 + d);
                uint32_t n0 = rem[d + 1];
                uint32_t n1 = rem[d + 2];
                uint32_t n2 = rem[d + 3];
                uint32_t n3 = (d + 4 < EW) ? rem[(d + 4) % EW] : 0u;
                rem[d + 0] = n0 ^ a.x ^ b.x ^ c.x ^ e.x;
                rem[d + 1] = n1 ^ a.y ^ b.y ^ c.y ^ e.y;
                rem[d + 2] = n2 ^ a.z ^ b.z ^ c.z ^ e.z;
                rem[d + 3] = n3 ^ a.w ^ b.w ^ c.w ^ e.w;
            }
        } else {
#pragma unroll
            for (int d = 0; d < EW; ++d) {
                uint32_t nx = (d + 1 < EW) ? rem[(d + 1) % EW] : 0u;
                rem[d] = nx ^ t3[d] ^ t2[d] ^ t1[d] ^ t0[d];
            }
        }
    } else if constexpr (S == 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint32_t x = rem[0] ^ (w >> (16 * h));
            const uint32_t *t1 = s_lut + (256 + (x & 0xff)) * lut_row(EW);
            const uint32_t *t0 = s_lut + ((x >> 8) & 0xff) * lut_row(EW);
#pragma unroll
            for (int d = 0; d < EW; ++d) {
                uint32_t nx = (d + 1 < EW) ? rem[(d + 1) % EW] : 0u;
                rem[d] = alignbyte(nx, rem[d], 2) ^ t1[d] ^ t0[d];
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t f = (rem[0] ^ (w >> (8 * j))) & 0xff;
            const uint32_t *t0 = s_lut + f * lut_row(EW);
#pragma unroll
            for (int d = 0; d < EW; ++d) {
                uint32_t nx = (d + 1 < EW) ? rem[(d + 1) % EW] : 0u;
                rem[d] = alignbyte(nx, rem[d], 1) ^ t0[d];
            }
        }
    }
}

// A remainder of the replicated layout (halfword-paired, halves swapped on swapped lanes)
// back in byte order.
__device__ inline void rep_to_bytes(uint32_t (&rem)[8])
{
    const bool swapped = lut_swapped();
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const uint32_t lo = rem[d], hi = rem[d + 4];
        rem[d] = swapped ? hi : lo;
        rem[d + 4] = swapped ? lo : hi;
    }
    // halfword-paired -> byte order: dword m = h_2m | h_(2m+1) << 16
    uint32_t o[8];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        o[m] = __builtin_amdgcn_perm(rem[2 * m + 1], rem[2 * m], 0x05040100u);
        o[m + 4] = __builtin_amdgcn_perm(rem[2 * m + 1], rem[2 * m], 0x07060302u);
    }
#pragma unroll
    for (int d = 0; d < 8; ++d)
        rem[d] = o[d];
}

// Remainder of x^ecc * M(x) by g(x) for the message of `k` bytes at byte offset
// `msg_off` of the buffer behind `rsrc` (whose range is `nrec` bytes). The message is fed in 4-byte words after
// `pad` leading zero bytes (leading zeros do not change the remainder).
template <int EW, int S>
__device__ inline void message_remainder(__amdgpu_buffer_rsrc_t rsrc, uint32_t nrec, int64_t msg_off, uint32_t k,
                                         const uint32_t *__restrict__ s_lut, uint32_t (&rem)[EW])
{
#pragma unroll
    for (int d = 0; d < EW; ++d)
        rem[d] = 0;
    uint32_t P = (k + 3) & ~3u;
    uint32_t pad = P - k;
    int64_t start = msg_off - (int64_t)pad;
    int64_t a = start >> 2;                 // floor division (start >= -3)
    uint32_t sh = (uint32_t)(start - a * 4);
    uint32_t off = (uint32_t)(a * 4);       // wraps below zero -> out of range -> reads 0
    uint32_t headmask = pad ? ~((1u << (8 * pad)) - 1u) : 0xffffffffu;
    uint32_t nw = P >> 2;
    uint32_t prev = buf_load_u32(rsrc, off);
    uint32_t t = 0;
    if constexpr (S == LUT_REP) {
        // 64 bytes per iteration, the 128 bytes after them already in flight while they are fed:
        // a lane visits each cache line of its message at most twice, and the loads have two
        // iterations of table reads to arrive
        constexpr int CW = 16;
        auto load = [&](uint32_t tt, uint32_t (&c)[CW]) {
            const uint32_t o = off + 4 * (tt + 1);
            if (o + 4 * CW <= nrec) {
#pragma unroll
                for (int q = 0; q < CW / 4; ++q) {
                    const uint4 v = buf_load_u32x4(rsrc, o + 16 * q);
                    c[4 * q] = v.x; c[4 * q + 1] = v.y; c[4 * q + 2] = v.z; c[4 * q + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int j = 0; j < CW; ++j)
                    c[j] = buf_load_u32(rsrc, o + 4 * j);
            }
        };
        // the last nw % CW words, loaded up front with the first chunks
        const uint32_t full = nw - nw % CW;
        uint32_t ct[CW];
        if (full < nw)
            load(full, ct);
        if (t + CW <= nw) {
            uint32_t c0[CW], c1[CW];
            load(t, c0);
            bool have1 = t + 2 * CW <= nw;
            if (have1)
                load(t + CW, c1);
            for (;;) {
                const bool have2 = t + 3 * CW <= nw;
                uint32_t c2[CW];
                if (have2) {
                    load(t + 2 * CW, c2);
                } else {
#pragma unroll
                    for (int j = 0; j < CW; ++j)
                        c2[j] = 0;
                }
                uint32_t w0 = alignbyte(c0[0], prev, sh);
                if (t == 0)
                    w0 &= headmask;
                lut_step<EW, S>(rem, w0, s_lut);
#pragma unroll
                for (int j = 1; j < CW; ++j)
                    lut_step<EW, S>(rem, alignbyte(c0[j], c0[j - 1], sh), s_lut);
                prev = c0[CW - 1];
                t += CW;
                if (!have1)
                    break;
#pragma unroll
                for (int j = 0; j < CW; ++j) {
                    c0[j] = c1[j];
                    c1[j] = c2[j];
                }
                have1 = have2;
            }
        }
        if (full < nw) {
#pragma unroll
            for (int j = 0; j < CW; ++j) {
                if (full + j < nw) {
                    uint32_t w = alignbyte(ct[j], j ? ct[j - 1] : prev, sh);
                    if (full + j == 0)
                        w &= headmask;
                    lut_step<EW, S>(rem, w, s_lut);
                }
            }
            t = nw;
        }
    }
    // 32 bytes per iteration, as two 16-byte loads where all of them are in range (each lane
    // reads its own row, so wider loads mean fewer L1 lines fetched per byte); single
    // words at the very end of the buffer
    for (; t + 8 <= nw; t += 8) {
        uint32_t c[8];
        const uint32_t o = off + 4 * (t + 1);
        if (o + 32 <= nrec) {
            const uint4 q0 = buf_load_u32x4(rsrc, o);
            const uint4 q1 = buf_load_u32x4(rsrc, o + 16);
            c[0] = q0.x; c[1] = q0.y; c[2] = q0.z; c[3] = q0.w;
            c[4] = q1.x; c[5] = q1.y; c[6] = q1.z; c[7] = q1.w;
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                c[j] = buf_load_u32(rsrc, o + 4 * j);
        }
        uint32_t w0 = alignbyte(c[0], prev, sh);
        if (t == 0)
            w0 &= headmask;
        lut_step<EW, S>(rem, w0, s_lut);
#pragma unroll
        for (int j = 1; j < 8; ++j)
            lut_step<EW, S>(rem, alignbyte(c[j], c[j - 1], sh), s_lut);
        prev = c[7];
    }
    for (; t < nw; ++t) {
        uint32_t c = buf_load_u32(rsrc, off + 4 * (t + 1));
        uint32_t w = alignbyte(c, prev, sh);
        if (t == 0)
            w &= headmask;
        lut_step<EW, S>(rem, w, s_lut);
        prev = c;
    }
    if constexpr (S == LUT_REP && EW == 8)
        rep_to_bytes(rem);
}

// Remainders of the 64 messages of a wave, cooperatively loaded (replicated layouts, k <= 254,
// so that a message with the dword of the word before it spans at most 65 dwords). Lane 16 g + p
// takes the message of row `base + 16 g + p`. Each message window (from the dword holding the
// word before the first fed one, as message_remainder reads it) is read by the sixteen lanes of
// its group, 16 bytes each, in one instruction: 256 contiguous bytes per group instead of 64
// scattered 16-byte reads per instruction (on MI355X, 6.1 against 2.6 TB/s for 247-byte rows).
// A 16 x 16 transpose of the 16-byte pieces within each group (DPP exchanges, no LDS) then gives
// every lane its own window, and the LFSR runs on registers.
__device__ inline uint32_t dpp_xor(uint32_t x, int b)
{
    if (b == 1)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    if (b == 2)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    if (b == 8)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, false);  // row_ror:8 (lane ^ 8)
    // lane ^ 4 = (lane ^ 3) ^ 7: quad_perm [3,2,1,0], then row_half_mirror
    const int y = __builtin_amdgcn_mov_dpp((int)x, 0x1B, 0xF, 0xF, false);
    return (uint32_t)__builtin_amdgcn_mov_dpp(y, 0x141, 0xF, 0xF, false);
}

// D[slot][lane] (slot = 16-byte piece index held by this lane, 4 dwords each) becomes D[lane][slot]
// within each 16-lane row: for every bit b, the pieces whose slot bit b differs from the lane's
// bit b trade places with the partner lane (lane ^ b, slot ^ b).
__device__ inline void transpose16(uint32_t (&d)[16][4], uint32_t lane)
{
#pragma unroll
    for (int b = 1; b < 16; b <<= 1) {
        const bool hi = (lane & (uint32_t)b) != 0;
#pragma unroll
        for (int s0 = 0; s0 < 16; ++s0) {
            if (s0 & b)
                continue;
            const int s1 = s0 | b;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t send = hi ? d[s0][w] : d[s1][w];
                const uint32_t recv = dpp_xor(send, b);
                d[s1][w] = hi ? d[s1][w] : recv;
                d[s0][w] = hi ? recv : d[s0][w];
            }
        }
    }
}

// Issues the cooperative window reads of the 64 messages of a wave from row `base` into d
// (lane 16 g + p reads 16-byte piece p of the windows of rows base + 16 g + 0..15); the caller
// waits for them where it first uses d.
__device__ inline void coop_issue(__amdgpu_buffer_rsrc_t rsrc, uint32_t base, uint32_t k, uint32_t (&d)[16][4])
{
    const uint32_t lane = threadIdx.x & 63, p = lane & 15, g0 = base + (lane & 48);
    const uint32_t pad = ((k + 3) & ~3u) - k;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        // window of row r: from the dword at floor((r k - pad) / 4) (wraps below zero)
        const uint32_t o = (uint32_t)(((int64_t)(g0 + i) * k - (int64_t)pad) >> 2) * 4u + 16 * p;
        // The window of the first message starts a dword before the buffer when pad > 0; that
        // dword only holds pad bytes, which the head mask clears. Such a piece is read from the
        // buffer start (a read at the wrapped offset is out of range as a whole) and shifted.
        const bool before = o >= 0xFFFFFFF0u;
        const uint4 v = buf_load_u32x4(rsrc, before ? o + 4 : o);
        d[i][0] = before ? 0u : v.x;
        d[i][1] = before ? v.x : v.y;
        d[i][2] = before ? v.y : v.z;
        d[i][3] = before ? v.z : v.w;
    }
}

// Remainder of this lane's message from the windows issued by coop_issue for row `base`
// (transposed here, so that every lane holds its own window) and the replicated tables.
template <int EW, int S>
__device__ inline void coop_remainder(__amdgpu_buffer_rsrc_t rsrc, uint32_t base, uint32_t k, uint32_t (&d)[16][4],
                                      const uint32_t *__restrict__ s_lut, uint32_t (&rem)[EW])
{
    static_assert(S == LUT_REP, "cooperative loads feed the replicated tables");
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t P = (k + 3) & ~3u, pad = P - k, nw = P >> 2;
    const int64_t start = (int64_t)(base + lane) * k - (int64_t)pad;
    const uint32_t sh = (uint32_t)(start - (start >> 2) * 4);
    const uint32_t off = (uint32_t)(start >> 2) * 4u;
    const uint32_t extra = nw >= 64 ? buf_load_u32(rsrc, off + 256) : 0u;  // the 65th dword
    transpose16(d, lane & 15);
    const uint32_t headmask = pad ? ~((1u << (8 * pad)) - 1u) : 0xffffffffu;
#pragma unroll
    for (int j = 0; j < EW; ++j)
        rem[j] = 0;
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        if ((uint32_t)t < nw) {
            const uint32_t hi = t + 1 < 64 ? d[(t + 1) >> 2][(t + 1) & 3] : extra;
            uint32_t w = alignbyte(hi, d[t >> 2][t & 3], sh);
            if (t == 0)
                w &= headmask;
            lut_step<EW, S>(rem, w, s_lut);
        }
    }
    if constexpr (EW == 8)
        rep_to_bytes(rem);
}

template <int EW>
__device__ inline uint8_t rem_byte(const uint32_t (&rem)[EW], uint32_t j)
{
    uint32_t v = 0;
#pragma unroll
    for (int d = 0; d < EW; ++d)
        v = ((j >> 2) == (uint32_t)d) ? rem[d] : v;
    return (uint8_t)(v >> (8 * (j & 3)));
}

// Message loads (LM): LOAD_LANE each lane its own message (message_remainder), LOAD_COOP sixteen
// lanes per message window (coop_issue / coop_remainder).
constexpr int LOAD_LANE = 0, LOAD_COOP = 1;

template <int EW, int S, int WG, int LM = LOAD_LANE>
__global__ __launch_bounds__(WG) void rs256_encode_kernel(const uint8_t *__restrict__ msgs, uint32_t msg_bytes,
                                                            uint32_t count, uint32_t k, uint8_t *__restrict__ out,
                                                            uint32_t ecc, const uint32_t *__restrict__ lut)
{
    extern __shared__ uint32_t s_dyn[];
    uint32_t *s_lut = lfsr_table<EW, S>(s_dyn);
    static_assert(lut_dwords(EW, S) % 4 == 0, "tables are copied in 16-byte units");
    for (uint32_t i = threadIdx.x; i < (uint32_t)lut_dwords(EW, S) / 4; i += blockDim.x)
        reinterpret_cast<uint4 *>(s_lut)[i] = reinterpret_cast<const uint4 *>(lut)[i];
    __syncthreads();

    const uint32_t nrec = (msg_bytes + 3u) & ~3u;
    __amdgpu_buffer_rsrc_t rsrc = make_rsrc(msgs, nrec);
    const uint32_t stride = gridDim.x * blockDim.x;
    auto store = [&](uint32_t cw, const uint32_t (&rem)[EW]) {
        uint8_t *o = out + (size_t)cw * ecc;
        if ((ecc & 3) == 0) {
            uint32_t *o4 = reinterpret_cast<uint32_t *>(o);
#pragma unroll
            for (int d = 0; d < EW; ++d)
                if ((uint32_t)d * 4 < ecc)
                    o4[d] = rem[d];
        } else {
            for (uint32_t j = 0; j < ecc; ++j)
                o[j] = rem_byte<EW>(rem, j);
        }
    };
    if constexpr (LM == LOAD_COOP) {
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        for (uint32_t base = (blockIdx.x * (WG / 64) + wave) * 64; base < count; base += stride) {
            uint32_t d[16][4], rem[EW];
            coop_issue(rsrc, base, k, d);
            coop_remainder<EW, S>(rsrc, base, k, d, s_lut, rem);
            const uint32_t cw = base + (threadIdx.x & 63);
            if (cw < count)
                store(cw, rem);
        }
    } else {
        for (uint32_t cw = blockIdx.x * blockDim.x + threadIdx.x; cw < count; cw += stride) {
            uint32_t rem[EW];
            message_remainder<EW, S>(rsrc, nrec, (int64_t)cw * k, k, s_lut, rem);
            store(cw, rem);
        }
    }
}

// ---- GF(2^8) helpers over the LDS copies of EXP2 / LOG2

struct GfLds {
    const uint8_t *exp2;
    const uint16_t *log2;
    __device__ inline uint32_t mul(uint32_t a, uint32_t b) const { return exp2[log2[a] + log2[b]]; }
    __device__ inline uint32_t mul_log(uint32_t la, uint32_t b) const { return exp2[la + log2[b]]; }
    // reference table conventions (galois.hpp:98-114): inv(0) = 1, div(a, 0) = a
    __device__ inline uint32_t inv(uint32_t a) const { return a ? exp2[255 - log2[a]] : 1u; }
    __device__ inline uint32_t div(uint32_t a, uint32_t b) const
    {
        if (!a)
            return 0;
        if (!b)
            return a;
        return exp2[log2[a] + 255 - log2[b]];
    }
};

// S_i = R(alpha^i) for i < E, where R(x) = sum_j r_j x^(E-1-j) is held in remainder
// order (r_0 = highest). Horner over j per syndrome; EXP2 is zero above index 509, so a
// zero accumulator (LOG2 = ZLOG) contributes nothing.
// Q syndromes advance together, so Q lookup chains are in flight at a time (fewer for wide
// remainders, whose words already fill the registers).
template <int EW>
__device__ inline void syndromes_from_remainder(const uint32_t (&rem)[EW], uint32_t E, const GfLds &gf,
                                                uint8_t *Sy)
{
    constexpr int Q = EW >= 32 ? 4 : 8;
    for (uint32_t i0 = 0; i0 < E; i0 += Q) {
        uint32_t acc[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q)
            acc[q] = 0;
#pragma unroll
        for (int d = 0; d < EW; ++d) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if ((uint32_t)(d * 4 + b) < E) {
                    const uint32_t r = (rem[d] >> (8 * b)) & 0xffu;
                    uint32_t la[Q];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        la[q] = gf.log2[acc[q]];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
                        acc[q] = (uint32_t)gf.exp2[la[q] + i0 + q] ^ r;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (i0 + q < E)
                Sy[i0 + q] = (uint8_t)acc[q];
    }
}

// Reads E parity bytes of one codeword and XORs them into the remainder words.
template <int EW>
__device__ inline uint32_t xor_parity(uint32_t (&rem)[EW], const uint8_t *pe, uint32_t E)
{
    uint32_t nz = 0;
    if ((E & 3) == 0) {
        const uint32_t *p4 = reinterpret_cast<const uint32_t *>(pe);
#pragma unroll
        for (int d = 0; d < EW; ++d)
            if ((uint32_t)d * 4 < E)
                rem[d] ^= p4[d];
    } else {
        for (uint32_t j = 0; j < E; ++j) {
            uint32_t v = (uint32_t)pe[j] << (8 * (j & 3));
#pragma unroll
            for (int d = 0; d < EW; ++d)
                rem[d] ^= ((j >> 2) == (uint32_t)d) ? v : 0u;
        }
    }
#pragma unroll
    for (int d = 0; d < EW; ++d)
        nz |= rem[d];
    return nz;
}

// Per-lane scratch rows in LDS: bytes [0,E) syndromes (later the root positions),
// [E,2E) locator, [2E,3E) B(x) (later Omega), then E u16 logs of the locator.
// The row stride is an odd number of dwords so that lanes spread over the banks.
__host__ __device__ inline uint32_t repair_row_stride(uint32_t ecc)
{
    uint32_t bytes = 3 * ecc + 1 + 2 * ecc;
    uint32_t dw = (bytes + 3) / 4;
    if ((dw & 1) == 0)
        dw += 1;
    return dw * 4;
}

// Lane-mode repair of one codeword (reed_solomon.hpp:559-595): message of k bytes at
// byte offset msg_off of `rsrc` (the same bytes as pm[0..k)), parity pe[0..E). Returns the
// status (1: no errors or repaired, 0: not repairable, buffers untouched). `row` is this
// lane's LDS scratch of repair_row_stride(E) bytes.
template <int EW, int S>
__device__ inline uint8_t repair_one(__amdgpu_buffer_rsrc_t rsrc, uint32_t nrec, int64_t msg_off, uint8_t *pm, uint8_t *pe,
                                     uint32_t k, uint32_t E, const uint32_t *__restrict__ s_lut, const GfLds &gf,
                                     uint8_t *row)
{
    const uint32_t n = k + E;
    uint8_t *Sy = row;
    uint8_t *lam = row + E;
    uint8_t *Bx = row + 2 * E;
    uint16_t *llam = reinterpret_cast<uint16_t *>(row + ((3 * E + 1) & ~1u));
    // ---- R(x) = C(x) mod g(x): remainder of the message XOR the received parity
    uint32_t rem[EW];
    message_remainder<EW, S>(rsrc, nrec, msg_off, k, s_lut, rem);
    if (xor_parity<EW>(rem, pe, E) == 0)  // all syndromes zero (reed_solomon.hpp:565-566)
        return 1;
    syndromes_from_remainder<EW>(rem, E, gf, Sy);

    // ---- Berlekamp-Massey (reed_solomon.hpp:655-702): locator capacity E, degree
    //      terms >= E dropped, L returned. Natural order lam[i] = Lambda_i; lenL / lenB
    //      bound the non-zero terms of Lambda and B, so updates touch only those.
    uint16_t *lS = llam;  // logs of the syndromes while BM runs
    for (uint32_t i = 0; i < E; ++i) {
        lam[i] = 0;
        Bx[i] = 0;
        lS[i] = gf.log2[Sy[i]];
    }
    lam[0] = 1;
    Bx[0] = 1;
    uint32_t L = 0, m = 1, b = 1, lenL = 1, lenB = 1;
    for (uint32_t nn = 0; nn < E; ++nn) {
        uint32_t d = Sy[nn];
        uint32_t top = L < lenL - 1 ? L : lenL - 1;
        for (uint32_t i = 1; i <= top; i += 4) {  // four terms' lookups in flight
            uint32_t lv[4], ls[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool in = i + q <= top;
                lv[q] = in ? (uint32_t)lam[i + q] : 0u;
                ls[q] = in ? (uint32_t)lS[nn - i - q] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                lv[q] = gf.log2[lv[q]];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (i + q <= top)
                    d ^= gf.exp2[lv[q] + ls[q]];
        }
        if (d == 0) {
            m += 1;
            continue;
        }
        uint32_t lc = gf.log2[gf.div(d, b)];
        bool grow = 2 * L <= nn;
        // L never decreases and the codeword fails once L > E / 2 (checked after the loop),
        // so stop there; while L <= E / 2, deg Lambda <= L and deg x^m B <= the new L, so
        // the coefficients above E / 2 stay zero and the update can skip them
        if (grow && nn + 1 - L > E / 2)
            return 0;
        uint32_t hiB = m + lenB < E ? m + lenB : E;  // x^m B truncated to degrees < E
        uint32_t hi = grow ? (hiB > lenL ? hiB : lenL) : hiB;
        hi = hi < E / 2 + 1 ? hi : E / 2 + 1;
        // lam <- lam - c x^m B ; if grow, B <- old lam. Downwards, four indices at a time: a
        // group reads B[ii - m] below every index written so far (m >= 1), so its reads can
        // all precede its writes.
        for (uint32_t g0 = hi; g0 > 0; g0 = g0 > 4 ? g0 - 4 : 0) {
            uint32_t old[4], sb[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint32_t ii = g0 - 1 - t;
                const bool in = g0 > (uint32_t)t;
                old[t] = in ? (uint32_t)lam[ii] : 0u;
                sb[t] = (in && ii >= m && ii < hiB) ? (uint32_t)Bx[ii - m] : 0u;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
                sb[t] = gf.mul_log(lc, sb[t]);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (g0 > (uint32_t)t) {
                    const uint32_t ii = g0 - 1 - t;
                    lam[ii] = (uint8_t)(old[t] ^ sb[t]);
                    if (grow)
                        Bx[ii] = (uint8_t)old[t];
                }
        }
        if (grow) {
            lenB = lenL;
            L = nn + 1 - L;
            b = d;
            m = 1;
        } else {
            m += 1;
        }
        if (hiB > lenL)
            lenL = hiB;
    }
    if (L > E / 2)
        return 0;

    // ---- Omega = S(x) Lambda(x) mod x^E (reed_solomon.hpp:708-728). The final LFSR
    //      (Lambda, L) generates all E syndromes, so Omega_k = 0 for L <= k < E; B is dead.
    uint8_t *om = Bx;
    for (uint32_t kk = 0; kk < L; ++kk) {
        uint32_t v = 0;
        for (uint32_t i = 0; i <= kk; i += 4) {  // four terms' lookups in flight
            uint32_t lv[4], ls[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool in = i + q <= kk;
                lv[q] = in ? (uint32_t)lam[i + q] : 0u;
                ls[q] = in ? (uint32_t)lS[kk - i - q] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                lv[q] = gf.log2[lv[q]];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (i + q <= kk)
                    v ^= gf.exp2[lv[q] + ls[q]];
        }
        om[kk] = (uint8_t)v;
    }
    for (uint32_t i = 0; i <= L; ++i)
        llam[i] = gf.log2[lam[i]];

    // ---- Chien search over [0, n): Lambda(alpha^-p) == 0, ascending p
    //      (reed_solomon.hpp:516-538); positions go to the dead syndrome row
    uint8_t *pos = Sy;
    uint32_t roots = 0;
    for (uint32_t p0 = 0; p0 < n; p0 += 4) {  // four positions per pass: one LDS read of each log
        uint32_t e[4], step[4], acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t p = p0 + q;
            step[q] = p ? 255 - p : 0;  // alpha^-p = alpha^(255 - p)
            e[q] = 0;
        }
        for (uint32_t i = 0; i <= L; i += 4) {  // four terms' logs read together
            uint32_t li[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
                li[t] = i + t <= L ? (uint32_t)llam[i + t] : (uint32_t)ZLOG;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    acc[q] ^= gf.exp2[li[t] + e[q]];
                    e[q] += step[q];
                    e[q] = e[q] >= 255 ? e[q] - 255 : e[q];
                }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (acc[q] == 0 && p0 + q < n) {
                if (roots < E)
                    pos[roots] = (uint8_t)(p0 + q);
                roots++;
            }
    }
    if (roots != L)
        return 0;

    // ---- Forney (reed_solomon.hpp:730-743): e_p = X Omega(X^-1) / Lambda'(X^-1),
    //      X = alpha^p, Lambda' = odd terms of Lambda (characteristic 2)
    // logs of Omega in the dead locator bytes (255 stands for a zero coefficient)
    uint8_t *lom = lam;
    for (uint32_t kk = 0; kk < L; ++kk)
        lom[kk] = om[kk] ? (uint8_t)gf.log2[om[kk]] : (uint8_t)255;
    for (uint32_t r = 0; r < L; ++r) {
        uint32_t p = pos[r];
        uint32_t pinv = p ? 255 - p : 0;
        uint32_t num = 0, den = 0, e = 0;
        for (uint32_t k0 = 0; k0 < L; k0 += 4) {  // Omega_k = 0 for k >= L; four terms at a time
            uint32_t lo[4], ld[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint32_t kk = k0 + t;
                lo[t] = kk < L ? (uint32_t)lom[kk] : 255u;
                ld[t] = (kk < L && (kk & 1) == 0) ? (uint32_t)llam[kk + 1] : (uint32_t)ZLOG;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                num ^= gf.exp2[(lo[t] == 255 ? (uint32_t)ZLOG : lo[t]) + e];
                den ^= gf.exp2[ld[t] + e];
                e += pinv;
                e = e >= 255 ? e - 255 : e;
            }
        }
        uint32_t X = gf.exp2[p];
        uint32_t mag = gf.mul(X, gf.div(num, den));
        uint32_t at = n - 1 - p;
        if (at < k)
            pm[at] ^= (uint8_t)mag;
        else
            pe[at - k] ^= (uint8_t)mag;
    }
    return 1;
}

template <int EW, int S, int WG>
__global__ __launch_bounds__(WG) void rs256_repair_kernel(uint8_t *__restrict__ msgs, uint32_t msg_bytes,
                                                           uint8_t *__restrict__ eccs, uint32_t count, uint32_t k,
                                                           uint32_t ecc, uint8_t *__restrict__ status,
                                                           Rs256Tables tab)
{
    extern __shared__ uint32_t s_mem[];
    uint32_t *s_lut = s_mem;
    uint8_t *s_exp2 = reinterpret_cast<uint8_t *>(s_mem + S * 256 * lut_row(EW));
    uint16_t *s_log2 = reinterpret_cast<uint16_t *>(s_exp2 + EXP2_LEN);
    uint8_t *s_rows = reinterpret_cast<uint8_t *>(s_log2 + 256);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(S * 256 * lut_row(EW)); i += blockDim.x)
        s_lut[i] = tab.lut[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += blockDim.x)
        s_exp2[i] = tab.exp2[i];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x)
        s_log2[i] = tab.log2[i];
    __syncthreads();
    GfLds gf{s_exp2, s_log2};
    uint8_t *row = s_rows + threadIdx.x * repair_row_stride(ecc);
    const uint32_t nrec = (msg_bytes + 3u) & ~3u;
    __amdgpu_buffer_rsrc_t rsrc = make_rsrc(msgs, nrec);
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t cw = blockIdx.x * blockDim.x + threadIdx.x; cw < count; cw += stride)
        status[cw] = repair_one<EW, S>(rsrc, nrec, (int64_t)cw * k, msgs + (size_t)cw * k, eccs + (size_t)cw * ecc, k, ecc,
                                       s_lut, gf, row);
}

// ======================================================================
// Wave-cooperative repair, one codeword per wave, for ecc_len 33..252 (the same steps and
// results as repair_one, reed_solomon.hpp:559-595, spread over the 64 lanes instead of run by
// one):
//   syndromes  S_i = sum_j c_j alpha^(i (n-1-j)), syndrome i by lane i % 64 (direct sums over
//              the codeword c = message then parity, whose logs sit in LDS); all zero -> 1
//   BM         lane l holds Lambda_i and x^m B (logs) for i = l + 64 c, c < NC; each step's
//              discrepancy is a wave XOR reduction; terms of Lambda above E/2 and of x^m B
//              above E/2 + 1 are dropped as in bm_fixed_one; L > E/2 -> 0
//   Omega      Omega_k = sum_{i <= k} Lambda_i S_(k-i), k < L, by lane k
//   Chien      positions p = lane + 64 q < n: Lambda(alpha^-p) == 0; roots != L -> 0
//   Forney     at each root e = X div(Omega(X^-1), Lambda_odd(X^-1)), X = alpha^p, XORed
//              into the message or parity byte n - 1 - p
// ======================================================================

// XOR of x over the 64 lanes, in every lane (row butterflies, then the four rows).
__device__ inline uint32_t wave_xor(uint32_t x)
{
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false);  // row_half_mirror
    x ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false);  // row_mirror
    return (uint32_t)(__builtin_amdgcn_readlane((int)x, 0) ^ __builtin_amdgcn_readlane((int)x, 16) ^
                      __builtin_amdgcn_readlane((int)x, 32) ^ __builtin_amdgcn_readlane((int)x, 48));
}

// Per-wave LDS of repair_wave (16-byte aligned: the logs are read eight and four at a time).
struct alignas(16) WaveRows {
    uint16_t lc[256];   // logs of the codeword bytes (ZLOG for 0)
    uint8_t cw[256];    // codeword bytes
    uint8_t sy[256];    // syndromes
    uint16_t ls[256];   // their logs
    uint16_t ll[128];   // logs of Lambda_0..L
    uint16_t lo[128];   // logs of Omega_0..L-1
};

// NC Lambda coefficients and NS syndromes per lane (E/2 + 1 <= 64 NC, E <= 64 NS).
template <int NC, int NS>
__device__ inline uint8_t repair_wave(uint8_t *pm, uint8_t *pe, uint32_t k, uint32_t E, const GfLds &gf, WaveRows &w)
{
    const uint32_t lane = threadIdx.x & 63, n = k + E, H = E / 2;
    // ---- codeword into LDS, with its logs
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t j = lane + 64 * q;
        const uint32_t c = j < k ? pm[j] : j < n ? pe[j - k] : 0u;
        w.cw[j] = (uint8_t)c;
        w.lc[j] = (uint16_t)gf.log2[c];
    }
    __builtin_amdgcn_wave_barrier();
    // ---- syndromes: lane l sums S_(l + 64 q) over the positions
    uint32_t sy[NS], e[NS], st[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const uint32_t i = lane + 64 * q;
        sy[q] = 0;
        st[q] = i % 255;
        e[q] = (i * (n - 1)) % 255;  // exponent of alpha^i at position j = 0
    }
    // eight positions per pass: their logs in one 16-byte read (positions past n hold ZLOG
    // logs of zero bytes, which add nothing), then the eight lookups of each syndrome together
    for (uint32_t j0 = 0; j0 < n; j0 += 8) {
        const uint4 l8 = *reinterpret_cast<const uint4 *>(&w.lc[j0]);
        const uint32_t lw[4] = {l8.x, l8.y, l8.z, l8.w};
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            uint32_t t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                t[u] = gf.exp2[((lw[u >> 1] >> (16 * (u & 1))) & 0xffffu) + e[q]];
                e[q] = e[q] >= st[q] ? e[q] - st[q] : e[q] + 255 - st[q];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                sy[q] ^= t[u];
        }
    }
    uint32_t any = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t i = lane + 64 * q;
        const uint32_t v = q < NS && i < E ? sy[q < NS ? q : 0] : 0u;
        any |= v;
        w.sy[i] = (uint8_t)v;
        w.ls[i] = (uint16_t)gf.log2[v];
    }
    if (__ballot(any != 0) == 0)
        return 1;  // all syndromes zero (reed_solomon.hpp:565-566)
    __builtin_amdgcn_wave_barrier();

    // ---- Berlekamp-Massey over lanes (logs of x^m B pre-shifted, as bm_fixed_one)
    uint32_t lam[NC], llam[NC], lB[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint32_t i = lane + 64 * c;
        lam[c] = i == 0 ? 1u : 0u;
        llam[c] = i == 0 ? 0u : ZLOG;
        lB[c] = i == 1 ? 0u : ZLOG;
    }
    uint32_t L = 0, lb = 0;
    // lsr[c] = log S_(r - i) for i = lane + 64 c (ZLOG for i > r): one lane up per step
    uint32_t lsr[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
        lsr[c] = lane == 0 && c == 0 ? (uint32_t)w.ls[0] : (uint32_t)ZLOG;
    for (uint32_t r = 0; r < E; ++r) {
        if (r > 0) {
            const uint32_t head = w.ls[r];
#pragma unroll
            for (int c = NC - 1; c >= 0; --c) {
                uint32_t v = (uint32_t)__builtin_amdgcn_update_dpp((int)ZLOG, (int)lsr[c], 0x138, 0xF, 0xF, false);
                const uint32_t in = c > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)lsr[c - 1], 63) : head;
                lsr[c] = lane == 0 ? in : v;
            }
        }
        uint32_t t = 0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const uint32_t i = lane + 64 * c;
            t ^= gf.exp2[llam[c] + ((i >= 1 && i <= H) ? lsr[c] : (uint32_t)ZLOG)];
        }
        const uint32_t d = w.sy[r] ^ wave_xor(t);
        uint32_t nlB[NC];
        if (d == 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
                nlB[c] = lB[c];
        } else {
            const bool grow = 2 * L <= r;
            if (grow && r + 1 - L > H)
                return 0;  // L exceeds E / 2 from here on
            const uint32_t ld = gf.log2[d];
            uint32_t lc = ld + 255 - lb;
            lc = lc >= 255 ? lc - 255 : lc;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const uint32_t i = lane + 64 * c;
                const uint32_t old = llam[c];
                if (i >= 1 && i <= H) {
                    lam[c] ^= gf.exp2[lc + lB[c]];
                    llam[c] = gf.log2[lam[c]];
                }
                nlB[c] = grow ? old : lB[c];
            }
            if (grow) {
                L = r + 1 - L;
                lb = ld;
            }
        }
        // x^m B <- x (x^m B), or x Lambda_old on growth: one lane up, index 0 empty, terms
        // above min(E/2 + 1, E - 1) dropped
        const uint32_t top = H + 1 < E - 1 ? H + 1 : E - 1;
#pragma unroll
        for (int c = NC - 1; c >= 0; --c) {
            uint32_t v = (uint32_t)__builtin_amdgcn_update_dpp((int)ZLOG, (int)nlB[c], 0x138, 0xF, 0xF, false);  // wave_shr:1
            if (c > 0) {
                const uint32_t carry = (uint32_t)__builtin_amdgcn_readlane((int)nlB[c - 1], 63);
                v = lane == 0 ? carry : v;
            } else {
                v = lane == 0 ? (uint32_t)ZLOG : v;
            }
            const uint32_t i = lane + 64 * c;
            lB[c] = i <= top ? v : (uint32_t)ZLOG;
        }
    }
    if (L > H)
        return 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint32_t i = lane + 64 * c;
        if (i <= L)
            w.ll[i] = (uint16_t)llam[c];
    }
    __builtin_amdgcn_wave_barrier();
    // ---- Omega_k, k < L (reed_solomon.hpp:708-728), logs into LDS
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint32_t kk = lane + 64 * c;
        uint32_t v = 0;
        for (uint32_t i = 0; i < L; ++i)
            if (i <= kk)
                v ^= gf.exp2[(uint32_t)w.ll[i] + w.ls[kk - i]];
        if (kk < L)
            w.lo[kk] = (uint16_t)gf.log2[v];
    }
    __builtin_amdgcn_wave_barrier();
    // ---- Chien over p = lane + 64 q < n (reed_solomon.hpp:516-538)
    // Lambda_(L+1) .. Lambda_(L+3) read as zero (ZLOG) so that terms go four at a time
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint32_t i = lane + 64 * c;
        if (i > L && i <= L + 3 && i < 128)
            w.ll[i] = (uint16_t)ZLOG;
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t zero = 0, acc[4] = {0, 0, 0, 0}, ex[4] = {0, 0, 0, 0}, step[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t p = lane + 64 * q;
        step[q] = p ? 255 - p % 255 : 0;
    }
    for (uint32_t i0 = 0; i0 <= L; i0 += 4) {
        const uint2 l4 = *reinterpret_cast<const uint2 *>(&w.ll[i0]);
        const uint32_t li[4] = {l4.x & 0xffffu, l4.x >> 16, l4.y & 0xffffu, l4.y >> 16};
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[q] ^= gf.exp2[li[t] + ex[q]];
                ex[q] += step[q];
                ex[q] = ex[q] >= 255 ? ex[q] - 255 : ex[q];
            }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        zero |= (acc[q] == 0 && lane + 64 * q < n) ? 1u << q : 0u;
    const uint32_t roots = (uint32_t)__builtin_popcountll(__ballot(zero & 1)) +
                           (uint32_t)__builtin_popcountll(__ballot(zero & 2)) +
                           (uint32_t)__builtin_popcountll(__ballot(zero & 4)) +
                           (uint32_t)__builtin_popcountll(__ballot(zero & 8));
    if (roots != L)
        return 0;
    // ---- Forney at the roots (reed_solomon.hpp:730-743)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (zero & (1u << q)) {
            const uint32_t p = lane + 64 * q;
            const uint32_t pinv = p ? 255 - p : 0;
            uint32_t num = 0, den = 0, ex = 0;
            for (uint32_t kk = 0; kk < L; ++kk) {
                num ^= gf.exp2[(uint32_t)w.lo[kk] + ex];
                if ((kk & 1) == 0)
                    den ^= gf.exp2[(uint32_t)w.ll[kk + 1] + ex];
                ex += pinv;
                ex = ex >= 255 ? ex - 255 : ex;
            }
            const uint32_t mag = gf.mul(gf.exp2[p], gf.div(num, den));
            const uint32_t at = n - 1 - p;
            const uint32_t v = (uint32_t)w.cw[at] ^ mag;
            if (at < k)
                pm[at] = (uint8_t)v;
            else
                pe[at - k] = (uint8_t)v;
        }
    }
    return 1;
}

// One codeword per wave over `count` codewords of message length k (rows of k and E bytes).
template <int NC, int NS, int WG>
__global__ __launch_bounds__(WG) void rs256_repair_wave_kernel(uint8_t *__restrict__ msgs, uint8_t *__restrict__ eccs,
                                                                uint32_t count, uint32_t k, uint32_t E,
                                                                uint8_t *__restrict__ status, Rs256Tables tab)
{
    __shared__ uint8_t s_exp2[EXP2_LEN];
    __shared__ uint16_t s_log2[256];
    __shared__ WaveRows s_rows[WG / 64];
    for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += WG)
        s_exp2[i] = tab.exp2[i];
    for (uint32_t i = threadIdx.x; i < 256u; i += WG)
        s_log2[i] = tab.log2[i];
    __syncthreads();
    const GfLds gf{s_exp2, s_log2};
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    WaveRows &w = s_rows[wave];
    for (uint32_t cw = blockIdx.x * (WG / 64) + wave; cw < count; cw += gridDim.x * (WG / 64)) {
        const uint8_t st = repair_wave<NC, NS>(msgs + (size_t)cw * k, eccs + (size_t)cw * E, k, E, gf, w);
        if ((threadIdx.x & 63) == 0)
            status[cw] = st;
        __builtin_amdgcn_wave_barrier();
    }
}

template <int EW, int S>
__global__ __launch_bounds__(256) void rs256_synds_kernel(const uint8_t *__restrict__ msgs, uint32_t msg_bytes,
                                                           const uint8_t *__restrict__ eccs, uint32_t count,
                                                           uint32_t k, uint32_t ecc, uint8_t *__restrict__ synds,
                                                           Rs256Tables tab)
{
    extern __shared__ uint32_t s_mem[];
    uint32_t *s_lut = s_mem;
    uint8_t *s_exp2 = reinterpret_cast<uint8_t *>(s_mem + S * 256 * lut_row(EW));
    uint16_t *s_log2 = reinterpret_cast<uint16_t *>(s_exp2 + EXP2_LEN);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(S * 256 * lut_row(EW)); i += blockDim.x)
        s_lut[i] = tab.lut[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += blockDim.x)
        s_exp2[i] = tab.exp2[i];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x)
        s_log2[i] = tab.log2[i];
    __syncthreads();
    GfLds gf{s_exp2, s_log2};
    const uint32_t nrec = (msg_bytes + 3u) & ~3u;
    __amdgpu_buffer_rsrc_t rsrc = make_rsrc(msgs, nrec);
    for (uint32_t cw = blockIdx.x * blockDim.x + threadIdx.x; cw < count; cw += gridDim.x * blockDim.x) {
        uint32_t rem[EW];
        message_remainder<EW, S>(rsrc, nrec, (int64_t)cw * k, k, s_lut, rem);
        xor_parity<EW>(rem, eccs + (size_t)cw * ecc, ecc);
        syndromes_from_remainder<EW>(rem, ecc, gf, synds + (size_t)cw * ecc);
    }
}

// ======================================================================
// Repair pipeline for ecc <= 32 (Lambda has at most 16 roots):
//   rs256_synd_kernel    lane per codeword: remainder + XOR parity; zero -> status 1;
//                        else syndromes from nibble tables ST[j][h][v] (E bytes each:
//                        (v << 4h) alpha^(i (E-1-j)), i < E) and status PENDING
//   rs256_bm_kernel      lane per codeword: Berlekamp-Massey; L > E/2 -> status 0; else
//                        a record {L, Lambda_0..L, Omega_0..L-1} (Omega = S Lambda mod x^E
//                        has degree < L once the LFSR of length L generates all E syndromes)
//                        (ecc 2, 4, 8, 16, 32: the Chien kernel runs this BM itself per tile
//                        and keeps the records in registers)
//   rs256_chien16_kernel Chien search, eight lanes per codeword over nibble tables, then
//                        Forney at the roots.
// ======================================================================

constexpr uint8_t RS_PENDING = 0xFE;
constexpr int CHIEN_LMAX = 16;
constexpr uint32_t CHIEN_TABLE_BYTES = CHIEN_LMAX * 2 * 16 * 256;

// BM record of one codeword (ecc <= 32): byte 0 L, bytes 4..19 Omega_0..15,
// bytes 20..36 Lambda_0..16, zero-filled, 48 bytes.
__host__ __device__ inline uint32_t rec_lam_off() { return 20; }
__host__ __device__ inline uint32_t rec_om_off(uint32_t) { return 4; }
__host__ __device__ inline uint32_t rec_stride(uint32_t) { return 48; }

// LM: message loads, as for rs256_encode_kernel.
template <int EW, int S, int WG, int LM = LOAD_LANE>
__global__ __launch_bounds__(WG) void rs256_synd_kernel(const uint8_t *__restrict__ msgs, uint32_t msg_bytes,
                                                          const uint8_t *__restrict__ eccs, uint32_t count,
                                                          uint32_t k, uint32_t E, uint8_t *__restrict__ status,
                                                          uint32_t *__restrict__ synd, const uint32_t *__restrict__ lut,
                                                          const uint32_t *__restrict__ stab)
{
    extern __shared__ uint32_t s_mem[];
    uint32_t *s_lut = lfsr_table<EW, S>(s_mem);
    uint32_t *s_st = s_mem + lfsr_dyn_dwords(EW, S);
    for (uint32_t i = threadIdx.x; i < (uint32_t)lut_dwords(EW, S) / 4; i += blockDim.x)
        reinterpret_cast<uint4 *>(s_lut)[i] = reinterpret_cast<const uint4 *>(lut)[i];
    for (uint32_t i = threadIdx.x; i < E * 32 * EW; i += blockDim.x)
        s_st[i] = stab[i];
    __syncthreads();
    const uint32_t nrec = (msg_bytes + 3u) & ~3u;
    __amdgpu_buffer_rsrc_t rsrc = make_rsrc(msgs, nrec);
    // whole-word parity rows, loaded before the message so that their latency overlaps the LFSR
    auto load_parity = [&](uint32_t cw, uint32_t (&par)[EW]) {
        const uint8_t *pe = eccs + (size_t)cw * E;
#pragma unroll
        for (int d = 0; d < EW; ++d)
            par[d] = (E & 3) == 0 && (uint32_t)d * 4 < E ? reinterpret_cast<const uint32_t *>(pe)[d] : 0u;
    };
    // remainder ^ parity: zero -> status 1; else the syndromes from the nibble tables, PENDING
    auto finish = [&](uint32_t cw, uint32_t (&rem)[EW], const uint32_t (&par)[EW]) {
        uint32_t nz = 0;
        if ((E & 3) == 0) {
#pragma unroll
            for (int d = 0; d < EW; ++d) {
                rem[d] ^= par[d];
                nz |= rem[d];
            }
        } else {
            nz = xor_parity<EW>(rem, eccs + (size_t)cw * E, E);
        }
        if (nz == 0) {
            status[cw] = 1;
            return;
        }
        uint32_t acc[EW];
#pragma unroll
        for (int d = 0; d < EW; ++d)
            acc[d] = 0;
#pragma unroll
        for (int j = 0; j < EW * 4; ++j) {
            if ((uint32_t)j < E) {
                uint32_t r = (rem[j >> 2] >> (8 * (j & 3))) & 0xffu;
                const uint32_t *lo = s_st + ((j * 2 + 0) * 16 + (r & 15)) * EW;
                const uint32_t *hi = s_st + ((j * 2 + 1) * 16 + (r >> 4)) * EW;
                if constexpr (EW % 4 == 0) {
#pragma unroll
                    for (int d = 0; d < EW; d += 4) {
                        uint4 a = *reinterpret_cast<const uint4 *>(lo + d);
                        uint4 b = *reinterpret_cast<const uint4 *>(hi + d);
                        acc[d] ^= a.x ^ b.x;
                        acc[d + 1] ^= a.y ^ b.y;
                        acc[d + 2] ^= a.z ^ b.z;
                        acc[d + 3] ^= a.w ^ b.w;
                    }
                } else {
#pragma unroll
                    for (int d = 0; d < EW; ++d)
                        acc[d] ^= lo[d] ^ hi[d];
                }
            }
        }
        uint32_t *o = synd + (size_t)cw * EW;
#pragma unroll
        for (int d = 0; d < EW; ++d)
            o[d] = acc[d];
        status[cw] = RS_PENDING;
    };
    const uint32_t stride = gridDim.x * blockDim.x;
    if constexpr (LM == LOAD_COOP) {
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        for (uint32_t base = (blockIdx.x * (WG / 64) + wave) * 64; base < count; base += stride) {
            const uint32_t cw = base + (threadIdx.x & 63);
            uint32_t d[16][4], par[EW], rem[EW];
            load_parity(cw < count ? cw : base, par);
            coop_issue(rsrc, base, k, d);
            coop_remainder<EW, S>(rsrc, base, k, d, s_lut, rem);
            if (cw < count)
                finish(cw, rem, par);
        }
    } else {
        for (uint32_t cw = blockIdx.x * blockDim.x + threadIdx.x; cw < count; cw += stride) {
            uint32_t par[EW], rem[EW];
            load_parity(cw, par);
            message_remainder<EW, S>(rsrc, nrec, (int64_t)cw * k, k, s_lut, rem);
            finish(cw, rem, par);
        }
    }
}

// Per-lane LDS row of the BM kernel: S[E], logS[E] (u16), Lambda[E], B[E].
__host__ __device__ inline uint32_t bm_row_stride(uint32_t E)
{
    uint32_t dw = (E + 2 * E + E + E + 3) / 4;
    return ((dw & 1) ? dw : dw + 1) * 4;
}

template <int WG>
__global__ __launch_bounds__(WG) void rs256_bm_kernel(uint32_t count, uint32_t E, uint32_t sstride_words,
                                                       uint8_t *__restrict__ status, const uint32_t *__restrict__ synd,
                                                       uint8_t *__restrict__ rec, Rs256Tables tab)
{
    extern __shared__ uint32_t s_mem[];
    uint8_t *s_exp2 = reinterpret_cast<uint8_t *>(s_mem);
    uint16_t *s_log2 = reinterpret_cast<uint16_t *>(s_exp2 + EXP2_LEN);
    uint8_t *s_rows = reinterpret_cast<uint8_t *>(s_log2 + 256);
    for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += blockDim.x)
        s_exp2[i] = tab.exp2[i];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x)
        s_log2[i] = tab.log2[i];
    __syncthreads();
    GfLds gf{s_exp2, s_log2};
    uint8_t *row = s_rows + threadIdx.x * bm_row_stride(E);
    uint8_t *Sy = row;
    uint16_t *lS = reinterpret_cast<uint16_t *>(row + ((E + 1) & ~1u));
    uint8_t *lam = row + ((E + 1) & ~1u) + 2 * E;
    uint8_t *Bx = lam + E;
    const uint32_t RS = rec_stride(E), LO = rec_lam_off(), OO = rec_om_off(E);

    for (uint32_t cw = blockIdx.x * blockDim.x + threadIdx.x; cw < count; cw += gridDim.x * blockDim.x) {
        if (status[cw] != RS_PENDING)
            continue;
        const uint32_t *sw = synd + (size_t)cw * sstride_words;
        for (uint32_t w = 0; w < (E + 3) / 4; ++w) {
            uint32_t v = sw[w];
            for (uint32_t b = 0; b < 4 && w * 4 + b < E; ++b) {
                uint32_t sb = (v >> (8 * b)) & 0xffu;
                Sy[w * 4 + b] = (uint8_t)sb;
                lS[w * 4 + b] = gf.log2[sb];
            }
        }
        for (uint32_t i = 0; i < E; ++i) {
            lam[i] = 0;
            Bx[i] = 0;
        }
        lam[0] = 1;
        Bx[0] = 1;
        // reed_solomon.hpp:655-702, natural order; lenL / lenB bound the non-zero terms
        uint32_t L = 0, m = 1, b = 1, lenL = 1, lenB = 1;
        for (uint32_t nn = 0; nn < E; ++nn) {
            uint32_t d = Sy[nn];
            uint32_t top = L < lenL - 1 ? L : lenL - 1;
            for (uint32_t i = 1; i <= top; ++i)
                d ^= gf.exp2[gf.log2[lam[i]] + lS[nn - i]];
            if (d == 0) {
                m += 1;
                continue;
            }
            uint32_t lc = gf.log2[gf.div(d, b)];
            bool grow = 2 * L <= nn;
            uint32_t hiB = m + lenB < E ? m + lenB : E;  // x^m B truncated to degrees < E
            uint32_t hi = grow ? (hiB > lenL ? hiB : lenL) : hiB;
            for (uint32_t ii = hi; ii-- > 0;) {
                uint32_t old = lam[ii];
                uint32_t sb = (ii >= m && ii < hiB) ? (uint32_t)Bx[ii - m] : 0u;
                lam[ii] = (uint8_t)(old ^ gf.mul_log(lc, sb));
                if (grow)
                    Bx[ii] = (uint8_t)old;
            }
            if (grow) {
                lenB = lenL;
                L = nn + 1 - L;
                b = d;
                m = 1;
            } else {
                m += 1;
            }
            if (hiB > lenL)
                lenL = hiB;
        }
        if (L > E / 2) {
            status[cw] = 0;
            continue;
        }
        uint8_t *r = rec + (size_t)cw * RS;
        for (uint32_t i = 0; i < RS; ++i)
            r[i] = 0;
        r[0] = (uint8_t)L;
        for (uint32_t i = 0; i <= E / 2; ++i)
            r[LO + i] = i <= L ? lam[i] : 0;
        // Omega_k = sum_{i <= k} Lambda_i S_{k-i}, k < L (reed_solomon.hpp:708-728)
        for (uint32_t kk = 0; kk < E / 2; ++kk) {
            uint32_t v = 0;
            if (kk < L)
                for (uint32_t i = 0; i <= kk; ++i)
                    v ^= gf.exp2[gf.log2[lam[i]] + lS[kk - i]];
            r[OO + kk] = (uint8_t)v;
        }
    }
}

// Log / exp tables of the register BM: EXP2 (1,024 bytes, zero above 509) and LOG2 (u16,
// log 0 = ZLOG = 511), so that a product of logs is one read at the sum.
struct LogExp1K {
    const uint8_t *ex;
    const uint16_t *lg;
    static constexpr uint32_t Z = ZLOG;
    __device__ inline uint32_t log(uint32_t a) const { return lg[a]; }
    __device__ inline uint32_t exp_sum(uint32_t s) const { return ex[s]; }
};

// Berlekamp-Massey with every polynomial in registers, fully unrolled for a fixed E.
// B is kept pre-shifted as x^m B (its logs only), so an update is Lambda_i ^= c B'_i and
// every iteration shifts B' by one: B' <- x Lambda_old when L grows, else x B'. At
// iteration n deg Lambda <= n + 1, which bounds the unrolled work. Same results as the
// reference's reversed-storage version (reed_solomon.hpp:655-702), whose truncation of
// degrees >= E never matters while L <= E / 2.
// Degree-bounded BM of one codeword from its packed syndromes sw[0..EW) (the tables in LDS).
// Returns false when L > E/2 (not repairable); otherwise the 48-byte record in words[]:
// L in byte 0, Omega_0..Omega_(E/2-1) from byte rec_om_off, Lambda_0..Lambda_(E/2) from byte
// rec_lam_off.
template <int E, class T>
__device__ inline bool bm_fixed_one(const uint32_t *__restrict__ sw, const T &tb, uint32_t (&words)[12])
{
    constexpr int EW = (E + 3) / 4;
    constexpr uint32_t ZL = T::Z;
    const uint32_t LO = rec_lam_off(), OO = rec_om_off(E);
    // the syndromes stay packed four to a register (S_n is extracted where it is used)
    uint32_t sv[EW], lS[E];
#pragma unroll
    for (int w = 0; w < EW; ++w) {
        sv[w] = sw[w];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (w * 4 + b < E)
                lS[w * 4 + b] = tb.log((sv[w] >> (8 * b)) & 0xffu);
    }
    uint32_t lam[E], llam[E], lBp[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        lam[i] = i == 0 ? 1u : 0u;
        llam[i] = i == 0 ? 0u : ZL;
        lBp[i] = i == 1 ? 0u : ZL;
    }
    // L never decreases, and a codeword whose L exceeds E/2 fails whatever follows; while
    // L <= E/2, deg Lambda <= L and every x^m B that enters an update has degree <= the new
    // L. So terms above E/2 (above E/2 + 1 for B') never reach a successful result.
    constexpr int H = E / 2;
    uint32_t L = 0, lb = 0;
#pragma unroll
    for (int n = 0; n < E; ++n) {
        uint32_t d = (sv[n >> 2] >> (8 * (n & 3))) & 0xffu;
#pragma unroll
        for (int i = 1; i <= (n < H ? n : H); ++i)
            d ^= tb.exp_sum(llam[i] + lS[n - i]);
        bool nz = d != 0;
        bool grow = nz && (2 * L <= (uint32_t)n);
        uint32_t ld = tb.log(d);
        uint32_t lc = ld + 255 - lb;
        lc = lc >= 255 ? lc - 255 : lc;
        // d == 0: the zero log puts every product below at zero, so no per-term select is needed
        lc = nz ? lc : ZL;
        const int hi = n + 1 < H ? n + 1 : H;
        uint32_t old_llam[E];
#pragma unroll
        for (int i = 0; i < E; ++i)
            old_llam[i] = llam[i];
#pragma unroll
        for (int i = 1; i <= hi; ++i) {
            lam[i] ^= tb.exp_sum(lc + lBp[i]);
            llam[i] = tb.log(lam[i]);
        }
        const int bh = n + 2 < H + 1 ? n + 2 : (H + 1 < E - 1 ? H + 1 : E - 1);
#pragma unroll
        for (int i = bh; i >= 1; --i)
            lBp[i] = grow ? old_llam[i - 1] : lBp[i - 1];
        lBp[0] = ZL;
        if (grow) {
            L = (uint32_t)n + 1 - L;
            lb = ld;
        }
    }
    if (L > (uint32_t)E / 2)
        return false;
#pragma unroll
    for (int w = 0; w < 12; ++w)
        words[w] = 0;
    words[0] = L;
#pragma unroll
    for (int i = 0; i <= E / 2; ++i)
        words[(LO + i) >> 2] |= (i <= (int)E ? lam[i < E ? i : 0] : 0u) << (8 * ((LO + i) & 3));
#pragma unroll
    for (int kk = 0; kk < E / 2; ++kk) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i <= kk; ++i)
            v ^= tb.exp_sum(llam[i] + lS[kk - i]);
        v = (uint32_t)kk < L ? v : 0u;
        words[(OO + kk) >> 2] |= v << (8 * ((OO + kk) & 3));
    }
    return true;
}

// XOR a byte in global memory without reading it: an atomic XOR on the aligned dword
// that holds it (the other three bytes get 0), issued without waiting for a result.
__device__ inline void xor_byte(uint8_t *p, uint32_t v)
{
    // p points into device memory: a global (not flat) atomic, which LDS waits do not
    // have to drain
    typedef __attribute__((address_space(1))) uint32_t global_u32;
    uintptr_t a = reinterpret_cast<uintptr_t>(p);
    global_u32 *w = (global_u32 *)(a & ~(uintptr_t)3);
    __hip_atomic_fetch_xor(w, v << (8 * (a & 3)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ======================================================================
// Chien search and Forney, G = 16 or 8 lanes per codeword (records of the BM kernels, L <= 16;
// reed_solomon.hpp:516-538, 704-743). With G = 16, lane j of a group holds positions
// 16j .. 16j+15 of its codeword as the sixteen bytes of four dwords; with G = 8, 32 positions
// in eight dwords (16j .. 16j+15 and 128+16j .. 128+16j+15). For a coefficient Lambda_i
// (i = 1..16) with nibbles v_0, v_1, a 16-byte read of row CT[i][h][v_h] per nibble and piece
// gives (v_h << 4h) alpha^(-ip) at sixteen of the lane's positions, so a wave evaluates 64 / G
// codewords per pass with no value read across lanes. A 64-codeword tile is taken in ascending
// L (the codewords of a pass have nearly the same degree and the pass stops at the largest).
// Where the zeros among positions [0, n) number exactly L, the pass's roots are listed one per
// slot in LDS, and each lane
// evaluates Omega at its root as a sum of log / exp reads (the logs of Omega's coefficients
// taken once per codeword) and applies the magnitude: with Lambda'(X^-1) = X Lambda_odd(X^-1)
// (characteristic 2), the reference's e = X div(Omega(X^-1), Lambda'(X^-1)) is
// div(Omega, Lambda_odd), or X Omega where Lambda_odd = 0 (its div(a, 0) = a).
// ======================================================================

// Nibble bytes of coefficients 1..16 of a polynomial whose coefficient c sits at byte
// `first + c` of the record words r: dword m holds the low and high nibbles of coefficient
// 2m+1, then of coefficient 2m+2.
template <int FIRST>
__device__ inline void record_nibbles(const uint32_t (&r)[12], uint32_t (&nib)[8])
{
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const uint32_t q = alignbyte(r[(FIRST + 4 * g + 4) / 4], r[(FIRST + 4 * g) / 4], (FIRST + 1) & 3);
#pragma unroll
        for (int hw = 0; hw < 2; ++hw) {
            const uint32_t t = (q >> (16 * hw)) & 0xffffu;
            const uint32_t b = (t & 0xffu) | ((t & 0xff00u) << 8);  // bytes a, 0, b, 0
            nib[2 * g + hw] = (b & 0x000f000fu) | ((b & 0x00f000f0u) << 4);
        }
    }
}

// acc ^= sum over coefficients i = 1..top (top <= IMAX, wave-uniform) of Lambda_i alpha^(-ip) at
// the lane's positions, odd coefficients into acc_odd when SPLIT. D = 4: sixteen positions, one
// 16-byte read per nibble at `off`; D = 8: thirty-two, a second read at off ^ 128 (the other
// half of the 256-byte row).
template <bool SPLIT, int IMAX, int D>
__device__ inline void chien_eval(const char *ct, const uint32_t (&nib)[8], uint32_t lo_off, uint32_t hi_off,
                                  uint32_t top, uint32_t (&acc)[D], uint32_t (&acc_odd)[D])
{
    static_assert(D == 4 || D == 8, "16 or 32 positions per lane");
#pragma unroll
    for (int gq = 0; gq < 4; ++gq) {
        if (4 * gq + 1 > (int)top)
            break;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int i = 4 * gq + jj + 1;
            if (i > IMAX)
                break;
            const uint32_t word = nib[2 * gq + (jj >> 1)];
            uint4 t[2][D / 4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int tbl = (i - 1) * 2 + h;  // 4 KB table index
                const uint32_t sel = 0x0c020000u | ((4u + (uint32_t)(2 * (jj & 1) + h)) << 8);
                const uint32_t a = __builtin_amdgcn_perm(word, tbl >= 16 ? hi_off : lo_off, sel);
                t[h][0] = *reinterpret_cast<const uint4 *>(ct + a + (tbl & 15) * 4096);
                if constexpr (D == 8)
                    t[h][1] = *reinterpret_cast<const uint4 *>(ct + (a ^ 128u) + (tbl & 15) * 4096);
            }
            uint32_t(&d)[D] = (SPLIT && (i & 1)) ? acc_odd : acc;
#pragma unroll
            for (int r = 0; r < D / 4; ++r) {
                d[4 * r + 0] = xor3(d[4 * r + 0], t[0][r].x, t[1][r].x);
                d[4 * r + 1] = xor3(d[4 * r + 1], t[0][r].y, t[1][r].y);
                d[4 * r + 2] = xor3(d[4 * r + 2], t[0][r].z, t[1][r].z);
                d[4 * r + 3] = xor3(d[4 * r + 3], t[0][r].w, t[1][r].w);
            }
        }
    }
}

// Position of byte q of position dword d of lane j in codeword group grp: G = 16 lanes hold 16
// consecutive positions each; G = 8 lanes hold two runs of 16, one in each half of the
// codeword, the half of dwords 0-3 alternating with the group so that one read instruction of
// the eight groups covers all 64 LDS banks.
template <int G>
__device__ inline uint32_t chien_pos(uint32_t j, uint32_t grp, int d, int q)
{
    if constexpr (G == 16)
        return 16 * j + 4 * (uint32_t)d + (uint32_t)q;
    else
        return 128 * (((uint32_t)d >> 2) ^ (grp & 1u)) + 16 * j + 4 * ((uint32_t)d & 3u) + (uint32_t)q;
}

// Sum over the G lanes of this lane's codeword group (DPP butterflies), in every lane of it.
template <int G>
__device__ inline uint32_t group_sum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false);  // row_half_mirror
    if constexpr (G == 16)
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false);  // row_mirror
    return x;
}

// Orders this wave's LDS writes before its later reads of them by other lanes.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}


// EB > 0: the BM records come from bm_fixed_one<EB> run per tile, one pending codeword per lane,
// on the syndromes `synd` (its LDS lookups overlap the Chien loops of other waves); EB == 0:
// from `rec`. G lanes per codeword in the Chien passes: 16 (sixteen positions each, four
// codewords per pass) or 8 (thirty-two each, eight per pass: half the per-pass work per
// codeword, for small ecc where that work dominates).
template <int WG, int EB, int G = 16>
__global__ __launch_bounds__(WG) void rs256_chien16_kernel(uint8_t *__restrict__ msgs, uint8_t *__restrict__ eccs,
                                                            uint32_t count, uint32_t k, uint32_t E,
                                                            uint8_t *__restrict__ status,
                                                            const uint32_t *__restrict__ rec,
                                                            const uint32_t *__restrict__ synd,
                                                            const uint32_t *__restrict__ ctab, Rs256Tables tab)
{
    // CT[i-1][h][v]: 256 position bytes (v << 4h) alpha^(-i p), at a known LDS address so that
    // the table part of every offset is an immediate
    __shared__ uint4 s_ct4[CHIEN_TABLE_BYTES / 16];
    __shared__ uint8_t s_exp2[EXP2_LEN];
    __shared__ uint16_t s_log2[256];
    static_assert(G == 16 || G == 8, "16 or 8 lanes per codeword");
    // with the BM in the kernel (EB > 0) L <= EB / 2: only the nibble words of coefficients
    // 1 .. EB / 2 travel between lanes and only that many degrees are ranked
    constexpr int LMX = EB > 0 ? EB / 2 : CHIEN_LMAX;
    constexpr int NW = EB > 0 ? (EB >= 8 ? EB / 4 : 1) : 8;
    constexpr int D = 64 / G;               // position dwords per lane, and codewords per pass
    constexpr int RS = D * LMX > 64 ? D * LMX : 64;  // root slots per wave (a pass has <= D LMX roots)
    __shared__ uint32_t s_roots[WG / 64 * RS];
    for (uint32_t i = threadIdx.x; i < CHIEN_TABLE_BYTES / 16; i += WG)
        s_ct4[i] = reinterpret_cast<const uint4 *>(ctab)[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += WG)
        s_exp2[i] = tab.exp2[i];
    for (uint32_t i = threadIdx.x; i < 256u; i += WG)
        s_log2[i] = tab.log2[i];
    __syncthreads();
    const LogExp1K tb{s_exp2, s_log2};
    const char *ct = reinterpret_cast<const char *>(s_ct4);
    const uint32_t lane = threadIdx.x & 63, j = lane % G, grp = lane / G;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n = k + E;
    // byte offset of the lane's first 16-byte piece within a table row
    const uint32_t lo_off = 16 * j + (G == 8 ? 128 * (grp & 1u) : 0u), hi_off = lo_off + 0x10000u;
    uint32_t *s_rt = s_roots + RS * wave;
    for (uint32_t base = (blockIdx.x * (WG / 64) + wave) * 64; base < count; base += gridDim.x * WG) {
        const uint32_t cw = base + lane;
        bool pend = cw < count && status[cw] == RS_PENDING;
        uint32_t r[12];
        if constexpr (EB > 0) {
#pragma unroll
            for (int w = 0; w < 12; ++w)
                r[w] = 0;
            if (pend && !bm_fixed_one<EB>(synd + (size_t)cw * ((EB + 3) / 4), tb, r)) {
                status[cw] = 0;
                pend = false;
            }
        }
        const uint64_t pm = __ballot(pend);
        if (!pm)
            continue;
        if (EB > 0) {
        } else if (pend) {
            const uint4 *r4 = reinterpret_cast<const uint4 *>(rec + (size_t)cw * 12);
            const uint4 a = r4[0], b = r4[1], c = r4[2];
            r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
            r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
            r[8] = c.x; r[9] = c.y; r[10] = c.z; r[11] = c.w;
        } else {
#pragma unroll
            for (int w = 0; w < 12; ++w)
                r[w] = 0;
        }
        const uint32_t L = r[0] & 0xffu;
        // bit 7 of byte q of dword d: position chien_pos(d, q) lies in the codeword (set up
        // after BM, whose registers it would otherwise add to)
        uint32_t vmask[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            vmask[d] = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                vmask[d] |= (chien_pos<G>(j, grp, d, q) < n) ? (0x80u << (8 * q)) : 0u;
        }
        uint32_t lnib[8];
        record_nibbles<20>(r, lnib);  // Lambda_1..Lambda_16 (Lambda_0 = 1 at byte 20)
        // logs of Omega_0 .. Omega_(LMX-1) (ZLOG for zero; Omega_k = 0 for k >= L), two per
        // word, for the magnitudes at the roots
        constexpr int NWO = (LMX + 1) / 2;
        uint32_t olog[NWO];
#pragma unroll
        for (int m = 0; m < NWO; ++m) {
            uint32_t lw = 0;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int kk = 2 * m + hh;
                const uint32_t b = kk < LMX ? (r[(4 + kk) >> 2] >> (8 * ((4 + kk) & 3))) & 0xffu : 0u;
                lw |= (uint32_t)s_log2[b] << (16 * hh);
            }
            olog[m] = lw;
        }
        // rank of each pending codeword in ascending (L, lane) order; others after them
        const uint32_t npend = (uint32_t)__builtin_popcountll(pm);
        uint32_t rank = 0, acc = 0;
#pragma unroll
        for (uint32_t v = 1; v <= (uint32_t)LMX; ++v) {
            const bool hit = pend && L == v;
            const uint64_t m = __ballot(hit);
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            rank = hit ? acc + below : rank;
            acc += (uint32_t)__builtin_popcountll(m);
        }
        if (!pend)
            rank = npend + __builtin_amdgcn_mbcnt_hi((uint32_t)(~pm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)~pm, 0u));
        // lane r receives the lane of the codeword ranked r
        const uint32_t src = (uint32_t)__builtin_amdgcn_ds_permute((int)(rank * 4), (int)lane);
        const uint32_t npass = (npend + D - 1) / D;
        for (uint32_t ps = 0; ps < npass; ++ps) {
            const uint32_t rr = D * ps + grp;
            const bool active = rr < npend;
            const uint32_t o = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((rr & 63) * 4), (int)src) & 63u;
            uint32_t ln[8];
#pragma unroll
            for (int m = 0; m < 8; ++m)
                ln[m] = m < NW ? (uint32_t)__builtin_amdgcn_ds_bpermute((int)(o * 4), (int)lnib[m]) : 0u;
            const uint32_t Lg = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(o * 4), (int)L);
            const uint32_t last = (npend - 1 < D * ps + D - 1 ? npend - 1 : D * ps + D - 1) - D * ps;
            const uint32_t Lmax = (uint32_t)__builtin_amdgcn_readlane((int)Lg, (int)(G * last));
            uint32_t ev[D], od[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                ev[d] = 0x01010101u;  // Lambda_0 = 1
                od[d] = 0;
            }
            chien_eval<true, LMX, D>(ct, ln, lo_off, hi_off, Lmax, ev, od);
            uint32_t z[D], cnt = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const uint32_t v = ev[d] ^ od[d];
                const uint32_t t = (v & 0x7f7f7f7fu) + 0x7f7f7f7fu;
                z[d] = __builtin_amdgcn_bitop3_b32(t, v, vmask[d], 0x02);  // ~t & ~v & mask: zero bytes
                cnt += (uint32_t)__builtin_popcount(z[d]);
            }
            cnt = group_sum<G>(cnt);
            const bool ok = active && cnt == Lg;
            if (j == 0 && active)
                status[base + o] = ok ? 1 : 0;
            if (!__ballot(ok))
                continue;
            // the pass's roots (at most D LMX) into this wave's slots, in lane order: bit 8q + d
            // of rm <-> position chien_pos(d, q); a slot holds the codeword's lane, the
            // position and Lambda_odd(X^-1) there
            uint32_t rm = 0;
#pragma unroll
            for (int d = 0; d < D; ++d)
                rm |= z[d] >> (7 - d);
            rm = ok ? rm : 0u;
            const uint32_t nr = (uint32_t)__builtin_popcount(rm);
            uint32_t slot = 0, total = 0;  // exclusive prefix sum of nr (<= 4 D) by bit planes
#pragma unroll
            for (int b = 0; b < (D == 8 ? 6 : 5); ++b) {
                const uint64_t m = __ballot((nr >> b) & 1u);
                slot += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << b;
                total += (uint32_t)__builtin_popcountll(m) << b;
            }
            while (rm) {
                const uint32_t t = (uint32_t)__builtin_ctz(rm);
                rm &= rm - 1;
                const uint32_t q = t >> 3, d = t & 7u;
                uint32_t ov = od[0], p = chien_pos<G>(j, grp, 0, 0) + q;
#pragma unroll
                for (int dd = 1; dd < D; ++dd) {
                    ov = d == (uint32_t)dd ? od[dd] : ov;
                    p = d == (uint32_t)dd ? chien_pos<G>(j, grp, dd, 0) + q : p;
                }
                s_rt[slot++] = o | (p << 8) | (((ov >> (8 * q)) & 0xffu) << 16);
            }
            wave_lds_sync();
            // one root per lane: Omega(X^-1) = sum_k Omega_k alpha^(-k p) from the logs held by
            // the codeword's lane, then Forney's magnitude (as the reference's e = X Omega(X^-1) /
            // Lambda'(X^-1), with Lambda' = X Lambda_odd in characteristic 2)
            for (uint32_t r0 = 0; r0 < total; r0 += 64) {
                const uint32_t e = r0 + lane < total ? s_rt[r0 + lane] : 0u;
                const uint32_t oc = e & 63u;
                uint32_t ol[NWO];
#pragma unroll
                for (int m = 0; m < NWO; ++m)
                    ol[m] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(oc * 4), (int)olog[m]);
                if (r0 + lane < total) {
                    const uint32_t p = (e >> 8) & 0xffu, dod = e >> 16;
                    // four terms at a time (their reads in flight together), up to the pass's
                    // largest degree; ek = -k p mod 255
                    uint32_t num = 0, ek = 0;
#pragma unroll
                    for (int g4 = 0; g4 < (LMX + 3) / 4; ++g4) {
                        if ((uint32_t)(4 * g4) >= Lmax)
                            break;
                        uint32_t x[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int kk = 4 * g4 + u;
                            x[u] = kk < LMX ? s_exp2[((ol[kk >> 1] >> (16 * (kk & 1))) & 0xffffu) + ek] : 0u;
                            ek = ek >= p ? ek - p : ek + 255 - p;
                        }
                        num ^= x[0] ^ x[1] ^ x[2] ^ x[3];
                    }
                    if (num) {
                        const uint32_t lnm = s_log2[num], ld = s_log2[dod];
                        const uint32_t mag = dod ? (uint32_t)s_exp2[lnm + 255 - ld] : (uint32_t)s_exp2[p + lnm];
                        const uint32_t at = n - 1 - p;
                        const size_t cwo = base + oc;
                        xor_byte(at < k ? msgs + cwo * k + at : eccs + cwo * E + (at - k), mag);
                    }
                }
            }
            wave_lds_sync();  // the next pass rewrites the slots
        }
    }
}


// ======================================================================
// Mixed-size batches: a plan holds a work list of codewords with their own message and
// ecc lengths, sorted by descending ecc then descending length and cut into chunks of at most
// WG codewords of one ecc. One persistent launch serves the whole list: each workgroup
// takes the next chunk from a device counter, loads that ecc's LFSR table into LDS when
// it differs from the one it holds, and runs one codeword per lane.
// ======================================================================

struct PlanItem {
    uint64_t msg_off;  // byte offset of the message in the message buffer
    uint64_t ecc_off;  // byte offset of the parity in the parity buffer
    uint32_t idx;      // position of the codeword in the caller's list (status index)
    uint16_t k;
    uint8_t E;
    uint8_t cls;
};

struct PlanChunk {
    uint32_t first;
    uint16_t count;
    uint16_t cls;
};

struct PlanClass {
    const uint32_t *lut;
    uint32_t ecc, EW, S;
};

__global__ void plan_reset_kernel(uint32_t *queue) { *queue = 0; }

template <int EW, int S, bool REPAIR>
__device__ inline void plan_item(__amdgpu_buffer_rsrc_t rsrc, uint32_t nrec, uint8_t *msgs, uint8_t *eccs, uint8_t *status,
                                 const PlanItem &it, const uint32_t *s_lut, const GfLds &gf, uint8_t *row)
{
    if constexpr (REPAIR) {
        status[it.idx] = repair_one<EW, S>(rsrc, nrec, (int64_t)it.msg_off, msgs + it.msg_off, eccs + it.ecc_off, it.k,
                                           it.E, s_lut, gf, row);
    } else {
        uint32_t rem[EW];
        message_remainder<EW, S>(rsrc, nrec, (int64_t)it.msg_off, it.k, s_lut, rem);
        uint8_t *o = eccs + it.ecc_off;
        if ((it.E & 3) == 0 && (it.ecc_off & 3) == 0) {
            uint32_t *o4 = reinterpret_cast<uint32_t *>(o);
#pragma unroll
            for (int d = 0; d < EW; ++d)
                if ((uint32_t)d * 4 < it.E)
                    o4[d] = rem[d];
        } else {
            for (uint32_t j = 0; j < it.E; ++j)
                o[j] = rem_byte<EW>(rem, j);
        }
    }
}

template <int MAXEW, int WG, bool REPAIR>
__global__ __launch_bounds__(WG) void rs256_plan_kernel(uint8_t *__restrict__ msgs, uint32_t msg_bytes,
                                                         uint8_t *__restrict__ eccs, uint8_t *__restrict__ status,
                                                         const PlanItem *__restrict__ items,
                                                         const PlanChunk *__restrict__ chunks, uint32_t nchunks,
                                                         const PlanClass *__restrict__ classes, uint32_t *queue,
                                                         uint32_t chunk0, uint32_t lut_words, uint32_t row_stride,
                                                         Rs256Tables tab)
{
    extern __shared__ uint32_t s_mem[];
    __shared__ uint32_t s_chunk;
    uint32_t *s_lut = s_mem;
    uint8_t *s_exp2 = reinterpret_cast<uint8_t *>(s_mem + lut_words);
    uint16_t *s_log2 = reinterpret_cast<uint16_t *>(s_exp2 + EXP2_LEN);
    uint8_t *s_rows = reinterpret_cast<uint8_t *>(s_log2 + 256);
    if constexpr (REPAIR) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += WG)
            s_exp2[i] = tab.exp2[i];
        for (uint32_t i = threadIdx.x; i < 256u; i += WG)
            s_log2[i] = tab.log2[i];
    }
    GfLds gf{s_exp2, s_log2};
    uint8_t *row = s_rows + threadIdx.x * row_stride;
    const uint32_t nrec = (msg_bytes + 3u) & ~3u;
    __amdgpu_buffer_rsrc_t rsrc = make_rsrc(msgs, nrec);
    uint32_t loaded = 0xffffffffu;
    for (;;) {
        if (threadIdx.x == 0)
            s_chunk = chunk0 + atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t c = s_chunk;
        if (c >= nchunks)
            break;
        const PlanChunk ch = chunks[c];
        const PlanClass cl = classes[ch.cls];
        if (ch.cls != loaded) {
            const uint32_t words = cl.S * 256u * (uint32_t)lut_row((int)cl.EW);
            for (uint32_t i = threadIdx.x; i < words; i += WG)
                s_lut[i] = cl.lut[i];
            loaded = ch.cls;
        }
        __syncthreads();
        if (threadIdx.x < ch.count) {
            const PlanItem it = items[ch.first + threadIdx.x];
            switch (cl.EW) {
            case 1: plan_item<1, 4, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row); break;
            case 2: plan_item<2, 4, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row); break;
            case 4: plan_item<4, 4, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row); break;
            case 8: plan_item<8, 4, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row); break;
            case 16: plan_item<16, 2, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row); break;
            default:
                if constexpr (MAXEW >= 32) {
                    if (cl.EW == 32)
                        plan_item<32, 1, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row);
                    else
                        plan_item<64, 1, REPAIR>(rsrc, nrec, msgs, eccs, status, it, s_lut, gf, row);
                }
                break;
            }
        }
        __syncthreads();
    }
}

// Repair of the plan items [0, nitems) (ecc_len > 32), one per wave (repair_wave).
template <int WG>
__global__ __launch_bounds__(WG) void rs256_plan_wave_kernel(uint8_t *__restrict__ msgs, uint8_t *__restrict__ eccs,
                                                              uint8_t *__restrict__ status,
                                                              const PlanItem *__restrict__ items, uint32_t nitems,
                                                              Rs256Tables tab)
{
    __shared__ uint8_t s_exp2[EXP2_LEN];
    __shared__ uint16_t s_log2[256];
    __shared__ WaveRows s_rows[WG / 64];
    for (uint32_t i = threadIdx.x; i < (uint32_t)EXP2_LEN; i += WG)
        s_exp2[i] = tab.exp2[i];
    for (uint32_t i = threadIdx.x; i < 256u; i += WG)
        s_log2[i] = tab.log2[i];
    __syncthreads();
    const GfLds gf{s_exp2, s_log2};
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    WaveRows &w = s_rows[wave];
    for (uint32_t j = blockIdx.x * (WG / 64) + wave; j < nitems; j += gridDim.x * (WG / 64)) {
        const PlanItem it = items[j];
        const uint32_t E = it.E, k = it.k;
        uint8_t *pm = msgs + it.msg_off, *pe = eccs + it.ecc_off;
        const uint8_t st = E <= 64    ? repair_wave<1, 1>(pm, pe, k, E, gf, w)
                           : E <= 124 ? repair_wave<1, 2>(pm, pe, k, E, gf, w)
                           : E <= 128 ? repair_wave<2, 2>(pm, pe, k, E, gf, w)
                           : E <= 192 ? repair_wave<2, 3>(pm, pe, k, E, gf, w)
                                      : repair_wave<2, 4>(pm, pe, k, E, gf, w);
        if ((threadIdx.x & 63) == 0)
            status[it.idx] = st;
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace ffrs

using namespace ffrs;

// ---------------------------------------------------------------- allocation-free repair

// Workspace of ffrs_rs256_repair_device_noalloc: a file-scope device array, one copy per GPU,
// held from the library's load until the process ends. Host code never hands this name to a
// kernel (on the host it is only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns with that GPU current, looked up once per device when the
// first codec is created there.
__device__ __align__(256) uint8_t scratch[FFRS_RS256_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: the device address of this GPU's copy of `scratch`.
int resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(FFRS_ERR_DEVICE, "device %d beyond the %d this build addresses", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device])
        return FFRS_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(FFRS_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                    hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return FFRS_OK;
}

void *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

// ---------------------------------------------------------------- C ABI

struct ffrs_rs256 {
    ffrs_ctx *ctx = nullptr;
    unsigned ecc = 0, EW = 0, S = 0;
    Gf256Host gf;
    uint8_t gen[256] = {0};
    uint8_t roots[256] = {0};
    uint32_t *d_lut = nullptr;
    uint32_t *d_lutr = nullptr;  // the same table in the replicated two-slice layout (EW <= 8)
    uint8_t *d_exp2 = nullptr;
    uint16_t *d_log2 = nullptr;
    uint32_t *d_stab = nullptr;  // syndrome nibble tables (ecc <= 32)
    uint32_t *d_ctab = nullptr;  // Chien nibble tables (ecc <= 32)
    void *d_work = nullptr;      // syndromes + BM records of the last repair
    size_t work_size = 0;
    int repair_mode = 0;         // 0 automatic, 1 pipeline (ecc <= 32), 2 one kernel, lane per codeword
};

namespace {

unsigned ew_class(unsigned ecc)
{
    unsigned w = (ecc + 3) / 4;
    unsigned c = 1;
    while (c < w)
        c <<= 1;
    return c;
}

unsigned slices_for(unsigned EW) { return EW <= 8 ? 4 : (EW == 16 ? 2 : 1); }

// Remainder-order bytes r_0..r_{E-1} (r_0 = coefficient of x^(E-1)) packed little-endian.
void lfsr_zero_step(const ffrs_rs256 &rs, uint8_t *r)
{
    unsigned E = rs.ecc;
    uint8_t f = r[0];
    memmove(r, r + 1, E - 1);
    r[E - 1] = 0;
    if (f)
        for (unsigned j = 0; j < E; ++j)
            r[j] ^= rs.gf.mul(f, rs.gen[j + 1]);
}

std::vector<uint32_t> build_lut(const ffrs_rs256 &rs)
{
    unsigned E = rs.ecc, EW = rs.EW, S = rs.S;
    std::vector<uint32_t> lut((size_t)S * 256 * lut_row((int)EW), 0);
    for (unsigned b = 0; b < 256; ++b) {
        uint8_t r[256] = {0};
        // b * x^E mod g: one LFSR step with feedback b
        r[0] = (uint8_t)b;
        lfsr_zero_step(rs, r);
        for (unsigned s = 0; s < S; ++s) {
            uint32_t *row = &lut[((size_t)s * 256 + b) * lut_row((int)EW)];
            for (unsigned j = 0; j < E; ++j)
                row[j / 4] |= (uint32_t)r[j] << (8 * (j % 4));
            lfsr_zero_step(rs, r);
        }
    }
    return lut;
}

// The replicated two-slice layout of lut_dwords(EW, LUT_REP) dwords (ecc_len <= 32).
std::vector<uint32_t> build_lut_rep(const ffrs_rs256 &rs)
{
    const unsigned E = rs.ecc, EW = rs.EW;
    std::vector<uint32_t> t((size_t)lut_dwords((int)EW, LUT_REP), 0);
    for (unsigned b = 0; b < 256; ++b) {
        uint8_t r[256] = {0};
        r[0] = (uint8_t)b;
        lfsr_zero_step(rs, r);
        for (unsigned s = 0; s < 2; ++s) {
            if (EW == 8) {
                // halfword-paired: dword j = symbols pair j | pair (j + 8) << 16
                uint32_t row[8] = {0};
                for (unsigned j = 0; j < E; ++j) {
                    const unsigned pair = j / 2;
                    row[pair % 8] |= (uint32_t)r[j] << (8 * (j % 2) + 16 * (pair / 8));
                }
                for (unsigned h = 0; h < 2; ++h)
                    for (unsigned c = 0; c < (unsigned)LUT_COPIES; ++c)
                        memcpy(&t[((((size_t)s * 256 + b) * 2 + h) * LUT_COPIES + c) * 4], row + 4 * h, 16);
            } else {
                // byte order, 64 / EW copies side by side
                uint32_t row[4] = {0};
                for (unsigned j = 0; j < E; ++j)
                    row[j / 4] |= (uint32_t)r[j] << (8 * (j % 4));
                for (unsigned c = 0; c < 64 / EW; ++c)
                    memcpy(&t[(((size_t)s * 256 + b) * 64) + c * EW], row, 4 * EW);
            }
            lfsr_zero_step(rs, r);
        }
    }
    return t;
}

// Workgroup sizes of the LFSR kernels. More resident waves do not help them: with 512- and
// 768-lane groups (twice and three times the waves per CU) the ecc-32 kernels ran 5% and
// 33% slower on MI355X, the scattered per-lane row reads contending for the L1 cache.
constexpr int lfsr_wg(int) { return 256; }
// The replicated table fills 128 KB of LDS, one workgroup per CU, which therefore takes 1024
// lanes (measured at ecc 32 on MI355X: encode 124 us with 1024 lanes, 126 with 512, 190 with
// 256; syndromes 150, 159 and 255 us).
constexpr int lfsr_wg(int EW, int S) { return S == LUT_REP ? 1024 : lfsr_wg(EW); }
template <int S>
const uint32_t *lut_of(const ffrs_rs256 *rs) { return S == LUT_REP ? rs->d_lutr : rs->d_lut; }
constexpr int synd_wg(int) { return 256; }

// Message loads of the LFSR kernels: cooperative windows where k <= 254 over the replicated
// tables, except for the syndromes of 32-byte rows (whose registers then spill: there one lane
// per message is faster, 131 against 148 us at ecc 32 on MI355X), else one lane per message.
template <int EW, int S>
int lfsr_load_mode(uint32_t k, bool synd)
{
    return S == LUT_REP && k <= 254 && !(synd && EW == 8) ? LOAD_COOP : LOAD_LANE;
}

template <int EW, int S>
int launch_encode(ffrs_rs256 *rs, const uint8_t *d_msg, size_t count, uint32_t k, uint8_t *d_ecc)
{
    ffrs_ctx *ctx = rs->ctx;
    const bool coop = lfsr_load_mode<EW, S>(k, false) == LOAD_COOP;
    constexpr int WG = lfsr_wg(EW, S), WGC = 1024;
    constexpr bool rep = S == LUT_REP;
    const int wg = coop ? WGC : WG;
    size_t lds = (size_t)lfsr_dyn_dwords(EW, S) * 4;
    size_t tiles = (count + wg - 1) / wg;
    int wgs_per_cu = (int)std::min<size_t>(2048 / wg, std::max<size_t>(1, (160 * 1024) / (lut_dwords(EW, S) * 4)));
    int grid = persistent_grid(ctx, wgs_per_cu, tiles);
    auto *kern = coop ? rs256_encode_kernel<EW, S, rep ? WGC : WG, rep ? LOAD_COOP : LOAD_LANE>
                      : rs256_encode_kernel<EW, S, WG, LOAD_LANE>;
    if (int r = allow_lds(kern, lds))
        return r;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(wg), lds, ctx->stream, d_msg, (uint32_t)(count * k), (uint32_t)count, k,
                       d_ecc, rs->ecc, lut_of<S>(rs));
    FFRS_HIP(hipGetLastError());
    mark_kernel(ctx);
    return FFRS_OK;
}

template <int EW, int S, int WG>
int launch_repair(ffrs_rs256 *rs, uint8_t *d_msg, uint8_t *d_ecc, size_t count, uint32_t k, uint8_t *d_status)
{
    ffrs_ctx *ctx = rs->ctx;
    size_t lds = (size_t)S * 256 * lut_row(EW) * 4 + EXP2_LEN + 512 + (size_t)WG * repair_row_stride(rs->ecc);
    if (lds > 160 * 1024)
        return fail(FFRS_ERR_INVALID, "repair workspace of %zu bytes exceeds LDS", lds);
    int per_cu = (int)std::max<size_t>(1, (160 * 1024) / lds);
    int grid = persistent_grid(ctx, per_cu * 2, (count + WG - 1) / WG);
    Rs256Tables tab{rs->d_lut, rs->d_exp2, rs->d_log2};
    if (int r = allow_lds(rs256_repair_kernel<EW, S, WG>, lds))
        return r;
    hipLaunchKernelGGL((rs256_repair_kernel<EW, S, WG>), dim3(grid), dim3(WG), lds, ctx->stream, d_msg,
                       (uint32_t)(count * k), d_ecc, (uint32_t)count, k, rs->ecc, d_status, tab);
    FFRS_HIP(hipGetLastError());
    mark_kernel(ctx);
    return FFRS_OK;
}

// The wave-cooperative repair (ecc_len 33..254).
int launch_repair_wave(ffrs_rs256 *rs, uint8_t *d_msg, uint8_t *d_ecc, size_t count, uint32_t k, uint8_t *d_status)
{
    ffrs_ctx *ctx = rs->ctx;
    const uint32_t E = rs->ecc;
    constexpr int WG = 256;
    Rs256Tables tab{rs->d_lut, rs->d_exp2, rs->d_log2};
    auto go = [&](auto kern) -> int {
        const int grid = persistent_grid(ctx, 8, (count + WG / 64 - 1) / (WG / 64));
        hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), 0, ctx->stream, d_msg, d_ecc, (uint32_t)count, k, E, d_status,
                           tab);
        FFRS_HIP(hipGetLastError());
        mark_kernel(ctx);
        return FFRS_OK;
    };
    if (E <= 64)
        return go(rs256_repair_wave_kernel<1, 1, WG>);
    if (E <= 124)
        return go(rs256_repair_wave_kernel<1, 2, WG>);
    if (E <= 128)
        return go(rs256_repair_wave_kernel<2, 2, WG>);
    if (E <= 192)
        return go(rs256_repair_wave_kernel<2, 3, WG>);
    return go(rs256_repair_wave_kernel<2, 4, WG>);
}

template <int EW, int S>
int launch_synds(ffrs_rs256 *rs, const uint8_t *d_msg, const uint8_t *d_ecc, size_t count, uint32_t k,
                 uint8_t *d_out)
{
    ffrs_ctx *ctx = rs->ctx;
    size_t lds = (size_t)S * 256 * lut_row(EW) * 4 + EXP2_LEN + 512;
    int grid = persistent_grid(ctx, 2, (count + 255) / 256);
    Rs256Tables tab{rs->d_lut, rs->d_exp2, rs->d_log2};
    if (int r = allow_lds(rs256_synds_kernel<EW, S>, lds))
        return r;
    hipLaunchKernelGGL((rs256_synds_kernel<EW, S>), dim3(grid), dim3(256), lds, ctx->stream, d_msg,
                       (uint32_t)(count * k), d_ecc, (uint32_t)count, k, rs->ecc, d_out, tab);
    FFRS_HIP(hipGetLastError());
    return FFRS_OK;
}

// ST[j][h][v]: E bytes (v << 4h) * alpha^(i (E-1-j)) for i < E, packed in EW dwords.
std::vector<uint32_t> build_synd_tables(const ffrs_rs256 &rs)
{
    unsigned E = rs.ecc, EW = rs.EW;
    std::vector<uint32_t> t((size_t)E * 32 * EW, 0);
    for (unsigned j = 0; j < E; ++j)
        for (unsigned h = 0; h < 2; ++h)
            for (unsigned v = 0; v < 16; ++v) {
                uint32_t *row = &t[(((size_t)j * 2 + h) * 16 + v) * EW];
                uint8_t c = (uint8_t)(v << (4 * h));
                for (unsigned i = 0; i < E; ++i) {
                    uint8_t a = rs.gf.exp_t[(i * (E - 1 - j)) % 255];
                    row[i / 4] |= (uint32_t)rs.gf.mul(c, a) << (8 * (i % 4));
                }
            }
    return t;
}

// CT[i-1][h][v]: 256 position bytes (v << 4h) * alpha^(-i p) for p < 255 (byte 255 unused).
std::vector<uint32_t> build_chien_tables(const ffrs_rs256 &rs)
{
    std::vector<uint32_t> t(CHIEN_TABLE_BYTES / 4, 0);
    for (unsigned i = 1; i <= (unsigned)CHIEN_LMAX; ++i)
        for (unsigned h = 0; h < 2; ++h)
            for (unsigned v = 0; v < 16; ++v) {
                uint32_t *row = &t[(((size_t)(i - 1) * 2 + h) * 16 + v) * 64];
                uint8_t c = (uint8_t)(v << (4 * h));
                for (unsigned p = 0; p < 255; ++p) {
                    uint8_t a = rs.gf.exp_t[(255 - (i * p) % 255) % 255];
                    row[p / 4] |= (uint32_t)rs.gf.mul(c, a) << (8 * (p % 4));
                }
            }
    return t;
}

int reserve_work(ffrs_rs256 *rs, size_t bytes)
{
    if (rs->work_size >= bytes)
        return FFRS_OK;
    if (rs->d_work)
        FFRS_HIP(hipFree(rs->d_work));
    rs->d_work = nullptr;
    rs->work_size = 0;
    size_t sz = bytes + bytes / 4;
    hipError_t e = hipMalloc(&rs->d_work, sz);
    if (e != hipSuccess)
        return fail(FFRS_ERR_NOMEM, "repair workspace of %zu bytes: %s", sz, hipGetErrorString(e));
    rs->work_size = sz;
    return FFRS_OK;
}

// Whether the Chien kernel runs Berlekamp-Massey itself (ecc 2, 4, 8, 16, 32), so that the
// pipeline keeps no BM records.
bool bm_in_chien(uint32_t E) { return E == 2 || E == 4 || E == 8 || E == 16 || E == 32; }

// Workspace of the repair pipeline for `count` codewords of ecc E (word class EW): the
// syndromes (EW dwords per codeword), then, for ecc outside bm_in_chien, the 48-byte BM
// records from the next 256-byte boundary.
size_t pipeline_work_bytes(uint32_t E, uint32_t EW, size_t count)
{
    const size_t synd_bytes = count * EW * 4;
    if (bm_in_chien(E))
        return synd_bytes;
    return ((synd_bytes + 255) & ~(size_t)255) + count * rec_stride(E);
}

// The pipeline over a workspace `work` of at least pipeline_work_bytes(E, EW, count) bytes.
// Nothing in it needs to start at zero: the syndrome kernel writes the syndromes of every
// codeword it marks pending, and only pending codewords' entries are read after it.
template <int EW, int S>
int launch_repair_pipeline(ffrs_rs256 *rs, uint8_t *d_msg, uint8_t *d_ecc, size_t count, uint32_t k,
                           uint8_t *d_status, void *work)
{
    ffrs_ctx *ctx = rs->ctx;
    const uint32_t E = rs->ecc;
    size_t synd_bytes = count * EW * 4;
    uint32_t *d_synd = (uint32_t *)work;
    uint8_t *d_rec = (uint8_t *)work + ((synd_bytes + 255) & ~(size_t)255);
    Rs256Tables tab{rs->d_lut, rs->d_exp2, rs->d_log2};

    const bool coop = lfsr_load_mode<EW, S>(k, true) == LOAD_COOP;
    constexpr int WG1 = S == LUT_REP ? lfsr_wg(EW, S) : synd_wg(EW), WG1C = 1024;
    constexpr bool rep = S == LUT_REP;
    const int wg1 = coop ? WG1C : WG1;
    size_t lds1 = (size_t)lfsr_dyn_dwords(EW, S) * 4 + (size_t)E * 32 * EW * 4;
    auto *k1 = coop ? rs256_synd_kernel<EW, S, rep ? WG1C : WG1, rep ? LOAD_COOP : LOAD_LANE>
                    : rs256_synd_kernel<EW, S, WG1, LOAD_LANE>;
    if (int r = allow_lds(k1, lds1))
        return r;
    int g1 = persistent_grid(ctx, (int)std::min<size_t>(2048 / wg1, std::max<size_t>(1, (160 * 1024) / (lds1 + (lut_dwords(EW, S) - lfsr_dyn_dwords(EW, S)) * 4))),
                             (count + wg1 - 1) / wg1);
    hipLaunchKernelGGL(k1, dim3(g1), dim3(wg1), lds1, ctx->stream, d_msg, (uint32_t)(count * k), d_ecc,
                       (uint32_t)count, k, E, d_status, d_synd, lut_of<S>(rs), rs->d_stab);
    FFRS_HIP(hipGetLastError());
    mark_kernel(ctx);

    // ecc 2, 4, 8, 16, 32: BM inside the Chien kernel, no records (512 threads for ecc 32,
    // whose BM would spill registers at 768); eight lanes per codeword in its passes, faster
    // than sixteen at every ecc measured (1 M codewords of ecc 32: 402 vs 416 us; 2 M of ecc
    // 16: 317 vs 348 us; 4 M of ecc 8: 384 vs 412 us)
    auto fused = [&](auto kern, int wg) -> int {
        int g = persistent_grid(ctx, 1, (count + wg - 1) / wg);
        hipLaunchKernelGGL(kern, dim3(g), dim3(wg), 0, ctx->stream, d_msg, d_ecc, (uint32_t)count, k, E, d_status,
                           (const uint32_t *)nullptr, (const uint32_t *)d_synd, rs->d_ctab, tab);
        FFRS_HIP(hipGetLastError());
        mark_kernel(ctx);
        return FFRS_OK;
    };
    switch (E) {
    case 2: return fused(rs256_chien16_kernel<1024, 2, 8>, 1024);
    case 4: return fused(rs256_chien16_kernel<1024, 4, 8>, 1024);
    case 8: return fused(rs256_chien16_kernel<1024, 8, 8>, 1024);
    case 16: return fused(rs256_chien16_kernel<1024, 16, 8>, 1024);
    case 32: return fused(rs256_chien16_kernel<512, 32, 8>, 512);
    default: break;
    }
    // other ecc: Berlekamp-Massey from the log tables, one lane per codeword, records in the
    // workspace
    size_t lds2 = EXP2_LEN + 512 + (size_t)256 * bm_row_stride(E);
    if (int r = allow_lds(rs256_bm_kernel<256>, lds2))
        return r;
    int g2 = persistent_grid(ctx, (int)std::max<size_t>(1, (160 * 1024) / lds2), (count + 255) / 256);
    hipLaunchKernelGGL((rs256_bm_kernel<256>), dim3(g2), dim3(256), lds2, ctx->stream, (uint32_t)count, E,
                       (uint32_t)EW, d_status, (const uint32_t *)d_synd, d_rec, tab);
    FFRS_HIP(hipGetLastError());
    mark_kernel(ctx);
    // Chien + Forney: the 128 KB table fills the CU's LDS, one workgroup of sixteen waves per CU
    constexpr int WGC = 1024;
    int gc = persistent_grid(ctx, 1, (count + WGC - 1) / WGC);
    hipLaunchKernelGGL((rs256_chien16_kernel<WGC, 0, 8>), dim3(gc), dim3(WGC), 0, ctx->stream, d_msg, d_ecc,
                       (uint32_t)count, k, E, d_status, (const uint32_t *)d_rec, (const uint32_t *)d_synd, rs->d_ctab,
                       tab);
    FFRS_HIP(hipGetLastError());
    mark_kernel(ctx);
    return FFRS_OK;
}

// The replicated tables (128 KB per workgroup to load) for launches of at least 8,192
// codewords; smaller ones take the slicing-by-S tables.
bool use_rep(const ffrs_rs256 *rs, size_t count) { return rs->d_lutr && count >= 8192; }

// Splits launches so that every launch addresses < 2 GiB of messages with 32-bit offsets.
size_t max_batch(size_t k) { return std::max<size_t>(1, ((size_t)1 << 31) / std::max<size_t>(k, 1)) & ~(size_t)255; }

int encode_device(ffrs_rs256 *rs, size_t message_len, const uint8_t *d_msg, size_t count, uint8_t *d_ecc)
{
    marks_begin(rs->ctx);
    if (count == 0)
        return FFRS_OK;
    uint32_t k = (uint32_t)message_len;
    size_t step = max_batch(message_len);
    for (size_t b = 0; b < count; b += step) {
        size_t c = std::min(step, count - b);
        const uint8_t *m = d_msg + b * message_len;
        uint8_t *e = d_ecc + b * rs->ecc;
        int r;
        const bool rep = use_rep(rs, c);
        switch (rs->EW) {
        case 1: r = rep ? launch_encode<1, LUT_REP>(rs, m, c, k, e) : launch_encode<1, 4>(rs, m, c, k, e); break;
        case 2: r = rep ? launch_encode<2, LUT_REP>(rs, m, c, k, e) : launch_encode<2, 4>(rs, m, c, k, e); break;
        case 4: r = rep ? launch_encode<4, LUT_REP>(rs, m, c, k, e) : launch_encode<4, 4>(rs, m, c, k, e); break;
        case 8: r = rep ? launch_encode<8, LUT_REP>(rs, m, c, k, e) : launch_encode<8, 4>(rs, m, c, k, e); break;
        case 16: r = launch_encode<16, 2>(rs, m, c, k, e); break;
        case 32: r = launch_encode<32, 1>(rs, m, c, k, e); break;
        default: r = launch_encode<64, 1>(rs, m, c, k, e); break;
        }
        if (r)
            return r;
    }
    return FFRS_OK;
}

// Whether repairs of this codec run the syndrome / BM / Chien pipeline (which needs a
// workspace) or the one-kernel lane mode (which does not).
bool uses_pipeline(const ffrs_rs256 *rs) { return rs->EW <= 8 && rs->repair_mode != 2; }

// `work`: the pipeline's workspace, at least pipeline_work_bytes(ecc, EW, min(count,
// max_batch)) bytes, or nullptr to use (and grow) the codec's own.
int repair_device(ffrs_rs256 *rs, size_t message_len, uint8_t *d_msg, uint8_t *d_ecc, size_t count,
                  uint8_t *d_status, void *work = nullptr)
{
    marks_begin(rs->ctx);
    if (count == 0)
        return FFRS_OK;
    uint32_t k = (uint32_t)message_len;
    size_t step = max_batch(message_len);
    // The pipeline is faster at every batch size measured (142 codewords of ecc 32:
    // 274 us against 373 us lane by lane); the lane kernel stays selectable.
    const bool pipe = uses_pipeline(rs);
    if (pipe && !work) {
        if (int r = reserve_work(rs, pipeline_work_bytes(rs->ecc, rs->EW, std::min(step, count))))
            return r;
        work = rs->d_work;
    }
    for (size_t b = 0; b < count; b += step) {
        size_t c = std::min(step, count - b);
        uint8_t *m = d_msg + b * message_len;
        uint8_t *e = d_ecc + b * rs->ecc;
        uint8_t *st = d_status + b;
        int r;
        switch (rs->EW) {
        case 1:
            r = !pipe ? launch_repair<1, 4, 256>(rs, m, e, c, k, st)
                : use_rep(rs, c) ? launch_repair_pipeline<1, LUT_REP>(rs, m, e, c, k, st, work)
                : launch_repair_pipeline<1, 4>(rs, m, e, c, k, st, work);
            break;
        case 2:
            r = !pipe ? launch_repair<2, 4, 256>(rs, m, e, c, k, st)
                : use_rep(rs, c) ? launch_repair_pipeline<2, LUT_REP>(rs, m, e, c, k, st, work)
                : launch_repair_pipeline<2, 4>(rs, m, e, c, k, st, work);
            break;
        case 4:
            r = !pipe ? launch_repair<4, 4, 256>(rs, m, e, c, k, st)
                : use_rep(rs, c) ? launch_repair_pipeline<4, LUT_REP>(rs, m, e, c, k, st, work)
                : launch_repair_pipeline<4, 4>(rs, m, e, c, k, st, work);
            break;
        case 8:
            r = !pipe ? launch_repair<8, 4, 128>(rs, m, e, c, k, st)
                : use_rep(rs, c) ? launch_repair_pipeline<8, LUT_REP>(rs, m, e, c, k, st, work)
                : launch_repair_pipeline<8, 4>(rs, m, e, c, k, st, work);
            break;
        case 16: r = rs->repair_mode == 2 ? launch_repair<16, 2, 128>(rs, m, e, c, k, st) : launch_repair_wave(rs, m, e, c, k, st); break;
        case 32: r = rs->repair_mode == 2 ? launch_repair<32, 1, 64>(rs, m, e, c, k, st) : launch_repair_wave(rs, m, e, c, k, st); break;
        default: r = rs->repair_mode == 2 ? launch_repair<64, 1, 64>(rs, m, e, c, k, st) : launch_repair_wave(rs, m, e, c, k, st); break;
        }
        if (r)
            return r;
    }
    return FFRS_OK;
}

int check_codec(const ffrs_rs256 *rs, size_t message_len)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (message_len == 0 || message_len + rs->ecc > 255)
        return fail(FFRS_ERR_INVALID, "message_len %zu invalid for ecc_len %u (block_len must be <= 255)",
                    message_len, rs->ecc);
    return FFRS_OK;
}

}  // namespace

extern "C" {

int ffrs_rs256_create(ffrs_ctx *ctx, unsigned ecc_len, unsigned primitive, unsigned polynomial, ffrs_rs256 **out)
{
    if (!ctx || !out)
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (ecc_len < 1 || ecc_len > 254)
        return fail(FFRS_ERR_INVALID, "ecc_len must be in [1, 254], got %u", ecc_len);
    auto *rs = new ffrs_rs256();
    rs->ctx = ctx;
    rs->ecc = ecc_len;
    rs->EW = ew_class(ecc_len);
    rs->S = slices_for(rs->EW);
    if (!gf256_build(primitive & 0xff, polynomial, rs->gf)) {
        delete rs;
        return fail(FFRS_ERR_INVALID,
                    "primitive %u does not generate GF(2^8) modulo polynomial 0x%x (unsupported field)", primitive,
                    polynomial);
    }
    // g(x) = prod_{i<ecc} (x - alpha^i), highest degree first (reed_solomon.hpp:65-86)
    uint8_t g[257] = {0};
    unsigned len = 1;
    g[0] = 1;
    for (unsigned i = 0; i < ecc_len; ++i) {
        uint8_t root = rs->gf.exp_t[i];
        rs->roots[i] = root;
        uint8_t ng[257] = {0};
        for (unsigned a = 0; a < len; ++a) {
            ng[a] ^= g[a];
            ng[a + 1] ^= rs->gf.mul(g[a], root);
        }
        len++;
        memcpy(g, ng, len);
    }
    memcpy(rs->gen, g, ecc_len + 1);

    std::vector<uint32_t> lut = build_lut(*rs);
    uint8_t exp2[EXP2_LEN];
    uint16_t log2[256];
    for (int x = 0; x < EXP2_LEN; ++x)
        exp2[x] = x <= 509 ? rs->gf.exp_t[x % 255] : 0;
    log2[0] = ZLOG;
    for (int a = 1; a < 256; ++a)
        log2[a] = rs->gf.log_t[a] % 255;

    int r = ctx_set_device(ctx);
    if (!r)
        r = resolve_scratch(ctx->device);
    if (r) {
        delete rs;
        return r;
    }
    hipError_t e = hipMalloc(&rs->d_lut, lut.size() * 4);
    if (e == hipSuccess)
        e = hipMalloc(&rs->d_exp2, EXP2_LEN);
    if (e == hipSuccess)
        e = hipMalloc(&rs->d_log2, 512);
    if (e == hipSuccess)
        e = hipMemcpy(rs->d_lut, lut.data(), lut.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && rs->EW <= 8) {
        std::vector<uint32_t> lr = build_lut_rep(*rs);
        e = hipMalloc(&rs->d_lutr, lr.size() * 4);
        if (e == hipSuccess)
            e = hipMemcpy(rs->d_lutr, lr.data(), lr.size() * 4, hipMemcpyHostToDevice);
    }

    if (e == hipSuccess)
        e = hipMemcpy(rs->d_exp2, exp2, EXP2_LEN, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(rs->d_log2, log2, 512, hipMemcpyHostToDevice);
    if (e == hipSuccess && rs->EW <= 8) {
        std::vector<uint32_t> st = build_synd_tables(*rs);
        std::vector<uint32_t> ct = build_chien_tables(*rs);
        e = hipMalloc(&rs->d_stab, st.size() * 4);
        if (e == hipSuccess)
            e = hipMalloc(&rs->d_ctab, ct.size() * 4);
        if (e == hipSuccess)
            e = hipMemcpy(rs->d_stab, st.data(), st.size() * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(rs->d_ctab, ct.data(), ct.size() * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        ffrs_rs256_destroy(rs);
        return fail(FFRS_ERR_DEVICE, "table upload failed: %s", hipGetErrorString(e));
    }
    *out = rs;
    return FFRS_OK;
}

void ffrs_rs256_destroy(ffrs_rs256 *rs)
{
    if (!rs)
        return;
    (void)hipSetDevice(rs->ctx->device);
    if (rs->d_lut)
        (void)hipFree(rs->d_lut);
    if (rs->d_lutr)
        (void)hipFree(rs->d_lutr);

    if (rs->d_exp2)
        (void)hipFree(rs->d_exp2);
    if (rs->d_log2)
        (void)hipFree(rs->d_log2);
    if (rs->d_stab)
        (void)hipFree(rs->d_stab);
    if (rs->d_ctab)
        (void)hipFree(rs->d_ctab);
    if (rs->d_work)
        (void)hipFree(rs->d_work);
    delete rs;
}

int ffrs_rs256_set_repair_mode(ffrs_rs256 *rs, int mode)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (mode < 0 || mode > 2)
        return fail(FFRS_ERR_INVALID, "repair mode %d (0 automatic, 1 pipeline, 2 lane per codeword)", mode);
    rs->repair_mode = mode;
    return FFRS_OK;
}

int ffrs_rs256_generator(const ffrs_rs256 *rs, uint8_t *generator, uint8_t *roots)
{
    if (!rs)
        return fail(FFRS_ERR_INVALID, "codec is NULL");
    if (generator)
        memcpy(generator, rs->gen, rs->ecc + 1);
    if (roots)
        memcpy(roots, rs->roots, rs->ecc);
    return FFRS_OK;
}

int ffrs_rs256_encode_device(ffrs_rs256 *rs, size_t message_len, const uint8_t *d_msg, size_t count, uint8_t *d_ecc)
{
    if (int r = check_codec(rs, message_len))
        return r;
    if (count && (!d_msg || !d_ecc))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = ctx_set_device(rs->ctx))
        return r;
    return encode_device(rs, message_len, d_msg, count, d_ecc);
}

int ffrs_rs256_repair_device(ffrs_rs256 *rs, size_t message_len, uint8_t *d_msg, uint8_t *d_ecc, size_t count,
                             uint8_t *d_status)
{
    if (int r = check_codec(rs, message_len))
        return r;
    if (count && (!d_msg || !d_ecc || !d_status))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = ctx_set_device(rs->ctx))
        return r;
    return repair_device(rs, message_len, d_msg, d_ecc, count, d_status);
}

int ffrs_rs256_repair_device_noalloc(ffrs_rs256 *rs, size_t message_len, uint8_t *d_msg, uint8_t *d_ecc,
                                     size_t count, uint8_t *d_status)
{
    if (int r = check_codec(rs, message_len))
        return r;
    if (count && (!d_msg || !d_ecc || !d_status))
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    if (int r = ctx_set_device(rs->ctx))
        return r;
    void *work = nullptr;
    if (count && uses_pipeline(rs)) {
        const size_t per_launch = std::min(max_batch(message_len), count);
        const size_t need = pipeline_work_bytes(rs->ecc, rs->EW, per_launch);
        if (need > (size_t)FFRS_RS256_NOALLOC_SCRATCH_BYTES) {
            size_t most = 0;  // the largest count whose workspace fits
            for (size_t lo = 0, hi = per_launch; lo <= hi;) {
                const size_t mid = lo + (hi - lo) / 2;
                if (pipeline_work_bytes(rs->ecc, rs->EW, mid) <= (size_t)FFRS_RS256_NOALLOC_SCRATCH_BYTES) {
                    most = mid;
                    lo = mid + 1;
                } else {
                    hi = mid - 1;
                }
            }
            return fail(FFRS_ERR_SIZE,
                        "allocation-free repair: %zu codewords of ecc_len %u need %zu bytes of workspace, the "
                        "static workspace holds %u (at most %zu codewords per call)",
                        count, rs->ecc, need, (unsigned)FFRS_RS256_NOALLOC_SCRATCH_BYTES, most);
        }
        work = scratch_ptr(rs->ctx->device);
        if (!work)
            return fail(FFRS_ERR_DEVICE, "no workspace address for device %d", rs->ctx->device);
    }
    return repair_device(rs, message_len, d_msg, d_ecc, count, d_status, work);
}

int ffrs_rs256_encode(ffrs_rs256 *rs, size_t message_len, const uint8_t *buf, size_t size, uint8_t *out,
                      size_t out_size)
{
    if (int r = check_codec(rs, message_len))
        return r;
    size_t full = size / message_len, tail = size - full * message_len;
    size_t need = (full + (tail ? 1 : 0)) * rs->ecc;
    if (out_size != need)
        return fail(FFRS_ERR_SIZE, "output holds %zu bytes, %zu needed", out_size, need);
    if (size == 0)
        return FFRS_OK;
    if (!buf || !out)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    void *d_in, *d_out;
    if (int r = stage_reserve(ctx, 0, size, &d_in))
        return r;
    if (int r = stage_reserve(ctx, 1, need, &d_out))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_in, buf, size, hipMemcpyHostToDevice, ctx->stream));
    if (int r = encode_device(rs, message_len, (const uint8_t *)d_in, full, (uint8_t *)d_out))
        return r;
    if (tail)
        if (int r = encode_device(rs, tail, (const uint8_t *)d_in + full * message_len, 1,
                                  (uint8_t *)d_out + full * rs->ecc))
            return r;
    FFRS_HIP(hipMemcpyAsync(out, d_out, need, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rs256_repair(ffrs_rs256 *rs, size_t message_len, uint8_t *msg, uint8_t *ecc, size_t count,
                      uint8_t *status)
{
    if (int r = check_codec(rs, message_len))
        return r;
    if (count == 0)
        return FFRS_OK;
    if (!msg || !ecc || !status)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = count * message_len, eb = count * rs->ecc;
    void *d_m, *d_e, *d_s;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    if (int r = stage_reserve(ctx, 2, count, &d_s))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_e, ecc, eb, hipMemcpyHostToDevice, ctx->stream));
    if (int r = repair_device(rs, message_len, (uint8_t *)d_m, (uint8_t *)d_e, count, (uint8_t *)d_s))
        return r;
    FFRS_HIP(hipMemcpyAsync(msg, d_m, mb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(ecc, d_e, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(status, d_s, count, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

int ffrs_rs256_synds(ffrs_rs256 *rs, size_t message_len, const uint8_t *msg, const uint8_t *ecc, size_t count,
                     uint8_t *synds)
{
    if (int r = check_codec(rs, message_len))
        return r;
    if (count == 0)
        return FFRS_OK;
    if (count > max_batch(message_len))
        return fail(FFRS_ERR_INVALID, "too many codewords in one syndrome call");
    if (!msg || !ecc || !synds)
        return fail(FFRS_ERR_INVALID, "NULL buffer");
    ffrs_ctx *ctx = rs->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    size_t mb = count * message_len, eb = count * rs->ecc;
    void *d_m, *d_e, *d_o;
    if (int r = stage_reserve(ctx, 0, mb, &d_m))
        return r;
    if (int r = stage_reserve(ctx, 1, eb, &d_e))
        return r;
    if (int r = stage_reserve(ctx, 2, eb, &d_o))
        return r;
    FFRS_HIP(hipMemcpyAsync(d_m, msg, mb, hipMemcpyHostToDevice, ctx->stream));
    FFRS_HIP(hipMemcpyAsync(d_e, ecc, eb, hipMemcpyHostToDevice, ctx->stream));
    uint32_t k = (uint32_t)message_len;
    int r;
    const uint8_t *m = (const uint8_t *)d_m, *e = (const uint8_t *)d_e;
    uint8_t *o = (uint8_t *)d_o;
    switch (rs->EW) {
    case 1: r = launch_synds<1, 4>(rs, m, e, count, k, o); break;
    case 2: r = launch_synds<2, 4>(rs, m, e, count, k, o); break;
    case 4: r = launch_synds<4, 4>(rs, m, e, count, k, o); break;
    case 8: r = launch_synds<8, 4>(rs, m, e, count, k, o); break;
    case 16: r = launch_synds<16, 2>(rs, m, e, count, k, o); break;
    case 32: r = launch_synds<32, 1>(rs, m, e, count, k, o); break;
    default: r = launch_synds<64, 1>(rs, m, e, count, k, o); break;
    }
    if (r)
        return r;
    FFRS_HIP(hipMemcpyAsync(synds, d_o, eb, hipMemcpyDeviceToHost, ctx->stream));
    FFRS_HIP(hipStreamSynchronize(ctx->stream));
    return FFRS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- mixed-size plans

struct ffrs_rs256_plan {
    ffrs_ctx *ctx = nullptr;
    std::vector<ffrs_rs256 *> codecs;  // one per distinct ecc_len (class)
    PlanItem *d_items = nullptr;
    PlanChunk *d_chunks = nullptr;
    PlanClass *d_classes = nullptr;
    uint32_t *d_queue = nullptr;
    uint32_t count = 0, nchunks = 0, max_ew = 1, max_e = 1, wg = 256, lut_words = 0;
    // repair: items [0, n_wave) one per wave, chunks from chunk0 one per lane
    uint32_t n_wave = 0, chunk0 = 0, lane_max_ew = 1, lane_max_e = 1, lane_lut_words = 0;
    uint64_t msg_end = 0, ecc_end = 0;
    // the caller's list (in its order) and the repair kernel of each class: 0 automatic (one
    // codeword per wave for ecc_len > 32, else per lane), 1 per lane, 2 per wave
    std::vector<PlanItem> list;
    int cls_of[256];
    int mode_of[256] = {0};
    bool use_graph = false;
    struct Graph {
        int mode;
        const void *m, *e, *st;
        size_t mb;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
};

namespace {

void plan_free(ffrs_rs256_plan *p)
{
    (void)hipSetDevice(p->ctx->device);
    for (auto &g : p->graphs)
        (void)hipGraphExecDestroy(g.exec);
    for (void *d : {(void *)p->d_items, (void *)p->d_chunks, (void *)p->d_classes, (void *)p->d_queue})
        if (d)
            (void)hipFree(d);
    for (auto *c : p->codecs)
        ffrs_rs256_destroy(c);
    delete p;
}

template <int MAXEW, int WG, bool REPAIR>
int plan_launch(ffrs_rs256_plan *p, uint8_t *msgs, size_t msg_bytes, uint8_t *eccs, uint8_t *status)
{
    ffrs_ctx *ctx = p->ctx;
    const ffrs_rs256 *c0 = p->codecs[0];
    Rs256Tables tab{nullptr, c0->d_exp2, c0->d_log2};
    // repair runs the lane kernel over the chunks of ecc_len <= 32 only
    const uint32_t chunk0 = REPAIR ? p->chunk0 : 0;
    const uint32_t max_e = REPAIR ? p->lane_max_e : p->max_e;
    const uint32_t lut_words = REPAIR ? p->lane_lut_words : p->lut_words;
    if (chunk0 >= p->nchunks)
        return FFRS_OK;
    uint32_t row_stride = REPAIR ? repair_row_stride(max_e) : 0;
    size_t lds = (size_t)lut_words * 4 + (REPAIR ? EXP2_LEN + 512 + (size_t)WG * row_stride : 0);
    if (lds > 160 * 1024)
        return fail(FFRS_ERR_INVALID, "plan workspace of %zu bytes exceeds LDS", lds);
    auto kern = rs256_plan_kernel<MAXEW, WG, REPAIR>;
    if (int r = allow_lds(kern, lds))
        return r;
    int per_cu = (int)std::max<size_t>(1, (160 * 1024) / std::max<size_t>(lds, 1));
    int grid = persistent_grid(ctx, std::min(per_cu, 2048 / WG), p->nchunks - chunk0);
    hipLaunchKernelGGL(plan_reset_kernel, dim3(1), dim3(1), 0, ctx->stream, p->d_queue);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), lds, ctx->stream, msgs, (uint32_t)msg_bytes, eccs, status,
                       p->d_items, p->d_chunks, p->nchunks, p->d_classes, p->d_queue, chunk0, lut_words, row_stride,
                       tab);
    FFRS_HIP(hipGetLastError());
    return FFRS_OK;
}

int plan_enqueue(ffrs_rs256_plan *p, int mode, uint8_t *msgs, size_t msg_bytes, uint8_t *eccs, uint8_t *status)
{
    if (mode == 0) {
        bool small = p->max_ew <= 16;
        return small ? plan_launch<16, 256, false>(p, msgs, msg_bytes, eccs, status)
                     : plan_launch<64, 64, false>(p, msgs, msg_bytes, eccs, status);
    }
    if (p->n_wave) {
        constexpr int WG = 256;
        const ffrs_rs256 *c0 = p->codecs[0];
        Rs256Tables tab{nullptr, c0->d_exp2, c0->d_log2};
        int grid = persistent_grid(p->ctx, 8, (p->n_wave + WG / 64 - 1) / (WG / 64));
        hipLaunchKernelGGL((rs256_plan_wave_kernel<WG>), dim3(grid), dim3(WG), 0, p->ctx->stream, msgs, eccs, status,
                           p->d_items, p->n_wave, tab);
        FFRS_HIP(hipGetLastError());
    }
    bool small = p->lane_max_ew <= 16;
    return small ? plan_launch<16, 256, true>(p, msgs, msg_bytes, eccs, status)
                 : plan_launch<64, 64, true>(p, msgs, msg_bytes, eccs, status);
}

// Whether the items of class ecc_len E are repaired one per wave.
bool plan_wave(const ffrs_rs256_plan *p, unsigned E)
{
    const int m = p->mode_of[E];
    return m == 2 || (m == 0 && E > 32);
}

// Sorts the work list (per-wave classes first, then per-lane ones; costliest first within each:
// largest ecc, then longest message, so that the longest chunks do not start last), cuts it into
// chunks of one class and uploads it. Drops the captured graphs, which hold the old layout.
int plan_layout(ffrs_rs256_plan *p)
{
    const uint32_t count = p->count;
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const PlanItem &x = p->list[a], &y = p->list[b];
        const bool wx = plan_wave(p, x.E), wy = plan_wave(p, y.E);
        if (wx != wy)
            return wx;
        if (x.E != y.E)
            return x.E > y.E;
        return x.k > y.k;
    });
    std::vector<PlanItem> items(count);
    std::vector<PlanChunk> chunks;
    p->lane_max_ew = p->lane_max_e = 1;
    p->lane_lut_words = 0;
    p->n_wave = 0;
    for (uint32_t j = 0; j < count; ++j) {
        items[j] = p->list[order[j]];
        if (chunks.empty() || chunks.back().cls != items[j].cls || chunks.back().count == p->wg)
            chunks.push_back({j, 0, (uint16_t)items[j].cls});
        chunks.back().count++;
        if (plan_wave(p, items[j].E)) {
            p->n_wave = j + 1;
        } else {
            const ffrs_rs256 *c = p->codecs[items[j].cls];
            p->lane_max_ew = std::max(p->lane_max_ew, c->EW);
            p->lane_max_e = std::max(p->lane_max_e, c->ecc);
            p->lane_lut_words = std::max(p->lane_lut_words, c->S * 256u * (uint32_t)lut_row((int)c->EW));
        }
    }
    p->nchunks = (uint32_t)chunks.size();
    p->chunk0 = 0;
    while (p->chunk0 < p->nchunks && chunks[p->chunk0].first < p->n_wave)
        p->chunk0++;
    for (auto &g : p->graphs)
        (void)hipGraphExecDestroy(g.exec);
    p->graphs.clear();
    if (count) {
        FFRS_HIP(hipMemcpy(p->d_items, items.data(), items.size() * sizeof(items[0]), hipMemcpyHostToDevice));
        FFRS_HIP(hipMemcpy(p->d_chunks, chunks.data(), chunks.size() * sizeof(chunks[0]), hipMemcpyHostToDevice));
    }
    return FFRS_OK;
}

int plan_run(ffrs_rs256_plan *p, int mode, uint8_t *msgs, size_t msg_bytes, uint8_t *eccs, uint8_t *status)
{
    ffrs_ctx *ctx = p->ctx;
    if (int r = ctx_set_device(ctx))
        return r;
    if (p->count == 0)
        return FFRS_OK;
    if (msg_bytes < p->msg_end || msg_bytes >= ((size_t)1 << 31))
        return fail(FFRS_ERR_SIZE, "message buffer of %zu bytes: the plan needs %llu (and < 2 GiB)", msg_bytes,
                    (unsigned long long)p->msg_end);
    marks_begin(ctx);
    if (!p->use_graph) {
        if (int r = plan_enqueue(p, mode, msgs, msg_bytes, eccs, status))
            return r;
        mark_kernel(ctx);
        return FFRS_OK;
    }
    for (auto &g : p->graphs)
        if (g.mode == mode && g.m == msgs && g.e == eccs && g.st == status && g.mb == msg_bytes) {
            FFRS_HIP(hipGraphLaunch(g.exec, ctx->stream));
            mark_kernel(ctx);
            return FFRS_OK;
        }
    // first run with these buffers: capture the reset and the persistent launch
    hipGraph_t graph = nullptr;
    FFRS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int r = plan_enqueue(p, mode, msgs, msg_bytes, eccs, status);
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
    p->graphs.push_back({mode, msgs, eccs, status, msg_bytes, exec});
    FFRS_HIP(hipGraphLaunch(exec, ctx->stream));
    mark_kernel(ctx);
    return FFRS_OK;
}

}  // namespace

extern "C" {

int ffrs_rs256_plan_create(ffrs_ctx *ctx, size_t count, const uint8_t *ecc_len, const uint8_t *message_len,
                           const uint64_t *msg_off, const uint64_t *ecc_off, ffrs_rs256_plan **out)
{
    if (!ctx || !out || (count && (!ecc_len || !message_len || !msg_off || !ecc_off)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (count > 0xffffffffu)
        return fail(FFRS_ERR_INVALID, "too many codewords in one plan");
    for (size_t i = 0; i < count; ++i)
        if (ecc_len[i] < 1 || message_len[i] < 1 || (unsigned)message_len[i] + ecc_len[i] > 255)
            return fail(FFRS_ERR_INVALID, "codeword %zu: message_len %u, ecc_len %u (block_len must be <= 255)", i,
                        message_len[i], ecc_len[i]);
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int r = ctx_set_device(ctx))
        return r;
    auto *p = new ffrs_rs256_plan();
    p->ctx = ctx;
    p->count = (uint32_t)count;
    int cls_of[256];
    std::fill(cls_of, cls_of + 256, -1);
    std::vector<PlanClass> classes;
    for (size_t i = 0; i < count; ++i) {
        unsigned e = ecc_len[i];
        if (cls_of[e] >= 0)
            continue;
        ffrs_rs256 *c = nullptr;
        if (int r = ffrs_rs256_create(ctx, e, 2, 0x11D, &c)) {
            plan_free(p);
            return r;
        }
        cls_of[e] = (int)p->codecs.size();
        p->codecs.push_back(c);
        classes.push_back({c->d_lut, e, c->EW, c->S});
        p->max_ew = std::max(p->max_ew, c->EW);
        p->max_e = std::max(p->max_e, e);
        p->lut_words = std::max(p->lut_words, c->S * 256u * (uint32_t)lut_row((int)c->EW));
    }
    p->wg = p->max_ew <= 16 ? 256 : 64;
    std::copy(cls_of, cls_of + 256, p->cls_of);
    p->list.resize(count);
    for (size_t i = 0; i < count; ++i) {
        p->list[i] = {msg_off[i], ecc_off[i], (uint32_t)i, message_len[i], ecc_len[i], (uint8_t)cls_of[ecc_len[i]]};
        p->msg_end = std::max<uint64_t>(p->msg_end, msg_off[i] + message_len[i]);
        p->ecc_end = std::max<uint64_t>(p->ecc_end, ecc_off[i] + ecc_len[i]);
    }
    hipError_t e = hipSuccess;
    auto up = [&](auto **dst, size_t bytes, const void *src) {
        if (e == hipSuccess)
            e = hipMalloc((void **)dst, std::max<size_t>(1, bytes));
        if (e == hipSuccess && src && bytes)
            e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    up(&p->d_items, count * sizeof(PlanItem), nullptr);
    up(&p->d_chunks, count * sizeof(PlanChunk), nullptr);  // at most one chunk per item
    up(&p->d_classes, classes.size() * sizeof(PlanClass), classes.data());
    if (e == hipSuccess)
        e = hipMalloc((void **)&p->d_queue, 4);
    if (e != hipSuccess) {
        plan_free(p);
        return fail(FFRS_ERR_NOMEM, "plan upload failed: %s", hipGetErrorString(e));
    }
    if (int r = plan_layout(p)) {
        plan_free(p);
        return r;
    }
    *out = p;
    return FFRS_OK;
}

void ffrs_rs256_plan_destroy(ffrs_rs256_plan *plan)
{
    if (plan)
        plan_free(plan);
}

int ffrs_rs256_plan_set_class_mode(ffrs_rs256_plan *plan, unsigned ecc_len, int mode)
{
    if (!plan)
        return fail(FFRS_ERR_INVALID, "plan is NULL");
    if (ecc_len > 255 || plan->cls_of[ecc_len] < 0)
        return fail(FFRS_ERR_INVALID, "the plan has no codewords of ecc_len %u", ecc_len);
    if (mode < 0 || mode > 2)
        return fail(FFRS_ERR_INVALID, "class mode %d (0 automatic, 1 one codeword per lane, 2 per wave)", mode);
    std::lock_guard<std::mutex> lk(plan->ctx->mu);
    if (int r = ctx_set_device(plan->ctx))
        return r;
    plan->mode_of[ecc_len] = mode;
    return plan_layout(plan);
}

int ffrs_rs256_plan_use_graph(ffrs_rs256_plan *plan, int enable)
{
    if (!plan)
        return fail(FFRS_ERR_INVALID, "plan is NULL");
    plan->use_graph = enable != 0;
    return FFRS_OK;
}

int ffrs_rs256_plan_encode(ffrs_rs256_plan *plan, const uint8_t *d_msgs, size_t msg_bytes, uint8_t *d_eccs,
                           size_t ecc_bytes)
{
    if (!plan || (plan->count && (!d_msgs || !d_eccs)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (ecc_bytes < plan->ecc_end)
        return fail(FFRS_ERR_SIZE, "parity buffer of %zu bytes: the plan needs %llu", ecc_bytes,
                    (unsigned long long)plan->ecc_end);
    return plan_run(plan, 0, const_cast<uint8_t *>(d_msgs), msg_bytes, d_eccs, nullptr);
}

int ffrs_rs256_plan_repair(ffrs_rs256_plan *plan, uint8_t *d_msgs, size_t msg_bytes, uint8_t *d_eccs,
                           size_t ecc_bytes, uint8_t *d_status)
{
    if (!plan || (plan->count && (!d_msgs || !d_eccs || !d_status)))
        return fail(FFRS_ERR_INVALID, "NULL argument");
    if (ecc_bytes < plan->ecc_end)
        return fail(FFRS_ERR_SIZE, "parity buffer of %zu bytes: the plan needs %llu", ecc_bytes,
                    (unsigned long long)plan->ecc_end);
    return plan_run(plan, 1, d_msgs, msg_bytes, d_eccs, d_status);
}

}  // extern "C"
