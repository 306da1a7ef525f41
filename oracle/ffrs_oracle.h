/*
 * ffrs_oracle.h -- CPU restatement of libffrs' RS256 / RSi16 encode and repair.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libffrs_amd) never links it.
 *
 * Every function restates one reference routine; the reference location is cited
 * next to each declaration (paths relative to the reference repository root).
 */
#ifndef FFRS_ORACLE_H
#define FFRS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- GF(2^8) */

/* Field tables, libffrs/galois.hpp:133-144 (loop i = 0..255 inclusive, so
 * exp[255] = 1 and log[1] = 255; log[0] stays 0). Multiplication used to build
 * the tables is the bit-serial one of galois.hpp:624-651. Returns 0 when the
 * primitive generates all 255 non-zero elements, -1 otherwise. */
int orc_gf256_tables(unsigned primitive, unsigned polynomial, uint8_t exp_t[256], uint8_t log_t[256]);

typedef struct orc_rs256 orc_rs256;

/* RS256 codec state: generator (reed_solomon.hpp:55-88), field tables. */
orc_rs256 *orc_rs256_new(unsigned ecc_len, unsigned primitive, unsigned polynomial);
void orc_rs256_free(orc_rs256 *rs);
/* generator[0..ecc], highest degree first (generator[0] = 1). */
void orc_rs256_generator(const orc_rs256 *rs, uint8_t *generator);

/* Systematic parity of one block of `len` bytes: x^ecc * M(x) mod g(x), byte 0 of
 * the message is the highest-degree coefficient (reed_solomon.hpp:169-297,
 * pinned to galois.hpp:461-499 by tests/test_lib_rs256.py:84-90). */
void orc_rs256_encode_block(const orc_rs256 *rs, const uint8_t *msg, size_t len, uint8_t *ecc_out);

/* PyRS256::py_encode (pyrs256.hpp:99-127): full blocks of message_len bytes plus a
 * shortened tail block. Returns the number of parity bytes written. */
size_t orc_rs256_encode(const orc_rs256 *rs, size_t message_len, const uint8_t *buf, size_t size, uint8_t *out);

/* Syndromes S_i = C(alpha^i), i < ecc (reed_solomon.hpp:412-453). */
void orc_rs256_synds(const orc_rs256 *rs, const uint8_t *msg, size_t k, const uint8_t *ecc, uint8_t *synds);

/* rs_decode::decode (reed_solomon.hpp:559-595): returns 1 (True) or 0 (False);
 * corrections are written in place only on success. */
int orc_rs256_decode(const orc_rs256 *rs, uint8_t *msg, size_t k, uint8_t *ecc);

/* Batched helpers over `count` codewords laid out back to back; split over threads. */
void orc_rs256_encode_batch(const orc_rs256 *rs, const uint8_t *msgs, size_t count, size_t k,
                            uint8_t *ecc_out, int threads);
void orc_rs256_decode_batch(const orc_rs256 *rs, uint8_t *msgs, uint8_t *eccs, size_t count, size_t k,
                            uint8_t *status, int threads);

/* ---------------------------------------------------------------- GF(65537) */

typedef struct orc_rsi16 orc_rsi16;

/* Repair status, libffrs/rsi16v.hpp:32-38. ORC_RSI16_ASSERT marks inputs on which
 * the reference raises RuntimeError from a py_assert inside repair. */
enum {
    ORC_NO_ERRORS = 0,
    ORC_NO_ERRORS_ZERO = 1,
    ORC_REPAIR_OK = 2,
    ORC_REPAIR_FAIL = 3,
    ORC_ERROR_LOCATION_FAIL = 4,
    ORC_RSI16_ASSERT = 255
};

/* RSi16 codec state: field tables (galois.hpp:93-146, 168-216), NTT tables
 * (ntt.hpp:55-123) and parity mix constants (rsi16v_impl.hpp:68-86).
 * Returns NULL where the reference constructor raises. */
orc_rsi16 *orc_rsi16_new(size_t block_len, size_t ecc_len, uint32_t primitive);
void orc_rsi16_free(orc_rsi16 *rs);
uint32_t orc_rsi16_root(const orc_rsi16 *rs);
size_t orc_rsi16_ntt_len(const orc_rsi16 *rs);
/* Copy of an internal table: 0 exp, 1 log, 2 roots_ntt, 3 pntt_shift, 4 pintt_shift,
 * 5 ecc_mix, 6 roots_ecc, 7 roots_i_ecc. Returns its length. */
size_t orc_rsi16_table(const orc_rsi16 *rs, int which, uint32_t *out, size_t cap);

/* Encode `count` messages of k u16 symbols (pyrsi16.hpp:128-140, 364-388 ->
 * rsi16v_impl.hpp:93-119). vector_lanes != 0 selects the arithmetic of the
 * SIMD template instantiations (W = 4/8/16); 0 the scalar W = 1 instantiation.
 * Parity is written as value & 0xFFFF. */
void orc_rsi16_encode(const orc_rsi16 *rs, const uint16_t *msgs, size_t count, uint16_t *ecc_out,
                      int vector_lanes, int threads);

/* Syndromes S_j, j < ecc, of one codeword (NTT::pntt, ntt.hpp:133-152). */
void orc_rsi16_synds(const orc_rsi16 *rs, const uint16_t *msg, const uint16_t *ecc, uint32_t *synds);

/* RSi16v<1>::_sugiyama (rsi16v_impl.hpp:271-440) on ecc syndromes; writes the
 * locator and evaluator (ecc values each) and returns the locator length, or
 * (uint32_t)-1 where the reference raises. */
uint32_t orc_rsi16_sugiyama(const orc_rsi16 *rs, const uint32_t *synds, uint32_t *locator, uint32_t *evaluator);

/* PyRSi16::py_repair with interleave == 1 and no error positions
 * (pyrsi16.hpp:588-631 -> rsi16v_impl.hpp:122-168). Returns a status above;
 * msg / ecc are updated in place exactly as the reference writes them back. */
int orc_rsi16_repair(const orc_rsi16 *rs, uint16_t *msg, uint16_t *ecc);
/* Interleaved repair with the SIMD words, W = 4/8/16 columns per coupled group
 * (pyrsi16.hpp:389-453, rsi16v_impl.hpp:122-168, 271-497 vector instantiations). */
int orc_rsi16_repair_interleaved_simd(const orc_rsi16 *rs, unsigned W, uint16_t *msg, uint16_t *ecc,
                                      size_t interleave);
void orc_rsi16_repair_batch(const orc_rsi16 *rs, uint16_t *msgs, uint16_t *eccs, size_t count,
                            uint8_t *status, int threads);

/* Known-position repair (rsi16v_impl.hpp:171-231), error positions in natural order. */
int orc_rsi16_repair_known(const orc_rsi16 *rs, uint16_t *msg, uint16_t *ecc, const size_t *pos, size_t npos);

#ifdef __cplusplus
}
#endif
#endif
