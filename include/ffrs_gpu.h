/*
 * ffrs_gpu.h -- C ABI of the MI355X (gfx950) Reed-Solomon encode / repair path.
 *
 * Plain pointers and sizes only. Every entry point returns an ffrs_status; on
 * failure ffrs_last_error() describes it (thread-local). Host-pointer entry
 * points stage through device memory owned by the context and return when the
 * result is in the caller's buffers; *_device entry points take device pointers,
 * are asynchronous on the context's stream and return after launch.
 *
 * Reference interfaces replaced (paths relative to the reference repository root):
 *   RS256  libffrs/pyrs256.hpp:58-210  (PyRS256: ctor, py_encode, py_repair, py_synds)
 *          libffrs/reed_solomon.hpp:55-88, 169-389, 412-453, 511-552, 555-744
 *   RSi16  libffrs/pyrsi16.hpp:34-692 (PyRSi16: ctor, py_encode, py_repair, py_synd,
 *          py_sugiyama, py_roots), libffrs/rsi16v_impl.hpp:68-636, libffrs/ntt.hpp:32-579
 */
#ifndef FFRS_GPU_H
#define FFRS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFRS_ABI_VERSION 2

typedef enum ffrs_status {
    FFRS_OK = 0,
    FFRS_ERR_INVALID = -1, /* argument value rejected (reference: ValueError)            */
    FFRS_ERR_SIZE = -2,    /* buffer size mismatch (reference: RuntimeError, py_assert)   */
    FFRS_ERR_DEVICE = -3,  /* HIP runtime failure                                         */
    FFRS_ERR_NOMEM = -4,   /* device or host allocation failed                            */
    FFRS_ERR_PARAM = -5    /* codec parameters rejected (reference: RuntimeError)         */
} ffrs_status;

/* Per-codeword RSi16 status values written by the repair entry points: the
 * RepairStatus enum of libffrs/rsi16v.hpp:32-38, plus FFRS_RSI16_RAISES for the
 * inputs on which the reference's repair raises RuntimeError from a py_assert
 * (libffrs/ntt.hpp:482). RS256 repair writes 1 (True) or 0 (False). */
enum {
    FFRS_RSI16_NO_ERRORS = 0,
    FFRS_RSI16_NO_ERRORS_ZERO = 1,
    FFRS_RSI16_REPAIR_OK = 2,
    FFRS_RSI16_REPAIR_FAIL = 3,
    FFRS_RSI16_ERROR_LOCATION_FAIL = 4,
    FFRS_RSI16_RAISES = 255
};

typedef struct ffrs_ctx ffrs_ctx;
typedef struct ffrs_rs256 ffrs_rs256;
typedef struct ffrs_rsi16 ffrs_rsi16;

int ffrs_abi_version(void);
const char *ffrs_last_error(void);

/* ---- context: one device, one stream, device staging memory --------------------- */
int ffrs_ctx_create(int device, ffrs_ctx **out);
void ffrs_ctx_destroy(ffrs_ctx *ctx);
int ffrs_ctx_synchronize(ffrs_ctx *ctx);
int ffrs_ctx_device_info(ffrs_ctx *ctx, char *name, size_t name_cap, int *cu_count, size_t *free_bytes,
                         size_t *total_bytes);
int ffrs_device_alloc(ffrs_ctx *ctx, size_t bytes, void **d_ptr);
int ffrs_device_free(ffrs_ctx *ctx, void *d_ptr);
/* Pinned host memory (hipHostMalloc) for staging, and copies that only enqueue on the
 * context's stream (pageable memory makes them synchronous). */
int ffrs_host_alloc(ffrs_ctx *ctx, size_t bytes, void **h_ptr);
int ffrs_host_free(ffrs_ctx *ctx, void *h_ptr);
int ffrs_memcpy_h2d_async(ffrs_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int ffrs_memcpy_d2h_async(ffrs_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int ffrs_memcpy_h2d(ffrs_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int ffrs_memcpy_d2h(ffrs_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int ffrs_memcpy_d2d(ffrs_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
/* HIP events recorded on the context stream; ffrs_timer_end waits and returns ms. */
int ffrs_timer_begin(ffrs_ctx *ctx);
int ffrs_timer_end(ffrs_ctx *ctx, float *ms);

/* Per-kernel timing (no counterpart in the reference, which has no device): when enabled,
 * every encode/repair call records a HIP event on the context stream after each kernel it
 * launches; ffrs_ctx_kernel_times waits for the last one and writes the durations of the
 * last call's kernels, in launch order, to ms[0..min(cap, *count)). Off by default. */
int ffrs_ctx_kernel_timing(ffrs_ctx *ctx, int enable);

/* Stream capture (no counterpart in the reference): the *_device and plan entry points
 * called between capture_begin and capture_end on this context are recorded, not run, and
 * replayed by ffrs_graph_launch on the context's stream. Codec workspaces must already be
 * large enough (run the same calls once before capturing). */
typedef struct ffrs_graph ffrs_graph;
int ffrs_ctx_capture_begin(ffrs_ctx *ctx);
int ffrs_ctx_capture_end(ffrs_ctx *ctx, ffrs_graph **out);
int ffrs_graph_launch(ffrs_graph *graph);
void ffrs_graph_destroy(ffrs_graph *graph);
int ffrs_ctx_kernel_times(ffrs_ctx *ctx, float *ms, int cap, int *count);

/* ---- RS256 over GF(2^8) ------------------------------------------------------------ */
/* Codec for one ecc_len (1..254) and field (primitive, polynomial); the parity and
 * syndrome tables are built on the host once and kept resident on the device.
 * Replaces the RS256<GF> composition of libffrs/pyrs256.hpp:33-55, 82-86. */
int ffrs_rs256_create(ffrs_ctx *ctx, unsigned ecc_len, unsigned primitive, unsigned polynomial,
                      ffrs_rs256 **out);
void ffrs_rs256_destroy(ffrs_rs256 *rs);
/* Repair kernels (no counterpart in the reference; results do not depend on it): 0 and 1
 * use the syndrome / Berlekamp-Massey / Chien pipeline where ecc_len <= 32 (default), 2
 * one kernel with one codeword per lane; ecc_len > 32 always uses the latter. */
int ffrs_rs256_set_repair_mode(ffrs_rs256 *rs, int mode);

/* generator[0..ecc_len], highest degree first (libffrs/reed_solomon.hpp:55-88) */
int ffrs_rs256_generator(const ffrs_rs256 *rs, uint8_t *generator, uint8_t *roots);

/* PyRS256::py_encode (libffrs/pyrs256.hpp:99-127): `size` bytes split into blocks of
 * message_len plus one shorter tail block; writes ceil(size / message_len) * ecc_len
 * parity bytes to out (out_size must be exactly that). */
int ffrs_rs256_encode(ffrs_rs256 *rs, size_t message_len, const uint8_t *buf, size_t size, uint8_t *out,
                      size_t out_size);
/* `count` codewords; msg holds count * message_len bytes, ecc count * ecc_len bytes,
 * both repaired in place; status[i] = 1 (True) / 0 (False). One call of
 * PyRS256::py_repair (libffrs/pyrs256.hpp:129-131 -> reed_solomon.hpp:559-595) per codeword. */
int ffrs_rs256_repair(ffrs_rs256 *rs, size_t message_len, uint8_t *msg, uint8_t *ecc, size_t count,
                      uint8_t *status);
/* PyRS256::py_synds (libffrs/pyrs256.hpp:133-140) for `count` codewords. */
int ffrs_rs256_synds(ffrs_rs256 *rs, size_t message_len, const uint8_t *msg, const uint8_t *ecc, size_t count,
                     uint8_t *synds);
/* Device-pointer forms: inputs already resident in HBM, asynchronous on the stream. */
int ffrs_rs256_encode_device(ffrs_rs256 *rs, size_t message_len, const uint8_t *d_msg, size_t count,
                             uint8_t *d_ecc);
int ffrs_rs256_repair_device(ffrs_rs256 *rs, size_t message_len, uint8_t *d_msg, uint8_t *d_ecc, size_t count,
                             uint8_t *d_status);
/* Allocation-free form of ffrs_rs256_repair_device (no counterpart in the reference, whose
 * repair has no device workspace): for callers that cannot allocate device memory, such as
 * code under HIP graph capture or a serving loop that forbids hipMalloc. Same arguments and
 * results; it calls no hipMalloc/hipFree or allocating library routine. The syndrome / BM
 * workspace of the repair pipeline lives in one static device array of
 * FFRS_RS256_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), which limits a call:
 * ecc_len 2, 4, 8, 16 and 32 take 4, 4, 8, 16 and 32 bytes per codeword (ecc_len 32: at most
 * 1,179,648 codewords per call), other ecc_len <= 32 take 4-32 bytes plus a 48-byte record,
 * ecc_len > 32 needs none. Larger calls fail with FFRS_ERR_SIZE and name the limit. Every call
 * on one GPU shares that array: calls that overlap in time on one GPU (other streams or host
 * threads) race and give wrong results; calls in order on one stream are safe. */
#define FFRS_RS256_NOALLOC_SCRATCH_BYTES (36u << 20)
int ffrs_rs256_repair_device_noalloc(ffrs_rs256 *rs, size_t message_len, uint8_t *d_msg, uint8_t *d_ecc,
                                     size_t count, uint8_t *d_status);

/* ---- RSi16 over GF(65537) ---------------------------------------------------------- */
/* Codec (block_len, ecc_len, primitive) with interleave 1: field, NTT and mix tables
 * resident on the device (libffrs/pyrsi16.hpp:41-125, libffrs/ntt.hpp:55-123,
 * libffrs/rsi16v_impl.hpp:68-86). Fails with FFRS_ERR_PARAM where the reference
 * constructor raises. */
int ffrs_rsi16_create(ffrs_ctx *ctx, size_t block_len, size_t ecc_len, uint32_t primitive, ffrs_rsi16 **out);
void ffrs_rsi16_destroy(ffrs_rsi16 *rs);
/* Lanes per codeword of the encode and repair kernels (no counterpart in the reference;
 * results do not depend on them): powers of two up to min(64, ecc_len); 0 keeps the
 * current value. Encode with 1 lane and ecc_len <= 32 runs the transforms in registers.
 * Defaults: encode 1 for ecc_len <= 32, else min(64, ecc_len / 2); repair min(64, ecc_len / 2).
 * libffrs_amd.autotune picks them per (block_len, ecc_len) by timing. */
int ffrs_rsi16_set_groups(ffrs_rsi16 *rs, unsigned g_encode, unsigned g_repair);
int ffrs_rsi16_get_groups(const ffrs_rsi16 *rs, unsigned *g_encode, unsigned *g_repair);
/* Encode kernels (no counterpart in the reference; results do not depend on them). The int8
 * MFMA contraction parity = message . G mod p applies where it is the encode of the words asked
 * for: interleave 1, ecc_len a multiple of 16 up to 256, message_len <= 16384, 16-byte aligned
 * buffers, and with the SIMD words only where no mix constant of rsi16v_impl.hpp:68-86 is 65536.
 * 0 automatic (the contraction where it applies and ecc_len <= 32, else the butterfly kernels),
 * 1 always the butterfly kernels, 2 the contraction wherever it applies. */
int ffrs_rsi16_set_encode_kernel(ffrs_rsi16 *rs, int kernel);
/* Where the repair computes its syndromes (no counterpart in the reference; results do not
 * depend on it): 0 automatic, 1 inside the repair kernel, 2 as an int8 MFMA contraction of the
 * codewords with the matrix w^(j rbo(m)) ahead of it, where its tables exist (ecc_len 16, 32 or
 * 64, message_len <= 16384; automatic uses it there). */
int ffrs_rsi16_set_repair_syndromes(ffrs_rsi16 *rs, int mode);
uint32_t ffrs_rsi16_root(const ffrs_rsi16 *rs);
size_t ffrs_rsi16_ntt_len(const ffrs_rsi16 *rs);

/* encode_blocks (libffrs/pyrsi16.hpp:128-140, 364-388 -> rsi16v_impl.hpp:93-119) over
 * `count` messages of block_len - ecc_len symbols. vector_lanes != 0 reproduces the
 * SIMD instantiations' lane arithmetic (W = 4/8/16, the reference's default on x86-64),
 * 0 the scalar W = 1 one. Parity symbols are written as value & 0xFFFF. */
int ffrs_rsi16_encode(ffrs_rsi16 *rs, const uint16_t *msg, size_t count, uint16_t *ecc, int vector_lanes);
/* Interleaved encode, RSi16(block_len, ecc_len, interleave) (libffrs/pyrsi16.hpp:143-150,
 * 324-361): `blocks` interleaved blocks of message_len * interleave symbols, symbol i of
 * column c at i * interleave + c; parity symbol j of column c at j * interleave + c of the
 * block's ecc_len * interleave outputs. Every block is encoded (the reference fills only the
 * first block of a multi-block call and leaves the others uninitialised). */
int ffrs_rsi16_encode_interleaved(ffrs_rsi16 *rs, const uint16_t *msg, size_t blocks, size_t interleave, uint16_t *ecc,
                                  int vector_lanes);
int ffrs_rsi16_encode_interleaved_device(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t blocks, size_t interleave,
                                         uint16_t *d_ecc, int vector_lanes);

/* `count` calls of PyRSi16::py_repair without error positions (libffrs/pyrsi16.hpp:588-631
 * -> rsi16v_impl.hpp:122-168); repaired in place, status[i] as above. */
int ffrs_rsi16_repair(ffrs_rsi16 *rs, uint16_t *msg, uint16_t *ecc, size_t count, uint8_t *status);
/* py_repair with error positions (natural order, < block_len, at most ecc_len):
 * rsi16v_impl.hpp:171-231. One codeword. */
int ffrs_rsi16_repair_known(ffrs_rsi16 *rs, uint16_t *msg, uint16_t *ecc, const uint32_t *positions,
                            size_t npos, uint8_t *status);
/* Interleaved repair without error positions with the SIMD words, RSi16(block_len, ecc_len,
 * interleave) with simd_x4/x8/x16 resolving to `width` = 4, 8 or 16 lanes
 * (libffrs/pyrsi16.hpp:389-453 -> the vector RSi16v<W>::_repair, rsi16v_impl.hpp:122-168,
 * 271-497): the columns (symbol i of column c at i * interleave + c, message and ecc alike)
 * are repaired in groups of `width`, the last one zero-filled past the columns, and the
 * columns of a group share the reference's vector control flow, so a column's result can
 * depend on its neighbours. Repaired in place; *status = the largest group status, or
 * FFRS_RSI16_RAISES where a group's repair raises in the reference (the groups before it are
 * written back, the others left as they were). With error positions the SIMD and scalar
 * words give the same result (column by column through ffrs_rsi16_repair_known). */
int ffrs_rsi16_repair_interleaved_simd(ffrs_rsi16 *rs, uint16_t *msg, uint16_t *ecc, size_t interleave,
                                       unsigned width, uint8_t *status);
/* py_synd (libffrs/pyrsi16.hpp:633-651): ecc_len syndromes per codeword. */
int ffrs_rsi16_synds(ffrs_rsi16 *rs, const uint16_t *msg, const uint16_t *ecc, size_t count, uint32_t *synds);
/* py_sugiyama<1> (libffrs/pyrsi16.hpp:653-674): locator and evaluator (ecc_len values
 * each) per syndrome vector; length[i] = locator length or 0xFFFFFFFF where the
 * reference raises. */
int ffrs_rsi16_sugiyama(ffrs_rsi16 *rs, const uint32_t *synds, size_t count, uint32_t *locator,
                        uint32_t *evaluator, uint32_t *length);
/* The reference's NTT class methods (libffrs/pyntt.hpp:28-84, rsi16v_impl.hpp:670-795) and
 * RSi16._sugiyama<W> (pyrsi16.hpp:653-674) on `width` = 1, 4, 8 or 16 vectors of this codec's
 * NTT (width 1: the scalar words; 4-16: one SIMD lane per vector, coupled where the reference's
 * vector code couples its lanes). Vectors are consecutive in a, b and out (width x len values,
 * len = ntt_len for ops 0-3, ecc_len for 4-7). op: 0 ntt (RBO in, natural out), 1 intt,
 * 2 nttr (first ecc_len values), 3 inttr, 4 poly_mul(a, b), 5 poly_div(a, b), 6 poly_inv(a, n)
 * (n <= ecc_len), 7 Sugiyama of the syndromes a (out: evaluator, out2: locator). result[l]:
 * 1 where the reference raises (ops 5, 6), the locator length or 0xFFFFFFFF (op 7), else 0.
 * Values are field elements (0..65536); the Python methods return them truncated to u16 as the
 * reference does. */
int ffrs_rsi16_ntt_op(ffrs_rsi16 *rs, int op, unsigned width, const uint32_t *a, const uint32_t *b, uint32_t n,
                      uint32_t *out, uint32_t *out2, uint32_t *result);
/* py_roots (libffrs/pyrsi16.hpp:676-691): is_root[m] = 1 where the partial inverse NTT
 * of the ecc_len coefficients vanishes, m < block_len. */
int ffrs_rsi16_roots(ffrs_rsi16 *rs, const uint32_t *poly, uint8_t *is_root);
int ffrs_rsi16_encode_device(ffrs_rsi16 *rs, const uint16_t *d_msg, size_t count, uint16_t *d_ecc,
                             int vector_lanes);
int ffrs_rsi16_repair_device(ffrs_rsi16 *rs, uint16_t *d_msg, uint16_t *d_ecc, size_t count,
                             uint8_t *d_status);

/* ---- CIRC16 ------------------------------------------------------------------------------
 * PyCIRC16 (libffrs/pycirc16.hpp:33-431): a cross-interleaved code over two RSi16 codecs of one
 * context and primitive, inner RSi16(inner_block_len, inner_ecc_len) and outer
 * RSi16(outer_block_len, outer_ecc_len) whose interleave is inner_message_len * interleave
 * (pycirc16.hpp:45-64). simd_width: the lanes of the reference's SIMD words that the simd_x*
 * flags select (16, 8, 4; 1 for the scalar words). The codecs must outlive the handle.
 * Message: outer_message_len rows of inner_message_len * interleave symbols; ecc: outer parity
 * (outer_ecc_len rows of the same width), inner parity of every message row block, inner parity
 * of every outer parity row block (message_offset / rso_ecc_offset / rsi_ecc_offset,
 * pycirc16.hpp:423-441). */
typedef struct ffrs_circ16 ffrs_circ16;
int ffrs_circ16_create(ffrs_rsi16 *inner, ffrs_rsi16 *outer, size_t interleave, unsigned simd_width,
                       ffrs_circ16 **out);
void ffrs_circ16_destroy(ffrs_circ16 *c);
size_t ffrs_circ16_message_len(const ffrs_circ16 *c);
size_t ffrs_circ16_ecc_len(const ffrs_circ16 *c);
/* py_encode / encode_block (pycirc16.hpp:356-381): `blocks` messages of message_len symbols to
 * blocks * ecc_len ecc symbols; the inner parity of the outer parity rows is computed from the
 * outer parity's field elements (a parity symbol of 65536 is stored as 0 but encoded as 65536). */
int ffrs_circ16_encode(ffrs_circ16 *c, const uint16_t *msg, size_t blocks, uint16_t *ecc);
/* py_repair (pycirc16.hpp:383-403) of one block, in place. *raised = 1 where the reference's
 * repair raises (a py_assert of the outer repair without positions): msg then holds what the
 * reference leaves (the interleave blocks before the raising one repaired, that one up to its
 * raising SIMD group) and ecc is unchanged. */
int ffrs_circ16_repair(ffrs_circ16 *c, uint16_t *msg, uint16_t *ecc, int *raised);
/* py_find_outer_error_locations (pycirc16.hpp:405-421): the outer rows of interleave block
 * `interleave` whose inner syndromes are not all zero after the zero repairs, ascending, into
 * rows[0 .. *count) (rows holds outer_block_len entries). */
int ffrs_circ16_find_outer_error_locations(ffrs_circ16 *c, const uint16_t *msg, const uint16_t *ecc,
                                           size_t interleave, uint32_t *rows, size_t *count);

/* ---- mixed-size batches --------------------------------------------------------------
 * No counterpart in the reference, whose codec objects serve one (block_len, ecc_len) each
 * (pyrs256.hpp:58-97, pyrsi16.hpp:41-125): a plan lists `count` codewords, each with its own
 * lengths and element offsets into one message buffer and one parity buffer. Running a plan
 * is one persistent kernel that takes chunks of codewords from a device work queue (plus a
 * one-thread reset of the queue); with use_graph set, the first run for a set of buffers is
 * captured in a HIP graph and later runs replay it. Runs are asynchronous on the context's
 * stream; status[i] belongs to the i-th codeword of the list. */
typedef struct ffrs_rs256_plan ffrs_rs256_plan;
typedef struct ffrs_rsi16_plan ffrs_rsi16_plan;

/* RS256: offsets in bytes; message_len[i] + ecc_len[i] <= 255. */
int ffrs_rs256_plan_create(ffrs_ctx *ctx, size_t count, const uint8_t *ecc_len, const uint8_t *message_len,
                           const uint64_t *msg_off, const uint64_t *ecc_off, ffrs_rs256_plan **out);
void ffrs_rs256_plan_destroy(ffrs_rs256_plan *plan);
int ffrs_rs256_plan_use_graph(ffrs_rs256_plan *plan, int enable);
/* Repair kernel of the plan's codewords of one ecc_len (no counterpart in the reference; results
 * do not depend on it): 0 automatic (one codeword per wave above ecc_len 32, else one per lane),
 * 1 one codeword per lane, 2 one codeword per wave. Lays the plan out again (graphs are
 * recaptured on the next run). libffrs_amd.autotune picks it by timing. */
int ffrs_rs256_plan_set_class_mode(ffrs_rs256_plan *plan, unsigned ecc_len, int mode);
int ffrs_rs256_plan_encode(ffrs_rs256_plan *plan, const uint8_t *d_msgs, size_t msg_bytes, uint8_t *d_eccs,
                           size_t ecc_bytes);
int ffrs_rs256_plan_repair(ffrs_rs256_plan *plan, uint8_t *d_msgs, size_t msg_bytes, uint8_t *d_eccs,
                           size_t ecc_bytes, uint8_t *d_status);

/* RSi16: offsets and sizes in u16 symbols; every (block_len, ecc_len) must be accepted by
 * ffrs_rsi16_create with the given primitive (0 stands for 3, the reference's default,
 * pyrsi16.hpp:269-287); block_len up to 65536. */
int ffrs_rsi16_plan_create(ffrs_ctx *ctx, size_t count, const uint32_t *block_len, const uint32_t *ecc_len,
                           const uint64_t *msg_off, const uint64_t *ecc_off, uint32_t primitive,
                           ffrs_rsi16_plan **out);
void ffrs_rsi16_plan_destroy(ffrs_rsi16_plan *plan);
int ffrs_rsi16_plan_use_graph(ffrs_rsi16_plan *plan, int enable);
/* Shape of the plan's RSi16(block_len, ecc_len) codewords (no counterpart in the reference;
 * results do not depend on it): l_encode lanes per codeword for the encode of ecc_len <= 32 (0:
 * unchanged), g_repair lanes per codeword for the repair (0 or min(64, ecc_len/2): the register
 * code; other powers of two: the LDS code). Lays the plan out again. libffrs_amd.autotune picks
 * them by timing. */
int ffrs_rsi16_plan_set_class_groups(ffrs_rsi16_plan *plan, uint32_t block_len, uint32_t ecc_len, unsigned l_encode,
                                     unsigned g_repair);
int ffrs_rsi16_plan_encode(ffrs_rsi16_plan *plan, const uint16_t *d_msgs, size_t msg_elems, uint16_t *d_eccs,
                           size_t ecc_elems, int vector_lanes);
int ffrs_rsi16_plan_repair(ffrs_rsi16_plan *plan, uint16_t *d_msgs, size_t msg_elems, uint16_t *d_eccs,
                           size_t ecc_elems, uint8_t *d_status);

#ifdef __cplusplus
}
#endif
#endif
