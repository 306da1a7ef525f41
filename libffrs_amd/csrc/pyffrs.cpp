// pyffrs.cpp -- pybind11 module `libffrs` of libffrs_amd.
//
// Mirrors the reference module's classes for the encode / repair path (names,
// argument meaning, properties, exceptions) and forwards every codec call to the
// C ABI of include/ffrs_gpu.h, which runs the HIP kernels. No codec arithmetic is
// done here; the scalar GF256 / GFi16 helpers are host utilities off the hot path.
//
// Reference interfaces mirrored (paths relative to the reference repository root):
//   module       libffrs/pyffrs.cpp:101-138
//   RS256        libffrs/pyrs256.hpp:58-210
//   RSi16        libffrs/pyrsi16.hpp:34-320, 568-691
//   GF256/GFi16  libffrs/pygf256.hpp:259-362, libffrs/pygfi16.hpp:40-79
//   buffers      libffrs/util.hpp:30-55, 98-102
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <sys/mman.h>

#include <cerrno>
#include <cstring>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ffrs_gpu.h"

namespace py = pybind11;

namespace {

// ---------------------------------------------------------------- errors / context

[[noreturn]] void raise_status(int code, const char *what)
{
    std::string msg = std::string(what) + ": " + ffrs_last_error();
    if (code == FFRS_ERR_INVALID)
        throw py::value_error(msg);
    throw std::runtime_error(msg);
}

inline void check(int code, const char *what)
{
    if (code != FFRS_OK)
        raise_status(code, what);
}

std::mutex g_ctx_mu;
ffrs_ctx *g_ctx = nullptr;

ffrs_ctx *default_ctx()
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx) {
        int dev = 0;
        if (const char *e = std::getenv("FFRS_DEVICE"))
            dev = std::atoi(e);
        int r = ffrs_ctx_create(dev, &g_ctx);
        if (r != FFRS_OK) {
            g_ctx = nullptr;
            throw std::runtime_error(std::string("libffrs_amd: no usable HIP device: ") + ffrs_last_error() +
                                     " (a process that also uses PyTorch on the GPU must import torch before "
                                     "libffrs_amd, so that both bind to one HIP runtime)");
        }
    }
    return g_ctx;
}

// ---------------------------------------------------------------- buffers (util.hpp:30-55)

template <typename T>
struct Buf {
    py::buffer_info info;
    size_t size = 0;
    T *data = nullptr;
    explicit Buf(const py::buffer &b) : info(b.request(false))
    {
        if (info.size % (py::ssize_t)sizeof(T) != 0)
            throw py::value_error("Invalid buffer size: " + std::to_string(info.size));
        size = (size_t)info.size / sizeof(T);
        data = reinterpret_cast<T *>(info.ptr);
    }
};

#define PY_ASSERT(cond, text)                                                                  \
    do {                                                                                       \
        if (!(cond))                                                                           \
            throw std::runtime_error(std::string("Assertion failed: " #cond ", ") + (text));  \
    } while (0)

// ---------------------------------------------------------------- GF(2^8) scalar helpers

struct GF256 {
    uint8_t primitive, poly1;
    uint8_t exp_t[256], log_t[256];
    GF256(uint8_t primitive_, uint16_t polynomial) : primitive(primitive_), poly1((uint8_t)(polynomial & 0xff))
    {
        memset(log_t, 0, sizeof log_t);
        uint8_t x = 1;
        for (int i = 0; i < 256; ++i) {  // galois.hpp:133-144
            exp_t[i] = x;
            log_t[x] = (uint8_t)i;
            x = mul_serial(x, primitive);
        }
    }
    uint8_t mul_serial(uint8_t a, uint8_t b) const
    {
        uint8_t r = 0;
        for (int i = 7; i >= 0; --i) {
            r = (r & 0x80) ? (uint8_t)((r << 1) ^ poly1) : (uint8_t)(r << 1);
            if (a & (1u << i))
                r ^= b;
        }
        return r;
    }
    uint8_t mul(uint8_t a, uint8_t b) const
    {
        if (!a || !b)
            return 0;
        size_t r = (size_t)log_t[a] + log_t[b];
        if (r >= 255)
            r -= 255;
        return exp_t[r];
    }
    uint8_t inv(uint8_t a) const { return exp_t[255 - log_t[a]]; }
    uint8_t div(uint8_t a, uint8_t b) const
    {
        if (!a)
            return 0;
        size_t r = (size_t)log_t[a] + 255 - log_t[b];
        if (r >= 255)
            r -= 255;
        return exp_t[r];
    }
    uint8_t pow(uint8_t a, uint8_t b) const { return exp_t[((size_t)log_t[a] * b) % 255]; }
};

// ---------------------------------------------------------------- GF(65537) scalar helpers

struct GFi16 {
    uint32_t primitive;
    std::vector<uint32_t> exp_t, log_t;
    explicit GFi16(uint32_t primitive_) : primitive(primitive_), exp_t(65537), log_t(65537, 0)
    {
        uint32_t x = 1;
        for (uint32_t i = 0; i < 65537; ++i) {  // galois.hpp:133-144 with gf_mul_mod
            exp_t[i] = x;
            log_t[x] = i;
            uint32_t r;
            if (__builtin_mul_overflow(x, primitive, &r))
                r = 1;
            x = r % 65537u;
        }
        if (primitive > 65536 || !(pow(primitive, 32768) == 65536 && pow(primitive, 65536) == 1))
            throw std::runtime_error(std::to_string(primitive) + " is not a primitive element of GF(65537)");
    }
    uint32_t add(uint32_t a, uint32_t b) const
    {
        uint32_t r = a + b;
        return r >= 65537u ? r - 65537u : r;
    }
    uint32_t sub(uint32_t a, uint32_t b) const
    {
        uint32_t r = a - b;
        return r >= 0x80000000u ? r + 65537u : r;
    }
    uint32_t mul(uint32_t a, uint32_t b) const
    {
        uint32_t r = a * b;
        if (r == 0 && a && b)
            return 1;
        r = (r & 0xffff) - (r >> 16);
        return r >= 0x80000000u ? r + 65537u : r;
    }
    uint32_t inv(uint32_t a) const { return exp_t.at(65536 - log_t.at(a)); }
    uint32_t div(uint32_t a, uint32_t b) const
    {
        if (!a)
            return 0;
        uint32_t r = log_t.at(a) + 65536 - log_t.at(b);
        if (r >= 65536)
            r -= 65536;
        return exp_t[r];
    }
    uint32_t pow(uint32_t a, uint32_t e) const { return exp_t[(uint32_t)(log_t.at(a) * e) % 65536u]; }
};

// ---------------------------------------------------------------- RS256

class PyRS256 {
public:
    size_t ecc_len = 0, block_len = 0, message_len = 0;
    std::unique_ptr<GF256> gf;
    ffrs_rs256 *codec = nullptr;
    std::vector<uint8_t> generator, roots;

    PyRS256(std::optional<uint8_t> block_len_, std::optional<uint8_t> message_len_, std::optional<uint8_t> ecc_len_,
            uint8_t primitive, uint16_t polynomial)
    {
        // pyrs256.hpp:63-97 and :195-209, in the reference's order of checks
        uint8_t e;
        if (ecc_len_) {
            e = *ecc_len_;
        } else if (message_len_ && block_len_) {
            if (*message_len_ >= *block_len_)
                throw py::value_error("block_len must be greater than message_len");
            e = (uint8_t)(*block_len_ - *message_len_);
        } else {
            throw py::value_error("Must specify either (block_len, message_len) or ecc_len");
        }
        ecc_len = e;
        set_block_len(block_len_.value_or(255));
        if (ecc_len_ && message_len_) {
            if (block_len_ && (size_t)*message_len_ + *ecc_len_ != *block_len_)
                throw py::value_error("block_len must be equal to message_len + ecc_len");
            else if (!block_len_)
                set_block_len((size_t)*message_len_ + (size_t)*ecc_len_);
        }
        if (ecc_len == 0)
            throw py::value_error("ecc_len must be at least 1");
        gf = std::make_unique<GF256>(primitive, polynomial);
        check(ffrs_rs256_create(default_ctx(), (unsigned)ecc_len, primitive, polynomial, &codec), "RS256");
        generator.resize(ecc_len + 1);
        roots.resize(ecc_len);
        check(ffrs_rs256_generator(codec, generator.data(), roots.data()), "RS256");
    }
    ~PyRS256()
    {
        if (codec)
            ffrs_rs256_destroy(codec);
    }

    void set_block_len(size_t bl)
    {
        if (bl <= ecc_len)
            throw py::value_error("block_len must be greater than ecc_len");
        if (bl > 255)
            throw py::value_error("block_len must be <= 255");
        block_len = bl;
        message_len = bl - ecc_len;
    }

    py::bytearray encode(py::buffer buffer)
    {
        Buf<const uint8_t> buf(buffer);
        size_t k = block_len - ecc_len;
        if (buf.size == 0)
            return py::bytearray();
        size_t full = buf.size / k, tail = buf.size - full * k;
        size_t out_size = (full + (tail ? 1 : 0)) * ecc_len;
        py::bytearray out;
        PyByteArray_Resize(out.ptr(), (Py_ssize_t)out_size);
        uint8_t *o = reinterpret_cast<uint8_t *>(PyByteArray_AsString(out.ptr()));
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rs256_encode(codec, k, buf.data, buf.size, o, out_size);
        }
        check(r, "RS256.encode");
        return out;
    }

    bool repair(py::buffer buffer, py::buffer ecc)
    {
        Buf<uint8_t> m(buffer), e(ecc);
        if (m.size < message_len || e.size < ecc_len)
            throw py::value_error("repair needs message_len message bytes and ecc_len parity bytes");
        uint8_t st = 0;
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rs256_repair(codec, message_len, m.data, e.data, 1, &st);
        }
        check(r, "RS256.repair");
        return st != 0;
    }

    // Extension: repair every codeword of a batch in one device pass.
    py::bytearray repair_batch(py::buffer buffer, py::buffer ecc)
    {
        Buf<uint8_t> m(buffer), e(ecc);
        size_t count = m.size / message_len;
        if (m.size != count * message_len || e.size != count * ecc_len)
            throw py::value_error("repair_batch needs count * message_len and count * ecc_len bytes");
        py::bytearray out;
        PyByteArray_Resize(out.ptr(), (Py_ssize_t)count);
        uint8_t *st = reinterpret_cast<uint8_t *>(PyByteArray_AsString(out.ptr()));
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rs256_repair(codec, message_len, m.data, e.data, count, st);
        }
        check(r, "RS256.repair_batch");
        return out;
    }

    py::bytearray synds(py::buffer buffer)
    {
        Buf<const uint8_t> b(buffer);
        if (b.size < ecc_len)
            return py::bytearray();
        if (b.size < block_len)
            throw py::value_error("_synds needs block_len bytes");
        py::bytearray out;
        PyByteArray_Resize(out.ptr(), (Py_ssize_t)ecc_len);
        uint8_t *o = reinterpret_cast<uint8_t *>(PyByteArray_AsString(out.ptr()));
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rs256_synds(codec, message_len, b.data, b.data + message_len, 1, o);
        }
        check(r, "RS256._synds");
        return out;
    }
};

// ---------------------------------------------------------------- RSi16

enum class RepairStatus : uint32_t { NoErrors, NoErrorsZero, RepairOk, RepairFail, ErrorLocationFail };

// The reference's NTT class (pyntt.hpp:28-84): the transforms and polynomial operations of an
// RSi16(block_len, ecc_len) code's NTT, on 1, 4, 8 or 16 vectors per call (the numbered
// variants run the reference's SIMD instantiations: one vector per SIMD lane). It holds its own
// codec (its device tables); RSi16.ntt returns one for the codec's shape.
class PyNTT {
public:
    size_t block_len, ecc_len, ntt_len;
    ffrs_rsi16 *codec = nullptr;

    PyNTT(const GFi16 &gf, size_t block_len_, size_t ecc_len_) : block_len(block_len_), ecc_len(ecc_len_)
    {
        int r = ffrs_rsi16_create(default_ctx(), block_len, ecc_len, gf.primitive, &codec);
        check(r, "NTT");
        ntt_len = ffrs_rsi16_ntt_len(codec);
    }
    ~PyNTT()
    {
        if (codec)
            ffrs_rsi16_destroy(codec);
    }

    // ntt.hpp:125-128: reverse the low log2(ntt_len) bits
    uint16_t rbo(uint16_t b) const
    {
        unsigned bits = 0;
        while ((1u << bits) < ntt_len)
            ++bits;
        uint32_t r = 0;
        for (unsigned i = 0; i < bits; ++i)
            r |= ((b >> i) & 1u) << (bits - 1 - i);
        return (uint16_t)r;
    }

    py::bytearray run(int op, unsigned W, const std::vector<py::buffer> &in, uint32_t n = 0)
    {
        const size_t len = op <= 3 ? ntt_len : ecc_len;
        std::vector<std::vector<uint32_t>> v;
        for (const auto &b : in) {
            Buf<const uint16_t> x(b);
            PY_ASSERT(x.size == W * len, std::to_string(x.size) + " != " + std::to_string(W * len));
            v.emplace_back(x.data, x.data + x.size);
        }
        std::vector<uint32_t> out(W * len), res(W);
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rsi16_ntt_op(codec, op, W, v[0].data(), v.size() > 1 ? v[1].data() : nullptr, n, out.data(), nullptr,
                                  res.data());
        }
        check(r, "NTT");
        for (uint32_t x : res)
            if (x)
                throw std::runtime_error(op == 5 ? "Assertion failed: !any((g_len == 0) && (f_len > 0))"
                                                 : "Assertion failed: g_len + l - 2 <= ecc_len (aliasing)");
        py::bytearray o;
        PyByteArray_Resize(o.ptr(), (Py_ssize_t)(out.size() * 2));
        uint16_t *d = reinterpret_cast<uint16_t *>(PyByteArray_AsString(o.ptr()));
        for (size_t i = 0; i < out.size(); ++i)
            d[i] = (uint16_t)(out[i] & 0xffffu);  // the reference copies GFT words into u16
        return o;
    }
};

class PyRSi16 {
public:
    size_t block_len, message_len, ecc_len, interleave;
    bool simd_x4, simd_x8, simd_x16;
    std::unique_ptr<GFi16> gf;
    ffrs_rsi16 *codec = nullptr;
    std::unique_ptr<PyNTT> ntt_obj;  // RSi16.ntt, made on first use

    PyRSi16(size_t block_len_, uint16_t ecc_len_, size_t interleave_, uint32_t primitive, std::optional<bool> x4,
            std::optional<bool> x8, std::optional<bool> x16)
        : block_len(block_len_), message_len(block_len_ - ecc_len_), ecc_len(ecc_len_), interleave(interleave_)
    {
        // pyrsi16.hpp:108-125: the SIMD defaults follow the host CPU
        simd_x4 = x4.value_or(__builtin_cpu_supports("sse2"));
        simd_x8 = x8.value_or(__builtin_cpu_supports("avx2"));
        simd_x16 = x16.value_or(__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq"));
        gf = std::make_unique<GFi16>(primitive);
        PY_ASSERT(ecc_len >= 2 && (ecc_len & (ecc_len - 1)) == 0, "ecc_len must be a power of 2: " + std::to_string(ecc_len));
        PY_ASSERT(block_len % ecc_len == 0, "block_len must be a multiple of ecc_len. block_len: " +
                                                std::to_string(block_len) + " ecc_len: " + std::to_string(ecc_len));
        PY_ASSERT(block_len > ecc_len, "block_len must exceed ecc_len");
        PY_ASSERT(interleave >= 1 && interleave <= 0xffff, "interleave must be in [1, 65535]");
        int r = ffrs_rsi16_create(default_ctx(), block_len, ecc_len, primitive, &codec);
        check(r, "RSi16");
    }
    ~PyRSi16()
    {
        if (codec)
            ffrs_rsi16_destroy(codec);
    }

    bool vector_lanes() const { return simd_x4 || simd_x8 || simd_x16; }
    // the lanes of the SIMD instantiation the reference dispatches to (pyrsi16.hpp:524-534)
    unsigned simd_width() const { return simd_x16 ? 16 : simd_x8 ? 8 : simd_x4 ? 4 : 1; }

    py::bytearray encode(py::buffer buffer)
    {
        Buf<const uint16_t> buf(buffer);
        if (buf.size == 0)
            return py::bytearray();
        size_t im = message_len * interleave;
        PY_ASSERT(buf.size % im == 0, std::to_string(buf.size));
        size_t count = buf.size / im;
        py::bytearray out;
        PyByteArray_Resize(out.ptr(), (Py_ssize_t)(count * ecc_len * interleave * 2));
        uint16_t *o = reinterpret_cast<uint16_t *>(PyByteArray_AsString(out.ptr()));
        int r;
        bool vec = vector_lanes();
        {
            py::gil_scoped_release nogil;
            r = interleave == 1 ? ffrs_rsi16_encode(codec, buf.data, count, o, vec ? 1 : 0)
                                : ffrs_rsi16_encode_interleaved(codec, buf.data, count, interleave, o, vec ? 1 : 0);
        }
        check(r, "RSi16.encode");
        return out;
    }

    RepairStatus status_of(uint8_t st)
    {
        if (st == FFRS_RSI16_RAISES)
            throw std::runtime_error("Assertion failed: g_len > 0 (the reference raises on this input)");
        return (RepairStatus)st;
    }

    RepairStatus repair(py::buffer message, py::buffer ecc, std::optional<std::vector<size_t>> const &error_pos)
    {
        Buf<uint16_t> m(message), e(ecc);
        if (error_pos && error_pos->empty())
            return RepairStatus::NoErrors;
        PY_ASSERT(m.size == message_len * interleave, std::to_string(m.size));
        PY_ASSERT(e.size == ecc_len * interleave, std::to_string(e.size));
        std::vector<uint32_t> pos;
        if (error_pos) {
            PY_ASSERT(error_pos->size() <= ecc_len, "too many error positions");
            pos.resize(error_pos->size());
            for (size_t i = 0; i < pos.size(); ++i) {
                if ((*error_pos)[i] >= block_len)
                    throw py::value_error("error position " + std::to_string((*error_pos)[i]) + " >= block_len");
                pos[i] = (uint32_t)(*error_pos)[i];
            }
        }
        if (interleave != 1 && vector_lanes() && !error_pos) {
            // the columns in coupled groups of the host's SIMD width (pyrsi16.hpp:389-453,
            // _simd_dispatch :524-534); with error positions the SIMD and scalar words agree
            uint8_t st = 0;
            int r;
            {
                py::gil_scoped_release nogil;
                r = ffrs_rsi16_repair_interleaved_simd(codec, m.data, e.data, interleave, simd_width(), &st);
            }
            check(r, "RSi16.repair");
            return status_of(st);
        }
        if (interleave != 1)
            return repair_columns(m.data, e.data, error_pos ? &pos : nullptr);
        uint8_t st = 0;
        int r;
        if (error_pos) {
            py::gil_scoped_release nogil;
            r = ffrs_rsi16_repair_known(codec, m.data, e.data, pos.data(), pos.size(), &st);
        } else {
            py::gil_scoped_release nogil;
            r = ffrs_rsi16_repair(codec, m.data, e.data, 1, &st);
        }
        check(r, "RSi16.repair");
        return status_of(st);
    }

    // Interleaved repair with the scalar words (pyrsi16.hpp:391-453 with SIMD_W = 1), and with
    // error positions in either words (there the vector instantiations give each column the
    // result of its own repair: the locator comes from the positions, the same in every lane,
    // and a column without errors gets zero corrections): column c
    // is the codeword of symbols i * interleave + c. The reference repairs the columns one
    // after another and writes each back after its repair; the call returns the largest
    // status, and a column whose repair raises ends it with the earlier columns written back.
    // Here all columns are repaired in one device call (known positions: one call per column)
    // and written back up to the first one that raises.
    RepairStatus repair_columns(uint16_t *msg, uint16_t *ecc, const std::vector<uint32_t> *pos)
    {
        const size_t il = interleave, k = message_len, E = ecc_len;
        std::vector<uint16_t> cm(il * k), ce(il * E);
        std::vector<uint8_t> st(il, 0);
        for (size_t i = 0; i < k; ++i)
            for (size_t c = 0; c < il; ++c)
                cm[c * k + i] = msg[i * il + c];
        for (size_t i = 0; i < E; ++i)
            for (size_t c = 0; c < il; ++c)
                ce[c * E + i] = ecc[i * il + c];
        int r = FFRS_OK;
        {
            py::gil_scoped_release nogil;
            if (pos) {
                for (size_t c = 0; c < il && r == FFRS_OK; ++c) {
                    r = ffrs_rsi16_repair_known(codec, &cm[c * k], &ce[c * E], pos->data(), pos->size(), &st[c]);
                    if (st[c] == FFRS_RSI16_RAISES)
                        break;
                }
            } else {
                r = ffrs_rsi16_repair(codec, cm.data(), ce.data(), il, st.data());
            }
        }
        check(r, "RSi16.repair");
        uint8_t res = FFRS_RSI16_NO_ERRORS;
        size_t done = 0;
        for (; done < il && st[done] != FFRS_RSI16_RAISES; ++done)
            res = std::max(res, st[done]);
        for (size_t i = 0; i < k; ++i)
            for (size_t c = 0; c < done; ++c)
                msg[i * il + c] = cm[c * k + i];
        for (size_t i = 0; i < E; ++i)
            for (size_t c = 0; c < done; ++c)
                ecc[i * il + c] = ce[c * E + i];
        return status_of(done < il ? (uint8_t)FFRS_RSI16_RAISES : res);
    }

    // Extension: repair `count` codewords (message / ecc buffers back to back) in one pass.
    // Returns one status byte per codeword: a RepairStatus value, or 255 where the
    // reference's repair raises RuntimeError on that codeword.
    py::bytearray repair_batch(py::buffer message, py::buffer ecc)
    {
        Buf<uint16_t> m(message), e(ecc);
        PY_ASSERT(interleave == 1, "repair_batch needs interleave 1");
        size_t count = m.size / message_len;
        PY_ASSERT(m.size == count * message_len && e.size == count * ecc_len, "batch sizes");
        py::bytearray out;
        PyByteArray_Resize(out.ptr(), (Py_ssize_t)count);
        uint8_t *st = reinterpret_cast<uint8_t *>(PyByteArray_AsString(out.ptr()));
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rsi16_repair(codec, m.data, e.data, count, st);
        }
        check(r, "RSi16.repair_batch");
        return out;
    }

    std::vector<uint32_t> synd(py::buffer message, py::buffer ecc)
    {
        Buf<const uint16_t> m(message), e(ecc);
        PY_ASSERT(m.size % message_len == 0, std::to_string(m.size));
        PY_ASSERT(e.size % ecc_len == 0, std::to_string(e.size));
        size_t count = m.size / message_len;
        PY_ASSERT(e.size / ecc_len == count, "count");
        std::vector<uint32_t> out(count * ecc_len);
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_rsi16_synds(codec, m.data, e.data, count, out.data());
        }
        check(r, "RSi16._synd");
        return out;
    }

    py::tuple sugiyama(std::vector<uint32_t> synd)
    {
        PY_ASSERT(synd.size() >= ecc_len, "synd must hold ecc_len values");
        std::vector<uint32_t> loc(ecc_len), ev(ecc_len);
        uint32_t len = 0;
        check(ffrs_rsi16_sugiyama(codec, synd.data(), 1, loc.data(), ev.data(), &len), "RSi16._sugiyama");
        if (len == 0xFFFFFFFFu)
            throw std::runtime_error("Assertion failed: g_len > 0 (the reference raises on this input)");
        return py::make_tuple(loc, ev);
    }

    // py_sugiyama<W> (pyrsi16.hpp:653-674): W syndrome vectors of ecc_len values, one after
    // another; the coupled vector Sugiyama of W lanes
    py::tuple sugiyama_w(std::vector<uint32_t> synd, unsigned W)
    {
        PY_ASSERT(synd.size() >= (size_t)W * ecc_len, "synd must hold " + std::to_string(W * ecc_len) + " values");
        std::vector<uint32_t> loc(W * ecc_len), ev(W * ecc_len), len(W);
        check(ffrs_rsi16_ntt_op(codec, 7, W, synd.data(), nullptr, 0, ev.data(), loc.data(), len.data()),
              ("RSi16._sugiyama" + std::to_string(W)).c_str());
        for (uint32_t l : len)
            if (l == 0xFFFFFFFFu)
                throw std::runtime_error("Assertion failed: g_len > 0 (the reference raises on this input)");
        return py::make_tuple(loc, ev);
    }

    std::vector<size_t> roots(std::vector<uint32_t> poly)
    {
        PY_ASSERT(poly.size() == ecc_len, std::to_string(poly.size()));
        std::vector<uint8_t> hit(block_len);
        check(ffrs_rsi16_roots(codec, poly.data(), hit.data()), "RSi16._roots");
        std::vector<size_t> out;
        for (size_t i = 0; i < block_len; ++i)
            if (hit[i])
                out.push_back(i);
        return out;
    }

    size_t message_offset(size_t row, size_t col)
    {
        PY_ASSERT(row < message_len, "row");
        PY_ASSERT(col < interleave, "col");
        return row * interleave + col;
    }

    size_t ecc_offset(size_t row, size_t col)
    {
        PY_ASSERT(row < ecc_len, "row");
        PY_ASSERT(col < interleave, "col");
        return row * interleave + col;
    }
};

// ---------------------------------------------------------------- CIRC16 (pycirc16.hpp:33-431)

// The reference's CIRC16: inner RSi16(inner_block_len, inner_ecc_len) and outer
// RSi16(outer_block_len, outer_ecc_len, inner_message_len * interleave) with the same primitive
// and SIMD flags (pycirc16.hpp:45-64); the cross-interleaved encode and repair run on the GPU
// through ffrs_circ16_*.
class PyCIRC16 {
public:
    py::object rsi_obj, rso_obj;
    PyRSi16 *rsi = nullptr, *rso = nullptr;
    size_t interleave;
    ffrs_circ16 *c = nullptr;

    PyCIRC16(size_t ibl, size_t iecc, size_t obl, size_t oecc, size_t interleave_, uint32_t primitive,
             std::optional<bool> x4, std::optional<bool> x8, std::optional<bool> x16)
        : interleave(interleave_)
    {
        PY_ASSERT(iecc <= 0xffff && oecc <= 0xffff, "ecc_len must fit 16 bits");
        auto ri = std::make_unique<PyRSi16>(ibl, (uint16_t)iecc, 1, primitive, x4, x8, x16);
        auto ro = std::make_unique<PyRSi16>(obl, (uint16_t)oecc, ri->message_len * interleave, primitive, x4, x8, x16);
        rsi = ri.get();
        rso = ro.get();
        rsi_obj = py::cast(std::move(ri));
        rso_obj = py::cast(std::move(ro));
        check(ffrs_circ16_create(rsi->codec, rso->codec, interleave, rsi->simd_width(), &c), "CIRC16");
    }
    ~PyCIRC16()
    {
        if (c)
            ffrs_circ16_destroy(c);
    }

    size_t block_len() const { return rsi->block_len * rso->block_len * interleave; }
    size_t message_len() const { return rso->message_len * rso->interleave; }
    size_t rsi_interleaved_ecc_len() const { return rsi->ecc_len * rso->message_len * interleave; }
    size_t rsio_ecc_len() const { return rsi->ecc_len * rso->ecc_len * interleave; }
    size_t ecc_len() const { return rso->ecc_len * rso->interleave + rsi_interleaved_ecc_len() + rsio_ecc_len(); }

    py::bytearray encode(py::buffer buffer)
    {
        Buf<const uint16_t> buf(buffer);
        const size_t ml = message_len();
        PY_ASSERT(buf.size % ml == 0, std::to_string(buf.size));
        const size_t blocks = buf.size / ml;
        py::bytearray out;
        PyByteArray_Resize(out.ptr(), (Py_ssize_t)(blocks * ecc_len() * 2));
        if (blocks) {
            uint16_t *o = reinterpret_cast<uint16_t *>(PyByteArray_AsString(out.ptr()));
            int r;
            {
                py::gil_scoped_release nogil;
                r = ffrs_circ16_encode(c, buf.data, blocks, o);
            }
            check(r, "CIRC16.encode");
        }
        return out;
    }

    void check_block(const Buf<uint16_t> &m, const Buf<uint16_t> &e)
    {
        PY_ASSERT(m.size == message_len(), std::to_string(m.size) + " != " + std::to_string(message_len()));
        PY_ASSERT(e.size == ecc_len(), std::to_string(e.size));
    }

    bool repair(py::buffer message, py::buffer ecc)
    {
        Buf<uint16_t> m(message), e(ecc);
        check_block(m, e);
        int raised = 0, r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_circ16_repair(c, m.data, e.data, &raised);
        }
        check(r, "CIRC16.repair");
        if (raised)
            throw std::runtime_error("Assertion failed: g_len > 0 (the reference raises on this input)");
        return false;  // py_repair returns false (pycirc16.hpp:402)
    }

    std::vector<size_t> find_outer_error_locations(py::buffer message, py::buffer ecc, size_t il)
    {
        Buf<uint16_t> m(message), e(ecc);
        check_block(m, e);
        PY_ASSERT(il < interleave, "interleave < this->interleave");
        std::vector<uint32_t> rows(rso->block_len);
        size_t n = 0;
        int r;
        {
            py::gil_scoped_release nogil;
            r = ffrs_circ16_find_outer_error_locations(c, m.data, e.data, il, rows.data(), &n);
        }
        check(r, "CIRC16._find_outer_error_locations");
        return std::vector<size_t>(rows.begin(), rows.begin() + n);
    }

    // pycirc16.hpp:423-445
    size_t message_offset(size_t il, size_t row, size_t col) const
    {
        PY_ASSERT(il < interleave, "interleave < this->interleave");
        PY_ASSERT(row < rso->message_len, "row < rso.message_len");
        PY_ASSERT(col < rsi->message_len, "col < rsi.message_len");
        return rso->interleave * row + il * rsi->message_len + col;
    }
    size_t rso_ecc_offset(size_t il, size_t row, size_t col) const
    {
        PY_ASSERT(il < interleave, "interleave < this->interleave");
        PY_ASSERT(row < rso->ecc_len, "row < rso.ecc_len");
        PY_ASSERT(col < rsi->message_len, "col < rsi.message_len");
        return rso->interleave * row + il * rsi->message_len + col;
    }
    size_t rsi_ecc_offset(size_t il, size_t row, size_t col) const
    {
        PY_ASSERT(il < interleave, "interleave < this->interleave");
        PY_ASSERT(row < rso->block_len, "row < rso.block_len");
        PY_ASSERT(col < rsi->ecc_len, "col < rsi.ecc_len");
        return rso->ecc_len * rso->interleave + (row * interleave + il) * rsi->ecc_len + col;
    }
};

// ---------------------------------------------------------------- create_buffer (pyffrs.cpp:41-98)

py::memoryview create_buffer(size_t requested)
{
    constexpr size_t huge = 2 * 1024 * 1024;
    size_t alloc = ((requested + huge - 1) / huge) * huge;
    if (alloc == 0)
        alloc = huge;
    void *ptr = mmap(nullptr, alloc, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (ptr == MAP_FAILED)
        throw std::runtime_error(std::string("mmap failed: ") + std::strerror(errno));
    madvise(ptr, alloc, MADV_HUGEPAGE);
    memset(ptr, 0, alloc);
    auto *mapping = new std::pair<void *, size_t>{ptr, alloc};
    py::capsule cap(mapping, "hugepage_array", [](PyObject *c) {
        auto *mp = reinterpret_cast<std::pair<void *, size_t> *>(PyCapsule_GetPointer(c, "hugepage_array"));
        if (mp) {
            munmap(mp->first, mp->second);
            delete mp;
        }
    });
    Py_buffer view;
    if (PyBuffer_FillInfo(&view, cap.ptr(), ptr, (Py_ssize_t)requested, 0, PyBUF_CONTIG) < 0)
        throw py::error_already_set();
    return py::reinterpret_steal<py::memoryview>(PyMemoryView_FromBuffer(&view));
}

py::object g_logger = py::none();

}  // namespace

PYBIND11_MODULE(libffrs, m)
{
    using namespace pybind11::literals;
    m.attr("__version__") = "0.1.0";
    m.attr("compiler_info") = "gcc" __VERSION__;
    m.attr("backend") = "hip-gfx950";
    m.doc() = "FFRS - Fairly Fast & Flexible Reed-Solomon coding, MI355X (gfx950) HIP backend";

    m.def("create_buffer", &create_buffer, "size"_a,
          "Create a memory buffer of the specified size, backed by hugepages if possible.");
    m.def("set_logger", [](py::object logger) { g_logger = std::move(logger); }, "logger"_a,
          "Logger object to be used by C++ library or ``None`` to disable logging.");
    m.def("abi_version", []() { return ffrs_abi_version(); });
    m.def("device_info", []() {
        char name[256] = {0};
        int cus = 0;
        size_t fr = 0, tot = 0;
        check(ffrs_ctx_device_info(default_ctx(), name, sizeof name, &cus, &fr, &tot), "device_info");
        return py::make_tuple(std::string(name), cus, fr, tot);
    });

    py::class_<GF256>(m, "GF256", "Finite-field operations optimized for GF(2^8)")
        .def(py::init<uint8_t, uint16_t>(), "primitive"_a = 2, "poly1"_a = 0x11d)
        .def_property_readonly("prime", [](GF256 &) { return 2; })
        .def_property_readonly("power", [](GF256 &) { return 8; })
        .def_property_readonly("primitive", [](GF256 &s) { return s.primitive; })
        .def_property_readonly("poly1", [](GF256 &s) { return s.poly1; })
        .def_property_readonly("field_elements", [](GF256 &) { return 256; })
        .def("mul", &GF256::mul, "lhs"_a, "rhs"_a)
        .def("add", [](GF256 &, uint8_t a, uint8_t b) { return (uint8_t)(a ^ b); }, "lhs"_a, "rhs"_a)
        .def("sub", [](GF256 &, uint8_t a, uint8_t b) { return (uint8_t)(a ^ b); }, "lhs"_a, "rhs"_a)
        .def("inv", &GF256::inv, "value"_a)
        .def("div", &GF256::div, "num"_a, "den"_a)
        .def("exp", [](GF256 &s, uint8_t a) { return s.exp_t[a]; }, "value"_a)
        .def("log", [](GF256 &s, uint8_t a) { return s.log_t[a]; }, "value"_a)
        .def("pow", &GF256::pow, "base"_a, "exponent"_a);

    py::class_<GFi16>(m, "GFi16", "Finite-field operations over GF(65537)")
        .def(py::init<uint32_t>(), "primitive"_a = 3)
        .def_property_readonly("prime", [](GFi16 &) { return 65537u; })
        .def_property_readonly("power", [](GFi16 &) { return 1u; })
        .def_property_readonly("primitive", [](GFi16 &s) { return s.primitive; })
        .def_property_readonly("field_elements", [](GFi16 &) { return 65537u; })
        .def("mul", &GFi16::mul, "lhs"_a, "rhs"_a)
        .def("add", &GFi16::add, "lhs"_a, "rhs"_a)
        .def("sub", &GFi16::sub, "lhs"_a, "rhs"_a)
        .def("inv", &GFi16::inv, "value"_a)
        .def("div", &GFi16::div, "num"_a, "den"_a)
        .def("exp", [](GFi16 &s, uint32_t a) { return s.exp_t.at(a); }, "value"_a)
        .def("log", [](GFi16 &s, uint32_t a) { return s.log_t.at(a); }, "value"_a)
        .def("pow", &GFi16::pow, "base"_a, "exponent"_a);

    {  // pyntt.hpp:28-84
        auto cls = py::class_<PyNTT>(m, "NTT", "NTT transforms and polynomial operations of an RSi16 code")
            .def(py::init<const GFi16 &, size_t, size_t>(), "gf"_a, "block_len"_a, "ecc_len"_a)
            .def_property_readonly("ntt_len", [](PyNTT &s) { return s.ntt_len; })
            .def_property_readonly("ntt_size", [](PyNTT &s) { return s.ntt_len * 2; })
            .def_property_readonly("ntt4_size", [](PyNTT &s) { return s.ntt_len * 8; })
            .def_property_readonly("ntt8_size", [](PyNTT &s) { return s.ntt_len * 16; })
            .def_property_readonly("ntt16_size", [](PyNTT &s) { return s.ntt_len * 32; })
            .def_property_readonly("block_len", [](PyNTT &s) { return s.block_len; })
            .def_property_readonly("ecc_len", [](PyNTT &s) { return s.ecc_len; })
            .def_property_readonly("ecc_size", [](PyNTT &s) { return s.ecc_len * 2; })
            .def("rbo", &PyNTT::rbo, "Reverse Bit Order for log2(ntt_len) bits", "i"_a);
        static const char *const names[] = {"ntt", "intt", "nttr", "inttr"};
        static const char *const docs[] = {"RBO input, normal order output", "Normal order input, RBO output",
                                           "Normal order input, RBO output", "RBO input, Normal order output"};
        static const char *const sfx[] = {"", "4", "8", "16"};
        static const unsigned widths[] = {1, 4, 8, 16};
        for (int w = 0; w < 4; ++w) {
            const unsigned W = widths[w];
            for (int op = 0; op < 4; ++op)
                cls.def((std::string(names[op]) + sfx[w]).c_str(),
                        [op, W](PyNTT &s, py::buffer in) { return s.run(op, W, {in}); }, docs[op], "input"_a);
            cls.def((std::string("poly_mul") + sfx[w]).c_str(),
                    [W](PyNTT &s, py::buffer a, py::buffer b) { return s.run(4, W, {a, b}); }, "Element wise multiplication",
                    "a"_a, "b"_a);
            cls.def((std::string("poly_div") + sfx[w]).c_str(),
                    [W](PyNTT &s, py::buffer a, py::buffer b) { return s.run(5, W, {a, b}); }, "Polynomial division",
                    "a"_a, "b"_a);
            cls.def((std::string("poly_inv") + sfx[w]).c_str(),
                    [W](PyNTT &s, py::buffer a, size_t n) {
                        PY_ASSERT(n <= s.ecc_len, "n must not exceed ecc_len: " + std::to_string(n));
                        return s.run(6, W, {a}, (uint32_t)n);
                    },
                    "Polynomial inverse", "a"_a, "n"_a);
        }
    }

    py::class_<PyRS256>(m, "RS256", "Reed-Solomon coding over GF(2^8)")
        .def(py::init<std::optional<uint8_t>, std::optional<uint8_t>, std::optional<uint8_t>, uint8_t, uint16_t>(),
             "block_len"_a = py::none(), "message_len"_a = py::none(), "ecc_len"_a = py::none(), "primitive"_a = 2,
             "polynomial"_a = 0x11d)
        .def_property_readonly("ecc_len", [](PyRS256 &s) { return s.ecc_len; })
        .def_property_readonly("ecc_size", [](PyRS256 &s) { return s.ecc_len; })
        .def_property("block_len", [](PyRS256 &s) { return s.block_len; }, &PyRS256::set_block_len)
        .def_property("block_size", [](PyRS256 &s) { return s.block_len; }, &PyRS256::set_block_len)
        .def_property_readonly("message_len", [](PyRS256 &s) { return s.block_len - s.ecc_len; })
        .def_property_readonly("message_size", [](PyRS256 &s) { return s.block_len - s.ecc_len; })
        .def_property_readonly("gf", [](PyRS256 &s) -> GF256 & { return *s.gf; }, py::return_value_policy::reference_internal)
        .def_property_readonly("generator", [](PyRS256 &s) {
            return py::bytes(reinterpret_cast<const char *>(s.generator.data()), s.generator.size()); })
        .def_property_readonly("generator_roots", [](PyRS256 &s) {
            return py::bytes(reinterpret_cast<const char *>(s.roots.data()), s.roots.size()); })
        .def("encode", &PyRS256::encode, "Encode message, return ecc", "buffer"_a)
        .def("repair", &PyRS256::repair, "Repair message + ecc", "buffer"_a, "ecc"_a)
        .def("repair_batch", &PyRS256::repair_batch,
             "Repair count back-to-back codewords (message and ecc buffers); returns one status byte each",
             "buffer"_a, "ecc"_a)
        .def("_synds", &PyRS256::synds, "Compute syndromes", "buffer"_a)
        .def("_set_repair_mode", [](PyRS256 &s, int mode) { check(ffrs_rs256_set_repair_mode(s.codec, mode), "RS256"); },
             "0: automatic, 1: pipeline (ecc_len <= 32), 2: one codeword per lane", "mode"_a);

    py::enum_<RepairStatus>(m, "RepairStatus", "Repair operation status")
        .value("NoErrors", RepairStatus::NoErrors)
        .value("NoErrorsZero", RepairStatus::NoErrorsZero)
        .value("RepairOk", RepairStatus::RepairOk)
        .value("RepairFail", RepairStatus::RepairFail)
        .value("ErrorLocationFail", RepairStatus::ErrorLocationFail);

    py::class_<PyRSi16>(m, "RSi16", "Reed-Solomon coding over GF(65537)")
        .def(py::init<size_t, uint16_t, size_t, uint32_t, std::optional<bool>, std::optional<bool>,
                      std::optional<bool>>(),
             "block_len"_a, "ecc_len"_a, "interleave"_a = 1, py::kw_only(), "primitive"_a = 3,
             "simd_x4"_a = py::none(), "simd_x8"_a = py::none(), "simd_x16"_a = py::none())
        .def_property_readonly("block_len", [](PyRSi16 &s) { return s.block_len * s.interleave; })
        .def_property_readonly("block_size", [](PyRSi16 &s) { return s.block_len * s.interleave * 2; })
        .def_property_readonly("message_len", [](PyRSi16 &s) { return s.message_len * s.interleave; })
        .def_property_readonly("message_size", [](PyRSi16 &s) { return s.message_len * s.interleave * 2; })
        .def_property_readonly("ecc_len", [](PyRSi16 &s) { return s.ecc_len * s.interleave; })
        .def_property_readonly("ecc_size", [](PyRSi16 &s) { return s.ecc_len * s.interleave * 2; })
        .def_property_readonly("rs_block_len", [](PyRSi16 &s) { return s.block_len; })
        .def_property_readonly("rs_block_size", [](PyRSi16 &s) { return s.block_len * 2; })
        .def_property_readonly("rs_message_len", [](PyRSi16 &s) { return s.message_len; })
        .def_property_readonly("rs_message_size", [](PyRSi16 &s) { return s.message_len * 2; })
        .def_property_readonly("rs_ecc_len", [](PyRSi16 &s) { return s.ecc_len; })
        .def_property_readonly("rs_ecc_size", [](PyRSi16 &s) { return s.ecc_len * 2; })
        .def_property_readonly("ntt_len", [](PyRSi16 &s) { return ffrs_rsi16_ntt_len(s.codec); })
        .def_property_readonly("ntt_size", [](PyRSi16 &s) { return ffrs_rsi16_ntt_len(s.codec) * 2; })
        .def_property_readonly("root", [](PyRSi16 &s) { return ffrs_rsi16_root(s.codec); })
        .def_property_readonly("gf", [](PyRSi16 &s) -> GFi16 & { return *s.gf; }, py::return_value_policy::reference_internal)
        .def_property_readonly("interleave", [](PyRSi16 &s) { return s.interleave; })
        .def_property_readonly("simd_x4", [](PyRSi16 &s) { return s.simd_x4; })
        .def_property_readonly("simd_x8", [](PyRSi16 &s) { return s.simd_x8; })
        .def_property_readonly("simd_x16", [](PyRSi16 &s) { return s.simd_x16; })
        .def("encode", &PyRSi16::encode, "Systematic encode", "buffer"_a)
        .def("repair", &PyRSi16::repair, "Repair a block with the given error locations", "message"_a, "ecc"_a,
             "error_pos"_a = py::none())
        .def("repair_batch", &PyRSi16::repair_batch,
             "Repair count back-to-back codewords (message and ecc buffers); returns one status byte each "
             "(RepairStatus value, 255 where the reference raises)",
             "message"_a, "ecc"_a)
        .def("_synd", &PyRSi16::synd, "Calculate syndromes for the given message and ecc buffers", "message"_a,
             "ecc"_a)
        .def("_sugiyama", &PyRSi16::sugiyama, "Compute error locator and evaluator polynomials", "synd"_a)
        .def("_sugiyama4", [](PyRSi16 &s, std::vector<uint32_t> v) { return s.sugiyama_w(std::move(v), 4); },
             "Compute error locator and evaluator polynomials", "synd"_a)
        .def("_sugiyama8", [](PyRSi16 &s, std::vector<uint32_t> v) { return s.sugiyama_w(std::move(v), 8); },
             "Compute error locator and evaluator polynomials", "synd"_a)
        .def("_sugiyama16", [](PyRSi16 &s, std::vector<uint32_t> v) { return s.sugiyama_w(std::move(v), 16); },
             "Compute error locator and evaluator polynomials", "synd"_a)
        .def_property_readonly("ntt", [](PyRSi16 &s) {  // one per codec, as the reference's member
            if (!s.ntt_obj)
                s.ntt_obj = std::make_unique<PyNTT>(*s.gf, s.block_len, s.ecc_len);
            return s.ntt_obj.get();
        }, py::return_value_policy::reference_internal, ":class:`NTT` instance")
        .def("message_offset", &PyRSi16::message_offset, "row"_a, "col"_a)
        .def("ecc_offset", &PyRSi16::ecc_offset, "row"_a, "col"_a)
        .def("_roots", &PyRSi16::roots, "Find locator polynomial roots", "synd"_a)
        .def("_set_groups", [](PyRSi16 &s, unsigned ge, unsigned gr) { check(ffrs_rsi16_set_groups(s.codec, ge, gr), "RSi16"); },
             "Lanes per codeword of the encode and repair kernels (0 keeps)", "g_encode"_a, "g_repair"_a);

    py::class_<PyCIRC16>(m, "CIRC16", "Cross-interleaved Reed-Solomon coding over GF(65537)")
        .def(py::init<size_t, size_t, size_t, size_t, size_t, uint32_t, std::optional<bool>, std::optional<bool>,
                      std::optional<bool>>(),
             "inner_block_len"_a, "inner_ecc_len"_a, "outer_block_len"_a, "outer_ecc_len"_a, "interleave"_a = 1,
             py::kw_only(), "primitive"_a = 3, "simd_x4"_a = py::none(), "simd_x8"_a = py::none(),
             "simd_x16"_a = py::none())
        .def_property_readonly("rsi", [](PyCIRC16 &s) { return s.rsi_obj; }, "Inner codec (RSi16)")
        .def_property_readonly("rso", [](PyCIRC16 &s) { return s.rso_obj; }, "Outer codec (RSi16)")
        .def_property_readonly("block_len", [](PyCIRC16 &s) { return s.block_len(); })
        .def_property_readonly("block_size", [](PyCIRC16 &s) { return s.block_len() * 2; })
        .def_property_readonly("message_len", [](PyCIRC16 &s) { return s.message_len(); })
        .def_property_readonly("message_size", [](PyCIRC16 &s) { return s.message_len() * 2; })
        .def_property_readonly("ecc_len", [](PyCIRC16 &s) { return s.ecc_len(); })
        .def_property_readonly("ecc_size", [](PyCIRC16 &s) { return s.ecc_len() * 2; })
        .def_property_readonly("inner_block_len", [](PyCIRC16 &s) { return s.rsi->block_len; })
        .def_property_readonly("inner_block_size", [](PyCIRC16 &s) { return s.rsi->block_len * 2; })
        .def_property_readonly("inner_message_len", [](PyCIRC16 &s) { return s.rsi->message_len; })
        .def_property_readonly("inner_message_size", [](PyCIRC16 &s) { return s.rsi->message_len * 2; })
        .def_property_readonly("inner_ecc_len", [](PyCIRC16 &s) { return s.rsi->ecc_len; })
        .def_property_readonly("inner_ecc_size", [](PyCIRC16 &s) { return s.rsi->ecc_len * 2; })
        .def_property_readonly("outer_block_len", [](PyCIRC16 &s) { return s.rso->block_len; })
        .def_property_readonly("outer_message_len", [](PyCIRC16 &s) { return s.rso->message_len; })
        .def_property_readonly("outer_ecc_len", [](PyCIRC16 &s) { return s.rso->ecc_len; })
        .def_property_readonly("outer_interleave", [](PyCIRC16 &s) { return s.rso->interleave; })
        .def_property_readonly("interleave", [](PyCIRC16 &s) { return s.interleave; })
        .def_property_readonly("primitive", [](PyCIRC16 &s) { return s.rsi->gf->primitive; })
        .def("__sizeof_cpp__", [](PyCIRC16 &) { return sizeof(PyCIRC16); })
        .def_property_readonly("simd_x4", [](PyCIRC16 &s) { return s.rsi->simd_x4; })
        .def_property_readonly("simd_x8", [](PyCIRC16 &s) { return s.rsi->simd_x8; })
        .def_property_readonly("simd_x16", [](PyCIRC16 &s) { return s.rsi->simd_x16; })
        .def("encode", &PyCIRC16::encode, "Encode data", "buffer"_a)
        .def("repair", &PyCIRC16::repair, "Repair data", "message"_a, "ecc"_a)
        .def("_find_outer_error_locations", &PyCIRC16::find_outer_error_locations,
             "Find outer codec error locations for a given interleaved block", "message"_a, "ecc"_a, "interleave"_a)
        .def("message_offset", &PyCIRC16::message_offset, "interleave"_a, "row"_a, "col"_a)
        .def("rso_ecc_offset", &PyCIRC16::rso_ecc_offset, "interleave"_a, "row"_a, "col"_a)
        .def("rsi_ecc_offset", &PyCIRC16::rsi_ecc_offset, "interleave"_a, "row"_a, "col"_a);
}
