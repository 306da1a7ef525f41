"""CPU-side tests of the boundary: the C-ABI library exports every entry point that
include/ffrs_gpu.h declares, argument checks that need no device, the scalar field
helpers of the pybind11 module against the oracle's tables, and loud failure of the
product path where no GPU is present (there is no CPU fallback).

Scalar GF tests follow tests/test_lib_gf256.py:30-110 and tests/test_lib_gfi16.py:38-104
of the reference (identities, inverses, commutativity, exp/log, pow, overflow wrap).
"""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ffrs_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(ffrs_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from libffrs_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libffrs_gpu.so not built (python libffrs_amd/build.py)")
    return capi.lib()


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ffrs_rs256_encode", "ffrs_rs256_repair", "ffrs_rsi16_encode", "ffrs_rsi16_repair",
              "ffrs_rsi16_repair_known", "ffrs_ctx_create", "ffrs_last_error"):
        assert s in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_signatures_match_header():
    from libffrs_amd import capi
    assert sorted(n for n, _, _ in capi.SIGNATURES) == declared_symbols()


def test_abi_version(lib):
    text = open(HEADER).read()
    want = int(re.search(r"#define\s+FFRS_ABI_VERSION\s+(\d+)", text).group(1))
    assert lib.ffrs_abi_version() == want


def test_null_arguments_are_rejected(lib):
    from libffrs_amd import capi
    h = ctypes.c_void_p()
    assert lib.ffrs_rs256_create(None, 8, 2, 0x11D, ctypes.byref(h)) == -1
    assert b"NULL" in lib.ffrs_last_error()
    assert lib.ffrs_rsi16_create(None, 256, 32, 3, ctypes.byref(h)) == -1
    assert lib.ffrs_rs256_encode(None, 1, None, 0, None, 0) == -1
    assert lib.ffrs_rsi16_repair(None, None, None, 0, None) == -1
    assert lib.ffrs_ctx_synchronize(None) == -1
    assert capi.FFRS_OK == 0


def _no_gpu():
    # decided without initialising any HIP runtime in the test process
    return not os.path.exists("/dev/kfd")


@pytest.mark.skipif(not _no_gpu(), reason="checks the behaviour without a GPU")
def test_context_without_device_fails_loudly(lib):
    h = ctypes.c_void_p()
    r = lib.ffrs_ctx_create(0, ctypes.byref(h))
    assert r == -3  # FFRS_ERR_DEVICE
    assert lib.ffrs_last_error()


@pytest.mark.skipif(not _no_gpu(), reason="checks the behaviour without a GPU")
def test_codec_without_device_raises():
    import libffrs_amd as L
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        L.RS256(ecc_len=8)
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        L.RSi16(256, 32)


# ------------------------------------------------------- parameter checks (no device)

@pytest.mark.parametrize("kw", [
    {},                                   # nothing given (pyrs256.hpp:93)
    {"block_len": 10, "ecc_len": 10},     # block_len <= ecc_len (pyrs256.hpp:75)
    {"block_len": 20, "message_len": 20},
    {"block_len": 20, "message_len": 5, "ecc_len": 4},  # inconsistent (pyrs256.hpp:90)
])
def test_rs256_parameter_errors(kw):
    import libffrs_amd as L
    with pytest.raises(ValueError):
        L.RS256(**kw)


def test_rs256_uint8_overflow_is_type_error():
    import libffrs_amd as L
    with pytest.raises(TypeError):
        L.RS256(ecc_len=256)


@pytest.mark.parametrize("n, ecc", [(256, 24), (256, 1), (100, 32), (32, 32)])
def test_rsi16_parameter_errors(n, ecc):
    import libffrs_amd as L
    with pytest.raises(RuntimeError, match="Assertion failed"):
        L.RSi16(n, ecc)


# ------------------------------------------------------- scalar fields

@pytest.fixture(scope="module")
def gf256():
    import libffrs_amd as L
    return L.GF256()


def test_gf256_tables_match_oracle(gf256):
    from oracle import oracle as O
    rc, exp_t, log_t = O.gf256_tables()
    assert rc == 0
    assert [gf256.exp(a) for a in range(256)] == list(exp_t)
    assert [gf256.log(a) for a in range(256)] == list(log_t)
    assert gf256.log(1) == 255 and gf256.log(0) == 0 and gf256.exp(255) == 1


def test_gf256_field_laws(gf256):
    for a in range(256):
        assert gf256.add(a, 0) == a and gf256.mul(a, 1) == a and gf256.div(a, 1) == a
    for a, b in itertools.product(range(256), range(1, 256)):
        assert gf256.div(gf256.mul(a, b), b) == a
        assert gf256.mul(a, b) == gf256.mul(b, a)
    for a in range(1, 256):
        assert gf256.mul(a, gf256.inv(a)) == 1
        assert gf256.inv(gf256.inv(a)) == a
        assert gf256.pow(2, gf256.log(a)) == a
    # zero conventions of galois.hpp:98-127: inv(0) = 1, div(a, 0) = a, div(0, b) = 0
    assert gf256.inv(0) == 1
    assert all(gf256.div(a, 0) == a for a in range(256))
    assert all(gf256.div(0, b) == 0 for b in range(256))


def test_gf256_mul_is_carryless_mod_0x11d(gf256):
    def slow(a, b):
        r = 0
        while b:
            if b & 1:
                r ^= a
            b >>= 1
            a <<= 1
            if a & 0x100:
                a ^= 0x11D
        return r
    rng = np.random.default_rng(5)
    for a, b in rng.integers(0, 256, (2000, 2)):
        assert gf256.mul(int(a), int(b)) == slow(int(a), int(b))


def test_gfi16_arithmetic():
    import libffrs_amd as L
    g = L.GFi16()
    p = 65537
    rng = np.random.default_rng(6)
    for a, b in rng.integers(0, p, (3000, 2)):
        a, b = int(a), int(b)
        assert g.add(a, b) == (a + b) % p
        assert g.sub(a, b) == (a - b) % p
        assert g.mul(a, b) == (a * b) % p
        if b:
            assert g.div(a, b) == a * pow(b, p - 2, p) % p
    # boundary cases of galois.hpp:258-395 (65536 * 65536 = 1 mod p)
    assert g.mul(65536, 65536) == 1
    assert g.add(65536, 65536) == 65535
    assert g.sub(0, 65536) == 1
    assert g.inv(0) == 1 and g.div(7, 0) == 7
    assert g.pow(3, 65536) == 1 and g.pow(3, 32768) == 65536
    assert g.exp(g.log(12345)) == 12345 and g.log(1) == 65536 and g.log(0) == 0


def test_gfi16_rejects_non_primitive():
    import libffrs_amd as L
    with pytest.raises(RuntimeError):
        L.GFi16(4)
