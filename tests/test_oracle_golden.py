"""The oracle against the golden fixtures written from the reference's own code."""
import pytest

import cases as C
import replay as R


@pytest.mark.parametrize("name", C.RS256_PLAIN_CASES)
def test_rs256_plain(oracle_backend, name):
    R.check_rs256_plain(oracle_backend, name)


@pytest.mark.parametrize("shape", C.RS256_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rs256_shapes(oracle_backend, shape):
    R.check_rs256_shape(oracle_backend, *shape)


def test_rs256_mixed(oracle_backend):
    R.check_rs256_mixed(oracle_backend)


@pytest.mark.parametrize("name", C.RSI16_PLAIN_CASES)
def test_rsi16_plain(oracle_backend, name):
    R.check_rsi16_plain(oracle_backend, name)


def test_rsi16_mixed(oracle_backend):
    R.check_rsi16_mixed(oracle_backend)


def test_rsi16_edges(oracle_backend):
    R.check_rsi16_edges(oracle_backend)


def test_rsi16_known(oracle_backend):
    R.check_rsi16_known(oracle_backend)


def test_rsi16_interleaved(oracle_backend):
    R.check_rsi16_interleaved(oracle_backend)


def test_rsi16_interleaved_repair(oracle_backend):
    """The reference's scalar-word interleaved repair equals its column-by-column repair."""
    R.check_rsi16_interleaved_repair(oracle_backend)


@pytest.mark.parametrize("shape", C.RS256_REF_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rs256_reference_test_shapes(oracle_backend, shape):
    R.check_rs256_reftest(oracle_backend, *shape)


@pytest.mark.parametrize("shape", C.RSI16_REF_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rsi16_reference_test_shapes(oracle_backend, shape):
    R.check_rsi16_reftest(oracle_backend, *shape)


@pytest.mark.parametrize("shape", C.RSI16_DIV_SHAPES, ids=lambda s: f"{s[0]}_{s[1]}")
def test_rsi16_simd_word_divergence(oracle_backend, shape):
    R.check_rsi16_divergence(oracle_backend, *shape)


def test_divergence_fixture_differs():
    """The SIMD-word and scalar-word parity of the reference differ on these fixtures (so the
    two encoders are different functions), and only on shapes whose mix holds 65536."""
    import numpy as np
    fx = R.fixture("rsi16_simd_div")
    for n, ecc in C.RSI16_DIV_SHAPES:
        t = f"{n}_{ecc}"
        assert C.rsi16_mix_minus_one(n, ecc)
        assert (fx["parity_" + t] != fx["parity_scalar_" + t]).any(1).sum() >= 4, t
        # the all-zero message: the scalar words give the zero parity
        assert not fx["parity_scalar_" + t][0].any() and fx["parity_" + t][0].any(), t
    for n, ecc in ((256, 32), (4096, 256), (1024, 128), (64, 8), (256, 16), (1024, 32)):
        assert not C.rsi16_mix_minus_one(n, ecc), (n, ecc)


@pytest.mark.parametrize("width", C.RSI16_ILSIMD_WIDTHS)
@pytest.mark.parametrize("case", C.RSI16_ILSIMD_CASES, ids=lambda c: f"{c[0]}_{c[1]}_{c[2]}")
def test_rsi16_interleaved_repair_simd_words(oracle_backend, case, width):
    """Coupled groups of 16, 8 or 4 columns, as the reference's vector instantiations repair them."""
    R.check_rsi16_ilsimd(oracle_backend, *case, width)
