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
