"""The NTT class (pyntt.hpp:28-84) and RSi16._sugiyama4/8/16 (pyrsi16.hpp:653-674) on the GPU.

Every entry of the reference-written fixture tests/golden/rsi16_ntt.npz is replayed through
libffrs_amd.NTT / RSi16 and must match bit for bit, raises included (the fixture's meaning is
pinned on the CPU by tests/test_ntt_golden.py). The reference's own NTT tests
(tests/test_lib_gfi16_ntt.py) follow with fresh random inputs: ntt/intt round trip, poly_mul
against the schoolbook product, poly_div by zero raising, a higher-degree denominator giving
zero, poly_inv against its defining property with the reference's aliasing predicate, and the
4/8/16-vector variants agreeing with the single-vector ones where the reference's do.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases as C  # noqa: E402

pytestmark = pytest.mark.gpu
P = 65537


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(HERE, "golden", "rsi16_ntt.npz"))


def u16(b):
    return np.frombuffer(bytes(b), np.uint16)


def call(fn, *args):
    try:
        return u16(fn(*args)), 0
    except RuntimeError:
        return None, 1


@pytest.mark.parametrize("bl,ecc,prim", C.NTT_SHAPES)
def test_ntt_fixture_replay(fx, bl, ecc, prim):
    import libffrs_amd as L
    gf = L.GFi16(prim)
    ntt = L.NTT(gf, bl, ecc)
    E = ecc
    assert ntt.ecc_len == E and ntt.block_len == bl and ntt.ntt_len == 1 << (bl - 1).bit_length()
    print(f"ntt replay {bl} {ecc} {prim}", flush=True)
    for w in C.NTT_WIDTHS:
        sfx = "" if w == 1 else str(w)
        tag = f"{bl}_{ecc}_{prim}_w{w}"
        if f"x_{tag}" in fx:
            x = fx[f"x_{tag}"]
            for op in ("ntt", "intt", "nttr", "inttr"):
                got = u16(getattr(ntt, op + sfx)(x.tobytes()))
                assert (got == fx[f"{op}_{tag}"]).all(), (op, w)
        for t, (a, b) in enumerate(zip(fx[f"mul_a_{tag}"], fx[f"mul_b_{tag}"])):
            got = u16(getattr(ntt, "poly_mul" + sfx)(a.tobytes(), b.tobytes()))
            assert (got == fx[f"mul_out_{tag}"][t]).all(), ("poly_mul", w, t)
        for t, (a, b) in enumerate(zip(fx[f"div_a_{tag}"], fx[f"div_b_{tag}"])):
            got, r = call(getattr(ntt, "poly_div" + sfx), a.tobytes(), b.tobytes())
            assert r == fx[f"div_raises_{tag}"][t], ("poly_div raises", w, t)
            if not r:
                assert (got == fx[f"div_out_{tag}"][t]).all(), ("poly_div", w, t)
        for t, (a, n) in enumerate(zip(fx[f"inv_a_{tag}"], fx[f"inv_n_{tag}"])):
            got, r = call(getattr(ntt, "poly_inv" + sfx), a.tobytes(), int(n))
            assert r == fx[f"inv_raises_{tag}"][t], ("poly_inv raises", w, t, int(n))
            if not r:
                assert (got == fx[f"inv_out_{tag}"][t]).all(), ("poly_inv", w, t, int(n))
    del gf


@pytest.mark.parametrize("n,ecc", C.SUGIYAMA_SHAPES)
def test_sugiyama_fixture_replay(fx, n, ecc):
    import libffrs_amd as L
    rs = L.RSi16(n, ecc)
    print(f"sugiyama replay {n} {ecc}", flush=True)
    for w in C.NTT_WIDTHS:
        tag = f"{n}_{ecc}_w{w}"
        fn = rs._sugiyama if w == 1 else getattr(rs, f"_sugiyama{w}")
        for t, syn in enumerate(fx[f"sug_synd_{tag}"]):
            try:
                loc, ev = fn([int(v) for v in syn])
                r = 0
            except RuntimeError:
                r = 1
            assert r == fx[f"sug_raises_{tag}"][t], (w, t)
            if not r:
                assert (np.array(loc[:w * ecc], np.uint32) == fx[f"sug_loc_{tag}"][t]).all(), (w, t)
                assert (np.array(ev[:w * ecc], np.uint32) == fx[f"sug_ev_{tag}"][t]).all(), (w, t)


# ------------------------------------------------ the reference's own NTT tests, fresh inputs

def poly_product(a, b):
    r = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] = (r[i + j] + x * y) % P
    while len(r) > 1 and r[-1] == 0:
        r.pop()
    return r


def coeffs(b):
    v = [int(x) for x in u16(b)]
    while len(v) > 1 and v[-1] == 0:
        v.pop()
    return v


def randbytes(rng, n):
    return bytearray(rng.integers(0, 256, n, dtype=np.uint8).tobytes())


@pytest.fixture(scope="module")
def gf():
    import libffrs_amd as L
    return L.GFi16(3)


@pytest.mark.parametrize("bl,ecc", [(4, 2), (8, 4), (16, 8)])  # the reference's shapes (its (32, 16) is off)
def test_reference_ntt_suite(gf, bl, ecc):
    import libffrs_amd as L
    ntt = L.NTT(gf, bl, ecc)
    rng = np.random.default_rng(bl * 7 + ecc)
    print(f"reference NTT tests {bl} {ecc}", flush=True)
    # test_inv
    data = randbytes(rng, ntt.ntt_size)
    d_ntt = ntt.ntt(data)
    assert bytes(d_ntt) != bytes(data) and bytes(ntt.intt(d_ntt)) == bytes(data)
    # test_poly_mul
    a, b = randbytes(rng, ntt.ecc_size), randbytes(rng, ntt.ecc_size)
    a[ntt.ecc_size // 2:] = bytearray(ntt.ecc_size // 2)
    b[ntt.ecc_size // 2:] = bytearray(ntt.ecc_size // 2)
    want = poly_product(coeffs(a), coeffs(b))
    assert [x & 0xFFFF for x in want] == coeffs(ntt.poly_mul(a, b))
    # test_poly_div_zero
    with pytest.raises(RuntimeError):
        ntt.poly_div(randbytes(rng, ntt.ecc_size), bytearray(ntt.ecc_size))
    with pytest.raises(RuntimeError):
        ntt.poly_div(bytearray(ntt.ecc_size), bytearray(ntt.ecc_size))
    # test_poly_div_higher_deg_denominator
    for size in range(2, ntt.ecc_size - 1, 2):
        pa, pb = bytearray(ntt.ecc_size), bytearray(ntt.ecc_size)
        pa[:size] = randbytes(rng, size)
        pb[:size + 2] = randbytes(rng, size + 2)
        if coeffs(pb) == [0] or len(coeffs(pb)) <= len(coeffs(pa)):
            continue  # random bytes made the denominator's top coefficient zero
        assert coeffs(ntt.poly_div(pa, pb)) == [0]
    # test_poly_inv_mod
    for mod in range(1, ntt.ecc_len + 1):
        for poly_size in range(2, ntt.ecc_size // 2 - 2, 2):
            mod_exp = math.ceil(math.log2(mod))
            aliasing = 2 ** (mod_exp - 1) + poly_size // 2 - 2 > ntt.ecc_len
            pa = bytearray(ntt.ecc_size)
            pa[:poly_size] = randbytes(rng, poly_size)
            if u16(pa)[0] == 0 or u16(pa)[poly_size // 2 - 1] == 0:
                continue  # keep g0 invertible and the length as drawn
            if not aliasing:
                inv = coeffs(ntt.poly_inv(pa, mod))
                prod = poly_product(coeffs(pa), inv)[:mod]
                assert prod + [0] * (mod - len(prod)) == [1] + [0] * (mod - 1)
            else:
                with pytest.raises(RuntimeError):
                    ntt.poly_inv(pa, mod)


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("bl,ecc", [(4, 2), (8, 4), (16, 8), (64, 32)])
def test_reference_ntt_lanes(gf, bl, ecc, n):
    """test_poly_mul_n, test_poly_div_n and test_poly_inv_n: the n-vector methods agree with
    the single-vector ones lane by lane."""
    import libffrs_amd as L
    ntt = L.NTT(gf, bl, ecc)
    rng = np.random.default_rng(bl + 100 * n)
    es = ntt.ecc_size
    a, b = randbytes(rng, es * n), randbytes(rng, es * n)
    for i in range(n):
        a[i * es + es // 2:(i + 1) * es] = bytearray(es // 2)
        b[i * es + es // 2:(i + 1) * es] = bytearray(es // 2)
    res = getattr(ntt, f"poly_mul{n}")(a, b)
    for i in range(n):
        s = slice(i * es, (i + 1) * es)
        assert bytes(res[s]) == bytes(ntt.poly_mul(a[s], b[s]))
    a, b = randbytes(rng, es * n), randbytes(rng, es * n)
    for i in range(n):
        b[i * es + es - 2:(i + 1) * es] = bytearray(2)
    res = getattr(ntt, f"poly_div{n}")(a, b)
    for i in range(n):
        s = slice(i * es, (i + 1) * es)
        assert bytes(res[s]) == bytes(ntt.poly_div(a[s], b[s]))
    mod = 1
    while mod < ntt.ecc_len // 2:
        a = randbytes(rng, es * n)
        for i in range(n):
            a[i * es + es // 2:(i + 1) * es] = bytearray(es // 2)
        res = getattr(ntt, f"poly_inv{n}")(a, mod)
        for i in range(n):
            s = slice(i * es, (i + 1) * es)
            assert bytes(res[s]) == bytes(ntt.poly_inv(a[s], mod))
        mod *= 2


def test_ntt_properties_and_rbo(gf):
    import libffrs_amd as L
    ntt = L.NTT(gf, 200, 8)
    assert ntt.ntt_len == 256 and ntt.ntt_size == 512 and ntt.ntt4_size == 2048
    assert ntt.ntt8_size == 4096 and ntt.ntt16_size == 8192 and ntt.ecc_size == 16
    assert [ntt.rbo(i) for i in (0, 1, 2, 3, 128, 255)] == [0, 128, 64, 192, 1, 255]
    with pytest.raises(RuntimeError):
        ntt.ntt(bytearray(10))
    rs = L.RSi16(256, 32)
    assert rs.ntt is rs.ntt and rs.ntt.ecc_len == 32 and rs.ntt.ntt_len == 256
