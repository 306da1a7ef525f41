"""Pins of the oracle that do not come from the reference's outputs: known answers
from the reference's own tests, brute force on small inputs, and invariants that
the mathematics fixes (a dropped term, wrong index or transposed operand fails one)."""
import numpy as np
import pytest

from oracle import oracle as O


def gf256_mul(a, b, poly=0x11D):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a <<= 1
        if a & 0x100:
            a ^= poly
        b >>= 1
    return r


def test_gf256_tables_conventions():
    ok, exp, log = O.gf256_tables()
    assert ok == 0
    assert exp[0] == 1 and exp[1] == 2 and exp[255] == 1   # loop runs i = 0..255 inclusive
    assert log[1] == 255 and log[0] == 0
    for i in range(255):
        assert exp[(i + 1) % 255] == gf256_mul(int(exp[i]), 2)


def test_rs256_generator_kat():
    # tests/test_reference.py:83-97 of the reference: 64 + 120x + 54x^2 + 15x^3 + x^4
    assert O.RS256(4).generator().tolist() == [1, 15, 54, 120, 64]


@pytest.mark.parametrize("ecc", [1, 2, 5, 8, 16, 32, 33])
def test_rs256_generator_roots(ecc):
    g = O.RS256(ecc).generator().tolist()
    for i in range(ecc):
        x = 1
        for _ in range(i):
            x = gf256_mul(x, 2)
        acc = 0
        for c in g:
            acc = gf256_mul(acc, x) ^ c
        assert acc == 0


def long_division_parity(msg, gen):
    """x^ecc * M(x) mod g(x) by plain long division over GF(2^8)."""
    ecc = len(gen) - 1
    r = list(msg) + [0] * ecc
    for i in range(len(msg)):
        c = r[i]
        if c:
            for j in range(1, ecc + 1):
                r[i + j] ^= gf256_mul(gen[j], c)
    return r[-ecc:]


@pytest.mark.parametrize("ecc,k", [(1, 3), (2, 9), (4, 17), (8, 30), (16, 40), (32, 100), (64, 191)])
def test_rs256_encode_is_long_division(ecc, k):
    rng = np.random.default_rng(ecc * 1000 + k)
    rs = O.RS256(ecc)
    gen = rs.generator().tolist()
    for _ in range(5):
        msg = rng.integers(0, 256, k, dtype=np.uint8)
        par = rs.encode(msg.tobytes(), k)
        assert list(par) == long_division_parity(msg.tolist(), gen)
        assert not rs.synds(msg, np.frombuffer(bytes(par), np.uint8)).any()


@pytest.mark.parametrize("ecc", [2, 8, 32])
def test_rs256_round_trip_up_to_t(ecc):
    rng = np.random.default_rng(ecc)
    rs = O.RS256(ecc)
    k = 255 - ecc
    for t in range(ecc // 2 + 1):
        msg = rng.integers(0, 256, k, dtype=np.uint8)
        par = np.frombuffer(bytes(rs.encode(msg.tobytes(), k)), np.uint8)
        cw = np.concatenate([msg, par])
        pos = rng.choice(255, t, replace=False)
        cw[pos] ^= rng.integers(1, 256, t, dtype=np.uint8)
        m, p = cw[:k].copy(), cw[k:].copy()
        assert rs.decode(m, p)
        assert (m == msg).all() and (p == par).all()


def rbo(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


@pytest.mark.parametrize("n,ecc", [(4, 2), (16, 4), (64, 8), (96, 32), (256, 32)])
def test_rsi16_parity_zeroes_syndromes(n, ecc):
    """The parity is the unique e making S_j = sum_m c_m w^(j*rbo(m)) zero, j < ecc."""
    p = 65537
    rs = O.RSi16(n, ecc)
    N = rs.ntt_len
    bits = N.bit_length() - 1
    w = rs.root
    assert pow(w, N, p) == 1 and pow(w, N // 2, p) == p - 1 and pow(3, 32768, p) == p - 1
    k = n - ecc
    rng = np.random.default_rng(n + ecc)
    msgs = rng.integers(0, 65536, (6, k), dtype=np.uint16)
    par = rs.encode(msgs, True)
    assert (par == rs.encode(msgs, False)).all()
    X = [pow(w, rbo(m, bits), p) for m in range(n)]
    for r in range(len(msgs)):
        # solve for the canonical parity directly: S_j(msg) + sum_i e_i X_{k+i}^j = 0
        A = [[pow(X[k + i], j, p) for i in range(ecc)] for j in range(ecc)]
        b = [(-sum(int(msgs[r, m]) * pow(X[m], j, p) for m in range(k))) % p for j in range(ecc)]
        e = solve_mod(A, b, p)
        assert [v & 0xFFFF for v in e] == par[r].tolist()
        syn = rs.synds(msgs[r], par[r])
        expect = [(sum(int(msgs[r, m]) * pow(X[m], j, p) for m in range(k))
                   + sum(int(par[r, i]) * pow(X[k + i], j, p) for i in range(ecc))) % p for j in range(ecc)]
        assert syn.tolist() == expect


def solve_mod(A, b, p):
    n = len(A)
    M = [row[:] + [bb] for row, bb in zip(A, b)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] % p)
        M[c], M[piv] = M[piv], M[c]
        inv = pow(M[c][c], p - 2, p)
        M[c] = [v * inv % p for v in M[c]]
        for r in range(n):
            if r != c and M[r][c]:
                f = M[r][c]
                M[r] = [(v - f * w) % p for v, w in zip(M[r], M[c])]
    return [M[i][n] for i in range(n)]


@pytest.mark.parametrize("n,ecc", [(64, 8), (256, 32)])
def test_rsi16_round_trip_up_to_t(n, ecc):
    rs = O.RSi16(n, ecc)
    k = n - ecc
    rng = np.random.default_rng(7 * n)
    for t in range(ecc // 2 + 1):
        msg = rng.integers(0, 65536, k, dtype=np.uint16)
        par = rs.encode(msg[None], True)[0]
        if (par == 0).any():
            continue        # a truncated 65536 parity is a built-in error
        cw = np.concatenate([msg, par]).astype(np.int64)
        pos = rng.choice(n, t, replace=False)
        v = (cw[pos] + rng.integers(1, 65536, t)) % 65537
        v[v == 65536] = 1
        cw[pos] = v
        m, pp = cw[:k].astype(np.uint16), cw[k:].astype(np.uint16)
        st = rs.repair(m, pp)
        if t == 0:
            assert st == O.NO_ERRORS
        else:
            assert st in (O.REPAIR_OK, O.ERROR_LOCATION_FAIL)
            if st == O.REPAIR_OK:
                assert (m == msg).all() and (pp == par).all()


def test_rsi16_field_conventions():
    rs = O.RSi16(256, 32)
    exp = rs.table("exp")
    log = rs.table("log")
    assert exp[0] == 1 and exp[65536] == 1 and log[1] == 65536 and log[0] == 0
    assert exp[1] == 3 and exp[32768] == 65536
