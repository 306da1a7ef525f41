"""The NTT fixture (tests/golden/rsi16_ntt.npz, written from the reference's own NTT class and
RSi16._sugiyama<W> by tests/golden/make_golden.py) checked on the CPU against what the
mathematics fixes, so that the fixture is known to hold the transforms it names:

  ntt    (ntt.hpp:596-599, RBO in, natural out): X[j] = sum_m x[rbo(m)] r^(m j), r = g^(65536/N)
  intt   (natural in, RBO out):                   out[rbo(m)] = N^-1 sum_j x[j] r^(-m j)
  nttr   (first ecc_len values, natural in, RBO out) with the E-th root g^(65536/E), rest kept
  inttr  its inverse
  poly_mul (rsi16v_impl.hpp:734-743): the cyclic convolution modulo x^E - 1
  poly_inv (rsi16v_impl.hpp:783-795): a * inv = 1 mod x^n wherever no product wraps
  poly_div (rsi16v_impl.hpp:745-781): zero quotient below the divisor's degree; the raise rules
           of poly_div_ntt (ntt.hpp:481-493)

Outputs are u16 as the reference returns them (a field element 65536 reads 0). The GPU replay
of the same fixture is tests/test_gpu_ntt.py."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = 65537

import sys  # noqa: E402

sys.path.insert(0, os.path.join(HERE, "golden"))
import cases as C  # noqa: E402


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(HERE, "golden", "rsi16_ntt.npz"))


def rbo(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def ntt_len(bl):
    return 1 << (bl - 1).bit_length()


def plen(v):
    nz = np.nonzero(v)[0]
    return int(nz[-1]) + 1 if nz.size else 0


def dft_matrix(root, n):
    e = (np.arange(n)[:, None] * np.arange(n)[None, :]) % n
    pw = np.ones(n, np.int64)
    for i in range(1, n):
        pw[i] = pw[i - 1] * root % P
    return pw[e]


def matvec(M, x):
    """(M @ x) mod p without overflow: x < 2^17, M < 2^17, split M in 2^9 limbs."""
    lo, hi = M & 511, M >> 9
    return ((lo @ x) % P + ((hi @ x) % P) * 512) % P


def cyc(a, b, E):
    r = np.zeros(E, np.int64)
    for i in np.nonzero(a)[0]:
        r = (r + int(a[i]) * np.roll(b.astype(np.int64), i)) % P
    return r


def mulmod_xn(a, b, n):
    r = [0] * n
    for i in range(min(len(a), n)):
        if a[i]:
            for j in range(min(len(b), n - i)):
                r[i + j] = (r[i + j] + int(a[i]) * int(b[j])) % P
    return r


SMALL = [s for s in C.NTT_SHAPES if ntt_len(s[0]) <= 256]


@pytest.mark.parametrize("bl,ecc,prim", SMALL)
def test_fixture_transforms_are_the_dft(fx, bl, ecc, prim):
    N, E = ntt_len(bl), ecc
    nb, eb = N.bit_length() - 1, E.bit_length() - 1
    r, re = pow(prim, 65536 // N, P), pow(prim, 65536 // E, P)
    F, Fi = dft_matrix(r, N), dft_matrix(pow(r, P - 2, P), N)
    Fe, Fei = dft_matrix(re, E), dft_matrix(pow(re, P - 2, P), E)
    Ni, Ei = pow(N, P - 2, P), pow(E, P - 2, P)
    rb = np.array([rbo(m, nb) for m in range(N)])
    rbe = np.array([rbo(m, eb) for m in range(E)])
    for w in C.NTT_WIDTHS:
        tag = f"{bl}_{ecc}_{prim}_w{w}"
        X = fx[f"x_{tag}"].astype(np.int64).reshape(w, N)
        for v in range(w):
            x = X[v]
            assert (fx[f"ntt_{tag}"].reshape(w, N)[v] == matvec(F, x[rb]) & 0xFFFF).all()
            y = matvec(Fi, x) * Ni % P
            assert (fx[f"intt_{tag}"].reshape(w, N)[v][rb] == y & 0xFFFF).all()
            out = fx[f"nttr_{tag}"].reshape(w, N)[v]
            assert (out[rbe] == matvec(Fe, x[:E]) & 0xFFFF).all() and (out[E:] == x[E:]).all()
            out = fx[f"inttr_{tag}"].reshape(w, N)[v]
            assert (out[:E] == matvec(Fei, x[:E][rbe]) * Ei % P & 0xFFFF).all() and (out[E:] == x[E:]).all()


@pytest.mark.parametrize("bl,ecc,prim", C.NTT_SHAPES)
def test_fixture_poly_mul_is_cyclic_convolution(fx, bl, ecc, prim):
    E = ecc
    for w in C.NTT_WIDTHS:
        tag = f"{bl}_{ecc}_{prim}_w{w}"
        A, B, O = fx[f"mul_a_{tag}"], fx[f"mul_b_{tag}"], fx[f"mul_out_{tag}"]
        wrapped = False
        for t in range(len(A)):
            for v in range(w):
                a, b = A[t].reshape(w, E)[v], B[t].reshape(w, E)[v]
                want = cyc(a, b, E)
                assert (O[t].reshape(w, E)[v] == want & 0xFFFF).all()
                lin = np.array(mulmod_xn(a, b, E))
                wrapped |= bool((lin != want).any())
        assert wrapped  # the full-length trials wrap: a truncated product would not match


@pytest.mark.parametrize("bl,ecc,prim", C.NTT_SHAPES)
def test_fixture_poly_inv_is_an_inverse(fx, bl, ecc, prim):
    E = ecc
    checked = 0
    for w in C.NTT_WIDTHS:
        tag = f"{bl}_{ecc}_{prim}_w{w}"
        A, NN, O, R = fx[f"inv_a_{tag}"], fx[f"inv_n_{tag}"], fx[f"inv_out_{tag}"], fx[f"inv_raises_{tag}"]
        for t in range(len(A)):
            if R[t]:
                continue
            n = int(NN[t])
            l_final = 2
            while l_final < n:
                l_final *= 2
            for v in range(w):
                a, o = A[t].reshape(w, E)[v], O[t].reshape(w, E)[v]
                if l_final + plen(a) - 2 > E:  # the Newton products wrap modulo x^E - 1
                    continue
                assert (o[n:] == 0).all()
                if n:
                    assert mulmod_xn(a, o, n) == [1] + [0] * (n - 1)
                checked += 1
    assert checked > 0


@pytest.mark.parametrize("bl,ecc,prim", C.NTT_SHAPES)
def test_fixture_poly_div_rules(fx, bl, ecc, prim):
    E = ecc
    raised = 0
    for w in C.NTT_WIDTHS:
        tag = f"{bl}_{ecc}_{prim}_w{w}"
        A, B, O, R = fx[f"div_a_{tag}"], fx[f"div_b_{tag}"], fx[f"div_out_{tag}"], fx[f"div_raises_{tag}"]
        raised += int(R.sum())
        assert not R.all()
        for t in range(len(A)):
            fl = np.array([plen(x) for x in A[t].reshape(w, E)])
            gl = np.array([plen(x) for x in B[t].reshape(w, E)])
            if w == 1:
                assert R[t] == (gl[0] == 0)  # py_assert(g_len > 0)
            elif ((gl == 0) & (fl > 0)).any():
                assert R[t] == 1
            if R[t]:
                continue
            for v in range(w):
                if fl[v] < gl[v]:
                    assert (O[t].reshape(w, E)[v] == 0).all()
    assert raised > 0


def test_fixture_sugiyama_shapes(fx):
    for n, ecc in C.SUGIYAMA_SHAPES:
        for w in C.NTT_WIDTHS:
            tag = f"{n}_{ecc}_w{w}"
            S, L, E = fx[f"sug_synd_{tag}"], fx[f"sug_loc_{tag}"], fx[f"sug_ev_{tag}"]
            assert S.shape == L.shape == E.shape == (6, w * ecc)
            assert (L.max() <= 65536) and (E.max() <= 65536)
