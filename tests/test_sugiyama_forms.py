"""The RSi16 repair kernels run the reference's Sugiyama solver (rsi16v_impl.hpp:271-440,
NTT::poly_div_ntt / poly_inv_mod_xn_rev / _reverse_ntt_vec, ntt.hpp:344-564) with every
polynomial in coefficient form: pointwise products of E-point transforms become cyclic
convolutions, and the usual step (a two-term quotient) becomes a two-tap update. This CPU
test states both forms in plain Python over GF(65537) and checks that they return the same
locator length, locator and evaluator (or both raise) on random, error-pattern, sparse,
small-valued, leading-zero and special-value syndrome vectors, including the steps that go
through the transforms. The GPU tests compare the kernels themselves."""
import random

import pytest

STATS = {"fast": 0, "slow": 0}

P = 65537
def add(a, b): return (a + b) % P
def sub(a, b): return (a - b) % P
def mul(a, b): return a * b % P
def inv(a): return 1 if a == 0 else pow(a, P - 2, P)
def rbo(i, bits): return int(format(i, f'0{bits}b')[::-1], 2) if bits else 0

class Code:
    def __init__(self, E, primitive=3):
        self.E = E
        er = pow(primitive, 65536 // E, P); eri = inv(er)
        self.re = [pow(er, i, P) for i in range(E)]
        self.ri = [pow(eri, i, P) for i in range(E)]
        bits = E.bit_length() - 1
        self.rbo = [rbo(i, bits) for i in range(E)]
        self.einv = inv(E)

def ct_full(x, roots):
    x = list(x); n = len(x); s = 1; ef = n >> 1
    while s < n:
        for b in range(n // 2):
            j = b & (s - 1); i = ((b - j) << 1) + j
            a, c = x[i], x[i + s]
            if j: c = mul(c, roots[ef * j])
            x[i], x[i + s] = add(a, c), sub(a, c)
        s <<= 1; ef >>= 1
    return x

def gs_full(x, roots):
    x = list(x); n = len(x); s = n >> 1; ef = 0
    while s > 0:
        for b in range(n // 2):
            j = b & (s - 1); i = ((b - j) << 1) + j
            a, c = x[i], x[i + s]
            x[i] = add(a, c); d = sub(a, c)
            x[i + s] = mul(d, roots[j << ef]) if j else d
        s >>= 1; ef += 1
    return x

def nttr(C, x): return gs_full(x, C.re)
def inttr(C, x): return [mul(v, C.einv) for v in ct_full(x, C.ri)]

def plen(r, mx):
    mx = min(mx, len(r))
    while mx > 0 and r[mx - 1] == 0: mx -= 1
    return mx

def reverse_ntt(C, v, shift):
    E = C.E; out = list(v)
    for i in range(1, E):
        j = C.rbo[i]; k = C.rbo[(E - j) % E]
        out[i] = mul(mul(v[k], C.ri[j]), C.re[(shift * j) % E])
    return out

def poly_inv(C, g, g_len, g0, a_len):
    E = C.E; a0 = inv(g0); a = [a0] * E; l = 2; a_max = a_len
    while l < a_len:
        if g_len + l - 2 > E: return None, None
        a = [mul(ai, sub(2, mul(ai, gi))) for ai, gi in zip(a, g)]
        a = inttr(C, a); a = [v if i < l else 0 for i, v in enumerate(a)]; a = nttr(C, a); l *= 2
    g = [mul(ai, sub(2, mul(ai, gi))) for ai, gi in zip(a, g)]
    g = inttr(C, g)
    out = []
    for i in range(E):
        src = a_max - 1 - i if i < a_max else i
        v = g[src] if src < l else 0
        out.append(v if i < a_len else 0)
    return nttr(C, out), True

def poly_div(C, f, f_len, g, g_len, f0, g0):
    E = C.E; q_len = f_len - g_len + 1
    if g_len == 0: return None, None
    if f_len < g_len: return [0] * E, 0
    t = reverse_ntt(C, g, g_len)
    q, ok = poly_inv(C, t, g_len, g0, q_len)
    if q is None: return None, None
    t = [mul(sub(fi, f0), qi) for fi, qi in zip(f, q)]
    t = inttr(C, t)
    right_n = min(q_len, E - g_len)
    qq = []
    for i in range(E):
        if i < right_n: v = t[g_len + i]
        elif i < q_len: v = t[i - right_n]
        else: v = 0
        qq.append(v)
    return nttr(C, qq), q_len

def sugiyama_ntt(C, r1):
    E = C.E; mask = E - 1
    q = [0] * E; t = [0] * E; a1 = [0] * E; a2 = [1] * E; r2 = list(r1)
    q_len = t_len = 0; a2_len = 1
    r1_len = plen(r1, E); r2_len = r1_len; r2l = r2[0]
    lanes = r1_len != 0
    r1h = r1[(r1_len - 1) & mask]; r1h2 = r1[(r1_len - 2) & mask]
    a11 = (P - inv(r1h)) % P; a10 = mul(r1h2, mul(a11, a11))
    a1[0] = a10; a1[1] = a11; a1_len = 2 if lanes else 0
    tmp = [add(mul(a10, r2[i]), mul(a11, r2[i - 1]) if i >= 1 else 0) for i in range(E)]
    r1_len = plen(tmp, E - 1) if lanes else 0
    r1 = [tmp[i] if i < r1_len else 0 for i in range(E)]
    a1 = nttr(C, a1); r2 = nttr(C, r2)
    if r1_len >= E // 2:
        for it in range(1, E // 2):
            r1h = r1[(r1_len - 1) & mask]; r1l = r1[0]
            r1 = nttr(C, r1)
            q, q_len = poly_div(C, r2, r2_len, r1, r1_len, r2l, r1h)
            if q is None: return None
            t = [sub(x, mul(qi, y)) for x, qi, y in zip(a2, q, a1)]
            q = [sub(x, mul(qi, y)) for x, qi, y in zip(r2, q, r1)]
            t_len = max(a2_len, q_len + a1_len - 1)
            nq_len = max(r2_len, q_len + r1_len - 1)
            a2 = a1; r2 = r1; a2_len = a1_len; r2_len = r1_len; r2l = r1l
            if not (r1_len >= E // 2):
                r1 = inttr(C, r1); break
            a1 = t; r1 = q; a1_len = t_len; r1_len = nq_len
            r1 = inttr(C, r1); r1_len = plen(r1, r1_len)
    a1 = inttr(C, a1)
    a1_len = plen(a1, a1_len if lanes else 0)
    i0 = inv(a1[0])
    a1 = [mul(v, i0) for v in a1]; r1 = [mul(v, i0) for v in r1]
    return (a1_len if lanes else 0), a1, r1

def div_ok(E, f_len, g_len):
    if g_len == 0: return False
    if f_len < g_len: return True
    q_len = f_len - g_len + 1; l = 2
    while l < q_len:
        if g_len + l - 2 > E: return False
        l *= 2
    return True

def sugiyama_coef(C, r1):
    E = C.E; mask = E - 1
    a1 = [0] * E; a2 = [1] + [0] * (E - 1); r2 = list(r1)   # a2 = NTT^-1(all ones) = delta
    a2_len = 1
    r1_len = plen(r1, E); r2_len = r1_len; r2l = r2[0]
    lanes = r1_len != 0
    r1h = r1[(r1_len - 1) & mask]; r1h2 = r1[(r1_len - 2) & mask]
    a11 = (P - inv(r1h)) % P; a10 = mul(r1h2, mul(a11, a11))
    a1[0] = a10; a1[1] = a11; a1_len = 2 if lanes else 0
    tmp = [add(mul(a10, r2[i]), mul(a11, r2[i - 1]) if i >= 1 else 0) for i in range(E)]
    r1_len = plen(tmp, E - 1) if lanes else 0
    r1 = [tmp[i] if i < r1_len else 0 for i in range(E)]
    if r1_len >= E // 2:
        for it in range(1, E // 2):
            if r1_len < E // 2:
                if not div_ok(E, r2_len, r1_len): return None
                break
            r1h = r1[(r1_len - 1) & mask]; r1l = r1[0]
            g_len = r1_len; f_len = r2_len
            q_len = f_len - g_len + 1
            if g_len >= 1 and f_len >= g_len and q_len == 2:
                STATS["fast"] += 1
                a0 = inv(r1h)
                aa = mul(a0, a0)
                A0 = sub(0, mul(aa, r1[(g_len - 2) & mask]))
                A1 = sub(mul(2, a0), mul(aa, r1h))
                def r2p(j): return sub(r2[0], r2l) if j == 0 else r2[j]
                def tv(j): return add(mul(A0, r2p(j)), mul(A1, r2p((j - 1) & mask)))
                right_n = min(2, E - g_len)
                idx = [(g_len + i) if i < right_n else (i - right_n) for i in range(2)]
                q0, q1 = tv(idx[0]), tv(idx[1])
                t = [sub(a2[i], add(mul(q0, a1[i]), mul(q1, a1[(i - 1) & mask]))) for i in range(E)]
                rn = [sub(r2[i], add(mul(q0, r1[i]), mul(q1, r1[(i - 1) & mask]))) for i in range(E)]
            else:
                STATS["slow"] += 1
                R1 = nttr(C, r1); R2 = nttr(C, r2); A1n = nttr(C, a1); A2n = nttr(C, a2)
                Q, q_len = poly_div(C, R2, r2_len, R1, r1_len, r2l, r1h)
                if Q is None: return None
                t = inttr(C, [sub(x, mul(qi, y)) for x, qi, y in zip(A2n, Q, A1n)])
                rn = inttr(C, [sub(x, mul(qi, y)) for x, qi, y in zip(R2, Q, R1)])
            t_len = max(a2_len, q_len + a1_len - 1)
            nq_len = max(r2_len, q_len + r1_len - 1)
            a2 = a1; r2 = r1; a2_len = a1_len; r2_len = r1_len; r2l = r1l
            a1 = t; r1 = rn; a1_len = t_len; r1_len = plen(rn, nq_len)
    a1_len = plen(a1, a1_len if lanes else 0)
    i0 = inv(a1[0])
    a1 = [mul(v, i0) for v in a1]; r1 = [mul(v, i0) for v in r1]
    return (a1_len if lanes else 0), a1, r1

def gen(E, rng):
    kind = rng.randrange(6)
    if kind == 0:
        return [rng.randrange(P) for _ in range(E)]
    if kind == 1:  # nu-error syndromes
        nu = rng.randrange(0, E // 2 + 3)
        X = [rng.randrange(1, P) for _ in range(nu)]; Y = [rng.randrange(1, P) for _ in range(nu)]
        return [sum(y * pow(x, j, P) for x, y in zip(X, Y)) % P for j in range(E)]
    if kind == 2:  # sparse
        v = [0] * E
        for _ in range(rng.randrange(1, 4)): v[rng.randrange(E)] = rng.randrange(1, P)
        return v
    if kind == 3:  # small values
        return [rng.randrange(3) for _ in range(E)]
    if kind == 4:  # leading zeros
        v = [rng.randrange(P) for _ in range(E)]; z = rng.randrange(1, E)
        return v[:E - z] + [0] * z
    v = [rng.choice([0, 1, P - 1, 65536, 2]) for _ in range(E)]
    return v



@pytest.mark.parametrize("E,cases", [(2, 300), (4, 300), (8, 300), (16, 200), (32, 60)])
def test_coefficient_form_matches_transform_form(E, cases):
    rng = random.Random(1000 + E)
    C = Code(E)
    STATS["fast"] = STATS["slow"] = 0
    for _ in range(cases):
        s = gen(E, rng)
        assert sugiyama_ntt(C, s) == sugiyama_coef(C, s), s
    if E >= 4:  # E = 2 has no division steps
        assert STATS["fast"] > 0 and STATS["slow"] > 0
