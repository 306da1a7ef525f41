"""CIRC16 (pycirc16.hpp:33-431) on the GPU.

Every case of the reference-written fixtures tests/golden/circ16.npz and circ16_planted.npz is
replayed through libffrs_amd.CIRC16 and must match bit for bit: the encode, each repair scenario
(outputs and whether it raises) and _find_outer_error_locations. The fixtures' meaning is pinned
on the CPU by tests/test_circ16_golden.py. The reference's own CIRC16 tests
(tests/test_lib_circ16.py) follow with fresh random inputs on a few of their shapes.
"""
import os
import random
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import cases as C  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(HERE, "golden", "circ16.npz"))


@pytest.fixture(scope="module")
def fp():
    return np.load(os.path.join(HERE, "golden", "circ16_planted.npz"))


def crc(a):
    return int(C.row_crc(np.asarray(a)[None])[0])


def run_repair(rs, m, e):
    mb, eb = bytearray(m.tobytes()), bytearray(e.tobytes())
    try:
        assert rs.repair(mb, eb) is False
        raised = 0
    except RuntimeError:
        raised = 1
    return raised, [crc(np.frombuffer(bytes(mb), np.uint16)), crc(np.frombuffer(bytes(eb), np.uint16))]


@pytest.mark.parametrize("i", range(len(C.CIRC16_CASES)), ids=[C.circ16_tag(c) for c in C.CIRC16_CASES])
def test_circ16_fixture_replay(fx, i):
    import libffrs_amd as L
    case = C.CIRC16_CASES[i]
    tag = C.circ16_tag(case)
    print(f"circ16 replay {tag}", flush=True)
    rs = L.CIRC16(*case[:5], **C.circ16_flags(case[5]))
    lay, msg = C.circ16_message(i, case)
    assert rs.message_len == lay.msg_len and rs.ecc_len == lay.ecc_len
    ecc = np.frombuffer(bytes(rs.encode(msg.tobytes())), np.uint16).copy()
    assert crc(ecc) == fx[f"{tag}_ecc_crc"][0]
    if f"{tag}_ecc" in fx:
        assert (ecc == fx[f"{tag}_ecc"]).all()
    for name, m, e in C.circ16_scenarios(i, case, lay, msg, ecc):
        assert (C.sha(m, e) == fx[f"{tag}_{name}_in_sha"]).all(), name
        raised, out = run_repair(rs, m, e)
        assert raised == fx[f"{tag}_{name}_raised"][0], name
        assert out == fx[f"{tag}_{name}_out_crc"].tolist(), name
    want = fx[f"{tag}_find"]
    for t, (il, m) in enumerate(C.circ16_find_inputs(i, case, lay, msg, ecc)):
        rows = rs._find_outer_error_locations(bytearray(m.tobytes()), bytearray(ecc.tobytes()), il)
        assert list(rows) == [int(x) for x in want[t] if x >= 0], (il, t)


@pytest.mark.parametrize("i", range(len(C.CIRC16_PLANT_CASES)), ids=[C.circ16_tag(c) for c in C.CIRC16_PLANT_CASES])
def test_circ16_planted_replay(fp, i):
    """Blocks whose outer or inner parity holds a 65536 (stored as 0): the encode keeps the field
    element for the inner parity of the outer parity rows, and the repair's zero steps restore it."""
    import libffrs_amd as L
    case = C.CIRC16_PLANT_CASES[i]
    tag = C.circ16_tag(case)
    print(f"circ16 planted {tag}", flush=True)
    rs = L.CIRC16(*case[:5], **C.circ16_flags(case[5]))
    lay = C.CircLayout(*case[:5])
    for kind in ("outer", "inner"):
        plant = fp[f"{tag}_{kind}_plant"]
        m = C.circ16_planted_message(i, kind, lay, int(plant[0]), plant[1:])
        ecc = np.frombuffer(bytes(rs.encode(m.tobytes())), np.uint16)
        assert (ecc == fp[f"{tag}_{kind}_ecc"]).all(), kind
        for name, mm, ee in C.circ16_planted_scenarios(i, kind, lay, m, ecc):
            raised, out = run_repair(rs, mm, ee)
            assert raised == fp[f"{tag}_{kind}_{name}_raised"][0], (kind, name)
            assert out == fp[f"{tag}_{kind}_{name}_out_crc"].tolist(), (kind, name)
        rows = rs._find_outer_error_locations(bytearray(m.tobytes()), bytearray(ecc.tobytes()), 0)
        assert list(rows) == [int(x) for x in fp[f"{tag}_{kind}_find0"] if x >= 0]


# ------------------------------------------------ the reference's own CIRC16 tests, fresh inputs

def randbytes(n):
    return bytearray(random.getrandbits(8) for _ in range(n))


def corrupt_outer_rows(rs, buf, ecc, n):
    for row in random.sample(range(rs.outer_block_len), n):
        if row < rs.outer_message_len:
            for col in range(rs.inner_message_len):
                o = 2 * rs.message_offset(0, row, col)
                buf[o] ^= random.randint(1, 255)
                buf[o + 1] ^= random.randint(1, 255)
        else:
            for col in range(rs.inner_message_len):
                o = 2 * rs.rso_ecc_offset(0, row - rs.outer_message_len, col)
                ecc[o] ^= random.randint(1, 255)
                ecc[o + 1] ^= random.randint(1, 255)
        for col in range(rs.inner_ecc_len):
            o = 2 * rs.rsi_ecc_offset(0, row, col)
            ecc[o] ^= random.randint(1, 255)
            ecc[o + 1] ^= random.randint(1, 255)


SUITE = [(16, 4, 16, 4, 1, 16), (32, 16, 4, 2, 1, 8), (128, 8, 128, 8, 4, 4), (100, 2, 100, 2, 3, 1)]


@pytest.mark.parametrize("case", SUITE, ids=[C.circ16_tag(c) for c in SUITE])
def test_reference_circ16_suite(case):
    import libffrs_amd as L
    random.seed(C.circ16_tag(case))
    rs = L.CIRC16(*case[:5], **C.circ16_flags(case[5]))
    print(f"reference CIRC16 tests {C.circ16_tag(case)}", flush=True)
    # test_properties
    assert rs.ecc_len == rs.inner_block_len * rs.outer_block_len * rs.interleave - rs.message_len
    assert rs.block_len == rs.message_len + rs.ecc_len
    assert rs.outer_interleave == rs.inner_message_len * rs.interleave
    # test_circ_encode
    buf = randbytes(rs.message_size)
    res = bytes(rs.encode(buf))
    assert len(res) == rs.ecc_size
    res_o = bytes(rs.rso.encode(buf))
    assert res[:len(res_o)] == res_o
    res_i = bytes(rs.rsi.encode(buf))
    assert res[len(res_o):][:len(res_i)] == res_i
    # test_circ_repair / test_circ_repair_multiple
    for _ in range(3):
        buf = randbytes(rs.message_size)
        ecc = bytearray(rs.encode(buf))
        b0, e0 = bytes(buf), bytes(ecc)
        corrupt_outer_rows(rs, buf, ecc, rs.outer_ecc_len)
        rs.repair(buf, ecc)
        assert bytes(buf) == b0 and bytes(ecc) == e0
    # test_repair_no_errors
    buf = randbytes(rs.message_size)
    ecc = bytearray(rs.encode(buf))
    b0, e0 = bytes(buf), bytes(ecc)
    rs.repair(buf, ecc)
    assert bytes(buf) == b0 and bytes(ecc) == e0
    # test_find_errors_diagonal
    for il in range(rs.interleave):
        assert rs._find_outer_error_locations(buf, ecc, il) == []
        for i in range(min(rs.inner_message_len, rs.outer_message_len)):
            o = 2 * rs.message_offset(il, i, i)
            buf[o] ^= random.randrange(1, 256)
            buf[o + 1] ^= random.randrange(1, 256)
        assert rs._find_outer_error_locations(buf, ecc, il) == list(range(min(rs.inner_message_len,
                                                                               rs.outer_message_len)))


def test_circ16_errors():
    import libffrs_amd as L
    rs = L.CIRC16(16, 4, 16, 4)
    with pytest.raises(RuntimeError):
        rs.encode(bytearray(rs.message_size + 2))
    with pytest.raises(RuntimeError):
        rs.repair(bytearray(rs.message_size - 2), bytearray(rs.ecc_size))
    with pytest.raises(RuntimeError):
        rs._find_outer_error_locations(bytearray(rs.message_size), bytearray(rs.ecc_size), 1)
    with pytest.raises(RuntimeError):
        rs.message_offset(0, rs.outer_message_len, 0)
    assert rs.rsi_ecc_offset(0, 0, 0) == rs.outer_ecc_len * rs.outer_interleave
    assert len(rs.encode(b"")) == 0
    two = bytes(range(256)) * (2 * rs.message_size // 256 + 1)
    two = two[:2 * rs.message_size]
    e2 = bytes(rs.encode(two))
    assert e2 == bytes(rs.encode(two[:rs.message_size])) + bytes(rs.encode(two[rs.message_size:]))
