"""The `.ffrs` format and the geometry choice of the par front end, on the CPU.

Pinned by what the reference states: the base-64 integer alphabet and the empty zero
(ffrs/util.py:21-48), the codec line of its own docstring (`circ16 ip:g01 ig:3 ib:x4 im:x0 ie:4
op:g01 og:3 ob:80 om:40 oe:40 i:1`, ffrs/__init__.py:45-50) and the consistency checks of
deserialize, the exact line layout of Writer (ffrs/par/format.py:31-106), and the size bounds
its CLI test puts on a backup of 8000 B .. 3 MiB (tests/test_lib_cli.py: the .ffrs within 15 %,
5 % or 2 % of the input)."""
import base64
import hashlib
import os
from types import SimpleNamespace

import pytest

import libffrs_amd as L
from libffrs_amd.par import Reader, Writer, b64hex, b64hex_dec
from libffrs_amd.par import opt


def test_b64hex_known_values():
    assert b64hex(0) == "" and b64hex_dec("") == 0
    assert b64hex(65537) == "g01" and b64hex_dec("g01") == 65537
    assert b64hex(63) == "_" and b64hex(64) == "10" and b64hex_dec("-") == 62
    for v in (1, 3, 4, 255, 256, 4095, 4096, 2 ** 40 + 12345, 10 ** 18):
        assert b64hex_dec(b64hex(v)) == v


def test_reference_codec_line():
    line = "circ16 ip:g01 ig:3 ib:x4 im:x0 ie:4 op:g01 og:3 ob:80 om:40 oe:40 i:1"
    p = {k: b64hex_dec(v) for k, v in (part.split(":") for part in line[7:].split())}
    assert p == dict(ip=65537, ig=3, ib=2116, im=2112, ie=4, op=65537, og=3, ob=512, om=256, oe=256, i=1)
    codec = SimpleNamespace(
        rsi=SimpleNamespace(gf=SimpleNamespace(field_elements=65537, primitive=3)),
        rso=SimpleNamespace(gf=SimpleNamespace(field_elements=65537, primitive=3)),
        inner_block_len=2116, inner_message_len=2112, inner_ecc_len=4,
        outer_block_len=512, outer_message_len=256, outer_ecc_len=256, interleave=1)
    assert L.CIRC16.serialize(codec) == line.encode()
    for bad in (line.replace("ie:4", "ie:8"), line.replace("oe:40", "oe:41"), line.replace("og:3", "og:5"),
                line.replace("ip:g01", "ip:g02"), "rsi16 " + line[7:]):
        with pytest.raises(AssertionError):  # the consistency checks run before any device work
            L.CIRC16.deserialize(bad.encode())


def test_writer_reader_layout(tmp_path):
    codec = SimpleNamespace(serialize=lambda: b"circ16 ip:g01 ig:3 ib:46 im:44 ie:2 op:g01 og:3 ob:w om:g oe:g i:1")
    digest = hashlib.blake2b(b"abc", digest_size=20).digest()
    ecc = bytes(range(40))
    out = tmp_path / "x.ffrs"
    with Writer(str(out)) as w:
        w.write_header(codec)
        w.write_block([("f", 1000, 1234567890123, digest, os.path.join("dir", "my file.bin")),
                       ("c", 300, 64, "other.bin"),
                       ("fc", 5000, 4096, 77, digest, "other.bin")], ecc)
    raw = out.read_bytes()
    h = base64.urlsafe_b64encode(digest).rstrip(b"=")
    want = (b"\x89ffrs\r\n" + codec.serialize() + b"\n"
            + b"p p:dir\n"
            + b"f s:" + b64hex(1000).encode() + b" m:" + b64hex(1234567890123).encode() + b" h:b2:" + h
            + b' p:"my file.bin"\n'
            + b"p p:\n"
            + b"c s:" + b64hex(300).encode() + b" o:" + b64hex(64).encode() + b" p:other.bin\n"
            + b"fc s:" + b64hex(5000).encode() + b" o:" + b64hex(4096).encode() + b" m:" + b64hex(77).encode()
            + b" h:b2:" + h + b" p:other.bin\n"
            + b"d s:" + b64hex(40).encode() + b"\n" + ecc + b"\n")
    assert raw == want
    with Reader(str(out)) as r:
        assert r.file.readline(16) == b"\x89ffrs\r\n" and r.file.readline(4096) == codec.serialize() + b"\n"
        buf = bytearray(40)
        blocks = list(r.blocks(buf))
    assert bytes(buf) == ecc
    assert blocks == [[("f", 1000, 1234567890123, digest.hex(), os.path.join("dir", "my file.bin")),
                       ("c", 300, 64, "other.bin"), ("fc", 5000, 4096, 77, digest.hex(), "other.bin")]]


@pytest.mark.parametrize("size", [8000, 8192, 65536, 1000000, 1 << 20, 1310720, 3072000, 3 << 20])
def test_geometry_meets_the_reference_bounds(size):
    ib, ie, ob, oe, I = opt.choose_geometry(size)
    ik, ok = ib - ie, ob - oe
    msg_bytes = 2 * ok * ik * I
    ecc_bytes = 2 * (oe * ik * I + ie * ob * I)
    assert msg_bytes >= size and msg_bytes <= 2 * max(size, 8192)
    assert ie & (ie - 1) == 0 and ie >= 2 and ib % ie == 0 and ob % oe == 0 and oe == ok
    assert 127 <= (2 * ik) // (2 * ie) <= 512 and 2 * ik <= 4096 and ik * I <= 0xFFFF
    limit = 0.15 if size <= 8192 else 0.05 if size <= 1 << 16 else 0.02
    assert abs(ecc_bytes / size - 1) < limit
