"""`par backup` / `par repair` (ffrs/par/backup.py:29-83, ffrs/par/repair.py:79-97) on the GPU
CIRC16, at the input sizes of the reference's CLI test (tests/test_cli.py:61-111): the parity file
stays within the size bound that test sets, a repair of an intact file changes neither its bytes
nor its mtime, and a file whose every byte was overwritten with random data (mtime put back) is
restored bit for bit with its mtime."""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [8000, 8192, 65536, 1000000, 1 << 20, 1310720, 3072000, 3 << 20]


def backup(tmp_path, size, seed):
    from libffrs_amd import par
    data = np.random.default_rng(seed).integers(0, 256, size, dtype=np.uint8).tobytes()
    src = tmp_path / "input.bin"
    src.write_bytes(data)
    print(f"par backup {size}", flush=True)
    out = par.backup([str(src)])
    assert out == [str(src) + ".ffrs"] and os.path.exists(out[0])
    limit = 0.15 if size <= 8192 else 0.05 if size <= 1 << 16 else 0.02
    assert abs(os.path.getsize(out[0]) / size - 1) < limit
    return src, out[0], data


@pytest.mark.parametrize("size", SIZES)
def test_backup_repair_no_errors(tmp_path, size):
    from libffrs_amd import par
    src, ffrs_file, data = backup(tmp_path, size, size)
    mtime = src.stat().st_mtime_ns
    assert par.repair(ffrs_file) == 0
    assert src.stat().st_mtime_ns == mtime and src.read_bytes() == data


@pytest.mark.parametrize("size", SIZES)
def test_backup_repair_corrupted(tmp_path, size):
    from libffrs_amd import par
    src, ffrs_file, data = backup(tmp_path, size, size + 1)
    mtime = src.stat().st_mtime_ns
    src.write_bytes(np.random.default_rng(size + 2).integers(0, 256, size, dtype=np.uint8).tobytes())
    os.utime(src, ns=(src.stat().st_atime_ns, mtime))
    assert src.read_bytes() != data
    print(f"par repair {size}", flush=True)
    assert par.repair(ffrs_file) == 0
    assert src.stat().st_mtime_ns == mtime and src.read_bytes() == data


def test_parity_file_reads_back(tmp_path):
    """The written parity file parses into the same codec and file entry, the hash is the
    20-byte BLAKE2b of the file, and a backup onto an existing parity file needs force."""
    from libffrs_amd import par
    src, ffrs_file, data = backup(tmp_path, 1000000, 5)
    with par.Reader(ffrs_file) as r:
        rs = r.read_header()
        ecc = bytearray(rs.ecc_size)
        blocks = list(r.blocks(ecc))
    assert rs.message_size >= len(data)
    st = src.stat()
    assert blocks == [[("f", len(data), st.st_mtime_ns, hashlib.blake2b(data, digest_size=20).hexdigest(), str(src))]]
    buf = bytearray(rs.message_size)
    buf[:len(data)] = data
    assert bytes(ecc) == bytes(rs.encode(buf))
    with pytest.raises(FileExistsError):
        par.backup([str(src)])
    assert par.backup([str(src)], force=True) == [ffrs_file]
    empty = tmp_path / "empty.bin"
    empty.write_bytes(b"")
    assert par.backup([str(empty)]) == []


def test_repair_rejects_changed_size(tmp_path):
    from libffrs_amd import par
    src, ffrs_file, data = backup(tmp_path, 8000, 6)
    src.write_bytes(data[:-1])
    with pytest.raises(AssertionError):
        par.repair(ffrs_file)


def test_codec_line_round_trip():
    """The reference package's own example line (ffrs/__init__.py:45-50) gives CIRC16(2116, 4,
    512, 256, 1) and serializes back to itself; the package reprs follow ffrs/__init__.py:24-40."""
    import libffrs_amd as L
    line = b"circ16 ip:g01 ig:3 ib:x4 im:x0 ie:4 op:g01 og:3 ob:80 om:40 oe:40 i:1"
    rs = L.CIRC16.deserialize(line + b"\n")
    assert (rs.inner_block_len, rs.inner_ecc_len, rs.outer_block_len, rs.outer_ecc_len, rs.interleave) == \
        (2116, 4, 512, 256, 1)
    assert rs.serialize() == line and repr(rs) == "CIRC(2116, 4, 512, 256, 1)"
    assert repr(L.RSi16(256, 32)) == "RSi16(block_len=256, ecc_len=32)"
    assert repr(L.RSi16(256, 32, 4)) == "RSi16(block_len=256, ecc_len=32, interleave=4)"
    assert L.CIRC16(128, 8, 128, 8, 4).serialize() == b"circ16 ip:g01 ig:3 ib:20 im:1U ie:8 op:g01 og:3 ob:20 om:1U oe:8 i:4"
