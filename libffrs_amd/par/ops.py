"""`par backup` and `par repair` (ffrs/par/backup.py:29-74, ffrs/par/repair.py:79-97) as
library calls."""
import hashlib
import os

from .format import Reader, Writer
from .fs import read_filelist_into, write_filelist_into
from .opt import choose_codec


def backup(paths, codec=None, force=False, **simd):
    """Write `<file>.ffrs` for each non-empty file: one CIRC16 block holding the whole file
    (zero-padded), its header line, and the file entry with size, mtime and a 20-byte BLAKE2b of
    the contents. `codec` defaults to `choose_codec(size)`. Returns the parity file names."""
    out = []
    for name in paths:
        if not os.path.isfile(name):
            raise FileNotFoundError(name)
        target = name + ".ffrs"
        if os.path.exists(target) and not force:
            raise FileExistsError(target)
        st = os.stat(name)
        if st.st_size == 0:
            continue
        rs = codec if codec is not None else choose_codec(st.st_size, **simd)
        assert rs.message_size >= st.st_size  # one block per file (backup.py:56-57)
        buf = bytearray(rs.message_size)
        with open(name, "rb") as f:
            got = f.readinto(buf)
        assert got == st.st_size
        digest = hashlib.blake2b(buf[:got], digest_size=20).digest()
        ecc = rs.encode(buf)
        with Writer(target) as w:
            w.write_header(rs)
            w.write_block([("f", st.st_size, st.st_mtime_ns, digest, name)], ecc)
        out.append(target)
    return out


def repair(path, **simd):
    """Repair the files a `.ffrs` file describes, block by block: read the block's byte ranges
    into a zeroed message buffer, CIRC16.repair it with the stored ecc, write the ranges back
    and restore their mtimes (repair.py:79-97)."""
    with Reader(path, **simd) as r:
        rs = r.read_header()
        ecc = bytearray(rs.ecc_size)
        for files in r.blocks(ecc):
            data = bytearray(rs.message_size)
            read_filelist_into(files, data)
            rs.repair(data, ecc)
            write_filelist_into(files, data)
    return 0
