"""Block file lists (ffrs/par/fs.py:118-245): which byte ranges of which files make up the
message bytes of a block, in order; the rest of the block's message buffer is zero."""
import os


def chunk_filelist(input_files, chunk_size):
    """Cut the files (in order, empty ones skipped) into blocks of chunk_size message bytes
    (fs.py:118-146): yields (bytes in the block, entries). A file that fits is ("f", size,
    mtime_ns, None, path); a file cut across blocks gives ("c", size, offset, path) pieces and
    ends with ("fc", file_size, offset, mtime_ns, None, path)."""
    files, total = [], 0
    for name in input_files:
        st = os.stat(name)
        if st.st_size == 0:
            continue
        total += st.st_size
        files.append(("f", st.st_size, st.st_mtime_ns, None, name))
        offset = 0
        piece = st.st_size - (total - chunk_size)
        while total > chunk_size:
            files[-1] = ("c", piece, offset, name)
            yield chunk_size, files
            total -= chunk_size
            offset += piece
            piece = min(total, chunk_size)
            files = [("c", piece, offset, name)]
        if files[-1][0] == "c":
            files[-1] = ("fc", st.st_size, offset, st.st_mtime_ns, None, name)
        if total == chunk_size:
            yield total, files
            files = []
    yield total, files


def _ranges(filelist):
    """(path, offset, size, file_size or None, mtime_ns or None) per entry."""
    for entry in filelist:
        kind = entry[0]
        if kind == "f":
            _, size, mtime_ns, _digest, path = entry
            yield path, 0, size, size, mtime_ns
        elif kind == "c":
            _, size, offset, path = entry
            yield path, offset, size, None, None
        elif kind == "fc":
            _, file_size, offset, mtime_ns, _digest, path = entry
            yield path, offset, file_size - offset, file_size, mtime_ns
        else:
            raise NotImplementedError(kind)


def read_filelist_into(filelist, buffer):
    """Read the block's byte ranges into the start of buffer, one after another (fs.py:149-211);
    a whole or final piece must still have its recorded size and an mtime no older than the
    recorded one. Returns the number of bytes read."""
    view = memoryview(buffer)
    pos = 0
    for path, offset, size, file_size, mtime_ns in _ranges(filelist):
        with open(path, "rb") as fd:
            st = os.fstat(fd.fileno())
            if file_size is not None:
                assert st.st_size == file_size, "file size mismatch for file '%s'" % path
                assert st.st_mtime_ns >= mtime_ns, "timestamp mismatch for file '%s'" % path
            assert size <= len(view) - pos
            fd.seek(offset)
            got = fd.readinto(view[pos:pos + size])
            assert got == size
            pos += size
    return pos


def write_filelist_into(filelist, buffer):
    """Write the block's byte ranges back from buffer (fs.py:214-245) and put back each whole or
    final file's recorded mtime (access time kept). Returns the number of bytes written."""
    view = memoryview(buffer)
    pos = 0
    for path, offset, size, _file_size, mtime_ns in _ranges(filelist):
        with open(path, "r+b") as fd:
            fd.seek(offset)
            assert size <= len(view) - pos
            assert fd.write(view[pos:pos + size]) == size
            pos += size
            atime_ns = os.fstat(fd.fileno()).st_atime_ns
        if mtime_ns is not None:
            os.utime(path, ns=(atime_ns, mtime_ns))
    return pos
