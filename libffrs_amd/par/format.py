"""`.ffrs` writer and reader (ffrs/par/format.py:31-204)."""
import base64
import json
import os
import re
import shlex

from ..util import b64hex, b64hex_dec

SIGNATURE = b"\x89ffrs\r\n"


def _path_token(name: str) -> str:
    """JSON string; the quotes dropped when it holds no whitespace (format.py:54-62)."""
    s = json.dumps(name, ensure_ascii=False)
    return s if re.search(r"\s", s) else s[1:-1]


def _hash_token(digest: bytes) -> str:
    return "b2:" + base64.urlsafe_b64encode(digest).rstrip(b"=").decode("ascii")


class Writer:
    """Writes a `.ffrs` file: `write_header(codec)`, then `write_block(filelist, ecc)` per block,
    where filelist entries are ("f", size, mtime_ns, digest, path), ("c", size, offset, path) or
    ("fc", size, offset, mtime_ns, digest, path)."""

    signature = SIGNATURE

    def __init__(self, path):
        self.path = path
        self._dir = ""

    def __enter__(self):
        self.file = open(self.path, "wb")
        return self

    def __exit__(self, *exc):
        self.file.close()

    def write_header(self, codec):
        self.file.write(SIGNATURE + codec.serialize() + b"\n")

    def _name(self, path):
        d, f = os.path.split(path)
        d = _path_token(d)
        if d != self._dir:  # a directory line when the directory changes (format.py:64-67)
            self._dir = d
            self.file.write(f"p p:{d}\n".encode("utf8"))
        return _path_token(f)

    def write_block(self, filelist, data):
        for entry in filelist:
            kind = entry[0]
            if kind == "f":
                _, size, mtime_ns, digest, path = entry
                name = self._name(path)
                line = f"f s:{b64hex(size)} m:{b64hex(mtime_ns)} h:{_hash_token(digest)} p:{name}\n"
            elif kind == "c":
                _, size, offset, path = entry
                name = self._name(path)
                line = f"c s:{b64hex(size)} o:{b64hex(offset)} p:{name}\n"
            elif kind == "fc":
                _, size, offset, mtime_ns, digest, path = entry
                name = self._name(path)
                line = f"fc s:{b64hex(size)} o:{b64hex(offset)} m:{b64hex(mtime_ns)} h:{_hash_token(digest)} p:{name}\n"
            else:
                raise NotImplementedError(kind)
            self.file.write(line.encode("utf8"))
        self.file.write(f"d s:{b64hex(len(data))}\n".encode("ascii"))
        self.file.write(data)
        self.file.write(b"\n")


class Reader:
    """Reads a `.ffrs` file: `read_header()` returns the codec, `blocks(ecc)` yields each block's
    file list after filling `ecc` with its ecc bytes."""

    signature = SIGNATURE

    def __init__(self, path, **simd):
        self.path = path
        self.simd = simd
        self._dir = ""

    def __enter__(self):
        self.file = open(self.path, "rb")
        self.file_size = os.fstat(self.file.fileno()).st_size
        return self

    def __exit__(self, *exc):
        self.file.close()

    def read_header(self):
        assert self.file.readline(16) == SIGNATURE
        from .. import CIRC16
        line = self.file.readline(4096)
        assert line[-1:] == b"\n", line
        return CIRC16.deserialize(line, **self.simd)

    @staticmethod
    def _path(token):
        if token.startswith('"'):
            assert len(token) >= 2 and token.endswith('"'), "invalid path encoding"
            return json.loads(token)
        return json.loads(f'"{token}"')

    @staticmethod
    def _hash(token):
        algo, value = token.split(":", 1)
        assert algo == "b2"
        return base64.urlsafe_b64decode(value + "==").hex()

    def _line(self, line):
        assert line.endswith("\n")
        kind, rest = line.split(" ", 1)
        params = {}
        for arg in shlex.split(rest):
            key, value = arg.split(":", 1)
            if key in ("m", "n", "o", "s"):
                params[key] = b64hex_dec(value)
            elif key == "p":
                params[key] = self._path(value)
            elif key == "h":
                params[key] = self._hash(value)
            else:
                raise NotImplementedError(key)
        return kind, params

    def read_block(self, ecc):
        files = []
        while True:
            raw = self.file.readline(4096)
            assert raw[-1:] == b"\n", raw
            if len(raw) == 1:
                return files
            kind, p = self._line(raw.decode("utf8"))
            if kind == "p":
                self._dir = p["p"]
            elif kind == "f":
                files.append(("f", p["s"], p["m"], p["h"], os.path.join(self._dir, p["p"])))
            elif kind == "c":
                files.append(("c", p["s"], p["o"], os.path.join(self._dir, p["p"])))
            elif kind == "fc":
                files.append(("fc", p["s"], p["o"], p["m"], p["h"], os.path.join(self._dir, p["p"])))
            elif kind == "d":
                assert p["s"] == len(ecc)
                assert self.file.readinto(ecc) == len(ecc)
                assert self.file.read(1) == b"\n"
                return files
            else:
                raise NotImplementedError(kind)

    def blocks(self, ecc):
        while self.file.tell() < self.file_size:
            yield self.read_block(ecc)
