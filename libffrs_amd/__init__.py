"""libffrs_amd -- MI355X (gfx950) Reed-Solomon encode / repair behind libffrs' API.

    from libffrs_amd import RS256, RSi16, RepairStatus

RS256 (GF(2^8)) and RSi16 (GF(65537)) keep the reference module's constructors,
properties, ``encode`` / ``repair`` semantics and exceptions; every codec call runs
HIP kernels through the C ABI declared in include/ffrs_gpu.h. Importing this
package loads the compiled pybind11 module ``libffrs_amd.libffrs``; if it is missing
the import fails (build it with ``python libffrs_amd/build.py``).
"""
from . import libffrs as _lib
from .libffrs import (  # noqa: F401
    GF256,
    GFi16,
    NTT,
    RS256,
    RepairStatus,
    __version__,
    abi_version,
    backend,
    compiler_info,
    create_buffer,
    device_info,
    set_logger,
)

from .util import b64hex, b64hex_dec  # noqa: E402


class RSi16(_lib.RSi16):
    """The extension's RSi16 with the Python package's repr (ffrs/__init__.py:24-28)."""

    __doc__ = _lib.RSi16.__doc__

    def __repr__(self):
        extra = f", interleave={self.interleave}" if self.interleave != 1 else ""
        return f"RSi16(block_len={self.rs_block_len}, ecc_len={self.rs_ecc_len}{extra})"


class CIRC16(_lib.CIRC16):
    """The extension's CIRC16 with the Python package's repr and codec line
    (ffrs/__init__.py:31-80)."""

    __doc__ = _lib.CIRC16.__doc__

    def __repr__(self):
        return (f"CIRC({self.inner_block_len}, {self.inner_ecc_len}, {self.outer_block_len}, "
                f"{self.outer_ecc_len}, {self.interleave})")

    def serialize(self):
        """The codec line of a `.ffrs` file, e.g.
        `circ16 ip:g01 ig:3 ib:x4 im:x0 ie:4 op:g01 og:3 ob:80 om:40 oe:40 i:1`."""
        fields = (("ip", self.rsi.gf.field_elements), ("ig", self.rsi.gf.primitive),
                  ("ib", self.inner_block_len), ("im", self.inner_message_len), ("ie", self.inner_ecc_len),
                  ("op", self.rso.gf.field_elements), ("og", self.rso.gf.primitive),
                  ("ob", self.outer_block_len), ("om", self.outer_message_len), ("oe", self.outer_ecc_len),
                  ("i", self.interleave))
        return ("circ16 " + " ".join(f"{k}:{b64hex(v)}" for k, v in fields)).encode("ascii")

    @staticmethod
    def deserialize(buf, **simd):
        """A CIRC16 from its codec line, with the reference's consistency checks (AssertionError
        when one fails). Keyword arguments go to the constructor (the SIMD-width knobs)."""
        text = buf.decode("ascii")
        assert text.startswith("circ16 ")
        p = {}
        for part in text[7:].split():
            key, value = part.split(":", 1)
            p[key] = b64hex_dec(value)
        assert p["ip"] == p["op"] == 0x10001
        assert p["ig"] == p["og"]
        assert p["ib"] == p["im"] + p["ie"]
        assert p["ob"] == p["om"] + p["oe"]
        return CIRC16(inner_block_len=p["ib"], inner_ecc_len=p["ie"], outer_block_len=p["ob"],
                      outer_ecc_len=p["oe"], interleave=p["i"], primitive=p["ig"], **simd)


__all__ = ["CIRC16", "GF256", "GFi16", "NTT", "RS256", "RSi16", "RepairStatus", "b64hex", "b64hex_dec",
           "create_buffer", "set_logger", "__version__"]
