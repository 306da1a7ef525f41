"""The `.ffrs` parity-file format and the `par` backup / repair front end on top of the GPU codecs.

A `.ffrs` file (ffrs/par/format.py:31-204) holds a signature line, one codec line
(`CIRC16.serialize()`: `circ16 ip:.. ig:.. ib:.. im:.. ie:.. op:.. og:.. ob:.. om:.. oe:.. i:..`),
then blocks: file-list lines (`p`, `f`, `c`, `fc`) describing where the block's message bytes
live on disk, and a `d s:<size>` line followed by the block's ecc bytes and a newline. Integers
are written in base 64 with the alphabet 0-9 a-z A-Z - _ (ffrs/util.py:21-48), hashes as
`b2:` + unpadded url-safe base64 of a 20-byte BLAKE2b digest.

    from libffrs_amd import par
    par.backup(["data.bin"])            # writes data.bin.ffrs
    par.repair("data.bin.ffrs")         # repairs data.bin in place, keeping its mtime

The command-line wrapper, the directory walker and the generic constraint solver of the
reference (ffrs/par/cli.py, fs.py:19-103, solver.py) are not reproduced: `choose_codec` searches
the reference's geometry constraints (ffrs/par/opt.py:29-75) directly.
"""
from ..util import b64hex, b64hex_dec  # noqa: F401
from .format import Reader, Writer  # noqa: F401
from .fs import chunk_filelist, read_filelist_into, write_filelist_into  # noqa: F401
from .opt import choose_codec  # noqa: F401
from .ops import backup, repair  # noqa: F401
