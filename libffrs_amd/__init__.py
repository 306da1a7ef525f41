"""libffrs_amd -- MI355X (gfx950) Reed-Solomon encode / repair behind libffrs' API.

    from libffrs_amd import RS256, RSi16, RepairStatus

RS256 (GF(2^8)) and RSi16 (GF(65537)) keep the reference module's constructors,
properties, ``encode`` / ``repair`` semantics and exceptions; every codec call runs
HIP kernels through the C ABI declared in include/ffrs_gpu.h. Importing this
package loads the compiled pybind11 module ``libffrs_amd.libffrs``; if it is missing
the import fails (build it with ``python libffrs_amd/build.py``).
"""
from .libffrs import (  # noqa: F401
    CIRC16,
    GF256,
    GFi16,
    NTT,
    RS256,
    RSi16,
    RepairStatus,
    __version__,
    abi_version,
    backend,
    compiler_info,
    create_buffer,
    device_info,
    set_logger,
)

__all__ = ["CIRC16", "GF256", "GFi16", "NTT", "RS256", "RSi16", "RepairStatus", "create_buffer", "set_logger", "__version__"]
