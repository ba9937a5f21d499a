"""fire_amd — MI355X-native Fire-X fixed-step solver (hand-written HIP for gfx950).

The compute path lives in libfirex_hip.so (C ABI: include/firex.h); this
package is its Python mirror.  See fire_amd.solver.
"""
from .solver import (  # noqa: F401
    FIELDS,
    REDUCTIONS,
    FireXAbortError,
    FireXError,
    FireXHipSolver,
    FireXLibraryError,
    abi_version,
    checksum_bytes,
    config_from_tier,
    device_count,
    estimate_memory,
    load_library,
    preflight,
    seeded_emitters,
    split_offline_delta,
)
