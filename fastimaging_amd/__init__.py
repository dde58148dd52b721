"""MI355X-native Slow Transients Pipeline imaging + source-finding hot path.

Host-side Python mirror of the reference's stp:: interface over the C-ABI
library fastimaging_amd/lib/libstp_hip.so (HIP kernels for gfx950).
"""
from . import capi  # noqa: F401
