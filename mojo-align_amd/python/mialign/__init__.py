"""mialign — MI355X (gfx950) seed -> extend -> sort path for linear WGBS reads.

Python host layer over libmialign.so (HIP kernels + C ABI, include/mialign.h).
Modules named after the reference's (linear_index, linear_gpu_locate,
linear_gpu_sort, linear_mapper, mojo_linear_pack) mirror its interfaces.
"""

from ._lib import MialignError, Params, default_params, device_count  # noqa: F401
from .core import (  # noqa: F401
    DeviceArray,
    DeviceRecs,
    Event,
    GpuIndex,
    Graph,
    Stream,
    Workspace,
    device_info,
    map_batch,
    map_batch_dev,
    map_forward,
    mate_rescue,
    mate_rescue_dev,
    seed_kmers,
    sort_markdup,
    sync,
)

__version__ = "0.1.0"
