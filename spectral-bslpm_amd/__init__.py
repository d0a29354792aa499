"""spectral-BSLPM hot path on MI355X (gfx950): cubed-sphere BLI FMM, direct
sums and the backward semi-Lagrangian BVE step, as hand-written HIP kernels
behind a C ABI (include/sbslpm.h, libsbslpm.so)."""
from ._lib import SbslpmError, exported_symbols, load  # noqa: F401
from .api import Context, HostIrregTree, HostTree, IrregContext, RunConfig, cubed_sphere_midpoints  # noqa: F401
from . import latlon, output, workloads  # noqa: F401
