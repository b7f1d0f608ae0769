"""sphugo_amd: the sphugo SPH step (kNN, density, force, leapfrog) as
hand-written HIP kernels for MI355X (gfx950) behind a C ABI
(include/sphugo.h).  `sphugo_amd.sim` mirrors the reference's Go `sim`
package API; `sphugo_amd._lib` is the raw ctypes binding."""
from . import ics  # noqa: F401

__all__ = ["ics", "sim", "_lib"]
