"""oxr — MI355X (gfx950) path tracer behind oxideav-render's PathTrace
backend: Python mirror of the reference API over liboxr.so's C ABI.

`import oxr` does not load the HIP library; the first PathTracer (or
parity call) does, and raises BackendError when liboxr.so is missing.
"""
from . import _abi  # noqa: F401
from ._lib import BackendError, InvalidOptions, device_count  # noqa: F401
from .pathtracer import (  # noqa: F401
    ACES, BSDF_ONLY, CLAMP, LIGHT_ONLY, MIS, REINHARD, PathTraceOptions, PathTraceRenderer, PathTracer,
    RenderOptions, bsdf_eval, bsdf_sample, make_renderer, sobol_owen_4d,
)
