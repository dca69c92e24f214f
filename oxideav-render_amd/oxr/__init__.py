"""oxr — MI355X (gfx950) path tracer behind oxideav-render's PathTrace
backend: Python mirror of the reference API over liboxr.so's C ABI."""
from . import _abi  # noqa: F401
