"""md_engine_amd: MI355X-native per-step force path of a DASH-style MD engine.

The compute lives in libmdx.so (HIP, gfx950) behind the C ABI in
include/mdx.h; ``_lib.Engine`` is the thin ctypes handle on one context.
"""
from . import _lib, systems  # noqa: F401
from ._lib import BONDED, E3B, ENERGY, PAIR, PPPM, SURROGATE, VIRIAL, VIRIAL_TRACE, Engine, MdxError, run_batch, tune_batch  # noqa: F401
