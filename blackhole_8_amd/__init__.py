"""blackhole_8_amd — MI355X (gfx950) renderer for the per-pixel photon march
of lackhole/blackhole_8, behind a C ABI (include/bh_api.h)."""
from . import abi  # noqa: F401
from .render import BHError, Path, Renderer, Ring  # noqa: F401

__all__ = ["abi", "Renderer", "Path", "Ring", "BHError"]
