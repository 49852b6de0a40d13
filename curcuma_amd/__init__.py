"""MI355X-native GFN-FF energy and gradient engine (drop-in for ``-method gfnff``)."""
from .gfnff import GFNFFMethod, Mol  # noqa: F401

__all__ = ["GFNFFMethod", "Mol"]
