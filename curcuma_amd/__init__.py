"""MI355X-native GFN-FF energy and gradient engine (drop-in for ``-method gfnff``)."""
from .gfnff import GFNFFBatch, GFNFFMethod, Mol  # noqa: F401

__all__ = ["GFNFFBatch", "GFNFFMethod", "Mol"]
