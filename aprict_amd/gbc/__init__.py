"""GBC (generalised backprojection) helpers (``aprict.gbc``)."""

from . import weightinteg

__all__ = ["weightinteg"]
