"""Iterative reconstructors (``aprict.iterators`` counterpart)."""

from .landweber_vram import Landweber_flatpanel
from .sirt_vram import SIRT_flatpanel

__all__ = ["Landweber_flatpanel", "SIRT_flatpanel"]
