"""Iterative reconstructors (``aprict.iterators`` counterpart)."""

from .batched import BatchedSIRT
from .landweber_vram import Landweber_flatpanel
from .sirt_vram import SIRT_flatpanel
from .tv_gd import TVGradientDescent, power_iteration

__all__ = ["BatchedSIRT", "Landweber_flatpanel", "SIRT_flatpanel", "TVGradientDescent", "power_iteration"]
