"""Cone-beam flat-panel operators (``aprict.cb`` counterpart)."""

from . import flatpanel, ramp_filter
from .ramp_filter import FBPFilterer, RampFilterer

__all__ = ["FBPFilterer", "RampFilterer", "flatpanel", "ramp_filter"]
