"""Volumes: voxel grids (``Volume``) and FORBILD phantoms (``FBVolume``)."""

from . import forbild
from .volume import FBVolume, Volume

__all__ = ["FBVolume", "Volume", "forbild"]
