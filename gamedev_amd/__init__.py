"""MI355X-native swarm collision operator (GameDev's evolve_local_collisions path) for gfx950.

The product is the HIP library built per model directory (gamedev_amd/lib/libgdx_col_<model>.so)
behind the C ABI of include/gdx_col.h; this package is its host-side mirror.
"""
from ._lib import GdxError, STATUS, load, lib_path  # noqa: F401
from .collision import CollisionOperator, Majorant, MortonIndex, evolve_batch, evolve_ensemble, model_info  # noqa: F401

__all__ = ["CollisionOperator", "Majorant", "MortonIndex", "evolve_batch", "evolve_ensemble", "model_info", "GdxError", "STATUS", "load", "lib_path"]
