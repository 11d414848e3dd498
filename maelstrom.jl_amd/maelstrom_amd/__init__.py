"""Maelstrom's hot path on MI355X: the Python mirror of the reference's Julia API.

`FlipBasin`, `initialize`, `step` (and `advance`), and the phases `resample`,
`particles2grid`, `project`, `grid2particles`, all running in the gfx950 HIP library
through its C ABI (include/maelstrom_hip.h).
"""
from ._lib import MaelstromError, Context, default_context, LIBPATH, lib
from .basin import Basin, narutobasin, surface, oseen, balanced
from .dense import (FlipBasin, flipgrid, initialize, resample, particles2grid, grid2particles, project, step, advance,
                    volume, maxspeed, classify, redistance)

__all__ = ["MaelstromError", "Context", "default_context", "LIBPATH", "lib", "Basin", "narutobasin", "surface",
           "oseen", "balanced", "FlipBasin", "flipgrid", "initialize", "resample", "particles2grid",
           "grid2particles", "project", "step", "advance", "volume", "maxspeed", "classify", "redistance"]
