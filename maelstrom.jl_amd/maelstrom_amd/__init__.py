"""Maelstrom's hot path on MI355X: the Python mirror of the reference's Julia API.

`FlipBasin`, `initialize`, `step` (and `advance`), and the phases `resample`,
`particles2grid`, `project`, `grid2particles`; `SparseBasin` and the same phases with
`refine` in `maelstrom_amd.sparse`. All run in the gfx950 HIP library through its C
ABI (include/maelstrom_hip.h).
"""
from ._lib import MaelstromError, Context, default_context, LIBPATH, lib
from .basin import Basin, Race, Stir, Wave, narutobasin, narutowaves, waves, surface, oseen, balanced
from .dense import (FlipBasin, flipgrid, initialize, resample, particles2grid, grid2particles, project, step, advance,
                    volume, maxspeed, classify, redistance, grid2particles_adjoint, project_adjoint,
                    particles2grid_adjoint, advect, filltrapped, forces, extrapolate)
from . import sparse
from .checkpoint import save_checkpoint, load_checkpoint
from .sparse import SparseBasin, sparsegrid

__all__ = ["MaelstromError", "Context", "default_context", "LIBPATH", "lib", "Basin", "Race", "Stir", "Wave",
           "narutobasin", "narutowaves", "waves", "surface",
           "oseen", "balanced", "FlipBasin", "flipgrid", "initialize", "resample", "particles2grid",
           "grid2particles", "project", "step", "advance", "volume", "maxspeed", "classify", "redistance",
           "grid2particles_adjoint", "project_adjoint", "particles2grid_adjoint", "advect", "filltrapped",
           "forces", "extrapolate", "sparse",
           "SparseBasin", "sparsegrid", "save_checkpoint", "load_checkpoint"]
