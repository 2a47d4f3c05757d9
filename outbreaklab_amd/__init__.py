"""MI355X-native model time step of OutbreakLab (gfx950 HIP kernels behind a C ABI).

The product is liboutbreak_hip.so (include/outbreak_hip.h); this package is the
host-side mirror of the reference's RegionalModel / PoissonDirect / Sounding
API over it.
"""
from .grid import GRID_PRESETS, custom_grid, grid_bytes, grid_pad, levels_for  # noqa: F401
from .model import PoissonDirect, RegionalModel  # noqa: F401
from .sounding import analyze, base_state, parse_sharppy, wk_sounding_text  # noqa: F401
from .uniforms import MODEL_FIELDS, UBlock, default_params, write_uniforms  # noqa: F401
