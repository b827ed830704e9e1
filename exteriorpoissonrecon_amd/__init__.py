"""MI355X-native (gfx950 HIP) ExteriorPoissonRecon solve path.

`Reconstructor` mirrors ExteriorPoisson::Reconstructor; the work runs in
libepr_hip.so (C ABI: include/epr.h).
"""
from ._lib import EXPORTS, EprError, LIB_PATH, default_params, load  # noqa: F401
from .reconstructor import (Reconstructor, autotune_batch, build_batch, set_tune_table, solve_batch,  # noqa: F401
                            time_batch_sweeps, tune_table)
from .levelset import level_set, subgrid_triangulation  # noqa: F401
from .mg import MGSolver  # noqa: F401
