"""MI355X-native Brownian-dynamics engine (Ermak-McCammon, bidisperse WCA).

The compute path is libbd_hip.so (hand-written gfx950 HIP kernels behind the C
ABI of include/bd_abi.h); this package only loads it and converts numpy arrays.
"""
from .abi import (BDParams, BDGeometry, BDObs, BDError, BD_POT_WCA, BD_POT_WCA_MODIFIED,
                  BD_POT_HARMONIC, BD_TEMP_LINEAR, BD_TEMP_SINE)
from .engine import Engine, make_params, box_side

__all__ = ["BDParams", "BDGeometry", "BDObs", "BDError", "Engine", "make_params", "box_side",
           "BD_POT_WCA", "BD_POT_WCA_MODIFIED", "BD_POT_HARMONIC", "BD_TEMP_LINEAR",
           "BD_TEMP_SINE"]
