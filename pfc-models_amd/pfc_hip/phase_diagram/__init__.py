"""Phase-diagram relaxations on the hip backend: the driver (`q0star`); the
model adapters live beside the models (`pfc_hip.models.stdpfc_hip.phase_diagram`,
`pfc_hip.models.logpfc_hip.phase_diagram`)."""
from .q0star import (FAILED_STOPS, FMIN_FILE, QSCAN_FILE, STATUS_ERROR, STATUS_INSUFFICIENT_POINTS,
                     STATUS_NO_LOCAL_MINIMUM, STATUS_OK, STOP_CONVERGED, STOP_MAX_DEPTH, STOP_MIN_AMPLITUDE,
                     STOP_NAN, STOP_POOR_FIT, STOP_TOTALT, Diagnostics, InsufficientPoints, MaxDepthReached,
                     MinimizeFreeEnergy, NoLocalMinimum, QScanSettings, RelaxSettings, scan_density,
                     status_from_exception)
