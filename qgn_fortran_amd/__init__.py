"""MI355X (gfx950) implementation of QGN's spectral right-hand side and time
steppers behind a C-ABI (include/qgn.h), with a host-side mirror of the
QGN_Module interface."""
from .model import PI, Batch, Dfti, Ensemble, Params, QGN, vertical_modes  # noqa: F401

__all__ = ["PI", "Batch", "Dfti", "Ensemble", "Params", "QGN", "vertical_modes"]
