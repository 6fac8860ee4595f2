"""MI355X-native semi-implicit spectral step of MEUMAPPS-SS.

Host-side mirror of the reference's driver around the C ABI of
``libmeumapps_ss.so`` (include/meumapps_ss.h).
"""
from .params import (MM_FP64, MM_MIXED, Inputs, MMDiag, MMParams, PhaseInput, SetupError,
                     read_input_txt, setup, write_input_txt)

__all__ = [
    "MM_FP64", "MM_MIXED", "Inputs", "MMDiag", "MMParams", "PhaseInput", "SetupError",
    "read_input_txt", "setup", "write_input_txt",
]
