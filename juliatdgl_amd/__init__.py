"""MI355X ladder-cycle stepper for the N-band TDGL solver of juliaTDGL release 2.0.

The device path is the HIP library built from csrc/ (C ABI in include/tdglhip.h); this package
holds the host side that mirrors TDGLGPU's interface (tdglgpu.py), the Stepper derivation the
kernels read (scheme.py), host States (state.py) and the ctypes binding (hip.py, abi.py).
"""
from .scheme import (Band, Grid, Masks, Model, Stepper, box_masks, hole_masks, make_stepper, notch_masks,
                     sloped_wedge_masks, two_band, wedge_masks)
from .state import State, initial_state, seeded_state

__all__ = ["Band", "Grid", "Masks", "Model", "Stepper", "box_masks", "hole_masks", "make_stepper",
           "notch_masks", "sloped_wedge_masks", "two_band", "wedge_masks", "State", "initial_state",
           "seeded_state"]
