"""Python side of the MI355X SPH step library: deck generation, dump I/O and ctypes bindings."""
from .deck import (Deck, column_collapse_2d, soil_water_column_3d, model_zoo_2d, moving_box_2d,  # noqa: F401
                   art_stress_block, slope_2d, ssi_block, tran_box, drag_box, air_water_box,
                   slope_pile_3d, vibration_box, rain_box, eob_box)
from .dumpio import read_dump  # noqa: F401
