"""MI355X-native REYES hider: Python side of the C ABI in include/aqsis_hider.h."""
from .abi import RhOptions, RhPrim, RhStats, default_options  # noqa: F401
