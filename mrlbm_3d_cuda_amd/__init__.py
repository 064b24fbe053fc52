"""MI355X-native D3Q27 MR-LBM solver (Python side of the C ABI in include/mrlbm.h)."""
from . import _abi, configs  # noqa: F401
