"""MI355X LIBOR Market Model Monte Carlo engine (HIP kernels behind a C ABI)."""
from . import _abi
from ._abi import (CalculationException, IllegalArgumentException, IllegalStateException,
                   UnsupportedOperationException, LmmContext, Product, device_info, library_path,
                   mt19937_doubles, mt19937_words)

__all__ = [
    "CalculationException", "IllegalArgumentException", "IllegalStateException",
    "UnsupportedOperationException", "LmmContext", "Product", "device_info", "library_path",
    "mt19937_doubles", "mt19937_words",
]
