"""MI355X-native CardioTrust epoch hot path (gfx950 HIP kernels behind a C ABI)."""
from .config import Algorithm, APDerivative, Optimizer  # noqa: F401
