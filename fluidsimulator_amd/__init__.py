"""MI355X-native Position Based Fluids backend (HIP, gfx950, FP32)."""
from .params import Params, State  # noqa: F401
