"""MI355X-native KAPSEL SPM hot path: C ABI over gfx950 HIP kernels."""
from .abi import FIELD, FIELDS, PFIELD, PFIELDS, make_params, make_time  # noqa: F401
