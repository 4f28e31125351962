"""padeops_amd -- MI355X-native compact-difference operators (cd06, cd10, cf90).

Drop-in for PadeOps' compact-operator path: the operator types
(``cd06``, ``cd10``, ``cf90``) and the ``derivatives`` / ``filters`` facades
keep the reference's call surface; the arithmetic runs in hand-written HIP
kernels for gfx950 behind the C ABI in ``include/padeops_hip.h``.
"""
from ._lib import PadeError  # noqa: F401
from .operators import cd06, cd06stagg, cd10, cf90, gaussian, fortran_empty, as_fortran  # noqa: F401
from .facades import derivatives, filters  # noqa: F401

__version__ = "0.1.0"
