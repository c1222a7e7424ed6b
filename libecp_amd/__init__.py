"""MI355X-native evaluator of semi-local ECP integrals (type 1 + type 2, FP64)
behind the reference libECP C API.  The numerical work runs in the HIP
library ``libecp_amd.so``; this package holds its ctypes bindings
(:mod:`libecp_amd.native`), the build script and seeded workloads."""
from .workloads import Molecule  # noqa: F401
