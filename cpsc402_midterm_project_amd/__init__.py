"""MI355X-native mixed-precision dense linear solver (drop-in for the mpsolve hot path of
aloupeep/CPSC402_Midterm_Project): HIP kernels for gfx950 behind a C ABI (include/mpsolve.h).

`from cpsc402_midterm_project_amd import Context` loads libmpsolve.so on first use and raises ImportError
when it is not built; there is no CPU fallback.  Submodules that do not need the library (build, problems)
import without it.
"""
_ABI_NAMES = {"Context", "DeviceBuffer", "MPSolveError", "Problem", "Server", "methods", "lib", "EXPORTS", "LIB_PATH"}

__all__ = ["Context", "DeviceBuffer", "MPSolveError", "Problem", "Server", "methods", "randsvd"]


def __getattr__(name):
    if name in _ABI_NAMES:
        from . import _abi
        return getattr(_abi, name)
    if name == "randsvd":
        from .problems import randsvd
        return randsvd
    raise AttributeError(name)
