"""Import the reference package in the development container only.

numba is not installed here; its decorators are replaced by pass-through
stubs.  The stubbed functions (numba DFTs) are not on the Crystal path.
"""
import sys
import types

REF = "/root/reference"


def import_reference():
    if "numba" not in sys.modules:
        nb = types.ModuleType("numba")

        def njit(*args, **kwargs):
            if len(args) == 1 and callable(args[0]) and not kwargs:
                return args[0]
            return lambda f: f

        nb.njit = njit
        nb.jit = njit
        nb.prange = range
        sys.modules["numba"] = nb
    if REF not in sys.path:
        sys.path.insert(0, REF)
    import khepri  # noqa: F401
    return khepri
