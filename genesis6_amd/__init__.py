"""genesis6_amd: MI355X-native rigid-water interaction step (Genesis6 seam).

The product is the HIP library lib/libg6gpu.so behind the C ABI in
include/g6gpu.h; `g6` is its ctypes front end used by Python callers, tests
and bench.py.  The Fortran 90 host layer lives in fortran/.
"""
from . import g6  # noqa: F401
