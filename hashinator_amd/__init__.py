"""hashinator_amd -- MI355X (gfx950) implementation of Hashinator's accelerated path.

The product is the C-ABI library ``libhashinator_amd.so`` (include/hashinator_amd.h)
and the drop-in C++ headers under include/hashinator and include/splitvector.
This package is the Python host side: the ctypes binding (``_lib``), a mirror
of the reference's ``Hashmap`` accelerated API (``hashmap``), ``split::tools``
(``tools``), and the benchmark workloads (``workloads``).
"""
from . import _lib
from ._lib import HnError, hn_info, hn_pred, hn_tune
from .hashmap import EMPTY, TOMBSTONE, Hashmap, as_device, status, targets, to_numpy

__all__ = [
    "_lib",
    "HnError",
    "hn_info",
    "hn_pred",
    "hn_tune",
    "EMPTY",
    "TOMBSTONE",
    "Hashmap",
    "as_device",
    "status",
    "targets",
    "to_numpy",
]


def library_path() -> str:
    return _lib.LIB_PATH
