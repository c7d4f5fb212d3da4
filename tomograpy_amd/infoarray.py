"""A numpy array that carries a FITS-like ``header`` attribute.

The reference passes maps and image stacks as fitsarray.InfoArray objects
(an unvendored dependency, tomograpy/siddon.py:66): arrays whose ``header``
is a dict (maps) or a list of per-view dicts (stacks).  This is the minimal
equivalent: the header follows views, slices and ufunc results.
"""
import numpy as np


class InfoArray(np.ndarray):
    """ndarray subclass with a ``header`` attribute."""

    def __new__(cls, data, header=None, dtype=None):
        obj = np.asarray(data, dtype=dtype).view(cls)
        obj.header = header
        return obj

    def __array_finalize__(self, obj):
        self.header = getattr(obj, "header", None)


def asinfoarray(x, header=None):
    """Views `x` as an InfoArray, keeping its header unless one is given."""
    out = np.asarray(x).view(InfoArray)
    out.header = header if header is not None else getattr(x, "header", None)
    return out
