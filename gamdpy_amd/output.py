"""Simulation output store.

The reference keeps ``sim.output`` in an HDF5 file (h5py, driver 'core' for
storage='memory').  This module provides the subset of the h5py File/Group
interface the simulation and ``extract_scalars`` use, backed by NumPy arrays,
and uses h5py for '*.h5' storage when it is importable.
"""
import numpy as np


class MemoryOutput:
    """Dict of NumPy datasets plus an ``attrs`` dict (h5py.File look-alike)."""

    def __init__(self):
        self._data = {}
        self.attrs = {}

    def keys(self):
        return self._data.keys()

    def __contains__(self, key):
        return key in self._data

    def __getitem__(self, key):
        return self._data[key]

    def __setitem__(self, key, value):
        self._data[key] = np.asarray(value)

    def __delitem__(self, key):
        del self._data[key]

    def create_dataset(self, name, shape=None, dtype=np.float32, data=None, chunks=None, **kw):
        if data is not None:
            arr = np.array(data, dtype=dtype)
            if shape is not None:
                arr = arr.reshape(shape)
        else:
            arr = np.zeros(shape, dtype=dtype)
        self._data[name] = arr
        return arr

    def close(self):
        pass


def open_output(storage, mode="w"):
    """'memory' -> MemoryOutput; '*.h5' -> h5py.File when h5py exists, else a
    MemoryOutput (the data then lives only in memory)."""
    if storage == 'memory':
        return MemoryOutput()
    if isinstance(storage, str) and storage.endswith('.h5'):
        try:
            import h5py
        except ImportError:
            return MemoryOutput()
        return h5py.File(storage, mode)
    return MemoryOutput()
