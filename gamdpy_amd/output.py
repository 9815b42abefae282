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


class NpzOutput(MemoryOutput):
    """MemoryOutput that writes itself to an .npz file on flush/close: the
    file stand-in for '*.h5' storage where h5py is not importable.  Datasets
    keep their names; the attributes travel as one JSON string under
    '__attrs__'.  load_output reads it back."""

    def __init__(self, path):
        super().__init__()
        self.path = path

    def flush(self):
        import json
        attrs = {k: (np.asarray(v).tolist() if isinstance(v, np.ndarray) or np.isscalar(v) else v)
                 for k, v in self.attrs.items()}
        np.savez(self.path, __attrs__=np.array(json.dumps(attrs)), **self._data)

    def close(self):
        self.flush()


def load_output(path):
    """An output written by NpzOutput, as a MemoryOutput (extract_scalars works on it)."""
    import json
    out = MemoryOutput()
    with np.load(path, allow_pickle=False) as f:
        for k in f.files:
            if k == '__attrs__':
                out.attrs = json.loads(str(f[k]))
            else:
                out._data[k] = f[k]
    return out


def open_output(storage, mode="w"):
    """'memory' -> MemoryOutput; '*.h5' -> h5py.File when h5py exists, else an
    NpzOutput writing '<name>.npz' beside it (with a warning); anything else
    -> MemoryOutput."""
    if storage == 'memory':
        return MemoryOutput()
    if isinstance(storage, str) and storage.endswith('.h5'):
        try:
            import h5py
        except ImportError:
            import warnings
            path = storage[:-3] + '.npz'
            warnings.warn(f"h5py is not importable: the output of {storage!r} is written to {path!r} "
                          "(gamdpy_amd.output.load_output reads it)", RuntimeWarning, stacklevel=3)
            return NpzOutput(path)
        return h5py.File(storage, mode)
    return MemoryOutput()
