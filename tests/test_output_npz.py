"""The '*.h5' storage stand-in when h5py is not importable: an .npz file
with the same dataset names and attributes (scalar_saver.py:64-70 layout),
readable by load_output and extract_scalars."""
import numpy as np
import pytest

import gamdpy_amd as gp
from gamdpy_amd import output


def test_npz_output_round_trip(tmp_path):
    path = str(tmp_path / "run.npz")
    out = output.NpzOutput(path)
    out.attrs['dt'] = 0.005
    out.attrs['scalar_names'] = ['U', 'W', 'K']
    out.attrs['simbox_initial'] = np.array([10.0, 10.0, 10.0], np.float32)
    out.create_dataset('scalars', shape=(2, 4, 3), dtype=np.float32)
    out['scalars'][...] = np.arange(24, dtype=np.float32).reshape(2, 4, 3)
    out.create_dataset('ptype', data=np.array([0, 1, 1], np.int32), shape=(3,), dtype=np.int32)
    out.close()
    back = output.load_output(path)
    assert back.attrs['dt'] == 0.005 and back.attrs['scalar_names'] == ['U', 'W', 'K']
    np.testing.assert_array_equal(back['ptype'], [0, 1, 1])
    U, K = gp.extract_scalars(back, ['U', 'K'])
    np.testing.assert_array_equal(U, np.arange(0, 24, 3))
    np.testing.assert_array_equal(K, np.arange(2, 24, 3))


def test_h5_storage_without_h5py_writes_npz(tmp_path, monkeypatch):
    pytest.importorskip("numpy")
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is importable here: '*.h5' storage uses it")
    except ImportError:
        pass
    path = str(tmp_path / "sim.h5")
    with pytest.warns(RuntimeWarning):
        out = output.open_output(path)
    assert isinstance(out, output.NpzOutput) and out.path.endswith("sim.npz")
