"""The HDF5 converter of the run loop's npz dumps writes the datasets and the
dimension order of SaveHDF5Parallel (src/3dim/saving.jl:57-146).  h5py is not
installed here, so a recording stand-in for its File object checks what would
be written, and the real error message is checked without it."""
import sys
import types

import numpy as np
import pytest

from verona_amd import driver


class _File:
    store = {}

    def __init__(self, path, mode):
        self.path, self.mode = path, mode
        if mode == "w":
            _File.store[path] = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False

    def create_dataset(self, name, data):
        _File.store[self.path][name] = np.array(data)

    def __getitem__(self, name):
        return _File.store[self.path][name]


@pytest.fixture
def fake_h5py(monkeypatch):
    mod = types.ModuleType("h5py")
    mod.File = _File
    monkeypatch.setitem(sys.modules, "h5py", mod)
    return _File


def _dump(tmp_path, shape):
    P = np.random.default_rng(1).random(shape)
    path = str(tmp_path / "dump3.npz")
    driver.save_dump(path, P, 0.25, (0.1, 0.2, 0.3)[: len(shape) - 1], step=17)
    return path, P


def test_layout_matches_the_reference_writer(tmp_path, fake_h5py):
    path, P = _dump(tmp_path, (5, 10, 9, 8))
    out = driver.dump_to_hdf5(path)
    f = fake_h5py.store[out]
    interior = P[:, 3:-3, 3:-3, 3:-3]
    # HDF5.jl stores the Julia (5, X, Y, Z) array with reversed dimensions
    assert f["data"].shape == (2, 3, 4, 5)
    assert f["data"][1, 2, 3, 4] == interior[4, 3, 2, 1]
    assert f["data"].tobytes() == np.asfortranarray(interior).tobytes(order="F")
    assert float(f["T"]) == 0.25 and tuple(f["grid"]) == (0.1, 0.2, 0.3) and int(f["step"]) == 17
    data, T, grid = driver.load_hdf5_dump(out)
    assert np.array_equal(data, interior) and T == 0.25 and grid == (0.1, 0.2, 0.3)


def test_2d_layout(tmp_path, fake_h5py):
    path, P = _dump(tmp_path, (4, 11, 9))
    out = driver.dump_to_hdf5(path, str(tmp_path / "x.h5"))
    assert fake_h5py.store[out]["data"].shape == (3, 5, 4)
    assert tuple(fake_h5py.store[out]["grid"]) == (0.1, 0.2)


def test_clear_error_without_h5py(tmp_path, monkeypatch):
    path, _ = _dump(tmp_path, (5, 8, 8, 8))
    monkeypatch.setitem(sys.modules, "h5py", None)
    with pytest.raises(ImportError, match="dump<i>.npz"):
        driver.dump_to_hdf5(path)
