"""kielflow's HDF5 layout (IO.cpp:60-504, 1175-1338) through
libkielflow_io.so, and axis::index_at; host code only, no GPU.

The layout is checked with h5dump, a tool of the HDF5 installation that
shares no code with the library under test."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import kielflow_amd as kf
from kielflow_amd import h5io

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _axes():
    return [kf.axis(kf.EQST, 8, -4.0, 8.0), kf.axis(kf.SIST, 12, -6.0, 12.0, 1.0), kf.axis(kf.HYST, 6, -3.0, 6.0, 2.0)]


def _write(path, rng):
    axes = _axes()
    h5io.file_create(path, "test-version")
    h5io.save_grid(axes, path)
    h5io.save_time(1.25, path)
    h5io.save_parameters({"M": 1.1, "tau": 0.2, "theta": 50.0, "mu": 0.01, "beta": 0.5}, path)
    f = rng.standard_normal((8, 12, 6))
    h5io.save_field_real("Ux", f, path)
    p = rng.standard_normal((8, 12))
    h5io.save_plane("Ph", p, path)
    return axes, f, p


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(0)
    path = str(tmp_path / "a.h5")
    axes, f, p = _write(path, rng)
    got = h5io.load_grid(path)
    for a, b in zip(axes, got):
        assert (a.type, a.n, a.l0, a.L, a.m) == (b.type, b.n, b.l0, b.L, b.m)
    assert h5io.load_time(path) == 1.25
    assert h5io.load_parameters(path) == {"M": 1.1, "tau": 0.2, "theta": 50.0, "mu": 0.01, "beta": 0.5}
    assert np.array_equal(h5io.load_field_real(path, "Ux"), f)
    assert np.array_equal(h5io.load_slice("Ph", path), p)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
def test_layout_matches_the_reference_names_and_types(tmp_path):
    rng = np.random.default_rng(1)
    path = str(tmp_path / "b.h5")
    _write(path, rng)
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    # group and dataset names of file_create / save_axis / save_parameters /
    # save_time / save_field_real / save_slice
    for g in ("program", "parameters", "domain", "data"):
        assert f'GROUP "{g}"' in out
    for name in ("version", "date", "path", "time", "mach number M", "damping ratio tau", "temperature ratio theta",
                 "velocity diffusion mu", "magnetization beta", "Ux", "Ph"):
        assert f'DATASET "{name}"' in out, name
    for c in "xyz":
        for pre in ("type_", "m_", "l0_", "N_", "L_"):
            assert f'DATASET "{pre}{c}"' in out, pre + c

    def block(name):
        i = out.index(f'DATASET "{name}"')
        return out[i:out.index("}", i)]

    for c in "xyz":
        assert "H5T_STD_I32LE" in block(f"type_{c}") and "H5T_STD_I32LE" in block(f"N_{c}")
        for pre in ("m_", "l0_", "L_"):
            assert "H5T_IEEE_F64LE" in block(pre + c)
    assert "SCALAR" in block("time") and "H5T_IEEE_F64LE" in block("time")
    assert "H5T_VARIABLE" in block("version")
    assert "SIMPLE { ( 8, 12, 6 ) / ( 8, 12, 6 )" in block("Ux")
    assert "SIMPLE { ( 8, 12 ) / ( 8, 12 )" in block("Ph")
    # the values, as h5dump prints them
    vals = subprocess.run([H5DUMP, "-d", "/domain/N_y", path], capture_output=True, text=True, check=True).stdout
    assert "(0): 12" in vals
    vals = subprocess.run([H5DUMP, "-d", "/parameters/temperature ratio theta", path], capture_output=True, text=True,
                          check=True).stdout
    assert "(0): 50" in vals


def test_errors_are_reported(tmp_path):
    path = str(tmp_path / "c.h5")
    with pytest.raises(h5io.IoError, match="no such file"):
        h5io.load_time(path)
    h5io.file_create(path)
    with pytest.raises(h5io.IoError, match="missing dataset time"):
        h5io.load_time(path)
    h5io.save_time(0.0, path)
    with pytest.raises(h5io.IoError, match="dataset exists"):
        h5io.save_time(1.0, path)
    h5io.save_grid(_axes(), path)
    h5io.save_field_real("ni", np.ones((8, 12, 6)), path)
    with pytest.raises(h5io.IoError, match="dataset exists"):
        h5io.save_field_real("ni", np.ones((8, 12, 6)), path)
    out = np.empty((8, 12, 5))
    lib = h5io.load_io()
    assert lib.kf_io_load_field_real(path.encode(), b"ni", out.ctypes.data, 8, 12, 5) != 0
    assert "extent 6 on axis 2, expected 5" in lib.kf_io_last_error().decode()
    with pytest.raises(h5io.IoError, match="rank 3, expected 2"):
        h5io.load_slice("ni", path)
    bad = kf.axis(7, 8, 0.0, 1.0)
    with pytest.raises(h5io.IoError, match="axis type"):
        h5io.save_grid([bad, bad, bad], str(tmp_path / "d.h5"))
    # file_create truncates
    h5io.file_create(path)
    with pytest.raises(h5io.IoError, match="missing dataset time"):
        h5io.load_time(path)


def test_save_grid_creates_the_file(tmp_path):
    """save_axis creates the file when it does not exist (IO.cpp:99-100)."""
    path = str(tmp_path / "e.h5")
    h5io.save_grid(_axes(), path)
    assert [a.n for a in h5io.load_grid(path)] == [8, 12, 6]


def test_i4():
    """I4 (IO.cpp:1281-1287): width 4, zero fill, right aligned."""
    assert h5io.I4(7) == "0007"
    assert h5io.I4(120) == "0120"
    assert h5io.I4(12345) == "12345"
    assert h5io.I4(-1) == "00-1"
    assert os.path.basename(h5io.slice_path(30)) == "slice_0030.h5"


@pytest.mark.parametrize("kind,m", [(kf.EQST, 0.0), (kf.SIST, 16 * (1 - 1 / 3) / (2 * np.pi)), (kf.HYST, 3.0)])
def test_index_at_inverts_val_at(kind, m):
    n, L = 32, 16.0
    a = kf.axis(kind, n, -L / 2, L, m)
    for i in range(n):
        x = kf.axis_values(a, i)[0]
        assert kf.axis_index_at(a, x) == i, (i, x)
    assert kf.axis_index_at(a, 0.0) == n // 2
    if kind == kf.HYST:
        # axis_HySt.cpp:126-146: whole periods add N
        assert kf.axis_index_at(a, kf.axis_values(a, 3)[0] + L) == 3 + n
        assert kf.axis_index_at(a, kf.axis_values(a, 3)[0] - 2 * L) == 3 - 2 * n
    if kind == kf.SIST:
        # axis_SiSt.cpp:102-111: wrapped into the box, N maps to 0
        assert kf.axis_index_at(a, kf.axis_values(a, 5)[0] + L) == 5
        assert kf.axis_index_at(a, L / 2 - 1e-9) == 0


def test_index_at_eqst_truncates():
    """axis_EqSt.cpp:63-66 converts N (x - l0) / L to int: toward zero,
    with no periodic wrap."""
    a = kf.axis(kf.EQST, 16, -8.0, 16.0)
    assert kf.axis_index_at(a, -8.0 + 0.99) == 0
    assert kf.axis_index_at(a, -8.0 + 1.99) == 1
    assert kf.axis_index_at(a, -8.5) == 0
    assert kf.axis_index_at(a, 8.5) == 16
    with pytest.raises(kf.KfError):
        kf.axis_index_at(a, float("nan"))
