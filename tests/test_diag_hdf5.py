"""QPAD-layout HDF5 files of the diagnostics writer (libqpad_diag.so), read
back with the HDF5 library itself.  The layout restates QPAD's writer:
wrattr_file / wrattr_dataset (hdf5io_class.f03:264-335), the 2D grid dataset
(:420-470), and the fixed character(len=100) fields of its hdf5file type
(:20-40), whose attributes are 100-character space-padded strings."""
import os

import numpy as np
import pytest

import h5read
from qpad_amd import QpadError, configs, diag

pytestmark = pytest.mark.skipif(not os.path.exists(diag.DIAG_LIB), reason="libqpad_diag.so not built (no HDF5)")


def test_grid_file_layout(tmp_path):
    box = diag.Box(0.0, 5.0, -5.0, 5.0, 10.0)
    data = np.arange(3 * 4, dtype=float).reshape(3, 4) * 0.25 - 1.0
    path = str(tmp_path / "Fields" / "Psi" / "Re1" / "psi_00000007.h5")
    diag.write_grid(path, "psi", data, 7, box, "m_ec^2/e", "\\Psi")
    with h5read.File(path) as f:
        # QPAD's merge_modes_tools read the second key: AXIS, then the dataset
        assert f.names("/") == ["AXIS", "psi"]
        assert f.names("/AXIS") == ["AXIS1", "AXIS2"]
        assert np.array_equal(f.dataset("psi"), data)
        assert np.array_equal(f.dataset("AXIS/AXIS1"), [0.0, 5.0])
        assert np.array_equal(f.dataset("AXIS/AXIS2"), [-5.0, 5.0])
        name, n, pad = f.attr("/", "NAME")
        assert (name.rstrip(), n, pad) == ("psi", 100, 2)          # H5T_STR_SPACEPAD
        assert f.attr("/", "TYPE")[0].rstrip() == "grid" and f.attr("/", "TYPE")[1] == 100
        assert f.attr("/", "TIME UNIT")[0].rstrip() == "1 / \\omega_p"
        assert f.attr("/", "ITER") == 7 and f.attr("/", "DT") == 10.0 and f.attr("/", "TIME") == 70.0
        assert f.attr("AXIS/AXIS1", "NAME")[:2] == ("r", 1)
        assert f.attr("AXIS/AXIS2", "LONG_NAME")[:2] == ("\\xi", 3)
        assert f.attr("AXIS/AXIS1", "UNITS")[0] == "c / \\omega_p" and f.attr("AXIS/AXIS1", "TYPE")[0] == "linear"
        assert f.attr("psi", "UNITS")[0].rstrip() == "m_ec^2/e" and f.attr("psi", "LONG_NAME")[0].rstrip() == "\\Psi"


def test_deck_diag_sections_are_read():
    d = configs.tri_gaussian()
    d["field"] = {"diag": [{"name": ["psi_cyl_m", "ez_cyl_m"], "ndump": 1}, {"name": ["bz_cyl_m"], "ndump": 0}]}
    d["species"][0]["diag"] = [{"name": ["charge_cyl_m"], "ndump": 2}, {"name": ["raw"], "ndump": 1, "psample": 10}]
    assert diag._entries(d["field"]) == [("psi_cyl_m", 1, 1), ("ez_cyl_m", 1, 1)]
    assert diag._entries(d["species"][0]) == [("charge_cyl_m", 2, 1), ("raw", 1, 10)]
    assert diag._entries(d["beam"][0]) == []


def test_bad_grid_arguments_are_refused(tmp_path):
    with pytest.raises(QpadError, match="bad grid-file"):
        diag.write_grid(str(tmp_path / "x.h5"), "psi", np.zeros((2, 2)), 0, diag.Box(1.0, 0.0, 0.0, 1.0, 1.0))


def test_beam_file_reader(tmp_path):
    """qpad_diag_read_beam_file reads QPAD's particle-file datasets, rank 1
    or rank 2 with one extent 1 (QPAD reads them as (n, 1) columns,
    fdist3d_file_class.f03:204-215), and refuses ragged or missing ones."""
    rng = np.random.default_rng(5)
    cols = {k: rng.standard_normal(37) for k in diag.BEAM_FILE_KEYS}
    p1 = str(tmp_path / "flat.h5")
    h5read.write_datasets(p1, cols)
    got = diag.read_beam_file(p1)
    for k in diag.BEAM_FILE_KEYS:
        assert np.array_equal(got[k], cols[k]), k
    p2 = str(tmp_path / "column.h5")
    h5read.write_datasets(p2, {k: v.reshape(1, -1) if k < "p" else v.reshape(-1, 1) for k, v in cols.items()})
    got = diag.read_beam_file(p2)
    for k in diag.BEAM_FILE_KEYS:
        assert np.array_equal(got[k], cols[k]), k
    p3 = str(tmp_path / "empty.h5")
    h5read.write_datasets(p3, {k: np.zeros(0) for k in diag.BEAM_FILE_KEYS})
    assert all(len(v) == 0 for v in diag.read_beam_file(p3).values())
    ragged = dict(cols, q=cols["q"][:-1])
    p4 = str(tmp_path / "ragged.h5")
    h5read.write_datasets(p4, ragged)
    with pytest.raises(QpadError, match="differ in length"):
        diag.read_beam_file(p4)
    p5 = str(tmp_path / "missing.h5")
    h5read.write_datasets(p5, {k: v for k, v in cols.items() if k != "p2"})
    with pytest.raises(QpadError, match="no dataset p2"):
        diag.read_beam_file(p5)
    p6 = str(tmp_path / "matrix.h5")
    h5read.write_datasets(p6, dict(cols, x1=np.zeros((3, 4))))
    with pytest.raises(QpadError, match="not a list"):
        diag.read_beam_file(p6)
    with pytest.raises(QpadError, match="cannot open"):
        diag.read_beam_file(str(tmp_path / "absent.h5"))
