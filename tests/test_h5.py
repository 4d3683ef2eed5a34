"""HDF5 field and particle dumps (libkapsel_h5.so, the reference's
hdf5_writer layout, output_writer.cxx:41-575): a writer round trip read back
through the same library, values checked against the arrays written (fields
exactly in double, particle data to float rounding, the QR tensor against the
rotation of a known quaternion), and the structure checked with HDF5's own
h5dump where it is installed.  Host only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from icep_kapsel_amd import abi

h5 = pytest.importorskip("icep_kapsel_amd.h5")
pytestmark = pytest.mark.skipif(not os.path.exists(h5.LIB_PATH), reason="libkapsel_h5.so not built (no HDF5)")


def _padded(real):
    nx, ny, nz = real.shape
    a = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    a[:, :, :nz] = real
    return a


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(h5.LIB_PATH)
    declared = abi.declared_symbols(os.path.join(os.path.dirname(abi.LIB_PATH), "..", "include", "kapsel_h5.h"))
    assert len(declared) == 8
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("crop", [None, ((2, 1, 3), (5, 5, 4), (2, 2, 3))])
def test_writer_round_trip(tmp_path, crop):
    shape, dx, dt, n = (16, 12, 20), 0.5, 0.05, 5
    rng = np.random.default_rng(2)
    u = [rng.standard_normal(shape) for _ in range(3)]
    phi = rng.uniform(0, 1, shape)
    st = rng.standard_normal((abi.P_NFIELDS, 3, n))
    q = rng.standard_normal((8, n))
    plist = [4, 0, 2]
    spec = [0, 1, 0, 1, 1]
    w = h5.Writer(str(tmp_path), "kapsel", shape, dx, dt, n, print_list=plist, spec=spec, crop=crop)
    for f in range(2):
        w.frame({"u_x": _padded(u[0] + f), "u_y": _padded(u[1]), "u_z": _padded(u[2]), "phi": _padded(phi)},
                st + f, q)
    w.close()
    path = str(tmp_path / "kapsel.h5")
    if crop is None:
        sl = tuple(slice(0, m) for m in shape)
        cnt = shape
        origin = (0.0, 0.0, 0.0)
    else:
        start, cnt, stride = crop
        sl = tuple(slice(start[d], start[d] + (cnt[d] - 1) * stride[d] + 1, stride[d]) for d in range(3))
        origin = tuple(start[d] * dx for d in range(3))
    assert h5.read(path, "/@dx")[0] == np.float32(dx)
    assert list(h5.read(path, "/@nxnynz")) == list(cnt)
    assert np.allclose(h5.read(path, "/@origin"), origin)
    assert h5.read(path, "/@nump")[0] == 3 and h5.read(path, "/@nump_total")[0] == n
    assert list(h5.read(path, "/system_data/particle/id")) == plist
    assert list(h5.read(path, "/system_data/particle/spec")) == [spec[i] for i in plist]
    gx = np.arange(shape[0])[:, None, None] * dx + 0 * u[0]
    assert np.array_equal(h5.read(path, "/system_data/field/x"), gx[sl].ravel())
    for f in range(2):
        fr = f"/trajectory_data/frame_{f}"
        assert h5.read(path, fr + "@time")[0] == np.float32(np.float32(f) * np.float32(dt))
        assert np.array_equal(h5.read(path, fr + "/field/u_x"), (u[0] + f)[sl].ravel())
        assert np.array_equal(h5.read(path, fr + "/field/phi"), phi[sl].ravel())
        for name, field in (("pos", "X"), ("pos_raw", "X_NOPBC"), ("vel", "V"), ("force_h", "F_HYDRO_PREVIOUS"),
                            ("force_r", "FR_PREVIOUS"), ("omega", "OMEGA"),
                            ("torque_h", "TORQUE_HYDRO_PREVIOUS"), ("torque_r", "TORQUE_R_PREVIOUS")):
            want = (st + f)[abi.PFIELD[field]][:, plist].T.astype(np.float32)
            assert np.array_equal(h5.read(path, fr + "/particle/" + name), want.ravel().astype(np.float64)), name
        qr = h5.read(path, fr + "/particle/QR").reshape(3, 3, 3)
        for j, i in enumerate(plist):
            s, x, y, z = q[:4, i] / np.linalg.norm(q[:4, i])
            # the rotation that q applies to a vector v: v + 2 s (w x v) + 2 w x (w x v), w = (x, y, z)
            wv = np.array([x, y, z])
            R = np.stack([e + 2 * s * np.cross(wv, e) + 2 * np.cross(wv, np.cross(wv, e)) for e in np.eye(3)], 1)
            assert np.abs(qr[j] - R.T).max() < 1e-6   # row c of the dataset = column c of R
    conf = open(tmp_path / "kapsel.conf").read()
    assert "[Grid Dimensions]\n %d %d %d" % tuple(cnt) in conf and "[Time Series]\n 0.0 0.050000 2" in conf


@pytest.mark.skipif(shutil.which("h5dump") is None and not os.path.exists("/opt/conda/bin/h5dump"),
                    reason="h5dump not installed")
def test_structure_with_h5dump(tmp_path):
    tool = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    shape = (8, 8, 8)
    w = h5.Writer(str(tmp_path), "run", shape, 1.0, 0.1, 2)
    w.frame({"u_x": np.zeros((8, 8, 10)), "phi": np.ones((8, 8, 10))}, np.zeros((abi.P_NFIELDS, 3, 2)))
    w.close()
    out = subprocess.run([tool, "-H", str(tmp_path / "run.h5")], capture_output=True, text=True, check=True).stdout
    for needle in ('GROUP "system_data"', 'GROUP "trajectory_data"', 'GROUP "frame_0"', 'DATASET "u_x"',
                   'ATTRIBUTE "nxnynz"', 'DATASET "QR"', 'DATASET "pos_raw"', 'H5T_IEEE_F32LE', 'H5T_IEEE_F64LE',
                   'DATASPACE  SIMPLE { ( 2, 3, 3 ) / ( 2, 3, 3 ) }', 'DATASPACE  SIMPLE { ( 8, 8, 8 ) / ( 8, 8, 8 ) }'):
        assert needle in out, needle
