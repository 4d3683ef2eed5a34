"""Restart files in the reference's UDF layout (resume.CONTINUE.Saved_Data,
UDF/define.udf:317-338): host-only writer and reader of the C ABI.

Pinned by: an exact write -> read round trip (17 significant digits); a
hand-written record in loose UDF syntax with known values
(tests/golden/udf_resume_tiny.udf); the resume record of a new run, laid out
as in the reference's UDF/input.udf, which holds no state; a template whose
other records survive the replacement of its resume record; size checks.
"""
import os

import numpy as np
import pytest

from icep_kapsel_amd import abi, restart

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _state(nx, ny, nz, np_, seed, conc=True):
    rng = np.random.default_rng(seed)
    nz_ = 2 * (nz // 2 + 1)
    z0 = rng.standard_normal((nx, ny, nz_)) * 10.0 ** rng.integers(-12, 3, (nx, ny, nz_))
    z1 = rng.standard_normal((nx, ny, nz_))
    ps = rng.standard_normal((abi.P_NFIELDS, 3, np_))
    ps[abi.PFIELD["X"]] = np.abs(ps[abi.PFIELD["X"]])
    q = rng.standard_normal((8, np_))
    q[:4] /= np.linalg.norm(q[:4], axis=0)
    q[4:] /= np.linalg.norm(q[4:], axis=0)
    c = [rng.standard_normal((nx, ny, nz_)), rng.standard_normal((nx, ny, nz_))] if conc else None
    return z0, z1, ps, q, c


@pytest.mark.parametrize("conc", [True, False])
def test_restart_round_trip_exact(tmp_path, conc):
    nx, ny, nz, np_ = 8, 4, 6, 3
    z0, z1, ps, q, c = _state(nx, ny, nz, np_, 3, conc)
    path = tmp_path / "restart.udf"
    restart.write_restart(path, z0, z1, [1e-3, -2.5e-7, 0.0], ps, q, 1234, 61.7000001, conc=c)
    r = restart.read_restart(path, nx, ny, nz, np_, 2 if conc else 0)
    assert np.array_equal(r["zeta0"], z0) and np.array_equal(r["zeta1"], z1)
    assert np.array_equal(r["uk_dc"], [1e-3, -2.5e-7, 0.0])
    assert r["ts"] == 1234 and r["time"] == 61.7000001
    assert np.allclose(r["orientations"], q, rtol=0, atol=2e-16)
    kept = [f for f in abi.PFIELDS if f != "X_PREVIOUS"]
    for f in kept:
        assert np.array_equal(r["particles"][abi.PFIELD[f]], ps[abi.PFIELD[f]]), f
    # not saved: x_previous = x (resume.cxx:409-410)
    assert np.array_equal(r["particles"][abi.PFIELD["X_PREVIOUS"]], ps[abi.PFIELD["X"]])
    if conc:
        assert np.array_equal(r["conc"][0], c[0]) and np.array_equal(r["conc"][1], c[1])
    text = path.read_text()
    assert text.startswith('\\include{"define.udf"}') and 'resume:{"CONTINUE",' in text


def test_restart_reads_hand_written_record():
    r = restart.read_restart(os.path.join(GOLDEN, "udf_resume_tiny.udf"), 1, 1, 2, 1, n_spec=2)
    assert r["ts"] == 12 and r["time"] == 0.6
    assert np.array_equal(r["uk_dc"], [0.25, -0.5, 0.125])
    z = np.array([1.0e-3, 3, 5, 7]), np.array([-2.0e-3, 4, 6, 8])
    assert np.array_equal(r["zeta0"][0, 0], z[0]) and np.array_equal(r["zeta1"][0, 0], z[1])
    assert np.array_equal(r["conc"][0][0, 0], [1, 2, 3, 4]) and np.array_equal(r["conc"][1][0, 0], [5, 6, 7, 8])
    p = r["particles"][:, :, 0]
    f = abi.PFIELD
    assert np.array_equal(p[f["X"]], [1.5, 2.5, 3.5]) and np.array_equal(p[f["X_NOPBC"]], [11.5, 12.5, 13.5])
    assert np.array_equal(p[f["X_PREVIOUS"]], [1.5, 2.5, 3.5])
    assert np.array_equal(p[f["V"]], [0.1, 0.2, 0.3]) and np.array_equal(p[f["V_OLD"]], [0.01, 0.02, 0.03])
    assert np.array_equal(p[f["F_HYDRO"]], [1, 2, 3]) and np.array_equal(p[f["F_HYDRO_PREVIOUS"]], [4, 5, 6])
    assert np.array_equal(p[f["FR"]], [7, 8, 9]) and np.array_equal(p[f["FR_PREVIOUS"]], [10, 11, 12])
    assert np.array_equal(p[f["OMEGA"]], [-0.1, -0.2, -0.3])
    assert np.array_equal(p[f["OMEGA_OLD"]], [-0.4, -0.5, -0.6])
    assert np.array_equal(p[f["TORQUE_HYDRO"]], [13, 14, 15])
    assert np.array_equal(p[f["TORQUE_HYDRO_PREVIOUS"]], [16, 17, 18])
    assert np.array_equal(p[f["TORQUE_R"]], [19, 20, 21]) and np.array_equal(p[f["TORQUE_R_PREVIOUS"]], [22, 23, 24])
    assert np.allclose(r["orientations"][:, 0], [0.6, 0, 0.8, 0, 0, 0, 0, 1], rtol=0, atol=1e-16)
    # concentrations are read only for an electrolyte
    r0 = restart.read_restart(os.path.join(GOLDEN, "udf_resume_tiny.udf"), 1, 1, 2, 1, n_spec=0)
    assert "conc" not in r0 and r0["ts"] == 12


def test_restart_template_and_new_run(tmp_path):
    tmpl = os.path.join(GOLDEN, "udf_template_resume_new.udf")
    with pytest.raises(OSError, match="no saved state"):
        restart.read_restart(tmpl, 4, 4, 4, 2)
    z0, z1, ps, q, _ = _state(4, 4, 4, 2, 9, conc=False)
    path = tmp_path / "continue.udf"
    restart.write_restart(path, z0, z1, [0.0, 0.0, 0.0], ps, q, 7, 0.35, template=tmpl)
    text = path.read_text()
    head = open(tmpl).read().split("resume:")[0]
    assert text.startswith(head)  # every other record kept as it was
    assert text.rstrip().endswith("\\end{data}") and text.count("resume:") == 1
    r = restart.read_restart(path, 4, 4, 4, 2)
    assert r["ts"] == 7 and np.array_equal(r["zeta1"], z1)


def test_restart_size_checks(tmp_path):
    z0, z1, ps, q, c = _state(4, 4, 4, 2, 5)
    path = tmp_path / "r.udf"
    restart.write_restart(path, z0, z1, [0, 0, 0], ps, q, 1, 0.1, conc=c)
    for args, msg in [((8, 4, 4, 2), "Zeta"), ((4, 4, 4, 3), "number of particles"),
                      ((4, 4, 8, 2), "Zeta"), ((4, 2, 4, 2), "Zeta"), ((2, 4, 4, 2), "Zeta")]:
        with pytest.raises(OSError, match=msg):
            restart.read_restart(path, *args)
    restart.write_restart(path, z0, z1, [0, 0, 0], ps, q, 1, 0.1)
    with pytest.raises(OSError, match="no ion concentrations"):
        restart.read_restart(path, 4, 4, 4, 2, n_spec=2)
