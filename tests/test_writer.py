"""Plot and checkpoint files in the reference's PDI/HDF5 layout
(src/io/WriterPDI.cpp, user/pdi_*_xdmf.yml), written from host arrays with no
device: names and shapes of the datasets (read back with h5dump), values of
the auxiliary variables, the XDMF description, and the checkpoint round trip."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import kfvm_amd
from helpers import config, init_text

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _text(tmp_path, **kw):
    c = config(**kw)
    return init_text(c) + f"io:datadir {tmp_path}\nio:basename run\n", c


def _h5(path, name):
    out = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", "-m", "%.17g", path], check=True,
                         capture_output=True, text=True).stdout
    dims = re.search(r"DATASPACE\s+(SIMPLE \{ \( ([0-9, ]+) \)|SCALAR)", out)
    shape = tuple(int(x) for x in dims.group(2).split(",")) if dims.group(2) else ()
    body = out[out.index("DATA {") + 6:]
    vals = np.array([float(x) for x in re.findall(r"-?[0-9.]+(?:e[-+]?[0-9]+)?", body.split("}")[0])])
    return shape, vals


def _state(rng, nv, shape):
    U = rng.uniform(-1, 1, (nv,) + shape)
    U[0] = rng.uniform(0.5, 2.0, shape)
    U[4] = rng.uniform(5.0, 8.0, shape)
    return U


@pytest.mark.skipif(H5DUMP is None, reason="needs h5dump")
def test_hydro_plot_file_datasets(tmp_path):
    text, c = _text(tmp_path, n=(6, 5, 4), hi=(1.0, 2.0, 0.5))
    rng = np.random.default_rng(1)
    U = _state(rng, 5, (4, 5, 6))
    F = np.stack([rng.integers(0, 2, (4, 5, 6)).astype(float), rng.uniform(0, 1, (4, 5, 6))])
    p = kfvm_amd.write_output_file(text, U, F, step=3, time=0.25)
    assert p == f"{tmp_path}/R2_NX6_NY5_NZ4/plot/run_0000003.h5"
    assert os.path.exists(p[:-3] + ".xmf")
    for v, name in enumerate(["dens", "momx", "momy", "momz", "etot"]):
        shape, vals = _h5(p, name)
        assert shape == (4, 5, 6) and np.array_equal(vals, U[v].ravel())
    assert np.array_equal(_h5(p, "weno")[1], F[0].ravel())
    assert np.array_equal(_h5(p, "fomx")[1], F[1].ravel())
    g = kfvm_amd.parse_init(text)["gamma"]  # the init file's float-rounded gamma
    ekin = 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]
    np.testing.assert_allclose(_h5(p, "velx")[1], (U[1] / U[0]).ravel(), rtol=1e-15)
    np.testing.assert_allclose(_h5(p, "eint")[1], (U[4] - ekin).ravel(), rtol=1e-14)
    np.testing.assert_allclose(_h5(p, "pres")[1], ((g - 1) * (U[4] - ekin)).ravel(), rtol=1e-14)
    shape, xc = _h5(p, "xcoord")
    assert shape == (7,) and np.allclose(xc, np.linspace(0, 1, 7))
    assert np.allclose(_h5(p, "ycoord")[1], np.linspace(0, 2, 6))
    assert _h5(p, "time") == ((), np.array([0.25])) or _h5(p, "time")[1][0] == 0.25
    assert _h5(p, "time_step")[1][0] == 3
    xmf = open(p[:-3] + ".xmf").read()
    assert 'NumberOfElements="5 6 7"' in xmf and "run_0000003.h5:/pres" in xmf
    assert 'Dimensions="4 5 6 3"' in xmf


@pytest.mark.skipif(H5DUMP is None, reason="needs h5dump")
def test_mhd_2d_plot_names_and_pressures(tmp_path):
    text, c = _text(tmp_path, dim=2, n=(8, 6, 1), equation="mhd_glm", riemann="llf", problem="orszag_tang",
                    user=(), gamma=5.0 / 3.0)
    rng = np.random.default_rng(2)
    U = _state(rng, 9, (1, 6, 8))
    F = np.zeros((2, 1, 6, 8))
    p = kfvm_amd.write_output_file(text, U, F, step=0, time=0.0)
    assert "/R2_NX8_NY6/plot/run_0000000.h5" in p
    for v, name in enumerate(["dens", "momx", "momy", "momz", "etot", "magx", "magy", "magz", "psi"]):
        shape, vals = _h5(p, name)
        assert shape == (1, 6, 8) and np.array_equal(vals, U[v].ravel())
    prsb = 0.5 * (U[5] ** 2 + U[6] ** 2 + U[7] ** 2)
    eint = U[4] - prsb - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] - 0.5 * U[8] ** 2
    np.testing.assert_allclose(_h5(p, "prsb")[1], prsb.ravel(), rtol=1e-15)
    g = kfvm_amd.parse_init(text)["gamma"]
    np.testing.assert_allclose(_h5(p, "prsg")[1], ((g - 1) * eint).ravel(), rtol=1e-13)
    assert _h5(p, "zcoord")[0] == (2,)


def test_checkpoint_round_trip(tmp_path):
    text, c = _text(tmp_path, dim=2, n=(7, 5, 1))
    rng = np.random.default_rng(3)
    U = _state(rng, 5, (1, 5, 7))
    F = np.stack([np.zeros((1, 5, 7)), rng.uniform(0.5, 1, (1, 5, 7))])
    p = kfvm_amd.write_output_file(text, U, F, step=12, time=0.125, dt=3.5e-3, err_est=0.75, checkpoint=True)
    assert p.endswith("/R2_NX7_NY5/ckpt/run_0000012.h5")
    U2, step, t, dt, e = kfvm_amd.read_checkpoint_file(text, p, U.shape)
    assert np.array_equal(U2, U) and (step, t, dt, e) == (12, 0.125, 3.5e-3, 0.75)
    xmf = open(p[:-3] + ".xmf").read()
    assert ":/pres" not in xmf and ":/etot" in xmf
    with pytest.raises(kfvm_amd.KfvmError):
        kfvm_amd.read_checkpoint_file(text, p + ".missing", U.shape)
