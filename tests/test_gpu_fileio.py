"""The reference's file entry points on a device context (src/file-io.c): OutputFluidHDF5
writes the context's moments, forcing and potential; OutputSwarmBinary its ions;
LoadFluidQuantities reads a "density-kji" field into the moments (and the next solve
uses it), or with no file switches the forcing to ComputeSinusoidalRHS3D.  Checked
against the context's own getters and the oracle's RHS, operator and sinusoidal forcing."""
import os

import numpy as np
import pytest

import oracle
from ecsprt_amd import configs, fileio, pic, _lib

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not fileio.hdf5_available()[0], reason="no HDF5 library on this host")]


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_output_fluid_and_swarm(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    wl = configs.small(12, 8)
    say("output fluid + swarm, 12^3 x 8")
    p = pic.setup_context(**wl.cli)
    ctx = pic.Context(p, pic.default_options(collide_mode=_lib.COLLIDE_OFF, ksp_rtol=1e-10))
    ctx.init_particles(wl.np, 7, 1e-3, 2)
    ctx.collect()
    ctx.compute_potential()
    ctx.step(1.0)
    ctx.output_fluid_hdf5("-0001")
    ctx.output_swarm_binary("-0001")
    m, phi = ctx.get_moments(), ctx.get_potential()
    pos, vel = ctx.get_particles()
    ctx.close()
    got = fileio.read_fluid_hdf5("fluid-0001.hdf")
    for f, name in enumerate(fileio.FLUID_FIELDS):
        np.testing.assert_array_equal(got[name], m[..., f])
    np.testing.assert_array_equal(got["potential"], phi)
    # the forcing written is the one the last solve used: ComputeFullRHS3D of the moments
    po = oracle.setup_context(**wl.cli)
    assert rel(got["rhs"], oracle.full_rhs(po, m)) <= 1e-12
    for name, ref in (("position", pos), ("velocity", vel)):
        v, bs = fileio.read_vec_binary(f"{name}-0001.bin")
        assert bs == 3
        np.testing.assert_array_equal(v.reshape(-1, 3), ref)


def test_load_density_then_solve(tmp_path):
    """A "density-kji" field with flux scale (0.5, -1, 2): the moments equal (n, 0.5 n,
    -n, 2 n) exactly, and the potential solved from them has the oracle's residual below
    rtol for the oracle's RHS and operator of those moments."""
    shape = (10, 12, 14)
    cli = dict(configs.base_cli(8, E0=0.05), Nx=shape[2], Ny=shape[1], Nz=shape[0])
    say(f"load density-kji {shape}")
    rng = np.random.default_rng(3)
    dens = np.exp(0.3 * rng.normal(size=shape))
    path = tmp_path / "density.hdf"
    fileio.write_dataset(path, "density-kji", dens)
    ctx = pic.Context(pic.setup_context(**cli), pic.default_options(ksp_rtol=1e-10))
    fs = np.array([0.5, -1.0, 2.0])
    ctx.load_fluid_quantities(fs, str(path))
    m = ctx.get_moments()
    np.testing.assert_array_equal(m[..., 0], dens)
    for f in range(3):
        np.testing.assert_array_equal(m[..., 1 + f], fs[f] * dens)
    ctx.compute_potential()
    phi = ctx.get_potential()
    ctx.close()
    po = oracle.setup_context(**cli)
    b = oracle.full_rhs(po, m)
    r = np.linalg.norm(b - oracle.lhs_apply(po, m, phi)) / np.linalg.norm(b)
    say(f"  ||b - A phi|| / ||b|| = {r:.2e}")
    assert r <= 1.01e-10


def test_load_rejects_wrong_shape_and_missing_dataset(tmp_path):
    cli = configs.base_cli(8)
    ctx = pic.Context(pic.setup_context(**cli))
    path = tmp_path / "d.hdf"
    fileio.write_dataset(path, "density-kji", np.ones((8, 8, 9)))
    with pytest.raises(pic.HipError, match="shape of the grid"):
        ctx.load_fluid_quantities([1.0, 1.0, 1.0], str(path))
    fileio.write_dataset(path, "density", np.ones((8, 8, 8)))
    with pytest.raises(pic.HipError, match="no dataset"):
        ctx.load_fluid_quantities([1.0, 1.0, 1.0], str(path))
    ctx.close()


def test_empty_path_selects_sinusoidal_forcing(tmp_path, monkeypatch):
    """No file: the forcing is n0 scale cos(2 pi x) cos(2 pi y) cos(2 pi z) at the cell
    centres, mean removed (the oracle's restatement to 1e-13), and the potential solves it."""
    monkeypatch.chdir(tmp_path)
    cli = dict(configs.base_cli(16, E0=0.05), dx=1.0 / 16, dy=1.0 / 16, dz=1.0 / 16)
    p = pic.setup_context(**cli)
    po = oracle.setup_context(**cli)
    ctx = pic.Context(p, pic.default_options(ksp_rtol=1e-10))
    m = np.zeros((16, 16, 16, 4))
    m[..., 0] = 1.0
    ctx.set_moments(m)
    ctx.load_fluid_quantities([0.0, 0.0, 0.0], "")
    ctx.compute_potential()
    ctx.output_fluid_hdf5("")
    ctx.close()
    got = fileio.read_fluid_hdf5("fluid.hdf")
    ref = oracle.sinusoidal_rhs(po)
    say(f"sinusoidal forcing: rel err {rel(got['rhs'], ref):.2e}")
    assert rel(got["rhs"], ref) <= 1e-13
    r = np.linalg.norm(ref - oracle.lhs_apply(po, m, got["potential"])) / np.linalg.norm(ref)
    assert r <= 1.01e-10
