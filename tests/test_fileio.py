"""On-disk formats of src/file-io.c on the host: HDF5 fluid files, the "density-kji" input
and PETSc binary vectors.  The byte layout of the binary vector is checked against
PETSc's VecView format directly; the HDF5 files are read back through the library and,
where the HDF5 command-line tools are installed, described by h5dump (an independent
reader of the same files).  The reference's tests hold no files of either kind."""
import os
import shutil
import struct
import subprocess
import sys

import numpy as np
import pytest

from ecsprt_amd import fileio

HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

needs_hdf5 = pytest.mark.skipif(not fileio.hdf5_available()[0], reason="no HDF5 library on this host")


@needs_hdf5
@pytest.mark.parametrize("shape", [(6, 5, 4), (7, 9)])
def test_fluid_file_round_trip(tmp_path, shape):
    rng = np.random.default_rng(1)
    m = rng.normal(size=shape + (4,))
    rhs = rng.normal(size=shape)
    phi = rng.normal(size=shape)
    path = tmp_path / "fluid-test.hdf"
    fileio.write_fluid_hdf5(path, m, rhs, phi)
    assert open(path, "rb").read(8) == HDF5_SIGNATURE
    got = fileio.read_fluid_hdf5(path)
    for f, name in enumerate(fileio.FLUID_FIELDS):
        assert got[name].shape == shape
        np.testing.assert_array_equal(got[name], m[..., f])
    np.testing.assert_array_equal(got["rhs"], rhs)
    np.testing.assert_array_equal(got["potential"], phi)


@needs_hdf5
@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_fluid_file_structure_by_h5dump(tmp_path):
    """Dataset names, [Nz][Ny][Nx] shape, IEEE little-endian doubles and the
    "timestepping" attribute that PETSc's VecView attaches, as h5dump describes them."""
    shape = (3, 4, 5)
    z = np.arange(60, dtype=np.float64).reshape(shape)
    path = tmp_path / "fluid.hdf"
    fileio.write_fluid_hdf5(path, np.stack([z, z + 1, z + 2, z + 3], axis=-1), z + 4, z + 5)
    out = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True).stdout
    for name in ("density", "x flux", "y flux", "z flux", "rhs", "potential"):
        assert f'DATASET "{name}"' in out
    assert out.count("H5T_IEEE_F64LE") == 6
    assert out.count("SIMPLE { ( 3, 4, 5 ) / ( 3, 4, 5 ) }") == 6
    assert out.count('ATTRIBUTE "timestepping"') == 6
    data = subprocess.run([H5DUMP, "-d", "/x flux", "-w", "0", str(path)], capture_output=True, text=True,
                          check=True).stdout
    assert "(0,0,0): 1, 2, 3, 4, 5," in data and "(2,3,0): 56, 57, 58, 59, 60" in data


@needs_hdf5
def test_dataset_errors(tmp_path):
    path = tmp_path / "d.hdf"
    fileio.write_dataset(path, "density-kji", np.ones((2, 3, 4)))
    assert fileio.read_dataset(path, "density-kji").shape == (2, 3, 4)
    with pytest.raises(fileio.FileIOError, match="no dataset"):
        fileio.read_dataset(path, "density")
    with pytest.raises(fileio.FileIOError, match="cannot open"):
        fileio.read_dataset(tmp_path / "missing.hdf", "density-kji")


def test_binary_vector_layout(tmp_path):
    """PETSc VecView binary: big-endian int32 classid 1211214 and length, then
    big-endian doubles; block size in the .info file."""
    v = np.array([1.5, -2.0, 3.25, 1e-300, np.pi, 0.0])
    path = tmp_path / "position-test.bin"
    fileio.write_vec_binary(path, v, bs=3)
    raw = open(path, "rb").read()
    assert len(raw) == 8 + 8 * v.size
    assert struct.unpack(">ii", raw[:8]) == (1211214, v.size)
    np.testing.assert_array_equal(np.frombuffer(raw[8:], dtype=">f8"), v)
    assert open(str(path) + ".info").read() == "-vecload_block_size 3\n"
    got, bs = fileio.read_vec_binary(path)
    np.testing.assert_array_equal(got, v)
    assert bs == 3


def test_binary_vector_rejects_other_files(tmp_path):
    path = tmp_path / "x.bin"
    path.write_bytes(struct.pack(">ii", 1211216, 2) + b"\0" * 16)
    with pytest.raises(fileio.FileIOError, match="not a PETSc binary vector"):
        fileio.read_vec_binary(path)
    path.write_bytes(struct.pack(">ii", 1211214, 5) + b"\0" * 16)
    with pytest.raises(fileio.FileIOError, match="shorter"):
        fileio.read_vec_binary(path)


def test_missing_hdf5_library_is_an_error(tmp_path):
    """With ECSPRT_HDF5_LIBRARY naming no library, the HDF5 calls return ECSPRT_ERR_STATE
    (8) and say why; binary vectors still work."""
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {os.path.dirname(os.path.dirname(os.path.abspath(__file__)))!r})\n"
        "from ecsprt_amd import fileio\n"
        "ok, why = fileio.hdf5_available()\n"
        "assert not ok and 'no HDF5' in why, why\n"
        "try:\n"
        f"    fileio.write_dataset({str(tmp_path / 'a.hdf')!r}, 'density-kji', np.ones((2, 2, 2)))\n"
        "except fileio.FileIOError as e:\n"
        "    assert str(e).startswith('code 8'), e\n"
        "else:\n"
        "    raise SystemExit('write succeeded without HDF5')\n"
        f"fileio.write_vec_binary({str(tmp_path / 'v.bin')!r}, np.ones(3))\n"
        "print('ok')\n")
    env = dict(os.environ, ECSPRT_HDF5_LIBRARY=str(tmp_path / "no-such-libhdf5.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
