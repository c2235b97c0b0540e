"""The C ABI boundary (CPU): the built libraries load and export every entry
point include/*.h declares.  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

import amrcheck as A
import miniamr_amd as M
from conftest import require


def header_symbols(header=M.INCLUDE, prefix="amr_dev_"):
    with open(header) as f:
        text = f.read()
    decl = r"^\s*(?:int|void|double|const char \*|amr_\w+ \*)\s*(%s\w+)\s*\(" % prefix
    return sorted(set(re.findall(decl, text, re.M)))


INC = os.path.dirname(M.INCLUDE)
HOST_SO = os.path.join(M.PKG, "lib", "libminiamr_host.so")


def device_symbols():
    """Entry points of the device library: amr_dev.h and the device half of
    amr_batch.h."""
    batch = [s for s in header_symbols(os.path.join(INC, "amr_batch.h"), "amr_batch_")
             if not s.startswith("amr_batch_run_")]
    return header_symbols() + batch


def adjoint_symbols():
    """include/amr_adjoint.h: the HIP library only (the CPU oracle has just
    amr_dev_tape, which refuses)."""
    h = os.path.join(INC, "amr_adjoint.h")
    return header_symbols(h, "amr_dev_") + header_symbols(h, "amr_adj_")


def host_symbols():
    """Entry points of the driver library: amr_run.h and the driver half of
    amr_batch.h."""
    return header_symbols(os.path.join(INC, "amr_run.h"), "amr_run_") + \
        header_symbols(os.path.join(INC, "amr_batch.h"), "amr_batch_run_")


def test_headers_declare_what_is_expected():
    assert len(device_symbols()) == 30 + 13
    assert len(adjoint_symbols()) == 11 + 10
    assert "amr_batch_timestep" in device_symbols()
    assert "amr_run_step" in host_symbols() and "amr_batch_run_step" in host_symbols()


def test_header_matches_python_binding():
    assert header_symbols() == sorted(M.ABI_SYMBOLS)


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_library_exports_every_symbol(which):
    path = M.LIB_PATH if which == "product" else A.ORACLE_SO
    require(path, "run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name in device_symbols():
        assert hasattr(lib, name), "%s does not export %s" % (path, name)
    for name in adjoint_symbols() if which == "product" else ["amr_dev_tape"]:
        assert hasattr(lib, name), "%s does not export %s" % (path, name)
    M.load_library(path)   # argtypes for every symbol bind


def test_driver_library_exports_every_symbol():
    require(HOST_SO, "run __graft_entry__.build()")
    lib = ctypes.CDLL(HOST_SO)
    for name in host_symbols():
        assert hasattr(lib, name), "libminiamr_host.so does not export %s" % name
    M.load_host_library(HOST_SO)


def test_config_struct_layout():
    assert ctypes.sizeof(M.Config) == 8 * 4
    assert M.FACE_JOB.itemsize == 24 and M.SPLIT_JOB.itemsize == 16


def test_product_fails_loudly_without_a_gpu():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present; covered by the GPU tests")
    require(M.LIB_PATH, "run __graft_entry__.build()")
    with pytest.raises(M.AmrError):
        M.Device(4, 4, 4, 1, 8)


def test_driver_binary_links_the_hip_library():
    require(M.EXE_PATH, "run __graft_entry__.build()")
    out = subprocess.run(["ldd", M.EXE_PATH], stdout=subprocess.PIPE).stdout.decode()
    assert "libminiamr_dev.so" in out and "oracle" not in out


def test_library_calls_return_errors_instead_of_exiting():
    """An invalid command line and (without a GPU) a failed device
    creation come back as exceptions with the reference's message; the
    calling process keeps running."""
    require(HOST_SO, "run __graft_entry__.build()")
    host = M.load_host_library(HOST_SO)
    with pytest.raises(M.AmrError, match="block size in x direction must be even"):
        M.Run("--nx 5 --num_tsteps 1", host=host)
    with pytest.raises(M.AmrError, match="Unknown input parameter --bogus"):
        M.Run("--nx 4 --bogus 3", host=host)
    with pytest.raises(M.AmrError, match="object number greater than num_objects"):
        M.Run("--num_objects 0 --object 2 0 0 0 0 0 0 0 0 0 0 0 0 0", host=host)
    with pytest.raises(M.AmrError, match="must be even"):
        M.BatchRun(["--nx 4 --num_tsteps 1", "--ny 3 --num_tsteps 1"], host=host)
    assert host.amr_run_last_error() != b""
    if not os.path.exists("/dev/kfd"):
        with pytest.raises(M.AmrError, match="device"):
            M.Run("--nx 4 --ny 4 --nz 4 --num_tsteps 1", host=host)
