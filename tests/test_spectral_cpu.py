"""The pseudo-spectral Laplacian's transform phases (geopvi_amd/csrc/
fwi2d_spectral.h) compiled for the CPU (tests/spectral_check.cpp runs every
phase for each thread of a workgroup in turn, over the kernels' own
workgroup tiling) against numpy's FFT and the PSM oracle's Laplacian.

Pins: each radix's butterfly against numpy's DFT in both directions; the
three-phase Laplacian (row pairs packed into one complex transform, column
transforms with the float |k|^2, Hermitian repack) against
oracle.psm_oracle.laplacian_unnormalised on even, odd and mixed sizes, with
partial last groups of rows and columns.  Test infrastructure only: nothing
here runs the product path."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import psm_oracle as P

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "geopvi_amd", "csrc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    out = str(tmp_path_factory.mktemp("spectral") / "libspectral_check.so")
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, "-o", out,
                    os.path.join(HERE, "spectral_check.cpp")], check=True)
    L = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    L.spec_plan.argtypes = [ctypes.c_int, ip]
    L.spec_fft.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    L.spec_laplacian.argtypes = [fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def plan(lib, n):
    r = (ctypes.c_int * 16)()
    k = lib.spec_plan(n, r)
    return None if k < 0 else list(r[:k])


def test_plans(lib):
    assert plan(lib, 420) == [4, 3, 5, 7]
    assert plan(lib, 220) == [4, 5, 11]
    assert plan(lib, 8) == [4, 2]
    assert plan(lib, 1) == []
    assert plan(lib, 2 * 3 * 13 * 13) == [2, 3, 13, 13]
    for n in (17, 34, 2 * 19, 23 * 4, 0):
        assert plan(lib, n) is None
    # LDS tiling: twiddle table (half length for even n) + two buffers within
    # 4096 complex values
    assert lib.spec_pairs_per_group(420) == 2 and lib.spec_cols_per_group(220) == 4
    assert lib.spec_pairs_per_group(1365) == 1 and lib.spec_pairs_per_group(1367) == 0
    assert lib.spec_pairs_per_group(1638) == 1 and lib.spec_pairs_per_group(1640) == 0


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 11, 13, 12, 60, 64, 220, 420, 2 * 3 * 5 * 7 * 11, 13 * 13 * 2])
@pytest.mark.parametrize("sgn", [1, -1])
def test_fft_against_numpy(lib, n, sgn):
    rng = np.random.default_rng(n)
    nseq = 3
    z = rng.standard_normal((nseq, n)) + 1j * rng.standard_normal((nseq, n))
    inp = np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(-1))
    out = np.empty_like(inp)
    assert lib.spec_fft(_dp(inp), n, nseq, sgn, 37, _dp(out)) == 0
    got = out.reshape(nseq, n, 2)
    got = got[..., 0] + 1j * got[..., 1]
    want = np.fft.fft(z, axis=1) if sgn > 0 else np.fft.ifft(z, axis=1, norm="forward")
    assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("ntz,ntx,np_,ncol", [
    (220, 420, 2, 4),      # the 200 x 400 model with 10-point absorbing bands
    (220, 420, 3, 5),      # partial last row group and column tile
    (35, 49, 2, 3),        # odd lengths: the last pair holds one row; no Nyquist column
    (26, 39, 4, 8),
    (121, 66, 1, 1),
    (8, 16, 4, 8),
    (1, 6, 1, 2),
])
def test_laplacian_against_the_oracle(lib, ntz, ntx, np_, ncol):
    rng = np.random.default_rng(ntz * 1000 + ntx)
    p1 = rng.standard_normal((ntz, ntx)).astype(np.float32)
    k = P.wavenumber(ntx, ntz, 10.0, 12.5)
    out = np.empty((ntz, ntx), np.float64)
    rc = lib.spec_laplacian(p1.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                            np.ascontiguousarray(k).ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                            ntz, ntx, np_, ncol, 64, _dp(out))
    assert rc == 0
    want = P.laplacian_unnormalised(p1, k)
    scale = np.abs(want).max()
    assert np.abs(out - want).max() <= 1e-13 * scale, np.abs(out - want).max() / scale


def test_laplacian_of_a_plane_wave(lib):
    """A sampled plane wave is an eigenvector: the unnormalised result is
    -ntx*ntz*|k|^2 times it (the float |k|^2 of its wavenumber)."""
    ntz, ntx = 26, 39
    iz, ix = np.meshgrid(np.arange(ntz), np.arange(ntx), indexing="ij")
    p1 = np.cos(2 * np.pi * (3 * iz / ntz + 5 * ix / ntx)).astype(np.float32)
    k = P.wavenumber(ntx, ntz, 10.0, 10.0)
    out = np.empty((ntz, ntx), np.float64)
    assert lib.spec_laplacian(p1.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              k.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ntz, ntx, 4, 8, 64, _dp(out)) == 0
    want = -float(k[3, 5]) * ntx * ntz * p1.astype(np.float64)
    assert np.abs(out - want).max() <= 1e-6 * np.abs(want).max()


def test_unsupported_lengths_are_refused(lib):
    out = np.empty(34 * 20)
    p1 = np.zeros((20, 34), np.float32)
    k = np.zeros((20, 34), np.float32)
    assert lib.spec_laplacian(p1.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              k.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 20, 34, 4, 8, 64, _dp(out)) == -1


def test_bench_reports_the_library_plans(lib):
    """The pseudo-spectral leg of bench.py names the passes of the
    library's plans (and hipFFT where there is none)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_for_plans", os.path.join(os.path.dirname(HERE), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for n in range(1, 700):
        assert bench._radices(n) == plan(lib, n), n
