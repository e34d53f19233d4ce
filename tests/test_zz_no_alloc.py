"""grw_calc_fru_flux_local_noalloc: the flux sums of calc_fru_flux_local.c
without any allocation by the library (its reduction scratch is a fixed device
array), against the context path grw_calc_fru_flux_local on the same FRUs:
bit for bit, on two calls in a row (the second smaller), plus the empty input
and the size limit.  Device buffers here are the test's own (hipMalloc through
the HIP runtime); the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from greenstein2002_amd import _lib
from greenstein2002_amd.fru import FruCell
from greenstein2002_amd.ic import load_stable

pytestmark = pytest.mark.gpu

MAX_FRU = 1_048_576


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # H2D
        return p

    def download(self, p, a: np.ndarray) -> np.ndarray:
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), 2) == 0  # D2H
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def frus(n, seed):
    rng = np.random.default_rng(seed)
    ic = load_stable(n)
    fru = np.ascontiguousarray(ic["fru"] * rng.uniform(0.5, 2.0, (n, 5)))
    lt = np.stack([rng.integers(1, 13, (n, 4)), rng.integers(1, 3, (n, 4))], -1).astype(np.int32)
    ryr = np.ascontiguousarray(ic["ryr"], np.int32)
    ito2 = rng.integers(1, 3, (n, 4)).astype(np.int32)
    return fru, np.ascontiguousarray(lt), ryr, ito2


def context_sums(fru, lt, ryr, ito2):
    n = len(fru)
    with FruCell(n) as c:
        c.seed(1234)
        st = c.parallel_get_FRUs()
        c.initialize_mpi_state(fru, lt, ryr, ito2, st["mti"], st["mt"])
        return c.calc_fru_flux_local()


def test_noalloc_flux_sums_equal_the_context_path():
    L = _lib.load()
    hip = Hip()
    try:
        out = hip.upload(np.zeros(5))
        for n, seed in ((3000, 1), (700, 2)):  # the second call smaller than the first
            fru, lt, ryr, ito2 = frus(n, seed)
            want = context_sums(fru, lt, ryr, ito2)
            d_lt, d_it, d_fru = hip.upload(lt), hip.upload(ito2), hip.upload(fru)
            rc = L.grw_calc_fru_flux_local_noalloc(n, d_lt, d_it, d_fru, out, None)
            assert rc == 0, L.grw_last_error()
            got = hip.download(out, np.zeros(5))
            assert np.array_equal(got, want), (n, got, want)
            assert got[3] == ((np.isin(lt[:, :, 0], (6, 12))) & (lt[:, :, 1] == 2)).sum()
            assert got[4] == (ito2 == 2).sum()
        # no FRUs: zeros
        assert L.grw_calc_fru_flux_local_noalloc(0, None, None, None, out, None) == 0
        assert np.array_equal(hip.download(out, np.ones(5)), np.zeros(5))
        # beyond the scratch array's capacity: refused before any launch
        rc = L.grw_calc_fru_flux_local_noalloc(MAX_FRU + 1, d_lt, d_it, d_fru, out, None)
        assert rc == -1 and b"allocation-free maximum" in L.grw_last_error()
    finally:
        hip.free()
