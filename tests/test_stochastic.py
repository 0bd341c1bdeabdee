"""Noise sources (lib/src/qcd_stochastic.c:8-58) and the random-number
interface (lib/src/qcd_rng.c:1-33).  The reference's numbers come from GSL's
ranlux389, which is absent here: the exact stream is "parity unpinned".
These tests pin what the reference fixes besides the stream: the seed
derivation per rank, the value set {+1, +i, -1, -i} in the order of
qcd_stochastic.c:10-14, the loop order in which draws fill the vector, the
time dilution, and uniform frequencies."""
import ctypes

import numpy as np
import pytest

import qcd_amd as Q

Z4 = np.array([1, 1j, -1, -1j])


def draws(lib, n):
    return np.array([lib.qcd_get_rand_4() for _ in range(n)])


def test_rng_stream_seeding_and_uniformity():
    lib = Q.lib()
    lib.qcd_rng_init(1234, 0)
    a = draws(lib, 40000)
    lib.qcd_rng_init(1234, 0)
    assert np.array_equal(draws(lib, 40000), a)
    lib.qcd_rng_init(1235, 0)
    assert not np.array_equal(draws(lib, 1000), a[:1000])
    assert set(np.unique(a)) <= {0, 1, 2, 3}
    counts = np.bincount(a, minlength=4)
    sigma = np.sqrt(40000 * 0.25 * 0.75)
    assert np.all(np.abs(counts - 10000) < 5 * sigma)
    # rank myid's seed is the myid-th rand() after srand(seed) (qcd_rng.c:11-13)
    libc = ctypes.CDLL(None)
    libc.srand(77)
    s1 = libc.rand()
    s2 = libc.rand()
    lib.qcd_rng_init(77, 2)
    b = draws(lib, 100)
    lib.qcd_rng_init(ctypes.c_uint(s2), 0)
    assert np.array_equal(draws(lib, 100), b)
    lib.qcd_rng_init(77, 1)
    c = draws(lib, 100)
    lib.qcd_rng_init(ctypes.c_uint(s1), 0)
    assert np.array_equal(draws(lib, 100), c)
    lib.qcd_rng_finalize()


@pytest.mark.parametrize("t_slice", [3, -1, 8])
def test_z2vector_fills_in_reference_order(t_slice, capfd):
    L = (8, 4, 2, 2)
    T, X, Y, Zz = L
    di = Q.DropIn(L)
    try:
        lib = Q.lib()
        v = di.vector(np.ones((T * X * Y * Zz, 4, 3), dtype=complex))
        lib.qcd_rng_init(99, 0)
        lib.qcd_z2Vector(ctypes.byref(v), t_slice)
        got = di.view(v, 12).reshape(T * X * Y * Zz, 4, 3)
        ctypes.CDLL(None).fflush(None)   # the library's printf goes through C stdio buffers
        out = capfd.readouterr().out
        assert ("time dilution" in out) == (t_slice >= 0) and ("full in time" in out) == (t_slice < 0)
        lib.qcd_rng_init(99, 0)
        exp = np.zeros_like(got)
        slices = range(T) if t_slice < 0 else ([t_slice] if t_slice < T else [])
        for t in slices:
            for z in range(Zz):
                for y in range(Y):
                    for x in range(X):
                        iv = t + T * (x + X * (y + Y * z))
                        for sp in range(4):
                            for col in range(3):
                                exp[iv, sp, col] = Z4[lib.qcd_get_rand_4()]
        assert np.array_equal(got, exp)
        if t_slice == 3:
            assert np.all(np.abs(got[t_slice::T]) == 1.0)
            assert np.count_nonzero(got) == X * Y * Zz * 12
    finally:
        di.close()
        Q.lib().qcd_rng_finalize()


@pytest.mark.gpu
def test_device_z4_source(gpu):
    Q = gpu
    L = (8, 8, 4, 4)
    T = L[0]
    V = int(np.prod(L))
    lat = Q.Lattice(L)
    f = Q.Spinor(lat, nrhs=2)
    f.upload(np.full((V, 4, 3), 2.0 + 0j), rhs=0)
    Q.z4_source(f, 1, 5, 42)
    a = f.download(1)
    assert np.array_equal(f.download(0), np.full((V, 4, 3), 2.0 + 0j))   # other rhs untouched
    on = np.zeros(V, dtype=bool)
    on[5::T] = True
    assert np.all(np.isin(a[on], Z4)) and np.all(a[~on] == 0)
    Q.z4_source(f, 1, 5, 42)
    assert np.array_equal(f.download(1), a)
    Q.z4_source(f, 1, 5, 43)
    assert not np.array_equal(f.download(1), a)
    Q.z4_source(f, 1, -1, 42)
    full = f.download(1)
    assert np.all(np.isin(full, Z4))
    counts = np.array([np.sum(full == z) for z in Z4])
    sigma = np.sqrt(full.size * 0.25 * 0.75)
    assert np.all(np.abs(counts - full.size / 4) < 5 * sigma)
    with pytest.raises(Q.QcdError):
        Q.z4_source(f, 1, T, 42)
