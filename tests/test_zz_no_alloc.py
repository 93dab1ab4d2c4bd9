"""pxh_matvec_noalloc (the FP64 product with no device allocation, working
memory in the library's per-GPU scratch array) against pxh_matvec through
the C ABI: two calls in a row on the same GPU, the second with a smaller
input, each equal to pxh_matvec's product to 1e-13 (the same kernels and
launch shapes), and a size beyond the scratch refused with PXH_ENOMEM.

The test owns its device buffers: it allocates them with the HIP runtime
libpxh itself links (libamdhip64.so.7), outside the entry point."""
import ctypes

import numpy as np
import pytest

from plasmonx_amd import _lib
from plasmonx_amd.model import TOBOHR, AtomSystem
from plasmonx_amd import geometry
from plasmonx_amd.workloads import c1_sodium_sphere, silver_type

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


class DeviceArray:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, a):
        assert self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = ctypes.c_void_p()


def noalloc(ctx, h, freq, i_freq, x):
    dx = DeviceArray(h, x.nbytes)
    dy = DeviceArray(h, x.nbytes)
    try:
        dx.put(x)
        st = ctx.lib.pxh_matvec_noalloc(ctx._h, float(freq), int(i_freq), x.shape[0], dx.p, dy.p)
        ctx._check(st)
        return dy.get(np.zeros_like(x))
    finally:
        dx.free()
        dy.free()


def wfqfmu_sphere(radius):
    coord = geometry.fcc_sphere(4.09, radius) * TOBOHR
    return AtomSystem(model="wfqfmu", kernel="gaussian", symbols=["Ag"] * len(coord), coord_bohr=coord,
                      types=[silver_type()], fermi_d=np.array([[12.0]]), fermi_s=np.array([[1.1]]),
                      r0=np.array([[2.885 * TOBOHR]]))


def relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("model", ["wfqfmu", "wfq"])
def test_noalloc_matches_matvec_twice(model):
    h = hip()
    with _lib.Context(0) as ctx:
        if model == "wfqfmu":
            s = wfqfmu_sphere(14.0)
        else:
            full, _ = c1_sodium_sphere()
            s = AtomSystem(model="wfq", kernel="gaussian", symbols=full.symbols, coord_bohr=full.coord_bohr,
                           types=full.types, fermi_d=full.fermi_d, fermi_s=full.fermi_s, r0=full.r0)
        freq = np.array([3.0, 3.4]) / 27.211386245981
        ctx.set_atoms(s)
        if s.model == "wfqfmu":
            ctx.set_alpha(s.alpha_table(freq))
        rng = np.random.default_rng(4)
        for n_vec in (9, 3):  # the second call smaller than the first
            print(f"{model}: {s.n_atoms} atoms, {n_vec} vectors", flush=True)
            x = rng.normal(size=(n_vec, s.n_var)) + 1j * rng.normal(size=(n_vec, s.n_var))
            ref = ctx.matvec(freq[1], 1, x)
            got = noalloc(ctx, h, freq[1], 1, x)
            assert relerr(got, ref) < 1e-13


def test_noalloc_refuses_sizes_beyond_the_scratch():
    h = hip()
    with _lib.Context(0) as ctx:
        s = wfqfmu_sphere(24.0)
        freq = np.array([3.0]) / 27.211386245981
        ctx.set_atoms(s)
        ctx.set_alpha(s.alpha_table(freq))
        # splits x 48 vectors x n_var x 16 B is far above 64 MiB here
        x = np.zeros((48, s.n_var), dtype=np.complex128)
        with pytest.raises(_lib.PxhError) as e:
            noalloc(ctx, h, freq[0], 0, x)
        assert e.value.code == _lib.PXH_ENOMEM and "64 MiB" in str(e.value)
