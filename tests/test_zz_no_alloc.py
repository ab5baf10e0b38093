"""fsibe_dev_coftxy_noalloc -- coftxy on device pointers with no context and
no allocation (the intermediate half spectrum and the twiddle tables in the
library's static device buffer) -- against fsibe_dev_coftxy, whose kernels it
shares: bit-identical results on consecutive calls of different sizes (the
second smaller than the first, then back), and sizes beyond the static
buffer rejected before any device call."""
import ctypes as C

import numpy as np
import pytest

FSIBE_ESIZE = 2


class _Hip:
    """device memory through the HIP runtime the library itself links
    (libamdhip64.so.7, already loaded by lib()); torch bundles another HIP
    runtime, which must not share the process's device state with it"""

    def __init__(self):
        from fsibe_2d_amd._lib import lib

        lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, d, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.L.hipMemcpy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device

    def get(self, d, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), d, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, d):
        assert self.L.hipFree(d) == 0


def test_sizes_beyond_the_static_buffer_are_rejected():
    """checked before any device call (runs without a GPU)"""
    from fsibe_2d_amd._lib import lib

    L = lib()
    buf = (C.c_double * 4)()
    p = C.cast(buf, C.c_void_p)
    assert L.fsibe_dev_coftxy_noalloc(4096, 4096, p, p, None) == FSIBE_ESIZE
    assert "static buffer" in L.fsibe_last_error().decode()
    assert L.fsibe_dev_coftxy_noalloc(8192, 256, p, p, None) == FSIBE_ESIZE
    assert L.fsibe_dev_coftxy_noalloc(100, 64, p, p, None) == FSIBE_ESIZE
    assert L.fsibe_dev_coftxy_noalloc(64, 64, None, p, None) == 1


@pytest.mark.gpu
def test_noalloc_coftxy_matches_the_context_entry_point(oracle_mod):
    from fsibe_2d_amd._lib import check, lib

    L = lib()
    hip = _Hip()

    def run(f, use_context):
        ny, nx = f.shape
        d_f, d_fk = hip.alloc(f.nbytes), hip.alloc(f.nbytes)
        hip.put(d_f, f)
        hip.put(d_fk, np.full(f.shape, np.nan))
        if use_context:
            h = C.c_void_p()
            check(L.fsibe_spectral_create(nx, ny, 4.0, 4.0, 0.0, C.byref(h)))
            check(L.fsibe_dev_coftxy(h, d_f, d_fk, None))
            check(L.fsibe_spectral_sync(h))
            check(L.fsibe_spectral_destroy(h))
        else:
            check(L.fsibe_dev_coftxy_noalloc(nx, ny, d_f, d_fk, None))
        hip.sync()
        out = hip.get(d_fk, f.shape)
        hip.free(d_f)
        hip.free(d_fk)
        return out

    rng = np.random.default_rng(11)
    for k, (nx, ny) in enumerate([(1024, 768), (256, 192), (1024, 768), (2048, 2048), (96, 64)]):
        f = rng.normal(size=(ny, nx))
        print(f"[coftxy noalloc {nx}x{ny}]", flush=True)
        a = run(f, False)
        b = run(f, True)
        print(f"[coftxy noalloc {nx}x{ny}] max |a - b| = {np.abs(a - b).max():.1e}", flush=True)
        assert np.array_equal(a, b), (nx, ny)
        if k == 1:  # and against the oracle's coftxy, to the plugin's bar
            o = oracle_mod.Oracle(nx=nx, ny=ny, iBeam=0, iSponge=0)
            ref = o.coftxy(f)
            assert np.abs(a - ref).max() <= 2e-14 * np.abs(ref).max()
