"""vpfft_strain_rate_evolution_noalloc (device arrays in and out, static
device scratch, no allocation) against vpfft_strain_rate_evolution (the
grid-owned buffers) on the same inputs: two calls in a row, the second on a
smaller grid, in both precisions.  Bit-identical: the same kernels run on the
same constants and twiddles."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(rng):
    A = rng.standard_normal((5, 5))
    return (4.0 * (A @ A.T / 5.0 + np.eye(5))).T.reshape(25)


class DeviceBuffer:
    """Device memory from the HIP runtime libvpfft.so itself uses (a second
    runtime in the process, such as torch's, cannot share the GPU with it)."""

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.ptr = hip, nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _hip():
    hip = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _noalloc(vpfft, dims, lengths, prec, L0, phi):
    lib = vpfft.load_library()
    hip = _hip()
    d_phi, d_out = DeviceBuffer(hip, phi.nbytes), DeviceBuffer(hip, phi.nbytes)
    d_phi.upload(phi)
    d_out.upload(np.full(phi.shape, np.nan))
    ln = np.ascontiguousarray(lengths, dtype=np.float64)
    L = np.ascontiguousarray(L0, dtype=np.float64)
    rc = lib.vpfft_strain_rate_evolution_noalloc(dims[0], dims[1], dims[2], ln.ctypes.data_as(C.POINTER(C.c_double)),
                                                 prec, L.ctypes.data_as(C.POINTER(C.c_double)), d_phi.ptr, d_out.ptr,
                                                 None)
    assert rc == 0, lib.vpfft_last_error()
    out = d_out.download(phi.shape)
    d_phi.free()
    d_out.free()
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_grid_entry_point(gpu, vpfft, prec):
    rng = np.random.default_rng(60 + prec)
    mat = vpfft.Material.uniform(1.0, 10)
    for dims, lengths in (((24, 20, 16), (1.0, 1.25, 0.8)), ((12, 10, 8), (0.9, 1.1, 1.0))):
        print(f"no-alloc transform {dims}, precision {prec}", flush=True)
        N = dims[0] * dims[1] * dims[2]
        L0 = _spd(rng)
        phi = rng.standard_normal((5, N))
        with vpfft.MaterialGrid(*dims, mat, precision=prec) as g:
            g.SetVoxelLength(*lengths)
            ref = g.StrainRateEvolutionWithFFT(L0, phi)
        out = _noalloc(vpfft, dims, lengths, prec, L0, phi)
        assert np.array_equal(out, ref), float(np.max(np.abs(out - ref)))


def test_noalloc_refuses_a_grid_beyond_its_scratch(gpu, vpfft):
    """A 128^3 grid needs more than the 32 MiB scratch: refused before any
    device work (the buffers are never touched)."""
    lib = vpfft.load_library()
    buf = DeviceBuffer(_hip(), 128)
    ln = (C.c_double * 3)(1.0, 1.0, 1.0)
    L0 = (C.c_double * 25)(*np.eye(5).reshape(25))
    rc = lib.vpfft_strain_rate_evolution_noalloc(128, 128, 128, ln, 0, L0, buf.ptr, buf.ptr, None)
    buf.free()
    assert rc == -1
    assert b"scratch" in lib.vpfft_last_error()
