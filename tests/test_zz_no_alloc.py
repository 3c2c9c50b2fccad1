"""The allocation-free time step (tbfm_gpu_compute_timestep_noalloc) against the
allocating entry point tbfm_gpu_compute_timestep, through the C-ABI.

Inputs and outputs live in device memory (buffers the test allocates with the HIP
runtime) in the library's device layout; the expected values come from the existing entry point on the
same state in the reference layout.  Two calls in a row on one GPU, the second on a
smaller grid, so the shared scratch array's leftovers from the first call cannot leak
into the second.  Agreement is bit-exact (the same time-step device code)."""
import ctypes as C

import numpy as np
import pytest

from cturbobfm_amd import Solver, cases, lib

pytestmark = pytest.mark.gpu


def to_device_layout(U):
    """reference layout (..., ni, nj, nk) -> device layout (..., nk, nj, ni)"""
    return np.ascontiguousarray(np.swapaxes(U, -1, -3))


def hip():
    import os
    L = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipFree.argtypes = [C.c_void_p]
    return L


def noalloc_timestep(s, U):
    H = hip()
    src = to_device_layout(U)
    out = np.full((s.nk, s.nj, s.ni), -1.0)
    dU, ddt = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(dU), src.nbytes) == 0 and H.hipMalloc(C.byref(ddt), out.nbytes) == 0
    try:
        assert H.hipMemcpy(dU, src.ctypes.data, src.nbytes, 1) == 0       # host to device
        assert H.hipMemcpy(ddt, out.ctypes.data, out.nbytes, 1) == 0
        rc = lib().tbfm_gpu_compute_timestep_noalloc(s.h, dU, ddt, None)
        assert rc == 0, lib().tbfm_gpu_last_error(s.h)
        s.sync()
        assert H.hipMemcpy(out.ctypes.data, ddt, out.nbytes, 2) == 0       # device to host
    finally:
        H.hipFree(dU)
        H.hipFree(ddt)
    return np.ascontiguousarray(np.swapaxes(out, 0, 2))


def perturbed_state(s, seed):
    U = s.state()
    rng = np.random.default_rng(seed)
    return U * (1.0 + 0.02 * rng.standard_normal(U.shape))


@pytest.mark.parametrize("time_step", ["global", "local"])
def test_noalloc_timestep_equals_allocating_entry_point(time_step):
    sizes = [(41, 13, 25), (17, 7, 9)]   # the second call on a smaller grid
    for q, (ni, nj, nk) in enumerate(sizes):
        print("no-alloc time step: %s, rotor %dx%dx%d" % (time_step, ni, nj, nk), flush=True)
        s = Solver.from_case(cases.rotor3d(ni, nj, nk, time_step=time_step))
        U = perturbed_state(s, 5 + q)
        ref = s.timestep(U)
        got = noalloc_timestep(s, U)
        assert np.array_equal(got, ref)
        if time_step == "global":
            assert np.all(got == got.flat[0])
        s.close()


def test_noalloc_timestep_refuses_null_arguments():
    s = Solver.from_case(cases.rotor3d(9, 5, 7))
    assert lib().tbfm_gpu_compute_timestep_noalloc(s.h, None, None, None) != 0
    s.close()
