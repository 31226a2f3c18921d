"""mgr_ng_acceleration_noalloc, the Ng step that allocates nothing, against
the context entry point mgr_iteration_using_Ng_acceleration on the same
populations and history: populations, history and convergence statistics
bit for bit, on two calls in a row, the second with a smaller model (the
scratch array keeps the first call's contents).  The device arrays are
allocated with the HIP runtime the library itself links (torch ships a
second one); the call goes through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from magritte_amd import hip
from magritte_amd.model import ConvStats
from magritte_amd.setup import synthetic

pytestmark = pytest.mark.gpu


class _Runtime:
    """hipMalloc / hipMemcpy / hipFree of the runtime libmagritte_hip.so uses."""

    def __init__(self):
        hip.load()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def download(self, p, n):
        a = np.empty(n)
        assert self.rt.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0   # device to host
        return a

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _history(m):
    """Populations and history after four normal FP64 ALI iterations."""
    ctx = hip.Context(m).setup()
    for _ in range(4):
        ctx.ali_iteration(False, "fp64")
    pop = ctx.populations()
    h = ctx.population_history()
    ctx.close()
    return pop, h


def test_noalloc_ng_matches_the_context_entry_point():
    lib = hip.load()
    rt = _Runtime()
    for n in (12, 7):   # the second call is smaller
        m = synthetic.envelope_model(n=n, n0=1.0e12, T0=1000.0, nside=1, seed=3, abundance=1.0e-3)
        print(f"\n[noalloc] envelope {n}^3: Ng step, context vs no-allocation entry point", flush=True)
        pop, (h1, h2, h3) = _history(m)
        ctx = hip.Context(m).setup()
        ctx.set_populations(pop)
        ctx.set_population_history(h1, h2, h3)
        conv_ctx = ctx.iteration_using_Ng_acceleration()
        P_ctx = ctx.populations()
        H_ctx = ctx.population_history()
        ctx.close()
        assert not np.array_equal(P_ctx, pop), "the Ng step must change the populations"

        N, V = m.ncells, m.linedata.nlev
        arrs = [rt.upload(a) for a in (pop, h1, h2, h3)]
        tot = rt.upload(m.abundance[:, m.linedata.num])
        conv = ConvStats()
        rc = lib.mgr_ng_acceleration_noalloc(N, V, m.pop_prec, *arrs, tot, None, C.byref(conv))
        assert rc == 0, lib.mgr_last_error(None)
        assert rt.rt.hipDeviceSynchronize() == 0
        P, G1, G2, G3 = [rt.download(a, N * V).reshape(N, V) for a in arrs]
        rt.free()
        assert np.array_equal(P, P_ctx)
        for a, b in zip((G1, G2, G3), H_ctx):
            assert np.array_equal(a, b)
        assert conv.fraction_not_converged == conv_ctx[0]
        assert conv.relative_change_mean == conv_ctx[1]
        assert conv.relative_change_max == conv_ctx[2]


def test_noalloc_ng_rejects_sizes_beyond_its_scratch():
    lib = hip.load()
    conv = ConvStats()
    one = C.c_void_p(16)   # never dereferenced: the size check comes first
    rc = lib.mgr_ng_acceleration_noalloc(4096 * 1024 + 1, 7, 1e-6, one, one, one, one, one, None, C.byref(conv))
    assert rc == -1
    assert b"exceeds the maximum of 4194304" in lib.mgr_last_error(None)
    rc = lib.mgr_ng_acceleration_noalloc(10, 7, 1e-6, None, one, one, one, one, None, C.byref(conv))
    assert rc == -1
