"""The C-ABI's promises about memory and ordering, called through ctypes as an elsa adapter
would call them.

  * Managed memory (elsa's DataContainer storage is thrust::universal_vector): FP32 (tiled and
    one-thread-per-ray) and FP64 operators read inputs from and write outputs to
    hipMallocManaged buffers, with the same results as in device memory. The kernels that
    scatter with global atomics go through the operator's staging buffer there.
  * Calls on one operator from two streams, or from two host threads, are ordered by the
    library (each call rewrites the operator's volume views and fixed-point scale), so
    their results equal calls made one after another.
"""
import ctypes
import threading

import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    """The HIP runtime torch uses (and libelsahip.so binds to)."""
    import torch

    assert torch.cuda.is_available()
    torch.zeros(1, device="cuda")
    from elsa_amd import _lib

    _lib.load()
    paths = sorted(_lib._hip_runtimes())
    assert paths, "no HIP runtime mapped"
    H = ctypes.CDLL(paths[0])
    H.hipMallocManaged.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    H.hipMallocManaged.restype = ctypes.c_int
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipFree.restype = ctypes.c_int
    H.hipDeviceSynchronize.restype = ctypes.c_int
    return H


class Managed:
    """A hipMallocManaged buffer (hipMemAttachGlobal) with a numpy view on the host."""

    def __init__(self, H, shape, dtype):
        self.H = H
        self.dtype = np.dtype(dtype)
        self.n = int(np.prod(shape))
        self.ptr = ctypes.c_void_p()
        rc = H.hipMallocManaged(ctypes.byref(self.ptr), self.n * self.dtype.itemsize, 1)
        assert rc == 0 and self.ptr.value, f"hipMallocManaged failed ({rc})"
        buf = (ctypes.c_char * (self.n * self.dtype.itemsize)).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=self.dtype).reshape(shape)

    def free(self):
        if self.ptr.value:
            self.H.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def case():
    return oracle.circle_geometry(3, (40, 24, 33), 13, 360, 60.0, 40.0, det=(61, 29))


# class, dtype, kernels, whether the back projection (Joseph, Siddon) or the forward (voxel)
# scatters with global atomics (results then agree to rounding, not bit for bit); the tiled
# kernels of both dtypes write every element once and are bitwise reproducible
OPS = [
    ("JosephsMethodCUDA", np.float32, "fast", False),
    ("JosephsMethodCUDA", np.float32, "simple", True),
    ("JosephsMethodCUDA", np.float64, "fast", False),
    ("JosephsMethodCUDA", np.float64, "simple", True),
    ("SiddonsMethodCUDA", np.float32, "fast", False),
    ("SiddonsMethodCUDA", np.float64, "fast", False),
    ("SiddonsMethodCUDA", np.float64, "simple", True),
    ("BlobVoxelProjectorCUDA", np.float32, "fast", False),
    ("BlobVoxelProjectorCUDA", np.float64, "fast", False),
]


@pytest.mark.parametrize("cls,dt,kernels,atomic", OPS,
                         ids=[f"{c}-{np.dtype(d).name}-{k}" for c, d, k, _ in OPS])
def test_managed_memory_inputs_and_outputs(hip, cls, dt, kernels, atomic):
    import torch

    import elsa_amd as E
    from elsa_amd import _lib

    g = case()
    vd, dd = geom_to_descriptors(g)
    kw = {} if kernels == "fast" or cls.startswith("Blob") else {"kernels": kernels}
    op = getattr(E, cls)(vd, dd, dtype=dt, **kw)
    L = _lib.load()
    rng = np.random.default_rng(4)
    x = rng.random(g.vol_shape).astype(dt)
    y = rng.random(g.sino_shape).astype(dt)
    ref_fwd = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    ref_adj = op.applyAdjoint(torch.from_numpy(y).cuda()).cpu().numpy()
    mx, my = Managed(hip, g.vol_shape, dt), Managed(hip, g.sino_shape, dt)
    mo_s, mo_v = Managed(hip, g.sino_shape, dt), Managed(hip, g.vol_shape, dt)
    try:
        mx.array[...] = x
        my.array[...] = y
        mo_s.array[...] = 7.0
        mo_v.array[...] = -3.0
        _lib.check(L.ehp_forward(op.handle, mx.ptr, mo_s.ptr, None), "forward")
        _lib.check(L.ehp_adjoint(op.handle, my.ptr, mo_v.ptr, None), "adjoint")
        assert hip.hipDeviceSynchronize() == 0
        fwd, adj = mo_s.array.copy(), mo_v.array.copy()
    finally:
        for m in (mx, my, mo_s, mo_v):
            m.free()
    tol = 1e-6 if dt == np.float32 else 1e-13
    scat_fwd = atomic and cls.startswith("Blob")
    scat_adj = atomic and not cls.startswith("Blob")
    for got, ref, scat in ((fwd, ref_fwd, scat_fwd), (adj, ref_adj, scat_adj)):
        if scat:
            err = np.linalg.norm((got - ref).astype(np.float64)) / np.linalg.norm(ref.astype(np.float64))
            assert err <= tol, err
        else:
            np.testing.assert_array_equal(got, ref)


def test_one_operator_on_two_streams_is_ordered():
    """Two forwards and two back projections of one operator issued back to back on two
    streams (no synchronisation between them) equal the same calls made in turn."""
    import torch

    import elsa_amd as E

    g = oracle.circle_geometry(3, (64, 64, 64), 90, 360, 6400.0, 320.0, det=(96, 96))
    vd, dd = geom_to_descriptors(g)
    op = E.JosephsMethodCUDA(vd, dd)
    gen = torch.Generator(device="cuda").manual_seed(1)
    xs = [torch.rand(g.vol_shape, device="cuda", generator=gen) for _ in range(2)]
    ys = [torch.rand(g.sino_shape, device="cuda", generator=gen) for _ in range(2)]
    ref_f = [op.apply(x) for x in xs]
    ref_b = [op.applyAdjoint(y) for y in ys]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]  # default priority
    outs_f = [torch.empty_like(r) for r in ref_f]
    outs_b = [torch.empty_like(r) for r in ref_b]
    for rep in range(3):
        for i, s in enumerate(streams):
            with torch.cuda.stream(s):
                op.apply(xs[i], outs_f[i])
        for i, s in enumerate(streams):
            with torch.cuda.stream(s):
                op.applyAdjoint(ys[i], outs_b[i])
        torch.cuda.synchronize()
        for i in range(2):
            assert torch.equal(outs_f[i], ref_f[i]), (rep, i)
            assert torch.equal(outs_b[i], ref_b[i]), (rep, i)


def test_one_operator_from_two_host_threads_is_ordered():
    """Two host threads apply one operator at once, each on its own stream."""
    import torch

    import elsa_amd as E

    g = oracle.circle_geometry(3, (48, 48, 48), 60, 360, 4800.0, 240.0, det=(72, 72))
    vd, dd = geom_to_descriptors(g)
    op = E.JosephsMethodCUDA(vd, dd)
    gen = torch.Generator(device="cuda").manual_seed(2)
    xs = [torch.rand(g.vol_shape, device="cuda", generator=gen) for _ in range(2)]
    refs = [op.applyAdjoint(op.apply(x)) for x in xs]
    torch.cuda.synchronize()
    results = [[None] * 5, [None] * 5]
    errors = []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for k in range(5):
                    results[i][k] = op.applyAdjoint(op.apply(xs[i]))
            s.synchronize()
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i in range(2):
        for k in range(5):
            assert torch.equal(results[i][k], refs[i]), (i, k)
