"""ehp_adjoint_noalloc (the back projection whose scratch lives in the library's static device
array) against ehp_adjoint on the same operator and input, through the C-ABI: bit for bit,
since both run the same kernels and sum in fixed point. Two calls in a row, the second on a
smaller operator, then the first again (the array keeps the last call's contents); FP64 Joseph
and FP32 Siddon; operators it does not take are refused."""
import ctypes

import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu

CASES = {
    "cone3d_48": dict(dim=3, vol=(48, 40, 44), views=40, det=(72, 64), s2c=4800.0, c2d=240.0),
    "fan2d_64": dict(dim=2, vol=(64, 64), views=60, det=(92,), s2c=6400.0, c2d=320.0),
}


def make(name, cls="JosephsMethodCUDA", dtype=np.float32, **kw):
    import elsa_amd as E

    c = CASES[name]
    g = oracle.circle_geometry(c["dim"], c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    return g, getattr(E, cls)(vd, dd, dtype=dtype, **kw)


def noalloc(op, y):
    """A^T y through ehp_adjoint_noalloc on torch's current stream."""
    import torch

    from elsa_amd import _lib

    L = _lib.load()
    out = torch.full(op.getDomainDescriptor().shape, float("nan"), dtype=y.dtype, device=y.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(y.device).cuda_stream)
    rc = L.ehp_adjoint_noalloc(op.handle, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()), stream)
    return rc, out


def seeded_sino(g, seed, dtype):
    import torch

    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.standard_normal(g.sino_shape).astype(dtype)).cuda()


@pytest.mark.parametrize("cls,dtype", [("JosephsMethodCUDA", np.float32), ("JosephsMethodCUDA", np.float64),
                                       ("SiddonsMethodCUDA", np.float32), ("SiddonsMethodCUDA", np.float64)],
                         ids=["joseph-f32", "joseph-f64", "siddon-f32", "siddon-f64"])
def test_noalloc_equals_adjoint_large_then_small(cls, dtype):
    import torch

    g1, op1 = make("cone3d_48", cls, dtype)
    g2, op2 = make("fan2d_64", cls, dtype)
    y1, y2 = seeded_sino(g1, 1, dtype), seeded_sino(g2, 2, dtype)
    for op, y in ((op1, y1), (op2, y2), (op1, y1)):
        print(f"noalloc {cls} {np.dtype(dtype).name} {op.getDomainDescriptor().shape}", flush=True)
        rc, out = noalloc(op, y)
        assert rc == 0, rc
        ref = op.applyAdjoint(y)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())


def test_noalloc_refuses_operators_without_a_tiled_back_projection():
    import torch

    from elsa_amd import _lib

    for kw in ({"cls": "JosephsMethodCUDA", "kernels": "simple"}, {"cls": "JosephsMethodCUDA", "dtype": "bfloat16"},
               {"cls": "BlobVoxelProjectorCUDA"}):
        cls = kw.pop("cls")
        g, op = make("fan2d_64", cls, **kw)
        dt = torch.bfloat16 if kw.get("dtype") == "bfloat16" else torch.float32
        y = seeded_sino(g, 3, np.float32).to(dt)
        rc, _ = noalloc(op, y)
        assert rc == _lib.EHP_ERR_UNSUPPORTED, (cls, kw, rc)
