"""bfloat16 storage (FP32 arithmetic) for the Joseph and voxel projectors, and iterative
refinement.

A bfloat16 operator reads bfloat16 volume views and sums each ray in FP32, so its forward
equals the FP32 operator's on the same (bfloat16-valued) volume rounded once to bfloat16; its
back projection is the FP32 fixed-point back projection of the widened sinogram rounded once
per voxel. Both are checked bit for bit against the FP32 operator and torch's own
round-to-nearest-even conversion."""
import numpy as np
import pytest

import oracle
from backends import geom_to_descriptors

pytestmark = pytest.mark.gpu

CASES = {
    "fan2d": dict(dim=2, vol=(80, 64), views=50, det=(110,), s2c=8000.0, c2d=400.0),
    "cone3d": dict(dim=3, vol=(40, 32, 36), views=30, det=(60, 48), s2c=4000.0, c2d=200.0),
}


def ops(name):
    import elsa_amd as E

    c = CASES[name]
    g = oracle.circle_geometry(c["dim"], c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    return g, E.JosephsMethodCUDA(vd, dd), E.JosephsMethodCUDA(vd, dd, dtype="bfloat16")


@pytest.mark.parametrize("name", list(CASES))
def test_bf16_projections_equal_fp32_rounded_once(name):
    import torch

    g, A32, A16 = ops(name)
    gen = torch.Generator(device="cuda").manual_seed(3)
    x = (torch.rand(g.vol_shape, device="cuda", generator=gen) * 3 - 1).to(torch.bfloat16)
    y = torch.rand(g.sino_shape, device="cuda", generator=gen).to(torch.bfloat16)
    f16 = A16.apply(x)
    b16 = A16.applyAdjoint(y)
    assert f16.dtype == torch.bfloat16 and b16.dtype == torch.bfloat16
    assert torch.equal(f16, A32.apply(x.float()).to(torch.bfloat16))
    assert torch.equal(b16, A32.applyAdjoint(y.float()).to(torch.bfloat16))
    # numpy in, float32 out (values rounded to bfloat16 on the way in)
    xn = x.float().cpu().numpy()
    np.testing.assert_array_equal(A16.apply(xn), f16.float().cpu().numpy())


VOXEL_KINDS = ["BlobVoxelProjectorCUDA", "BSplineVoxelProjectorCUDA", "PhaseContrastBlobVoxelProjectorCUDA",
               "PhaseContrastBSplineVoxelProjectorCUDA"]


@pytest.mark.parametrize("kind", VOXEL_KINDS)
@pytest.mark.parametrize("name", list(CASES))
def test_bf16_voxel_projections_equal_fp32_rounded_once(name, kind):
    """The voxel projectors on bfloat16 storage: the tiled forward's fixed-point pixel sums and
    the gather's FP32 voxel sums are those of the FP32 operator on the same (bfloat16-valued)
    data, each rounded to bfloat16 once."""
    import torch

    import elsa_amd as E

    c = CASES[name]
    g = oracle.circle_geometry(c["dim"], c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    A32 = getattr(E, kind)(vd, dd)
    A16 = getattr(E, kind)(vd, dd, dtype="bfloat16")
    assert A16.lut.dtype == np.float32 and np.array_equal(A16.lut, A32.lut)
    gen = torch.Generator(device="cuda").manual_seed(4)
    x = (torch.rand(g.vol_shape, device="cuda", generator=gen) * 3 - 1).to(torch.bfloat16)
    y = torch.rand(g.sino_shape, device="cuda", generator=gen).to(torch.bfloat16)
    f16 = A16.apply(x)
    b16 = A16.applyAdjoint(y)
    assert f16.dtype == torch.bfloat16 and b16.dtype == torch.bfloat16
    assert torch.equal(f16, A32.apply(x.float()).to(torch.bfloat16))
    assert torch.equal(b16, A32.applyAdjoint(y.float()).to(torch.bfloat16))
    assert float(f16.float().abs().max()) > 0 and float(b16.float().abs().max()) > 0


def test_bf16_operator_is_fast_kernels_only():
    import elsa_amd as E

    c = CASES["fan2d"]
    g = oracle.circle_geometry(2, c["vol"], c["views"], 360, c["s2c"], c["c2d"], det=c["det"])
    vd, dd = geom_to_descriptors(g)
    with pytest.raises(NotImplementedError):
        E.JosephsMethodCUDA(vd, dd, dtype="bfloat16", kernels="simple")
    with pytest.raises(NotImplementedError):
        E.BlobVoxelProjectorCUDA(vd, dd, dtype="bfloat16", kernels="simple")
    with pytest.raises(NotImplementedError):
        E.SiddonsMethodCUDA(vd, dd, dtype="bfloat16")


@pytest.mark.parametrize("name", list(CASES))
def test_bf16_cgls_and_refinement(name):
    """CGLS on bfloat16 storage reduces the residual like FP32 CGLS to bfloat16 accuracy, and
    refinement (FP32 residual in FP64, bfloat16 inner solves) drives the FP32 residual below
    what the bfloat16 solve alone reaches."""
    import torch

    import elsa_amd as E

    g, A32, A16 = ops(name)
    grids = np.meshgrid(*[np.linspace(-1, 1, s) for s in g.vol_shape], indexing="ij")
    x_true = torch.from_numpy((sum(q * q for q in grids) < 0.5).astype(np.float32)).cuda()
    b = A32.apply(x_true)
    nb = float(torch.linalg.norm(b))
    ref = E.CGLS(A32, b, tol=0.0)
    x32 = ref.solve(30)
    cg16 = E.CGLS(A16, b.to(torch.bfloat16), tol=0.0, graph=True)
    x16 = cg16.solve(30)
    r16 = float(torch.linalg.norm(A32.apply(x16.float()) - b)) / nb
    r32 = ref.residuals[-1] / nb
    assert r16 < 0.1 * float(cg16.residuals[0]) / nb
    assert float(torch.linalg.norm(x16.float() - x32) / torch.linalg.norm(x32)) < 0.05
    rf = E.CGLSRefined(A32, A16, b, inner=10, graph=True)
    xr = rf.solve(3)
    assert np.all(np.diff(rf.residuals) < 0), rf.residuals
    assert rf.residuals[-1] / nb < r16, (rf.residuals[-1] / nb, r16, r32)
    rr = float(torch.linalg.norm(A32.apply(xr) - b)) / nb
    assert abs(rr - rf.residuals[-1] / nb) <= 1e-6
