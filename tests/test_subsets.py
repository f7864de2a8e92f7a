"""SubsetSampler (elsa/projectors/tests/test_SubsetSampler.cpp): the sampling strategies'
known answers on the CPU; on the GPU, subset projectors give the full projector's rows."""
import numpy as np
import pytest


def circle(n=32, poses=32):
    import elsa_amd as E

    vd = E.VolumeDescriptor([n, n])
    return vd, E.CircleTrajectoryGenerator.createTrajectory(poses, vd, 360, 100.0 * n, float(n))


def test_round_robin_known_answers():
    """test_SubsetSampler.cpp:41-78: 32 poses into 4 subsets of 8 and into 5 of 7/6, subset i
    holding poses j * n + i."""
    from elsa_amd import SubsetSampler

    for n, sizes in ((4, [8] * 4), (5, [7, 7, 6, 6, 6])):
        m = SubsetSampler.splitRoundRobin(list(range(32)), n)
        assert [len(s) for s in m] == sizes
        for i, s in enumerate(m):
            assert s == [j * n + i for j in range(len(s))]


def test_rotational_clustering_known_answer():
    """test_SubsetSampler.cpp:80-92 (circular trajectory, 32 poses, 4 subsets)."""
    from elsa_amd import SubsetSampler

    vd, dd = circle()
    m = SubsetSampler.splitRotationalClustering(dd, 4)
    assert m == [[15, 11, 7, 3, 0, 27, 23, 19], [14, 10, 6, 2, 30, 26, 22, 18], [13, 9, 5, 1, 29, 25, 21, 17],
                 [12, 8, 4, 31, 28, 24, 20, 16]]


def test_partitioned_data_and_descriptors():
    """test_SubsetSampler.cpp:166-260: the subsets cover every pose once, each subset's
    descriptor has the full detector size and the subset's poses."""
    from elsa_amd import SubsetSampler

    vd, dd = circle(16, 20)
    sino = np.arange(20 * 22, dtype=np.float32).reshape(20, 22)
    for strategy in (SubsetSampler.ROUND_ROBIN, SubsetSampler.ROTATIONAL_CLUSTERING):
        s = SubsetSampler(vd, dd, 4, strategy)
        parts = s.getPartitionedData(sino)
        assert sorted(i for idx in s.indexMapping for i in idx) == list(range(20))
        for idx, part, d in zip(s.indexMapping, parts, s.detectorDescriptors):
            np.testing.assert_array_equal(part, sino[idx])
            assert tuple(d.getNumberOfCoefficientsPerDimension()) == (22, len(idx))
            assert d.getGeometry() == [dd.getGeometry()[i] for i in idx]
    with pytest.raises(ValueError):
        SubsetSampler(vd, dd, 1)
    with pytest.raises(ValueError):
        SubsetSampler(vd, dd, 4, "random")


@pytest.mark.gpu
@pytest.mark.parametrize("cls", ["JosephsMethodCUDA", "SiddonsMethodCUDA"])
def test_subset_projectors_give_the_full_projectors_rows(cls):
    import elsa_amd as E
    from elsa_amd import SubsetSampler

    vd, dd = circle(48, 30)
    x = np.random.default_rng(0).random(vd.shape, dtype=np.float32)
    s = SubsetSampler(vd, dd, 4, SubsetSampler.ROTATIONAL_CLUSTERING)
    full = s.getProjector(getattr(E, cls)).apply(x)
    for part, op in zip(s.getPartitionedData(full), s.getSubsetProjectors(getattr(E, cls))):
        np.testing.assert_array_equal(op.apply(x), part)
