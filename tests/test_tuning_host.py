"""The BVH leaf size on the host (no device needed): trees rebuilt with at
most 1, 2, 4 or 8 triangles per leaf, and invalid knobs refused."""
import pytest

from puzzletunneldiscovery_amd import pyosr


@pytest.fixture()
def world(torus_meshes):
    env, rob = torus_meshes
    return pyosr.create_unit_world_from_arrays(env, rob)


def test_leaf_size_rebuilds_the_trees(world):
    assert world.tuning == {"leaf_size": 1, "pool_chunk": 64, "queue_depth": 1}
    d1, n1 = world.bvh_info(0)
    assert n1 == 2 * 2048 - 1
    prev = (d1, n1)
    for leaf in (2, 4, 8):
        world.tuning = {"leaf_size": leaf}
        d, n = world.bvh_info(0)
        # a median split halves the triangle count per level, so each doubling
        # of the leaf size takes one level and about half the nodes away
        assert d == prev[0] - 1 and n < prev[1] and n >= 2 * (2048 // leaf) - 1
        prev = (d, n)
    world.tuning = {"leaf_size": 1}
    assert world.bvh_info(0) == (d1, n1)


def test_bad_tuning_is_refused(world):
    before = world.tuning
    for bad in ({"leaf_size": 0}, {"leaf_size": 9}, {"pool_chunk": 8}, {"pool_chunk": 2048},
                {"queue_depth": 0}, {"queue_depth": 17}):
        with pytest.raises(RuntimeError):
            world.tuning = bad
    assert world.tuning == before
