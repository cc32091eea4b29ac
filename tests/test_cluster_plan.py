"""The cluster plan of the compressed inner operator (plan_hmat, wb_setup.cpp), on the CPU:
every device index in exactly one cluster of at most 256 dofs, each cluster made of at most
4 runs of consecutive device indices (so that block reads stay coalesced), the geometric
'near' flags symmetric with the diagonal set, and most blocks far at 10k and 40k dofs
(else the compression would gain little)."""
import numpy as np
import pytest

from tests import helpers as T
from wavebem_amd import bem as B
from wavebem_amd import mesh as M


def _runs(idx):
    return 1 + int(np.count_nonzero(np.diff(idx) != 1))


@pytest.mark.parametrize("kind", ["wigley_small", "2.5k", "10k", "sphere", "mixed"])
def test_clusters_partition_the_dofs(kind):
    if kind == "wigley_small":
        m = T.wigley_small()
    elif kind in ("2.5k", "10k"):
        m = T.wigley_preset(kind)
    elif kind == "sphere":
        m = M.cube_sphere_mesh(n=9)
    else:
        m = M.merge_meshes(M.box_mesh(n=5), M.tri_block_mesh(n=4, centre=(3.0, 0, 0)))
    p = B.host_cluster_plan(m)
    n = m.n_dofs
    size, idx, near = p["size"], p["idx"], p["near"]
    assert np.all(size > 0) and np.all(size <= 256)
    seen = np.zeros(n, dtype=np.int64)
    for c in range(len(size)):
        ids = idx[c, :size[c]]
        assert np.all(idx[c, size[c]:] == -1)
        assert _runs(ids) <= 4
        seen[ids] += 1
    assert np.all(seen == 1)
    assert np.array_equal(near, near.T) and np.all(np.diag(near))
    if kind == "10k":
        # clusters are compact patches (on these graded meshes, large where the mesh is
        # coarse): each touches a few others, so most blocks can be compressed
        assert near.mean() < 0.15, near.mean()


def test_most_blocks_far_at_40k():
    m = T.wigley_preset("40k")
    near = B.host_cluster_plan(m)["near"]
    assert near.mean() < 0.05, near.mean()
