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


@pytest.mark.parametrize("ncl,seed", [(1, 0), (7, 1), (97, 2), (175, 3)])
def test_product_work_items_cover_every_block_once(ncl, seed):
    """The compressed product's work lists (hm_plan_items): every low-rank block is one
    item and every dense block splits into column parts covering its columns, each listed
    exactly once, in the chunks of its row cluster, low-rank items first; chunk loads are
    balanced to within one item's weight of the mean (largest-first packing)."""
    rng = np.random.default_rng(seed)
    size = rng.integers(1, 257, ncl).astype(np.int32)
    u = rng.random((ncl, ncl))
    rank = np.where(u < 0.05, 0, np.where(u < 0.15, 255, rng.integers(1, 13, (ncl, ncl))))
    rank = rank.astype(np.uint8)
    w = B.host_hm_work_items(rank, size)
    ch, ptr, items, dc, sh, lr = (w["chunks"], w["item_ptr"], w["items"], w["dcols"],
                                  w["shift"], w["lowrank"])
    assert ptr[0] == 0 and np.all(np.diff(ptr) >= 0) and ptr[-1] == items.size
    expect = 0
    for I in range(ncl):
        seen = {}
        loads = []
        for k in range(ch):
            its = items[ptr[I * ch + k]:ptr[I * ch + k + 1]]
            kinds = its & ((1 << sh) - 1)
            J = its >> sh
            low = kinds == lr
            # low-rank first, then dense parts
            assert not np.any(low[1:] & ~low[:-1])
            load = 0
            for j, q in zip(J, kinds):
                key = (int(j), int(q))
                seen[key] = seen.get(key, 0) + 1
                r = int(rank[I, j])
                if q == lr:
                    assert 0 < r < 255
                    load += 8 + 16 * r
                else:
                    assert r == 255 and q * dc < size[j]
                    load += 8 * dc
            loads.append(load)
        for j in range(ncl):
            r = int(rank[I, j])
            if r == 255:
                for q in range(-(-int(size[j]) // dc)):
                    assert seen.pop((j, q)) == 1
                    expect += 1
            elif r > 0:
                assert seen.pop((j, lr)) == 1
                expect += 1
        assert not seen  # nothing else listed
        assert max(loads) <= np.mean(loads) + max(8 * dc, 8 + 16 * 12)
    assert items.size == expect
