"""Seed orders the Q-map pipelines trace in (fastqsl_amd.qmap.block_order):
a permutation of the grid, blocks in C order with C order inside a block,
partial blocks at the far edges."""
import numpy as np
import pytest

from fastqsl_amd.qmap import _block_order_host


@pytest.mark.parametrize("shape,block", [((52, 70), (16, 16)), ((64, 64), (16, 16)), ((8, 12, 16), (4, 4, 4)),
                                         ((9, 10, 7), (4, 4, 4)), ((5,), (4,))])
def test_permutation_in_blocks(shape, block):
    p = _block_order_host(shape, block)
    n = int(np.prod(shape))
    np.testing.assert_array_equal(np.sort(p), np.arange(n))
    coords = np.stack(np.unravel_index(p, shape), axis=1)
    blk = coords // np.array(block)
    # block indices are visited in C order and never revisited
    nb = [-(-s // b) for s, b in zip(shape, block)]
    flat = np.ravel_multi_index(blk.T, nb)
    assert np.all(np.diff(flat) >= 0)
    # inside a block, C order
    for b in np.unique(flat)[:5]:
        inside = coords[flat == b]
        lin = np.ravel_multi_index(inside.T, shape)
        assert np.all(np.diff(lin) > 0)


def test_first_block_of_a_plane():
    p = _block_order_host((32, 48), (16, 16))
    np.testing.assert_array_equal(p[:18], list(range(16)) + [48, 49])
