"""The oracle's slice iteration against the reference's own known answers (CPU only).

CPPQEDutils/Testing/MultiIndexIterator.cc:12-40 fixes the odometer order (last index fastest);
CPPQEDutils/Testing/SliceIterator.cc:21-105 fixes, for retained axes V = <3,6,1,9,7> of a rank-11
array, which axes are dummies (filterOut), that a slice's axes follow V's order (the transposer
permutation (0,3,2,6,4,5,1,9,8,7,10)) and the extents of the sliced view.
"""
import numpy as np

import oracle_lib as O


def test_odometer_order():
    """MultiIndexIterator<3> from (-2,0,1) to (-1,1,3): the reference's twelve indices in order."""
    lb = np.array([-2, 0, 1])
    seq = [(-2, 0, 1), (-2, 0, 2), (-2, 0, 3), (-2, 1, 1), (-2, 1, 2), (-2, 1, 3),
           (-1, 0, 1), (-1, 0, 2), (-1, 0, 3), (-1, 1, 1), (-1, 1, 2), (-1, 1, 3)]
    dims = (2, 2, 3)
    want = [np.ravel_multi_index(tuple(np.array(i) - lb), dims) for i in seq]
    off, _ = O.slices(dims, [])                 # no retained axis: every index is a dummy
    np.testing.assert_array_equal(off, want)


V = [3, 6, 1, 9, 7]
DIMS = (22, 2, 5, 1, 12, 4, 4, 3, 24, 2, 8)     # ArraySlicingTest's array11


def test_slice_axes_follow_V():
    strides = np.array([int(np.prod(DIMS[i + 1:])) for i in range(len(DIMS))])
    off, vs = O.slices(DIMS, V)
    # TransposerMeta (SliceIterator.cc:44-57): view axis k is array axis TM[p_k] over the retained
    # positions p_k in increasing order
    TM = [0, 3, 2, 6, 4, 5, 1, 9, 8, 7, 10]
    retained = sorted(V)
    np.testing.assert_array_equal(vs, [strides[TM[p]] for p in retained])
    # the view's extents: (2,1,4,3,2) in array order (ArraySlicingTest), V's order after transposing
    assert tuple(DIMS[p] for p in retained) == (2, 1, 4, 3, 2)
    assert tuple(DIMS[TM[p]] for p in retained) == (1, 4, 2, 2, 3)
    # the dummies are the remaining axes in increasing order (filterOut, SliceIterator.cc:66-69):
    # the slice of idx = (21,2,4,10,11,3,5,9,23,22,7) is reached at dummy index (21,4,11,3,23,7),
    # at the odometer position of that index over the dummy extents
    dummies = [i for i in range(len(DIMS)) if i not in V]
    filtered = (21, 4, 11, 3, 23, 7)
    assert dummies == [0, 2, 4, 5, 8, 10]
    assert len(off) == int(np.prod([DIMS[i] for i in dummies]))
    k = np.ravel_multi_index(filtered, [DIMS[i] for i in dummies])
    assert off[k] == sum(f * strides[i] for f, i in zip(filtered, dummies))


def test_slices_cover_the_array_once():
    """Slices x retained ranges reach every element of a rank-11 array exactly once."""
    dims = (3, 2, 2, 1, 2, 2, 3, 2, 3, 2, 2)
    strides = np.array([int(np.prod(dims[i + 1:])) for i in range(len(dims))])
    off, vs = O.slices(dims, V)
    np.testing.assert_array_equal(vs, strides[V])
    grid = np.array(np.meshgrid(*[np.arange(dims[v]) * strides[v] for v in V], indexing="ij")).sum(0).ravel()
    reach = (off[:, None] + grid[None, :]).ravel()
    assert len(reach) == int(np.prod(dims)) and len(np.unique(reach)) == len(reach)
