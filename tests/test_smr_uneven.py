"""MeshBlocks of uneven sizes on refinement levels (smr.uneven_leaves, split_uneven,
merge_uneven), host side: the leaves tile the root grid as a valid tree (2:1 balanced),
splitting a block's arrays into its leaves and merging them back is the identity, and a
leaf's cells are the block's cells at the leaf's offset."""
import numpy as np

from tests.conftest import load_smr

NX, NG = 4, 2
FACE_EXT = {"b1f": (0, 0, 1), "b2f": (0, 1, 0), "b3f": (1, 0, 0)}


def three_level_blocks():
    """root grid 2^3 of 2x-size blocks (4^3 base blocks); one root block refined into
    level-1 blocks, seven of 2x size and eight of base size, one of which is refined
    again into eight level-2 base blocks"""
    blocks = [(0, 2 * x, 2 * y, 2 * z, 1) for z in range(2) for y in range(2) for x in range(2)
              if (x, y, z) != (1, 1, 1)]
    blocks += [(1, x, y, z, 1) for z in (4, 6) for y in (4, 6) for x in (4, 6) if (x, y, z) != (4, 4, 4)]
    blocks += [(1, 4 + a, 4 + b, 4 + c, 0) for c in range(2) for b in range(2) for a in range(2)
               if (a, b, c) != (1, 1, 1)]
    blocks += [(2, 10 + a, 10 + b, 10 + c, 0) for c in range(2) for b in range(2) for a in range(2)]
    return blocks


def test_leaves_form_a_balanced_tree():
    smr = load_smr()
    blocks = three_level_blocks()
    leaves, owner = smr.uneven_leaves(blocks)
    assert len(leaves) == 7 * 8 + 7 * 8 + 7 + 8 == len(owner)
    tree = smr.SMRTree((4, 4, 4), leaves)   # raises unless the leaves tile the root grid, 2:1
    assert tree.maxlev == 2


def test_split_merge_identity_and_offsets():
    smr = load_smr()
    blocks = three_level_blocks()
    leaves, owner = smr.uneven_leaves(blocks)
    rng = np.random.default_rng(3)
    for name, nv in (("u", 5), ("bcc", 3), ("b1f", None), ("b2f", None), ("b3f", None)):
        e = FACE_EXT.get(name, (0, 0, 0))
        arrs = []
        for (_, _, _, _, p) in blocks:
            n = (NX << p) + 2 * NG
            arrs.append(rng.standard_normal(((nv,) if nv else ()) + (n + e[0], n + e[1], n + e[2])))
        sub = smr.split_uneven(blocks, owner, name, arrs, NX, NG)
        back = smr.merge_uneven(blocks, owner, name, sub, NX, NG)
        for a, b in zip(arrs, back):
            assert np.array_equal(a, b), name
        for g, (bi, a, b, c) in enumerate(owner):
            # the leaf's first own cell (or face) is the block's cell at the leaf's offset
            lead = (slice(None),) if nv else ()
            want = arrs[bi][lead + (NG + c * NX, NG + b * NX, NG + a * NX)]
            assert np.array_equal(sub[g][lead + (NG, NG, NG)], want), (name, g)


def test_corner_not_aligned_refused():
    smr = load_smr()
    try:
        smr.uneven_leaves([(0, 1, 0, 0, 1)])
    except ValueError as e:
        assert "multiple of 2" in str(e)
    else:
        raise AssertionError("misaligned block accepted")
