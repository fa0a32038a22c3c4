"""Solid interior of surface-only trees (yx_scene_fill_interior): the empty
space a 6-connected flood from the domain's faces does not reach becomes
FullNode. Pinned against scipy's connected-component labelling of the
synthetic recipe's outside: the filled solid must be the recipe's inside set
plus every empty pocket it encloses, while leaves and colours stay as they
were. An edit of the filled tree then opens solid walls."""
import numpy as np
import pytest
from scipy import ndimage

import svo_helpers as H
from test_scene_tree_edit import check_edit


def expected_solid(inside):
    lab, _ = ndimage.label(~inside, structure=ndimage.generate_binary_structure(3, 1))
    faces = np.unique(np.concatenate([lab[0].ravel(), lab[-1].ravel(), lab[:, 0].ravel(), lab[:, -1].ravel(),
                                      lab[:, :, 0].ravel(), lab[:, :, -1].ravel()]))
    outside = np.isin(lab, faces[faces > 0])
    return ~outside


CASES = [
    dict(depth=5, num_spheres=6, radius_min=2.0, radius_max=7.0, heightfield=False, hf_octaves=5,
         hf_amplitude=0.2, hf_base=0.3, seed=1),
    dict(depth=6, num_spheres=40, radius_min=3.0, radius_max=9.0, heightfield=True, hf_octaves=4,
         hf_amplitude=0.2, hf_base=0.25, seed=7),
]


@pytest.mark.parametrize("desc", CASES)
def test_fill_matches_labelled_outside(yx, desc):
    syn = yx.SVOData.build_synthetic(**desc)
    s = yx.SVOData.from_nodes(syn.nodes(), syn.GetRoot())
    depth = desc["depth"]
    st0, col0 = H.decode_tree(s.nodes(), s.GetRoot(), depth)
    filled = s.fill_interior()
    st1, col1 = H.decode_tree(s.nodes(), s.GetRoot(), depth)
    exp = expected_solid(H.synthetic_inside(**desc))
    assert filled > 0
    assert np.array_equal(st1 > 0, exp)
    assert np.array_equal(st1 == 2, st0 == 2) and np.array_equal(col1, col0)  # leaves untouched
    assert s.info()["num_leaves"] == syn.info()["num_leaves"]
    # filling again finds nothing more
    assert s.fill_interior() == 0


def test_edits_after_fill_open_solid_walls(yx):
    desc = CASES[1]
    syn = yx.SVOData.build_synthetic(**desc)
    s = yx.SVOData.from_nodes(syn.nodes(), syn.GetRoot())
    s.fill_interior()
    for e in [(1, 6, (32.0, 32.0, 10.0), 7.0, 0xFF808080), (0, 6, (20.0, 40.0, 50.0), 5.0, 0xFF0000FF),
              (1, 5, (8.0, 8.0, 5.0), 3.0, 0xFF204060)]:
        check_edit(s, 6, e)
    st, _ = H.decode_tree(s.nodes(), s.GetRoot(), 6)
    assert (st == 1).sum() > 10000


def test_fill_errors(yx):
    syn = yx.SVOData.build_synthetic(**CASES[0])
    with pytest.raises(yx.YoxelError):
        syn.fill_interior()  # a synthetic scene keeps its solid in the recipe
    e = yx.SVOData.from_nodes(np.zeros((0, 10), np.uint32), H.EMPTY)
    with pytest.raises(yx.YoxelError):
        e.fill_interior()


def test_fill_closes_a_cavity(yx):
    """A hollow ball (shell between radii 8 and 13) as a surface-only tree:
    the shell and the cavity it encloses both become solid."""
    n = 32
    c = np.arange(n) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    d2 = (X - 16) ** 2 + (Y - 15) ** 2 + (Z - 17) ** 2
    shell = (d2 <= 13 ** 2) & (d2 > 8 ** 2)
    state = np.where(H.surface(shell), 2, 0).astype(np.uint8)
    color = np.where(state == 2, 0xFF336699, 0).astype(np.uint32)
    nodes, root = H.encode_tree(state, color)
    s = yx.SVOData.from_nodes(nodes, root)
    assert s.fill_interior() > 0
    st, col = H.decode_tree(s.nodes(), s.GetRoot(), 5)
    exp = expected_solid(shell)
    assert np.array_equal(st > 0, exp)
    assert (exp & ~shell).sum() > 1500  # the cavity
    assert np.array_equal(st == 2, state == 2)


def test_fill_on_word_rows(yx):
    """Depth 7: rows of two 64-bit words, runs across word edges, pockets."""
    desc = dict(depth=7, num_spheres=60, radius_min=3.0, radius_max=14.0, heightfield=True, hf_octaves=4,
                hf_amplitude=0.25, hf_base=0.3, seed=21)
    syn = yx.SVOData.build_synthetic(**desc)
    s = yx.SVOData.from_nodes(syn.nodes(), syn.GetRoot())
    s.fill_interior()
    st, _ = H.decode_tree(s.nodes(), s.GetRoot(), 7)
    assert np.array_equal(st > 0, expected_solid(H.synthetic_inside(**desc)))
