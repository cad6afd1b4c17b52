"""Node numbering, naming, unrooting and post-order of the Newick parser
against tables derived by hand from the reference's rules (CPU), applied to
both the product's Tree (_hyphy3) and the oracle's restatement
(oracle/data_ref.py).

The rules (include/hyphy/core/tree.hpp:273-585, 591-690, 881-892):
  * ids in creation order: '(' creates an internal node (child of the node
    on top of the stack, then pushed); ',' or ')' after a non-empty token
    creates a leaf (child of the stack top); an empty token creates nothing;
  * after ')' the text up to the next , ) ( ; labels the closed node
    (name, ':' length, '{tag}'); quotes protect characters and '' escapes a
    quote; [comments] outside quotes and whitespace are dropped;
  * names: with the named leaves' names taken, nodes in id order get
    "Node<k>" (internal: unnamed or name already taken) or "Leaf<k>"
    (unnamed leaf), k one counter over both, skipping taken names;
  * unrooting (bifurcating root, more than three nodes): an internal child of
    the root is absorbed (the first one, unless only the second is untagged
    while the first has a tag), its length added to the other child's, its
    children spliced into the root's child list in its place (before or after
    the other child), and ids compacted;
  * post-order: children in list order, then the node.

Each table row: (id, name, parent, children, is_leaf, branch length, tag)."""
import pytest

from oracle import data_ref as dr

h3 = pytest.importorskip("_hyphy3")

CASES = [
    # rooted, bifurcating root kept
    ("((A:0.1,B:0.2)x:0.3,(C:0.4,D:0.5):0.6);", True, 0, [2, 3, 1, 5, 6, 4, 0], [
        (0, "Node1", -1, [1, 4], False, 0.0, ""),
        (1, "x", 0, [2, 3], False, 0.3, ""),
        (2, "A", 1, [], True, 0.1, ""),
        (3, "B", 1, [], True, 0.2, ""),
        (4, "Node2", 0, [5, 6], False, 0.6, ""),
        (5, "C", 4, [], True, 0.4, ""),
        (6, "D", 4, [], True, 0.5, ""),
    ]),
    # the same tree unrooted: node 1 ("x") absorbed, its children first
    ("((A:0.1,B:0.2)x:0.3,(C:0.4,D:0.5):0.6);", False, 0, [1, 2, 4, 5, 3, 0], [
        (0, "Node1", -1, [1, 2, 3], False, 0.0, ""),
        (1, "A", 0, [], True, 0.1, ""),
        (2, "B", 0, [], True, 0.2, ""),
        (3, "Node2", 0, [4, 5], False, 0.6 + 0.3, ""),
        (4, "C", 3, [], True, 0.4, ""),
        (5, "D", 3, [], True, 0.5, ""),
    ]),
    # polytomy, quoted labels (space, escaped quote), tags, a comment
    ("(('Homo sapiens':0.1,B{FG}:0.2,'it''s':0.3)'anc one'{FG}:0.4,[note]C:0.5,D:0.6);", False, 0,
     [2, 3, 4, 1, 5, 6, 0], [
         (0, "Node1", -1, [1, 5, 6], False, 0.0, ""),
         (1, "anc one", 0, [2, 3, 4], False, 0.4, "FG"),
         (2, "Homo sapiens", 1, [], True, 0.1, ""),
         (3, "B", 1, [], True, 0.2, "FG"),
         (4, "it's", 1, [], True, 0.3, ""),
         (5, "C", 0, [], True, 0.5, ""),
         (6, "D", 0, [], True, 0.6, ""),
     ]),
    # internal labels that collide with taken names are renamed
    ("((A:1,B:2)A:3,(C:4,D:5)Node1:6,E:7);", False, 0, [2, 3, 1, 5, 6, 4, 7, 0], [
        (0, "Node1", -1, [1, 4, 7], False, 0.0, ""),
        (1, "Node2", 0, [2, 3], False, 3.0, ""),
        (2, "A", 1, [], True, 1.0, ""),
        (3, "B", 1, [], True, 2.0, ""),
        (4, "Node3", 0, [5, 6], False, 6.0, ""),
        (5, "C", 4, [], True, 4.0, ""),
        (6, "D", 4, [], True, 5.0, ""),
        (7, "E", 0, [], True, 7.0, ""),
    ]),
    # unnamed leaves (length only) share the counter with internal names
    ("((:0.1,:0.2):0.3,X:0.4,(Y:0.5,:0.6)lab:0.7);", False, 0, [2, 3, 1, 4, 6, 7, 5, 0], [
        (0, "Node1", -1, [1, 4, 5], False, 0.0, ""),
        (1, "Node2", 0, [2, 3], False, 0.3, ""),
        (2, "Leaf3", 1, [], True, 0.1, ""),
        (3, "Leaf4", 1, [], True, 0.2, ""),
        (4, "X", 0, [], True, 0.4, ""),
        (5, "lab", 0, [6, 7], False, 0.7, ""),
        (6, "Y", 5, [], True, 0.5, ""),
        (7, "Leaf5", 5, [], True, 0.6, ""),
    ]),
    # unrooting with a leaf child: the internal child is absorbed, the other
    # child first; the absorbed node's name disappears with it
    ("(A:0.1,(B:0.2,(C:0.3,D:0.4):0.5):0.6);", False, 0, [1, 2, 4, 5, 3, 0], [
        (0, "Node1", -1, [1, 2, 3], False, 0.0, ""),
        (1, "A", 0, [], True, 0.1 + 0.6, ""),
        (2, "B", 0, [], True, 0.2, ""),
        (3, "Node3", 0, [4, 5], False, 0.5, ""),
        (4, "C", 3, [], True, 0.3, ""),
        (5, "D", 3, [], True, 0.4, ""),
    ]),
    # unrooting keeps a tagged clade: the untagged second child is absorbed
    ("((A,B){FG}:0.1,(C,D):0.2);", False, 0, [2, 3, 1, 4, 5, 0], [
        (0, "Node1", -1, [1, 4, 5], False, 0.0, ""),
        (1, "Node2", 0, [2, 3], False, 0.1 + 0.2, "FG"),
        (2, "A", 1, [], True, 0.0, ""),
        (3, "B", 1, [], True, 0.0, ""),
        (4, "C", 0, [], True, 0.0, ""),
        (5, "D", 0, [], True, 0.0, ""),
    ]),
    # an empty token creates no leaf: "(,C)" is an internal node with one child
    ("((,C:1)u:2,D:3,E:4);", True, 0, [2, 1, 3, 4, 0], [
        (0, "Node1", -1, [1, 3, 4], False, 0.0, ""),
        (1, "u", 0, [2], False, 2.0, ""),
        (2, "C", 1, [], True, 1.0, ""),
        (3, "D", 0, [], True, 3.0, ""),
        (4, "E", 0, [], True, 4.0, ""),
    ]),
]


def rows_product(t):
    return [(n.id, n.name, n.parent_id, list(n.children), bool(n.is_leaf), n.branch_length, n.model_tag)
            for n in t.nodes]


def rows_oracle(r):
    return [(n.id, n.name, n.parent, list(n.children), bool(n.is_leaf), n.bl, n.tag or "") for n in r.nodes]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_product_tree_matches_hand_table(k):
    nwk, keep, root, post, table = CASES[k]
    t = h3.Tree.from_newick(nwk, keep)
    assert t.root_id == root
    assert list(t.post_order) == post
    assert rows_product(t) == table


@pytest.mark.parametrize("k", range(len(CASES)))
def test_oracle_tree_matches_hand_table(k):
    nwk, keep, root, post, table = CASES[k]
    r = dr.TreeRef.from_newick(nwk, keep_rooted=keep)
    assert r.root == root
    assert list(r.post_order) == post
    assert rows_oracle(r) == table


def test_cd2_reduced_tree_counts():
    """tests/test_core.cpp:37-52 on the cd2_reduced tree."""
    nwk = ("((((Pig:0.147969,Cow:0.213430):0.085099,Horse:0.165787,Cat:0.264806):0.058611,((RhMonkey:0.002015,"
           "Baboon:0.003108):0.022733,(Human:0.004349,Chimp:0.000799):0.011873):0.101856):0.340802,Rat:0.050958,"
           "Mouse:0.097950);")
    for t in (h3.Tree.from_newick(nwk, True), h3.Tree.from_newick(nwk, False)):
        assert t.num_leaves() == 10
        assert t.root_id >= 0
        assert len(t.post_order) == t.num_nodes()
        assert t.post_order[-1] == t.root_id
        for name in ("Human", "Chimp", "Pig"):
            assert name in t.leaf_name_to_id
    r = dr.TreeRef.from_newick(nwk, keep_rooted=True)
    assert sum(1 for n in r.nodes if n.is_leaf) == 10 and r.post_order[-1] == r.root
