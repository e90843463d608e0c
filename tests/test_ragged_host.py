"""Host-side checks of the ragged-list entry point (pdlhip_ragged_prepare) in
plan-only mode: argument validation, the per-node PDL trans types and the
output types it creates, without a device."""
import pytest

from pdl_amd import core as C
from pdl_amd._lib import PDLError
from oracle import pdlcpu as O


@pytest.fixture
def plan_only():
    C.set_plan_only(True)
    yield
    C.set_plan_only(False)


def vec(t, n, state=0):
    return C.shell(t, [n], state=state)


def scalar(t):
    return C.shell(t, [])


def expect(msg, items, chain, final, outputs=None):
    with pytest.raises(PDLError) as e:
        C.Ragged(items, chain, final, outputs)
    assert msg in str(e.value), str(e.value)


def test_output_types_follow_pdl(plan_only):
    chain = [("minus", 0, 1), ("power", ("node", 0), 2), ("mult", ("node", 1), 3)]
    items = [[vec(O.F, n), vec(O.F, n), scalar(O.SB), vec(O.B, n)] for n in (1, 5, 9000)]
    for final, want in (("sumover", O.F), ("dsumover", O.D), ("average", O.F), ("daverage", O.D),
                        ("minimum", O.F), ("maximum", O.F)):
        r = C.Ragged(items, chain, final)
        assert [o.dtype for o in r.outputs] == [want] * 3, final
        assert all(o.dims == [] for o in r.outputs)
    # an int chain scans into int+ (max(long, T)), with the item's length
    for t, want in ((O.SB, O.L), (O.US, O.L), (O.L, O.L), (O.UL, O.UL), (O.LL, O.LL)):
        r = C.Ragged([[vec(t, 7)], [vec(t, 0)]], [], "cumusumover")
        assert [o.dtype for o in r.outputs] == [want, want]
        assert [o.dims for o in r.outputs] == [[7], [0]]


def test_double_scalar_promotes_the_chain(plan_only):
    # float minus a Perl float (double) scalar computes in double, as PDL's eager op would
    r = C.Ragged([[vec(O.F, 4), scalar(O.D)]], [("minus", 0, 1)], "sumover")
    assert r.outputs[0].dtype == O.D


def test_mixed_chains_take_each_nodes_type(plan_only):
    """a node in long feeding a float node: each node keeps PDL's trans
    type, the chain's (and the reduction's) is the last node's"""
    r = C.Ragged([[vec(O.L, 4), vec(O.L, 4), vec(O.F, 4)]], [("plus", 0, 1), ("mult", ("node", 0), 2)], "sumover")
    assert r.outputs[0].dtype == O.F
    r = C.Ragged([[vec(O.F, 4), vec(O.F, 4), vec(O.D, 4)]],
                 [("or2", 0, 1), ("plus", ("node", 0), 2)], "minimum", values=True)
    assert r.outputs[0].dtype == O.D and r.values[0].dtype == O.D


def test_refusals(plan_only):
    f4, f5 = vec(O.F, 4), vec(O.F, 5)
    expect("lengths 4 and 5 differ", [[f4, f5]], [("plus", 0, 1)], "sumover")
    expect("has bad values", [[vec(O.F, 4, state=O.BADVAL)]], [], "sumover")
    expect("unsupported final op", [[f4]], [], "prodover")
    expect("takes one input and no chain", [[vec(O.L, 4)]], [("plus", 0, 0)], "cumusumover")
    expect("integer types only", [[f4]], [], "cumusumover")
    expect("float or double only", [[vec(O.L, 4)]], [], "sumover")
    # a mixed chain whose last node is an integer type: no float reduction
    expect("float or double only", [[vec(O.F, 4), vec(O.F, 4)]], [("plus", 0, 1), ("or2", ("node", 0), 0)],
           "sumover")
    expect("differs in type or rank", [[vec(O.F, 4)], [vec(O.D, 4)]], [], "sumover")
    expect("is not 1-D", [[C.shell(O.F, [2, 2])]], [], "sumover")
    expect("must exist with type", [[f4]], [], "sumover", outputs=[C.shell(O.D, [])])
    expect("wrong shape", [[f4]], [], "sumover", outputs=[C.shell(O.F, [2])])
    expect("node 0 operand", [[f4]], [("plus", 0, 3)], "sumover")
    expect("node 0 operand", [[f4]], [("plus", 0, ("node", 0))], "sumover")


def test_chain_shapes_follow_broadcasting(plan_only):
    a = C.shell(O.F, [64, 50, 3])
    b = C.shell(O.F, [64, 1, 3])
    w = C.shell(O.SB, [])
    ch = C.Chain([a, b, w], [("minus", 0, 1), ("mult", ("node", 0), 2)], "sumover", value=True)
    assert ch.value.dims == [64, 50, 3] and ch.value.dtype == O.F
    assert ch.result.dims == [50, 3] and ch.result.dtype == O.F
    ch = C.Chain([C.shell(O.L, [7, 4])], [], "cumusumover")
    assert ch.result.dims == [7, 4] and ch.result.dtype == O.L
    with pytest.raises(PDLError) as e:
        C.Chain([C.shell(O.F, [4, 5]), C.shell(O.F, [4, 6])], [("plus", 0, 1)], "sumover")
    assert "mismatched broadcast dim 1: size 5 vs 6" in str(e.value)
    with pytest.raises(PDLError) as e:
        C.Chain([C.shell(O.F, [4])], [], "none")
    assert "needs a value output" in str(e.value)


def test_prepared_matmult_list_refuses_entries_that_cannot_join(plan_only):
    ok = C.MatmultList([C.shell(O.F, [3, 4]), C.shell(O.F, [5, 2])], [C.shell(O.F, [6, 3]), C.shell(O.F, [7, 5])])
    assert [o.dims for o in ok.outputs] == [[6, 4], [7, 2]]
    with pytest.raises(PDLError) as e:
        C.MatmultList([C.shell(O.F, [3, 4]), C.shell(O.D, [5, 2])], [C.shell(O.F, [6, 3]), C.shell(O.D, [7, 5])])
    assert "another type than the first" in str(e.value)


def test_explicit_broadcast_ids_are_refused(plan_only):
    """the ragged and chain planners take implicit broadcasting only"""
    a = C.shell(O.F, [4, 3]).broadcast(1)
    with pytest.raises(PDLError) as e:
        C.Chain([a, C.shell(O.F, [4, 3])], [("plus", 0, 1)], "sumover")
    assert "explicit broadcast ids" in str(e.value)
    with pytest.raises(PDLError) as e:
        C.Ragged([[C.shell(O.F, [4]).broadcast(0)]], [], "sumover")
    assert "explicit broadcast ids" in str(e.value)
