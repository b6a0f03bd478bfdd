"""NetworkBlueprint parsing and validation on the host (netfile.py:48-527;
test_netfile.py parser / ORDER / slot tests)."""

import pytest

from tenax_amd.network.netfile import NetworkBlueprint, _parse_order, parse_netfile


def test_parse_basic_and_comments():
    spec = parse_netfile("# c\nA: i, j\n\nB: j, k\nTOUT: i, k\nORDER: (A, B)\n")
    assert spec == {"tensors": {"A": ["i", "j"], "B": ["j", "k"]}, "tout": ["i", "k"], "order": "(A, B)"}


def test_parse_empty_tout_is_scalar():
    assert parse_netfile("A: i\nB: i\nTOUT:")["tout"] == []
    assert parse_netfile("A: i\nB: i")["tout"] is None


@pytest.mark.parametrize("text,msg", [
    ("A i j", "expected 'Name: labels'"),
    ("A: i\nA: j", "duplicate tensor name"),
    ("A: i\nTOUT: i\nTOUT: i", "duplicate TOUT"),
    ("A: i\nORDER: (A,A)\nORDER: (A,A)", "duplicate ORDER"),
    ("A:", "has no labels"),
    ("# only a comment", "No tensor declarations"),
    ("A: i\nB: i\nC: i", "appears on 3 tensors"),
])
def test_parse_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        parse_netfile(text)


def test_order_flattening():
    assert _parse_order("((A,B),C)", {"A", "B", "C"}) == [("A", "B"), ("_0", "C")]
    assert _parse_order("(A,(B,C))", {"A", "B", "C"}) == [("B", "C"), ("A", "_0")]
    assert _parse_order("((A,B),(C,D))", set("ABCD")) == [("A", "B"), ("C", "D"), ("_0", "_1")]


@pytest.mark.parametrize("order,msg", [("(A,X)", "unknown tensor name"), ("(A,B", "expected '\\)'"),
                                       ("(A B)", "expected ','"), ("(A,B))", "unexpected tokens"),
                                       ("(", "unexpected end")])
def test_order_errors(order, msg):
    with pytest.raises(ValueError, match=msg):
        _parse_order(order, {"A", "B"})


def test_blueprint_properties_and_subscripts():
    bp = NetworkBlueprint("L: a, b, c\nM1: a, p, q, d\nA: b, p, s, e\nM2: e, q, t, f\nR: d, f, g\nTOUT: c, s, t, g")
    assert bp.tensor_names == ["L", "M1", "A", "M2", "R"]
    assert bp.subscripts == "abc,ahid,bhje,eikf,dfg->cjkg"
    assert bp.output_labels == ["c", "s", "t", "g"]
    assert not bp.is_ready()
    with pytest.raises(RuntimeError, match="missing tensors"):
        bp.launch()
    with pytest.raises(KeyError, match="Unknown tensor name"):
        bp.put_tensor("X", None)


def test_tout_must_be_free():
    with pytest.raises(ValueError, match="not a free label"):
        NetworkBlueprint("A: i, j\nB: j, k\nTOUT: j")
