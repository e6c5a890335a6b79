"""CPU: EDB ingestion (TSV / CSV / JSON with the reference's column
conventions, crates/strata-k/src/input.rs) and canonical rendering
(crates/strata-cli/src/render.rs), including the routing workload's pinned
output lines (crates/strata-cli/tests/workloads.rs:25-32) through the
oracle."""
import os

import pytest

from oracle import oracle as O
from strata_k.core_ir import BOOL, INF, TROP, CoreProgram, IntV, Sym, atom, pos, pred, rule, v
from strata_k import io as IO
from strata_k.core_ir import Agg, tc_program
from strata_k.io import LoadError, SymbolDict, load_input, render_db, split_csv_line

HERE = os.path.dirname(os.path.abspath(__file__))


def test_csv_splitter_reference_cases():
    # input.rs:339-358
    assert split_csv_line("a,b") == ["a", "b"]
    assert split_csv_line('"a","b"') == ["a", "b"]
    assert split_csv_line('"a""b",c') == ['a"b', "c"]
    assert split_csv_line('"",b') == ["", "b"]
    for bad in ('"a"junk,b', 'a"b,c', '"abc'):
        with pytest.raises(LoadError):
            split_csv_line(bad)


def test_symbol_dict_first_seen():
    d = SymbolDict()
    assert [d.intern(s) for s in ("b", "a", "b", "c")] == [0, 1, 0, 2]
    assert d.get("a") == 1 and d.get("z") is None
    assert d.resolve(2) == "c" and d.resolve(3) is None and len(d) == 3


def test_tsv_csv_json_rows(tmp_path):
    d = SymbolDict()
    (tmp_path / "w.tsv").write_text("x\ty\t7\n\ny\tz\t-3\r\n")
    facts = load_input(str(tmp_path / "w.tsv"), "w", 2, TROP, ["node", "node"], d)
    assert facts == [("w", (Sym(0), Sym(1)), 7), ("w", (Sym(1), Sym(2)), -3)]
    (tmp_path / "n.csv").write_text('"a,b", 5\n"c""d",12\n')
    facts = load_input(str(tmp_path / "n.csv"), "n", 2, BOOL, ["node", "int"], d)
    assert facts == [("n", (Sym(3), IntV(5)), None), ("n", (Sym(4), IntV(12)), None)]
    assert d.resolve(3) == "a,b" and d.resolve(4) == 'c"d'
    (tmp_path / "j.json").write_text('[["x", 4, 10], ["q", "5", 2]]')
    facts = load_input(str(tmp_path / "j.json"), "j", 2, TROP, ["node", "int"], d)
    assert facts == [("j", (Sym(0), IntV(4)), 10), ("j", (Sym(5), IntV(5)), 2)]


@pytest.mark.parametrize("name,text,msg", [
    ("a.tsv", "x\t\n", "empty value in column 2"),
    ("b.tsv", "x\ty\tz\n", "expects 2 column(s), found 3"),
    ("c.json", '[["x", 1.5]]', "is not an integer"),
    ("d.json", '[["x", 3]]', "is a symbol column, got the number 3"),
    ("e.json", '{"x": 1}', "expected a JSON array of rows"),
    ("f.csv", 'x,"y"z\n', "trailing characters"),
    ("g.txt", "x\ty\n", "unsupported input format"),
])
def test_load_errors_name_the_line(tmp_path, name, text, msg):
    (tmp_path / name).write_text(text)
    with pytest.raises(LoadError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        load_input(str(tmp_path / name), "p", 2, BOOL, ["node", "node"], SymbolDict())


@pytest.mark.parametrize("cell", ["\u00b2", "\u0663", "1\u0663", "+", "-", "", "9223372036854775808", "1_000"])
def test_int_cells_are_ascii_i64(tmp_path, cell):
    """`s.trim().parse::<i64>()` (input.rs) takes an optional sign and ASCII
    digits within i64 only: superscripts, other scripts' digits, a bare sign
    and out-of-range values are load errors that name the line."""
    (tmp_path / "n.tsv").write_text(f"a\t{cell}\n", encoding="utf-8")
    with pytest.raises(LoadError, match="n.tsv:1:"):
        load_input(str(tmp_path / "n.tsv"), "n", 2, BOOL, ["node", "int"], SymbolDict())
    (tmp_path / "w.tsv").write_text(f"a\tb\t{cell}\n", encoding="utf-8")
    with pytest.raises(LoadError, match="w.tsv:1:"):
        load_input(str(tmp_path / "w.tsv"), "w", 2, TROP, ["node", "node"], SymbolDict())


def test_json_integers_outside_i64_are_not_integers(tmp_path):
    """serde_json's as_i64 fails outside i64, so such a number is a float
    cell (input.rs:316-321): refused in an int column and as a weight."""
    for big in (2**63, 2**64, -(2**63) - 1):
        (tmp_path / "j.json").write_text(f'[["a", {big}]]')
        with pytest.raises(LoadError, match="j.json:1:"):
            load_input(str(tmp_path / "j.json"), "j", 2, BOOL, ["node", "int"], SymbolDict())
        (tmp_path / "k.json").write_text(f'[["a", "b", {big}]]')
        with pytest.raises(LoadError, match="k.json:1:"):
            load_input(str(tmp_path / "k.json"), "k", 2, TROP, ["node", "node"], SymbolDict())
    (tmp_path / "e.json").write_text(f'[["a", {2**63 - 1}], ["b", {-(2**63)}]]')
    facts = load_input(str(tmp_path / "e.json"), "e", 2, BOOL, ["node", "int"], SymbolDict())
    assert [f[1][1].value for f in facts] == [2**63 - 1, -(2**63)]


def test_trop_needs_integer_weight(tmp_path):
    (tmp_path / "w.tsv").write_text("x\ty\tcheap\n")
    with pytest.raises(LoadError, match="w.tsv:1: `w` needs a trailing integer weight"):
        load_input(str(tmp_path / "w.tsv"), "w", 2, TROP, ["node", "node"], SymbolDict())


class _OracleDb:
    """The oracle's result in the engine's Db shape (predicates, rows)."""

    class _Rel:
        def __init__(self, r):
            self.r = r

        def rows(self):
            conv = lambda x: Sym(x[1]) if x[0] == "sym" else IntV(x[1])
            return [(tuple(conv(x) for x in t), INF if a == "inf" else a) for t, a in self.r.rows()]

    def __init__(self, res):
        self.res = res

    def predicates(self):
        return sorted(self.res.relations)

    def relation(self, name):
        return self._Rel(self.res.relations[name])


def routing_program():
    return CoreProgram([pred("link", 2, TROP), pred("route", 2, TROP)],
                       [rule(atom("route", v(0), v(1)), [pos("link", v(0), v(1))], 2),
                        rule(atom("route", v(0), v(2)), [pos("link", v(0), v(1)), pos("route", v(1), v(2))], 3)], 1)


def test_routing_workload_pinned_lines():
    d = SymbolDict()
    facts = load_input(os.path.join(HERE, "golden", "routing_links.tsv"), "link", 2, TROP, ["node", "node"], d)
    assert len(facts) == 228
    res = O.run(routing_program(), facts)
    assert res.ok, res.status
    queries = [("route", (Sym(d.get("n0_0")), Sym(d.get("n7_7")))), ("route", (Sym(d.get("n0_0")), Sym(d.get("n4_4")))),
               ("route", (Sym(d.get("n7_7")), Sym(d.get("n0_0"))))]
    out = render_db(_OracleDb(res), d, filter=queries).splitlines()
    for line in ("route(n0_0, n7_7) = 5", "route(n0_0, n4_4) = 2", "route(n7_7, n0_0) = 44"):
        assert line in out, (line, out)
    assert len(out) == 3
    full = render_db(_OracleDb(res), d)
    assert full.startswith("link(n0_0, n0_1) = 1\n")
    assert full.count("\n") == len(res.relations["link"]) + len(res.relations["route"])


def test_render_inf_and_ints():
    class R:
        def rows(self):
            return [((Sym(0), IntV(-4)), INF), ((Sym(1), IntV(3)), 12)]

    class D:
        def predicates(self):
            return ["p"]

        def relation(self, n):
            return R()

    d = SymbolDict()
    d.intern("a")
    d.intern("b")
    assert render_db(D(), d) == "p(a, -4) = inf\np(b, 3) = 12\n"


# ---------------------------------------------------------------- IR JSON

def test_weight_json_reference_cases():
    """crates/strata-ir/src/trop.rs:199-209: 42 <-> 42, +inf <-> "inf",
    "nan" refused; serde also refuses non-i64 numbers and booleans."""
    import json
    assert json.dumps(IO.weight_to_json(42)) == "42"
    assert json.dumps(IO.weight_to_json(INF)) == '"inf"'
    assert IO.weight_from_json(json.loads("42")) == 42
    assert IO.weight_from_json(json.loads('"inf"')) is INF
    for bad in ['"nan"', '"5"', "1.5", "true", str(2**63), "null"]:
        with pytest.raises(LoadError):
            IO.weight_from_json(json.loads(bad))
    assert IO.weight_from_json(-(2**63)) == -(2**63)


# The Core-IR program of crates/strata-ir/tests/roundtrip.rs:155-186, written
# out in serde's encoding (snake_case fields, {"kind","data"} tags).
CORE_TC_JSON = """{
  "predicates": [
    {"name": "edge", "arity": 2, "semiring": "bool", "stratum": 0},
    {"name": "path", "arity": 2, "semiring": "bool", "stratum": 0}
  ],
  "rules": [
    {"head": {"pred": "path", "args": [{"kind": "var", "data": {"slot": 0}}, {"kind": "var", "data": {"slot": 2}}]},
     "body": [
       {"kind": "pos", "data": {"pred": "edge", "args": [{"kind": "var", "data": {"slot": 0}},
                                                         {"kind": "var", "data": {"slot": 1}}]}},
       {"kind": "pos", "data": {"pred": "path", "args": [{"kind": "var", "data": {"slot": 1}},
                                                         {"kind": "var", "data": {"slot": 2}}]}}],
     "stratum": 0, "var_count": 3}
  ],
  "num_strata": 1
}"""


def test_core_program_json_roundtrip():
    import json
    prog = IO.core_program_from_json(CORE_TC_JSON)
    assert prog.num_strata == 1 and len(prog.rules_in_stratum(0)) == 1
    assert prog.predicates[1].semiring == BOOL
    assert prog.rules[0].head.args[1].slot == 2 and prog.rules[0].var_count == 3
    assert IO.core_program_to_json(prog) == json.loads(CORE_TC_JSON)
    # every shape this engine runs survives the round trip
    for p in (tc_program(BOOL), tc_program(TROP), IO.core_program_from_json(CORE_TC_JSON)):
        assert IO.core_program_from_json(json.dumps(IO.core_program_to_json(p))) == p
    agg = {"kind": "agg", "data": {"op": "min", "slot": 1}}
    assert IO._core_term(agg, "t") == Agg("min", 1) and IO._term_json(Agg("min", 1)) == agg


@pytest.mark.parametrize("mutate,msg", [
    (lambda d: d["predicates"][0].update(semiring="prob"), "semiring"),
    (lambda d: d["rules"][0]["head"]["args"][0].update(kind="compound"), "compound"),
    (lambda d: d["rules"][0]["body"][0].update(kind="maybe"), "literal kind"),
    (lambda d: d.pop("num_strata"), "missing field"),
    (lambda d: d["rules"][0]["head"]["args"][0]["data"].update(slot=-1), "u32"),
])
def test_core_program_json_errors(mutate, msg):
    import json
    d = json.loads(CORE_TC_JSON)
    mutate(d)
    with pytest.raises(LoadError, match=msg):
        IO.core_program_from_json(d)


def _const(name):
    return {"kind": "const", "data": {"name": name}}


def _fact(p, args, **extra):
    return {"kind": "fact", "data": dict({"atom": {"pred": p, "args": args}}, **extra)}


def test_high_ir_facts_sssp_with_inf():
    """The sssp() document of crates/strata-ir/tests/roundtrip.rs:67-110: a
    finite and a +inf Trop fact; constants interned first seen."""
    doc = {"ir_version": "0.1.0", "items": [
        {"kind": "domain", "data": {"name": "node"}},
        _fact("edge", [_const("a"), _const("b")], weight=5),
        _fact("edge", [_const("b"), _const("c")], weight="inf"),
        _fact("hop", [_const("c"), {"kind": "int", "data": {"value": -3}}])]}
    d = SymbolDict()
    facts = IO.high_ir_facts(doc, d)
    assert facts == [("edge", (Sym(0), Sym(1)), 5), ("edge", (Sym(1), Sym(2)), INF), ("hop", (Sym(2), IntV(-3)), None)]
    assert [d.resolve(i) for i in range(3)] == ["a", "b", "c"]


@pytest.mark.parametrize("version,ok", [("0.1.0", True), ("0.1.9", True), ("0.2.0", False), ("1.1.0", False)])
def test_high_ir_version_gate(version, ok):
    doc = {"ir_version": version, "items": [_fact("e", [_const("a")])]}
    if ok:
        assert IO.high_ir_facts(doc, SymbolDict()) == [("e", (Sym(0),), None)]
    else:
        with pytest.raises(LoadError, match="incompatible"):
            IO.high_ir_facts(doc, SymbolDict())
    for bad in ("0.1", "0.1.0.0", "a.b.c"):
        with pytest.raises(LoadError, match="bad ir_version"):
            IO.high_ir_facts({"ir_version": bad, "items": []}, SymbolDict())


def test_high_ir_facts_refusals():
    for fact in (_fact("e", [_const("a")], prob=0.5), _fact("e", [{"kind": "var", "data": {"name": "X"}}]),
                 _fact("e", [_const("a")], weight="nan")):
        with pytest.raises(LoadError):
            IO.high_ir_facts({"ir_version": "0.1.0", "items": [fact]}, SymbolDict())


def test_json_program_and_facts_through_the_oracle():
    """A Core-IR document (the program above plus its base rule) and High-IR
    facts evaluate to the reference's interp.rs:84-114 answer: a->b->c->d
    gives 6 path tuples."""
    import json
    d = json.loads(CORE_TC_JSON)
    var = lambda i: {"kind": "var", "data": {"slot": i}}  # noqa: E731
    d["rules"].insert(0, {"head": {"pred": "path", "args": [var(0), var(1)]},
                          "body": [{"kind": "pos", "data": {"pred": "edge", "args": [var(0), var(1)]}}],
                          "stratum": 0, "var_count": 2})
    prog = IO.core_program_from_json(json.dumps(d))
    doc = {"ir_version": "0.1.0",
           "items": [_fact("edge", [_const(x), _const(y)]) for x, y in ("ab", "bc", "cd")]}
    r = O.run(prog, IO.high_ir_facts(doc, SymbolDict()))
    assert r.ok and len(r.relations["path"]) == 6


# ------------------------------------------- the library's JSON decoder (C)

def test_native_json_decoder_matches_python():
    """sk_program_from_json (no device) decodes the same programs as the
    Python decoder, including every shape this engine runs and aggregates."""
    import json
    import strata_k as sk
    from strata_k.core_ir import CoreProgram, cost_program, neg, triangle_program
    agg = CoreProgram([pred("e", 2, BOOL), pred("n", 1, BOOL), pred("deg", 2, BOOL, 1), pred("iso", 1, BOOL, 1)],
                      [rule(atom("deg", v(0), Agg("count", 1)), [pos("e", v(0), v(1))], 2, 1),
                       rule(atom("iso", v(0)), [pos("n", v(0)), neg("e", v(0), v(0))], 1, 1)], 2)
    for p in (IO.core_program_from_json(CORE_TC_JSON), tc_program(BOOL), tc_program(TROP), cost_program(),
              triangle_program(), agg):
        text = json.dumps(IO.core_program_to_json(p))
        assert sk.program_from_json(text) == IO.core_program_from_json(text) == p
    assert sk.program_from_json(CORE_TC_JSON) == IO.core_program_from_json(CORE_TC_JSON)


def test_native_json_decoder_serde_rules():
    """Unknown fields ignored and neg_weight_cycle_check defaulted, as serde
    does; string escapes (a surrogate pair included) decoded."""
    import json
    import strata_k as sk
    d = json.loads(CORE_TC_JSON)
    d["comment"] = {"nested": [1, 2, {"x": None}]}
    d["rules"][0]["neg_weight_cycle_check"] = True
    d["predicates"][0]["name"] = "edgé\U0001F600"
    d["rules"][0]["body"][0]["data"]["pred"] = "edgé\U0001F600"
    text = json.dumps(d)  # ASCII with \\u escapes
    assert "\\ud83d\\ude00" in text
    p = sk.program_from_json(text)
    assert p.predicates[0].name == "edgé\U0001F600" == p.rules[0].body[0].atom.pred
    q = IO.core_program_from_json(text)
    # the C layout has no place for the checker's flag: the engine checks the
    # universe + 2 round cap of every Trop stratum anyway
    assert q.rules[0].neg_weight_cycle_check and not p.rules[0].neg_weight_cycle_check
    q.rules[0].neg_weight_cycle_check = False
    assert p == q


@pytest.mark.parametrize("text,msg", [
    ('{"predicates": [], "rules": [], "num_strata": 1,}', "expected a field name"),
    ('{"predicates": [], "rules": []}', "missing field `num_strata`"),
    ('{"predicates": [], "rules": [], "num_strata": 1.5}', "fraction or exponent"),
    ('{"predicates": [], "rules": [], "num_strata": -1}', "num_strata: expected a u32"),
    ('{"predicates": [{"name": "e", "arity": 2, "semiring": "prob", "stratum": 0}], "rules": [], '
     '"num_strata": 1}', "predicates[0].semiring"),
    ('{"predicates": [], "rules": [{"head": {"pred": "p", "args": [{"kind": "var", "data": {"slot": 0}}]}, '
     '"body": [{"kind": "maybe", "data": {"pred": "e", "args": []}}], "stratum": 0, "var_count": 1}], '
     '"num_strata": 1}', "rules[0].body[0]: unknown literal kind"),
    ('{"predicates": [], "rules": [], "num_strata": 1} x', "trailing characters"),
    ('{"predicates": "', "unterminated string"),
    ("[" * 300 + "]" * 300, "nesting too deep"),
])
def test_native_json_decoder_errors(text, msg):
    import strata_k as sk
    with pytest.raises(sk.EngineError, match=msg.replace("[", r"\[").replace("]", r"\]").replace(".", r"\.")):
        sk.program_from_json(text)
