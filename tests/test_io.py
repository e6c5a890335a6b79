"""CPU: EDB ingestion (TSV / CSV / JSON with the reference's column
conventions, crates/strata-k/src/input.rs) and canonical rendering
(crates/strata-cli/src/render.rs), including the routing workload's pinned
output lines (crates/strata-cli/tests/workloads.rs:25-32) through the
oracle."""
import os

import pytest

from oracle import oracle as O
from strata_k.core_ir import BOOL, INF, TROP, CoreProgram, IntV, Sym, atom, pos, pred, rule, v
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
