"""EDB ingestion and canonical text output around the engine.

- ``SymbolDict``: first-seen interning of constants to u32 symbol ids
  (crates/strata-ir/src/dict.rs:20-60).
- ``load_input``: the rows of one ``input pred from "file"`` declaration,
  TSV (Souffle convention), CSV (quoted fields, ``""`` escapes) or JSON (an
  array of rows), with the reference's column conventions: a Trop predicate
  takes one trailing integer weight column, an ``int`` column is an integer
  whatever the format, a symbol column interns, empty cells and malformed
  rows are errors naming ``file:line`` (crates/strata-k/src/input.rs:53-330).
  Rows are validated before any is returned, so a failed load adds nothing.
- ``render_db``: the canonical output — relations in name order, tuples in
  ``BTreeMap`` order, symbols resolved through the dictionary, `` = w`` for
  Trop rows with ``inf`` for +inf, optionally filtered by plain queries
  (crates/strata-cli/src/render.rs:29-92, crates/strata-ir/src/output.rs:11).

- JSON forms of the IR (crates/strata-ir, docs/ir-encoding.md): a Trop
  weight is a bare integer or the string ``"inf"`` (trop.rs:69-107);
  ``core_program_from_json`` / ``core_program_to_json`` read and write a
  serialised ``CoreProgram`` (core/mod.rs:26-94: snake_case fields,
  ``{"kind", "data"}`` tags for literals and terms); ``high_ir_facts``
  takes the facts of a High-IR program document through the same version
  gate as ``load_program`` (schema/mod.rs:54-80, version.rs).

Neural (probabilistic) inputs and compound terms are outside this engine's
scope and are refused.
"""
import json
import os
import re
from typing import Dict, List, Optional, Sequence, Tuple

from .core_ir import (AGG_OPS, BOOL, INF, SEMIRINGS, TROP, Agg, Const, CoreAtom, CorePred, CoreProgram, CoreRule,
                      Int, IntV, Neg, Pos, Sym, Var)

POS_INF_TOKEN = "inf"


class LoadError(ValueError):
    """A file that cannot be loaded as EDB rows; the message names file:line."""


class SymbolDict:
    """Constants interned to dense ids in order of first sight."""

    def __init__(self):
        self._by_id: List[str] = []
        self._by_str: Dict[str, int] = {}

    def intern(self, s: str) -> int:
        i = self._by_str.get(s)
        if i is None:
            i = len(self._by_id)
            self._by_id.append(s)
            self._by_str[s] = i
        return i

    def get(self, s: str) -> Optional[int]:
        return self._by_str.get(s)

    def resolve(self, i: int) -> Optional[str]:
        return self._by_id[i] if 0 <= i < len(self._by_id) else None

    def __len__(self):
        return len(self._by_id)


# ------------------------------------------------------------------ reading

def split_csv_line(line: str) -> List[str]:
    """One CSV line: quoted fields with doubled-quote escapes; a quote must
    wrap a whole field and nothing may follow its closing quote."""
    cells, cur = [], []
    in_quotes = quoted = closed = False
    i, n = 0, len(line)
    while i < n:
        c = line[i]
        if not in_quotes and c == ",":
            cells.append("".join(cur))
            cur, quoted, closed = [], False, False
        elif not in_quotes and c == '"' and not cur and not quoted:
            in_quotes = quoted = True
        elif not in_quotes and c == '"':
            raise LoadError('unexpected `"` in CSV field (quotes must wrap the whole field, and `"` inside '
                            "must be doubled)")
        elif not in_quotes and closed:
            raise LoadError("trailing characters after a closing CSV quote")
        elif in_quotes and c == '"':
            if i + 1 < n and line[i + 1] == '"':
                cur.append('"')
                i += 1
            else:
                in_quotes, closed = False, True
        else:
            cur.append(c)
        i += 1
    if in_quotes:
        raise LoadError("unterminated quoted CSV field")
    cells.append("".join(cur))
    return cells


def _delimited_rows(path: str, delim: str):
    try:
        with open(path, encoding="utf-8") as f:
            text = f.read()
    except OSError as e:
        raise LoadError(f"cannot read {path}: {e}")
    rows = []
    for i, line in enumerate(text.split("\n")):
        line = line.rstrip("\r")
        if not line:
            continue
        if delim == ",":
            try:
                cells = split_csv_line(line)
            except LoadError as e:
                raise LoadError(f"{path}:{i + 1}: {e}")
        else:
            cells = line.split(delim)
        rows.append((i + 1, cells))
    return rows


def _json_rows(path: str):
    try:
        with open(path, encoding="utf-8") as f:
            value = json.load(f)
    except OSError as e:
        raise LoadError(f"cannot read {path}: {e}")
    except ValueError as e:
        raise LoadError(f"{path}: invalid JSON: {e}")
    if not isinstance(value, list):
        raise LoadError(f"{path}: expected a JSON array of rows")
    rows = []
    for i, row in enumerate(value):
        if not isinstance(row, list):
            raise LoadError(f"{path}:{i + 1}: a row must be an array")
        cells = []
        for v in row:
            if isinstance(v, bool) or not isinstance(v, (str, int, float)):
                raise LoadError(f"{path}:{i + 1}: unsupported JSON value {json.dumps(v)} (strings and numbers only)")
            if isinstance(v, int) and not -(2**63) <= v < 2**63:
                # serde's as_i64 fails outside i64 and the number is read as a
                # float cell (input.rs:316-321)
                v = float(v)
            cells.append(v)
        rows.append((i + 1, cells))
    return rows


def read_rows(path: str):
    """(line number, cells) of a .tsv (or extension-less), .csv or .json file.
    TSV/CSV cells are text; JSON cells are str, int or float."""
    ext = os.path.splitext(path)[1]
    if ext in (".tsv", ""):
        return _delimited_rows(path, "\t")
    if ext == ".csv":
        return _delimited_rows(path, ",")
    if ext == ".json":
        return _json_rows(path)
    raise LoadError(f"{path}: unsupported input format `{ext}` (tsv, csv, or json)")


_INT_TEXT = re.compile(r"[+-]?[0-9]+")


def _parse_int(s: str) -> Optional[int]:
    """`s.trim().parse::<i64>()`: an optional sign and ASCII digits only,
    within i64."""
    s = s.strip()
    if not _INT_TEXT.fullmatch(s):
        return None
    v = int(s)
    return v if -(2**63) <= v < 2**63 else None


def load_input(path: str, pred: str, arity: int, semiring: str, col_types: Optional[Sequence[str]],
               dct: SymbolDict) -> List[tuple]:
    """Facts ``(pred, tuple of Sym/IntV, weight or None)`` of one input file.

    ``col_types[c]`` is ``"int"`` for an ``int`` column, anything else for a
    symbol (domain) column; a Trop predicate's rows carry one more column,
    its integer weight."""
    if semiring not in (BOOL, TROP):
        raise LoadError(f"`{pred}`: only Bool and Trop inputs are loaded here")
    col_types = list(col_types or [])
    is_trop = semiring == TROP
    ncols = arity + (1 if is_trop else 0)
    facts = []
    for lineno, row in read_rows(path):
        def at(msg, _l=lineno):
            return LoadError(f"{path}:{_l}: {msg}")
        if len(row) != ncols:
            raise at(f"`{pred}` expects {ncols} column(s), found {len(row)}")
        args = []
        for c in range(arity):
            cell = row[c]
            want_int = c < len(col_types) and col_types[c] == "int"
            if isinstance(cell, float):
                raise at(f"number {cell} is not an integer (floats belong only in a neural predicate's "
                         "trailing probability column)")
            if want_int:
                if isinstance(cell, int):
                    args.append(IntV(cell))
                else:
                    v = _parse_int(cell)
                    if v is None:
                        raise at(f"column {c + 1} of `{pred}` is `int`, got {json.dumps(cell)}")
                    args.append(IntV(v))
            else:
                if isinstance(cell, int):
                    raise at(f"column {c + 1} of `{pred}` is a symbol column, got the number {cell}")
                if cell == "":
                    raise at(f"empty value in column {c + 1} of `{pred}`")
                args.append(Sym(dct.intern(cell)))
        w = None
        if is_trop:
            cell = row[arity]
            w = cell if isinstance(cell, int) else (_parse_int(cell) if isinstance(cell, str) else None)
            if w is None:
                raise at(f"`{pred}` needs a trailing integer weight")
        facts.append((pred, tuple(args), w))
    return facts


# ---------------------------------------------------------------- rendering

def render_value(v, dct: SymbolDict) -> str:
    if isinstance(v, Sym):
        s = dct.resolve(v.id)
        return s if s is not None else "?"
    if isinstance(v, IntV):
        return str(v.value)
    raise TypeError(f"cannot render {v!r}")


def render_weight(w) -> str:
    return POS_INF_TOKEN if w is INF else str(int(w))


def _matches(pattern: Sequence, tup: tuple) -> bool:
    return len(pattern) == len(tup) and all(p is None or p == t for p, t in zip(pattern, tup))


def render_db(db, dct: SymbolDict, filter: Optional[Sequence[Tuple[str, Sequence]]] = None) -> str:
    """Canonical text of a result: every relation in name order, its tuples in
    order, `` = w`` after Trop tuples.  ``filter``: plain queries
    ``(pred, pattern)`` whose pattern holds a ground Sym/IntV or None per
    position; when given, only queried predicates print, and only tuples
    matching one of their queries."""
    out = []
    for pred in db.predicates():
        qs = [p for q, p in (filter or []) if q == pred]
        if filter and not qs:
            continue
        rel = db.relation(pred)
        for tup, ann in rel.rows():
            if filter and not any(_matches(p, tup) for p in qs):
                continue
            line = f"{pred}({', '.join(render_value(v, dct) for v in tup)})"
            if ann is not None:
                line += f" = {render_weight(ann)}"
            out.append(line + "\n")
    return "".join(out)


# ------------------------------------------------------------- IR JSON forms

IR_VERSION = (0, 1, 0)


def weight_from_json(v):
    """A Trop weight's JSON value: an i64, or ``"inf"`` for +inf
    (crates/strata-ir/src/trop.rs:78-107)."""
    if isinstance(v, int) and not isinstance(v, bool):
        if -(2**63) <= v < 2**63:
            return v
        raise LoadError(f"weight {v} is outside i64")
    if v == POS_INF_TOKEN:
        return INF
    raise LoadError(f'invalid weight {json.dumps(v)}: an i64 weight or the string "{POS_INF_TOKEN}"')


def weight_to_json(w):
    return POS_INF_TOKEN if w is INF else int(w)


def _tagged(v, where):
    if not isinstance(v, dict) or "kind" not in v or "data" not in v:
        raise LoadError(f"{where}: expected a {{\"kind\", \"data\"}} object")
    return v["kind"], v["data"]


def _u32(v, where):
    if isinstance(v, bool) or not isinstance(v, int) or not 0 <= v < 2**32:
        raise LoadError(f"{where}: expected a u32, got {json.dumps(v)}")
    return v


def _fields(d, names, optional=(), where=""):
    """Required ``names`` present; other fields are ignored, as serde does
    for a struct without ``deny_unknown_fields``."""
    if not isinstance(d, dict):
        raise LoadError(f"{where}: expected an object")
    missing = [n for n in names if n not in d]
    if missing:
        raise LoadError(f"{where}: missing field(s) {missing}")
    return d


def _core_term(t, where):
    kind, d = _tagged(t, where)
    if kind == "var":
        return Var(_u32(_fields(d, ["slot"], where=where)["slot"], where))
    if kind == "const":
        return Const(_u32(_fields(d, ["sym"], where=where)["sym"], where))
    if kind == "int":
        val = _fields(d, ["value"], where=where)["value"]
        if isinstance(val, bool) or not isinstance(val, int) or not -(2**63) <= val < 2**63:
            raise LoadError(f"{where}: expected an i64, got {json.dumps(val)}")
        return Int(val)
    if kind == "agg":
        _fields(d, ["op", "slot"], where=where)
        if d["op"] not in AGG_OPS:
            raise LoadError(f"{where}: unknown aggregate {json.dumps(d['op'])}")
        return Agg(d["op"], _u32(d["slot"], where))
    if kind == "compound":
        raise LoadError(f"{where}: compound terms (@terms) are outside this engine")
    raise LoadError(f"{where}: unknown term kind {json.dumps(kind)}")


def _core_atom(a, where):
    _fields(a, ["pred", "args"], where=where)
    if not isinstance(a["pred"], str) or not isinstance(a["args"], list):
        raise LoadError(f"{where}: an atom is {{pred: string, args: [term]}}")
    return CoreAtom(a["pred"], [_core_term(t, f"{where}.args[{i}]") for i, t in enumerate(a["args"])])


def core_program_from_json(doc) -> CoreProgram:
    """A serialised CoreProgram (a JSON string or its parsed object), as
    strata-ir's serde derives write it (core/mod.rs:26-94)."""
    if isinstance(doc, (str, bytes)):
        try:
            doc = json.loads(doc)
        except ValueError as e:
            raise LoadError(f"invalid JSON: {e}")
    _fields(doc, ["predicates", "rules", "num_strata"], where="program")
    preds = []
    for i, p in enumerate(doc["predicates"]):
        w = f"predicates[{i}]"
        _fields(p, ["name", "arity", "semiring", "stratum"], where=w)
        if p["semiring"] not in SEMIRINGS:
            raise LoadError(f"{w}: semiring must be one of {list(SEMIRINGS)}")
        preds.append(CorePred(p["name"], _u32(p["arity"], w), p["semiring"], _u32(p["stratum"], w)))
    rules = []
    for i, r in enumerate(doc["rules"]):
        w = f"rules[{i}]"
        _fields(r, ["head", "body", "stratum", "var_count"], optional=["neg_weight_cycle_check"], where=w)
        body = []
        for j, lit in enumerate(r["body"]):
            kind, a = _tagged(lit, f"{w}.body[{j}]")
            if kind not in ("pos", "neg"):
                raise LoadError(f"{w}.body[{j}]: unknown literal kind {json.dumps(kind)}")
            at = _core_atom(a, f"{w}.body[{j}]")
            body.append(Pos(at) if kind == "pos" else Neg(at))
        flag = r.get("neg_weight_cycle_check", False)
        if not isinstance(flag, bool):
            raise LoadError(f"{w}: neg_weight_cycle_check must be a boolean")
        rules.append(CoreRule(_core_atom(r["head"], f"{w}.head"), body, _u32(r["stratum"], w),
                              _u32(r["var_count"], w), flag))
    return CoreProgram(preds, rules, _u32(doc["num_strata"], "num_strata"))


def _term_json(t):
    if isinstance(t, Var):
        return {"kind": "var", "data": {"slot": t.slot}}
    if isinstance(t, Const):
        return {"kind": "const", "data": {"sym": t.sym}}
    if isinstance(t, Int):
        return {"kind": "int", "data": {"value": t.value}}
    if isinstance(t, Agg):
        return {"kind": "agg", "data": {"op": t.op, "slot": t.slot}}
    raise TypeError(f"cannot encode term {t!r}")


def _atom_json(a):
    return {"pred": a.pred, "args": [_term_json(t) for t in a.args]}


def core_program_to_json(prog: CoreProgram) -> dict:
    """The JSON object serde writes for a CoreProgram (field order as
    declared; ``neg_weight_cycle_check`` only when set)."""
    rules = []
    for r in prog.rules:
        d = {"head": _atom_json(r.head),
             "body": [{"kind": "pos" if isinstance(l, Pos) else "neg", "data": _atom_json(l.atom)} for l in r.body],
             "stratum": r.stratum, "var_count": r.var_count}
        if r.neg_weight_cycle_check:
            d["neg_weight_cycle_check"] = True
        rules.append(d)
    return {"predicates": [{"name": p.name, "arity": p.arity, "semiring": p.semiring, "stratum": p.stratum}
                           for p in prog.predicates],
            "rules": rules, "num_strata": prog.num_strata}


def ir_version_compatible(s: str) -> bool:
    """The version gate of load_program (schema/mod.rs:56-80, version.rs):
    exactly three numeric parts; before 1.0 the minor must match, from 1.0 on
    the same major with a minor no newer than this build's."""
    parts = s.split(".") if isinstance(s, str) else []
    if len(parts) != 3 or not all(p.isdigit() for p in parts):
        raise LoadError(f"bad ir_version {json.dumps(s)}")
    major, minor, _ = (int(p) for p in parts)
    if major != IR_VERSION[0]:
        return False
    return minor == IR_VERSION[1] if IR_VERSION[0] == 0 else minor <= IR_VERSION[1]


def high_ir_facts(doc, dct: SymbolDict) -> List[tuple]:
    """The facts of a High-IR program document, in document order, as
    ``(pred, tuple of Sym/IntV, weight or None)``: constants interned first
    seen, a present ``weight`` (integer or ``"inf"``) makes a Trop fact.
    Facts with a probability, and non-ground arguments, are refused."""
    if isinstance(doc, (str, bytes)):
        try:
            doc = json.loads(doc)
        except ValueError as e:
            raise LoadError(f"invalid JSON: {e}")
    if not isinstance(doc, dict) or "ir_version" not in doc or "items" not in doc:
        raise LoadError("a High-IR program is {ir_version, items}")
    if not ir_version_compatible(doc["ir_version"]):
        raise LoadError(f"incompatible ir_version {doc['ir_version']} (this build reads "
                        f"{'.'.join(map(str, IR_VERSION))})")
    facts = []
    for i, item in enumerate(doc["items"]):
        if not isinstance(item, dict) or item.get("kind") != "fact":
            continue
        w = f"items[{i}]"
        d = _fields(item["data"], ["atom"], optional=["weight", "prob"], where=w)
        if "prob" in d:
            raise LoadError(f"{w}: probabilistic facts are outside this engine")
        a = _fields(d["atom"], ["pred", "args"], where=w)
        args = []
        for j, t in enumerate(a["args"]):
            kind, td = _tagged(t, f"{w}.args[{j}]")
            if kind == "const":
                args.append(Sym(dct.intern(_fields(td, ["name"], where=w)["name"])))
            elif kind == "int":
                val = _fields(td, ["value"], where=w)["value"]
                if isinstance(val, bool) or not isinstance(val, int) or not -(2**63) <= val < 2**63:
                    raise LoadError(f"{w}.args[{j}]: expected an i64")
                args.append(IntV(val))
            else:
                raise LoadError(f"{w}.args[{j}]: a fact's arguments are constants or integers, not {kind}")
        weight = weight_from_json(d["weight"]) if "weight" in d else None
        facts.append((a["pred"], tuple(args), weight))
    return facts
