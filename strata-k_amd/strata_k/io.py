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

Neural (probabilistic) inputs and compound terms are outside this engine's
scope and are refused.
"""
import json
import os
from typing import Dict, List, Optional, Sequence, Tuple

from .core_ir import BOOL, INF, TROP, IntV, Sym

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


def _parse_int(s: str) -> Optional[int]:
    s = s.strip()
    body = s[1:] if s[:1] in "+-" else s
    if not body.isdigit():
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
