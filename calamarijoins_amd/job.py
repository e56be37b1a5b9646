"""JOB front end: the IMDB schema, the SQL subset of the Join Order Benchmark
queries, filter statements, and the EXPLAIN-JSON -> Plan loader.

This is the caller side of execute() in the reference's harness
(tests/read_sql.cpp): PostgreSQL's EXPLAIN (FORMAT JSON) plan of a query
(plans.json) fixes the join tree; the query's SQL gives the tables and their
aliases, the single-table filters and the join conditions; the plan loader
turns both into a Plan whose inputs are the filtered base tables.

  - IMDB schema ............................ read_sql.cpp:28-145 (attributes_map)
  - SQL -> aliases, filters, join classes, select list
                                             read_sql.cpp:240-299, 302-672, 674-866
                                             (the reference parses with the
                                             third-party Hyrise sql-parser; this
                                             module parses the subset JOB uses)
  - filter statements and their text ...... include/data_model/statement.h
                                             (Comparison, LogicalOperation,
                                             pretty_print) and src/statement.cpp
                                             (eval on a record)
  - EXPLAIN tree -> Plan .................. read_sql.cpp:867-1233 (load_join_pipeline)

The text of a filter (pretty_print) names the .tbl cache file of a filtered
scan: "<table>_<FNV-1a 64 of the text>.tbl" (read_sql.cpp:1154-1158).
"""
from __future__ import annotations

import re

from . import abi
from .host import Plan

INT32, INT64, FP64, VARCHAR = abi.CJ_INT32, abi.CJ_INT64, abi.CJ_FP64, abi.CJ_VARCHAR
_I, _S = INT32, VARCHAR

# Column names and types of the IMDB tables in schema order (the index of a
# column here is the column index of its base table).
SCHEMA = {
    "aka_name": [("id", _I), ("person_id", _I), ("name", _S), ("imdb_index", _S), ("name_pcode_cf", _S),
                 ("name_pcode_nf", _S), ("surname_pcode", _S), ("md5sum", _S)],
    "aka_title": [("id", _I), ("movie_id", _I), ("title", _S), ("imdb_index", _S), ("kind_id", _I),
                  ("production_year", _I), ("phonetic_code", _S), ("episode_of_id", _I), ("season_nr", _I),
                  ("episode_nr", _I), ("note", _S), ("md5sum", _S)],
    "cast_info": [("id", _I), ("person_id", _I), ("movie_id", _I), ("person_role_id", _I), ("note", _S),
                  ("nr_order", _I), ("role_id", _I)],
    "char_name": [("id", _I), ("name", _S), ("imdb_index", _S), ("imdb_id", _I), ("name_pcode_nf", _S),
                  ("surname_pcode", _S), ("md5sum", _S)],
    "comp_cast_type": [("id", _I), ("kind", _S)],
    "company_name": [("id", _I), ("name", _S), ("country_code", _S), ("imdb_id", _I), ("name_pcode_nf", _S),
                     ("name_pcode_sf", _S), ("md5sum", _S)],
    "company_type": [("id", _I), ("kind", _S)],
    "complete_cast": [("id", _I), ("movie_id", _I), ("subject_id", _I), ("status_id", _I)],
    "info_type": [("id", _I), ("info", _S)],
    "keyword": [("id", _I), ("keyword", _S), ("phonetic_code", _S)],
    "kind_type": [("id", _I), ("kind", _S)],
    "link_type": [("id", _I), ("link", _S)],
    "movie_companies": [("id", _I), ("movie_id", _I), ("company_id", _I), ("company_type_id", _I), ("note", _S)],
    "movie_info_idx": [("id", _I), ("movie_id", _I), ("info_type_id", _I), ("info", _S), ("note", _S)],
    "movie_keyword": [("id", _I), ("movie_id", _I), ("keyword_id", _I)],
    "movie_link": [("id", _I), ("movie_id", _I), ("linked_movie_id", _I), ("link_type_id", _I)],
    "name": [("id", _I), ("name", _S), ("imdb_index", _S), ("imdb_id", _I), ("gender", _S), ("name_pcode_cf", _S),
             ("name_pcode_nf", _S), ("surname_pcode", _S), ("md5sum", _S)],
    "role_type": [("id", _I), ("role", _S)],
    "title": [("id", _I), ("title", _S), ("imdb_index", _S), ("kind_id", _I), ("production_year", _I),
              ("imdb_id", _I), ("phonetic_code", _S), ("episode_of_id", _I), ("season_nr", _I), ("episode_nr", _I),
              ("series_years", _S), ("md5sum", _S)],
    "movie_info": [("id", _I), ("movie_id", _I), ("info_type_id", _I), ("info", _S), ("note", _S)],
    "person_info": [("id", _I), ("person_id", _I), ("info_type_id", _I), ("info", _S), ("note", _S)],
}


class JobError(ValueError):
    pass


def column_index(table: str, column: str) -> int:
    for i, (name, _) in enumerate(SCHEMA[table]):
        if name == column:
            return i
    raise JobError(f"Cannot find column {column} in table {table}")


# ------------------------------------------------------------- statements
class Comparison:
    """column (schema index) op value; op is one of = != < > <= >= LIKE
    NOT LIKE IS NULL (statement.h: Comparison)."""

    def __init__(self, column: int, op: str, value):
        self.column, self.op, self.value = column, op, value

    def pretty(self, indent: int = 0) -> str:
        if self.op in ("IS NULL", "IS NOT NULL"):
            v = ""
        elif isinstance(self.value, str):
            v = f"'{self.value}'"
        else:
            v = str(self.value)
        return f"{' ' * indent}{self.column} {self.op} {v}"

    def eval(self, record, types=None) -> bool:
        """Comparison::eval(record) (src/statement.cpp:166-249): NULL fails
        every comparison; numbers compare as numbers, strings as strings,
        LIKE with % and _ over the whole string.  With the column types
        given, the column-wise evaluation the reference's scans use
        (statement.cpp:47-164): an INT32 column compares with the literal cast
        to int32, INT64 exactly, FP64 as doubles, VARCHAR byte-wise."""
        x = record[self.column]
        if types is not None and x is not None and self.op not in ("IS NULL", "IS NOT NULL", "LIKE", "NOT LIKE"):
            t = types[self.column]
            if t == INT32:
                a, b = int(x), ((int(self.value) + 2 ** 31) % 2 ** 32) - 2 ** 31
            elif t == INT64:
                a, b = int(x), int(self.value)
            elif t == FP64:
                a, b = float(x), float(self.value)
            else:
                a = x.encode("utf-8") if isinstance(x, str) else x
                b = self.value.encode("utf-8") if isinstance(self.value, str) else self.value
            return {"=": a == b, "!=": a != b, "<": a < b, ">": a > b, "<=": a <= b, ">=": a >= b}[self.op]
        if self.op == "IS NULL":
            return x is None
        if self.op == "IS NOT NULL":
            return x is not None
        if self.op in ("LIKE", "NOT LIKE"):
            if not isinstance(x, (str, bytes)) or not isinstance(self.value, str):
                return False
            m = like_match(x, self.value)
            return m if self.op == "LIKE" else not m
        if x is None:
            return False
        num = lambda v: isinstance(v, (int, float)) and not isinstance(v, bool)
        if num(x) and num(self.value):
            a, b = float(x), float(self.value)
        elif isinstance(x, (str, bytes)) and isinstance(self.value, str):
            a = x.encode("utf-8") if isinstance(x, str) else x
            b = self.value.encode("utf-8")
        else:
            return False
        return {"=": a == b, "!=": a != b, "<": a < b, ">": a > b, "<=": a <= b, ">=": a >= b}[self.op]


class Logical:
    """[AND] / [OR] / [NOT] over children (statement.h: LogicalOperation)."""

    def __init__(self, op: str, children):
        self.op, self.children = op, list(children)

    def pretty(self, indent: int = 0) -> str:
        s = f"{' ' * indent}[{self.op}]\n" + "".join(c.pretty(indent + 2) + "\n" for c in self.children)
        return s[:-1] if self.children else s

    def eval(self, record, types=None) -> bool:
        if self.op == "AND":
            return all(c.eval(record, types) for c in self.children)
        if self.op == "OR":
            return any(c.eval(record, types) for c in self.children)
        return len(self.children) == 1 and not self.children[0].eval(record, types)


def _text(v) -> str:
    b = v if isinstance(v, bytes) else v.encode("utf-8", "surrogateescape")
    return b.decode("utf-8", "surrogateescape")


def like_match(s, pattern) -> bool:
    """SQL LIKE over the whole string as the reference's regular expression
    reads it (statement.h: like_match, RE2 with default options): % any run
    of characters, _ one UTF-8 character, neither matching a newline, every
    other character literal."""
    s, pattern = _text(s), _text(pattern)
    rx = "".join(".*" if c == "%" else "." if c == "_" else re.escape(c) for c in pattern)
    return re.fullmatch(rx, s) is not None


_OPS = {"=": abi.CJ_OP_EQ, "!=": abi.CJ_OP_NEQ, "<": abi.CJ_OP_LT, ">": abi.CJ_OP_GT, "<=": abi.CJ_OP_LEQ,
        ">=": abi.CJ_OP_GEQ, "LIKE": abi.CJ_OP_LIKE, "NOT LIKE": abi.CJ_OP_NOT_LIKE, "IS NULL": abi.CJ_OP_IS_NULL,
        "IS NOT NULL": abi.CJ_OP_IS_NOT_NULL}


def filter_program(stmt, types) -> list:
    """A statement as the postfix program of cj_filter (children before their
    operator); `types` are the table's column types.  Literal types follow
    the reference's column comparisons: integers for INT32 / INT64 columns,
    floats for FP64, strings for VARCHAR."""
    out, keep = [], []

    def rec(s):
        if isinstance(s, Logical):
            for c in s.children:
                rec(c)
            kind = {"AND": abi.CJ_F_AND, "OR": abi.CJ_F_OR, "NOT": abi.CJ_F_NOT}[s.op]
            out.append(abi.FilterNode(kind, 0, 0, abi.CJ_LIT_NONE, 0, 0.0, None, 0))
            return
        t = types[s.column]
        n = abi.FilterNode(abi.CJ_F_CMP, _OPS[s.op], s.column, abi.CJ_LIT_NONE, 0, 0.0, None, 0)
        if s.op not in ("IS NULL", "IS NOT NULL"):
            v = s.value
            if t == VARCHAR:
                if not isinstance(v, str):
                    raise JobError("a VARCHAR column compares with a string literal")
                b = v.encode("utf-8")
                keep.append(b)
                n.lit_kind, n.sval, n.slen = abi.CJ_LIT_STR, b, len(b)
            elif t == FP64:
                if not isinstance(v, float):
                    raise JobError("an FP64 column compares with a float literal")
                n.lit_kind, n.fval = abi.CJ_LIT_FLOAT, v
            else:
                if not isinstance(v, int):
                    raise JobError("an integer column compares with an integer literal")
                n.lit_kind, n.ival = abi.CJ_LIT_INT, v
        out.append(n)

    rec(stmt)
    return _Program(out, keep)


class _Program(list):
    """Filter nodes plus the string literals they point at (kept alive)."""

    def __init__(self, nodes, keep):
        super().__init__(nodes)
        self.keep = keep


# ----------------------------------------------------------------- SQL subset
_TOKEN = re.compile(r"\s*(?:(?P<str>'(?:[^']|'')*')|(?P<num>-?\d+(?:\.\d+)?)|(?P<op><>|!=|<=|>=|=|<|>|\(|\)|,|;)"
                    r"|(?P<id>[A-Za-z_][A-Za-z0-9_]*(?:\.[A-Za-z_][A-Za-z0-9_]*)?))")
_KEYWORDS = {"SELECT", "FROM", "WHERE", "AS", "AND", "OR", "NOT", "LIKE", "IN", "BETWEEN", "IS", "NULL"}


def _tokens(sql: str):
    pos, out = 0, []
    while pos < len(sql):
        m = _TOKEN.match(sql, pos)
        if not m or m.end() == pos:
            if sql[pos:].strip() == "":
                break
            raise JobError(f"cannot tokenize SQL near: {sql[pos:pos + 30]!r}")
        pos = m.end()
        if m.group("str") is not None:
            out.append(("str", m.group("str")[1:-1].replace("''", "'")))
        elif m.group("num") is not None:
            t = m.group("num")
            out.append(("num", float(t) if "." in t else int(t)))
        elif m.group("op") is not None:
            out.append(("op", m.group("op")))
        else:
            t = m.group("id")
            out.append(("kw", t.upper()) if t.upper() in _KEYWORDS else ("id", t))
    return out


class _Parser:
    """Recursive descent over the JOB subset; the expression tree has the
    shape the reference's parser builds (AND / OR left-associative, NOT above
    comparisons, x IS NOT NULL as NOT(x IS NULL), BETWEEN as a two-item list,
    IN as a list)."""

    def __init__(self, sql: str):
        self.t = _tokens(sql)
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else ("eof", None)

    def take(self, kind=None, value=None):
        tok = self.peek()
        if (kind and tok[0] != kind) or (value is not None and tok[1] != value):
            raise JobError(f"expected {value or kind}, found {tok[1]!r}")
        self.i += 1
        return tok

    def accept(self, kind, value):
        if self.peek() == (kind, value):
            self.i += 1
            return True
        return False

    def statement(self):
        self.take("kw", "SELECT")
        select = [self.select_item()]
        while self.accept("op", ","):
            select.append(self.select_item())
        self.take("kw", "FROM")
        tables = [self.table_item()]
        while self.accept("op", ","):
            tables.append(self.table_item())
        where = None
        if self.accept("kw", "WHERE"):
            where = self.expr()
        self.accept("op", ";")
        if self.peek()[0] != "eof":
            raise JobError(f"unexpected {self.peek()[1]!r} after the statement")
        return select, tables, where

    def select_item(self):
        tok = self.take("id")
        if self.accept("op", "("):  # aggregate over one column reference
            arg = self.take("id")[1]
            self.take("op", ")")
            col = arg
        else:
            col = tok[1]
        if self.accept("kw", "AS"):
            self.take("id")
        return col

    def table_item(self):
        name = self.take("id")[1]
        alias = None
        if self.accept("kw", "AS"):
            alias = self.take("id")[1]
        elif self.peek()[0] == "id":
            alias = self.take("id")[1]
        return name, alias

    def expr(self):
        e = self.and_expr()
        while self.accept("kw", "OR"):
            e = ("OR", e, self.and_expr())
        return e

    def and_expr(self):
        e = self.not_expr()
        while self.accept("kw", "AND"):
            e = ("AND", e, self.not_expr())
        return e

    def not_expr(self):
        if self.accept("kw", "NOT"):
            return ("NOT", self.not_expr())
        return self.predicate()

    def operand(self):
        tok = self.peek()
        if tok[0] == "str":
            self.i += 1
            return ("lit", tok[1])
        if tok[0] == "num":
            self.i += 1
            return ("lit", tok[1])
        if tok[0] == "id":
            self.i += 1
            return ("col", tok[1])
        raise JobError(f"unexpected {tok[1]!r} in an expression")

    def predicate(self):
        if self.accept("op", "("):
            e = self.expr()
            self.take("op", ")")
            return e
        left = self.operand()
        tok = self.peek()
        if tok[0] == "op" and tok[1] in ("=", "!=", "<>", "<", ">", "<=", ">="):
            self.i += 1
            return ("cmp", "!=" if tok[1] == "<>" else tok[1], left, self.operand())
        if self.accept("kw", "IS"):
            neg = self.accept("kw", "NOT")
            self.take("kw", "NULL")
            e = ("isnull", left)
            return ("NOT", e) if neg else e
        neg = self.accept("kw", "NOT")
        if self.accept("kw", "LIKE"):
            return ("like", "NOT LIKE" if neg else "LIKE", left, self.operand())
        if self.accept("kw", "IN"):
            self.take("op", "(")
            items = [self.operand()]
            while self.accept("op", ","):
                items.append(self.operand())
            self.take("op", ")")
            e = ("in", left, items)
            return ("NOT", e) if neg else e
        if not neg and self.accept("kw", "BETWEEN"):
            lo = self.operand()
            self.take("kw", "AND")
            return ("between", left, lo, self.operand())
        raise JobError(f"unsupported predicate near {tok[1]!r}")


class ParsedSQL:
    """Tables (entities), aliases, per-entity filters, join equivalence
    classes and the select list of one query (read_sql.cpp:674-866)."""

    def __init__(self, sql: str):
        select, tables, where = _Parser(sql).statement()
        self.table_counts = {}
        self.alias_map = {}
        self.entities = []  # FROM order
        for name, alias in tables:
            if name not in SCHEMA:
                raise JobError(f"No table: {name} in schema")
            k = self.table_counts.get(name, 0)
            self.table_counts[name] = k + 1
            ent = (name, k)
            self.entities.append(ent)
            if alias:
                self.alias_map[alias] = ent
        # columns of every entity, numbered in FROM order x schema order (DSU items)
        self.column_vec = [(e, c) for e in self.entities for c, _ in SCHEMA[e[0]]]
        self.column_id = {ec: i for i, ec in enumerate(self.column_vec)}
        self.output_attrs = [self.resolve(c) for c in select]
        self.filters = {}
        self._dsu = list(range(len(self.column_vec)))
        if where is not None:
            stmt, ent = self._expr(where, 0)
            if stmt is not None:
                self._insert(ent, stmt)
        self.join_graph = self._join_graph()

    # -- column references
    def resolve(self, ref: str):
        if "." in ref:
            t, c = ref.split(".", 1)
            if t in self.alias_map:
                ent = self.alias_map[t]
            elif self.table_counts.get(t) == 1:
                ent = (t, 0)
            elif t in self.table_counts:
                raise JobError(f"Ambiguous table: {t}")
            else:
                raise JobError(f"Unknown table name: {t}")
        else:
            c = ref
            owners = [e for e in self.entities if any(n == c for n, _ in SCHEMA[e[0]])]
            tabs = sorted({e[0] for e in owners})
            if len(tabs) != 1:
                raise JobError(f"Ambiguous or unknown column: {c}")
            if self.table_counts[tabs[0]] != 1:
                raise JobError(f"Ambiguous table: {tabs[0]}")
            ent = (tabs[0], 0)
        column_index(ent[0], c)
        return ent, c

    # -- WHERE clause (parse_expr_impl, read_sql.cpp:332-672)
    def _insert(self, ent, stmt):
        if ent in self.filters:
            self.filters[ent] = Logical("AND", [self.filters[ent], stmt])
        else:
            self.filters[ent] = stmt

    def _find(self, i):
        while self._dsu[i] != i:
            self._dsu[i] = self._dsu[self._dsu[i]]
            i = self._dsu[i]
        return i

    def _col(self, operand):
        if operand[0] != "col":
            raise JobError("the left side of a comparison must be a column")
        return self.resolve(operand[1])

    def _expr(self, e, level):
        kind = e[0]
        if kind in ("AND", "OR"):
            add = 0 if kind == "AND" else 1
            ls, le = self._expr(e[1], level + add)
            rs, re_ = self._expr(e[2], level + add)
            if level == 0 and kind == "AND":
                if ls is not None:
                    self._insert(le, ls)
                if rs is not None:
                    self._insert(re_, rs)
                return None, None
            if ls is None or rs is None:
                raise JobError("Non top level contains join condition instead of filter")
            if le != re_:
                raise JobError("Filter can not be pushed down")
            return Logical(kind, [ls, rs]), le
        if kind == "NOT":
            cs, ce = self._expr(e[1], level + 1)
            if cs is None:
                raise JobError("NOT over a join condition")
            return Logical("NOT", [cs]), ce
        if kind == "cmp":
            _, op, left, right = e
            ent, col = self._col(left)
            if right[0] == "col":
                if op != "=":
                    raise JobError("Non-EuqalJoins are not supported")
                rent, rcol = self.resolve(right[1])
                a, b = self._find(self.column_id[(ent, col)]), self._find(self.column_id[(rent, rcol)])
                self._dsu[a] = b
                return None, None
            return Comparison(column_index(ent[0], col), op, right[1]), ent
        if kind == "like":
            _, op, left, right = e
            ent, col = self._col(left)
            if right[0] != "lit" or not isinstance(right[1], str):
                raise JobError("LIKE needs a string literal")
            return Comparison(column_index(ent[0], col), op, right[1]), ent
        if kind == "between":
            _, left, lo, hi = e
            ent, col = self._col(left)
            ci = column_index(ent[0], col)
            return Logical("AND", [Comparison(ci, ">=", lo[1]), Comparison(ci, "<=", hi[1])]), ent
        if kind == "in":
            _, left, items = e
            ent, col = self._col(left)
            ci = column_index(ent[0], col)
            out = None
            for it in items:
                if it[0] != "lit":
                    raise JobError("IN needs literals")
                c = Comparison(ci, "=", it[1])
                out = c if out is None else Logical("OR", [out, c])
            return out, ent
        if kind == "isnull":
            ent, col = self._col(e[1])
            return Comparison(column_index(ent[0], col), "IS NULL", None), ent
        raise JobError(f"unsupported expression {kind}")

    def _join_graph(self):
        classes = {}
        for i in range(len(self.column_vec)):
            classes.setdefault(self._find(i), []).append(i)
        graph = {}
        for items in classes.values():
            for x in range(len(items) - 1):
                for y in range(x + 1, len(items)):
                    (le, lc), (re_, rc) = self.column_vec[items[x]], self.column_vec[items[y]]
                    if re_ in graph.get(le, {}):
                        raise JobError("At least two conditions between a same pair of tables.")
                    graph.setdefault(le, {})[re_] = (lc, rc)
                    graph.setdefault(re_, {})[le] = (rc, lc)
        return graph

    def join_class(self, ent, col):
        return self._find(self.column_id[(ent, col)])


# ------------------------------------------------------- EXPLAIN -> Plan
_OTHER = ("Aggregate", "Gather")
_SCANS = ("Seq Scan", "Index Only Scan")


def _scan_entity(node, parsed: ParsedSQL):
    if "Alias" in node:
        if node["Alias"] not in parsed.alias_map:
            raise JobError(f"Cannot find alias: {node['Alias']}")
        return parsed.alias_map[node["Alias"]]
    rel = node.get("Relation Name")
    if rel is None:
        raise JobError('Neither "Alias" nor "Relation Name" exists in scan node')
    if parsed.table_counts.get(rel) != 1:
        raise JobError(f"Table {rel} is used at least twice or not at all")
    return (rel, 0)


def _hash_children(node):
    lt, rt = node["Plans"][0]["Node Type"], node["Plans"][1]["Node Type"]
    if lt == "Hash" and rt != "Hash":
        return True, node["Plans"][0]["Plans"][0], node["Plans"][1]
    if lt != "Hash" and rt == "Hash":
        return False, node["Plans"][0], node["Plans"][1]["Plans"][0]
    raise JobError("Hash Join should have at least one Hash child")


def entities_of(node, parsed: ParsedSQL) -> set:
    t = node["Node Type"]
    if t in _OTHER:
        return entities_of(node["Plans"][0], parsed)
    if t == "Hash Join":
        _, l, r = _hash_children(node)
        return entities_of(l, parsed) | entities_of(r, parsed)
    if t in ("Nested Loop", "Merge Join"):
        raise JobError("Not Hash Join")
    if t in _SCANS:
        return {_scan_entity(node, parsed)}
    raise JobError(f"Not supported node type: {t}")


def load_plan(explain: dict, parsed: ParsedSQL, table_for) -> tuple:
    """load_join_pipeline (read_sql.cpp:867-1233): the EXPLAIN tree gives
    the join tree and build sides (the Hash child is the build side), the
    SQL gives each join's condition (a pair of columns of one equivalence
    class across the two sides) and the required columns, which flow down
    the tree.  table_for(entity, filter) returns the filtered base table of a
    scan.  Returns (plan, [(entity, column, type)] of the root).

    Where several column pairs connect the two sides, the reference keeps the
    last one its hash-map iteration meets (an unspecified order); this picks
    the last in sorted order.  Pairs of one equivalence class give the same
    join; tests/test_job.py checks that every join of the 113 JOB plans is
    connected by a single class."""
    plan = Plan()

    def rec(node, required):
        t = node["Node Type"]
        if t in _OTHER:
            return rec(node["Plans"][0], required)
        if t == "Hash Join":
            build_left, pl, pr = _hash_children(node)
            le, re_ = entities_of(pl, parsed), entities_of(pr, parsed)
            cond = None
            for ent in sorted(le):
                for adj, cols in sorted(parsed.join_graph.get(ent, {}).items()):
                    if adj in re_:
                        cond = (ent, adj, cols)
            if cond is None:
                raise JobError("Cannot find join condition")
            lent, rent, (lcol, rcol) = cond
            lreq, rreq = [], []
            lin = rin = False
            for ent, col in required:
                if ent in le:
                    lin |= (ent, col) == (lent, lcol)
                    lreq.append((ent, col))
                elif ent in re_:
                    rin |= (ent, col) == (rent, rcol)
                    rreq.append((ent, col))
                else:
                    raise JobError("Required attributes cannot be found in neither left child nor right child.")
            if not lin:
                lreq.append((lent, lcol))
            if not rin:
                rreq.append((rent, rcol))
            left, lcols = rec(pl, lreq)
            right, rcols = rec(pr, rreq)
            lattr = next(i for i, (e, c, _) in enumerate(lcols) if (e, c) == (lent, lcol))
            rattr = next(i for i, (e, c, _) in enumerate(rcols) if (e, c) == (rent, rcol))
            both = lcols + rcols
            out_cols, out_attrs = [], []
            for ent, col in required:
                k = next((i for i, (e, c, _) in enumerate(both) if (e, c) == (ent, col)), None)
                if k is None:
                    raise JobError(f"Cannot found the required attr: {ent}.{col} in children's output")
                out_cols.append(both[k])
                out_attrs.append((k, abi.TYPE_IDS[both[k][2]]))
            node_id = plan.new_join_node(build_left, left, right, lattr, rattr, out_attrs)
            return node_id, out_cols
        if t in ("Nested Loop", "Merge Join"):
            raise JobError("Not Hash Join")
        if t in _SCANS:
            ent = _scan_entity(node, parsed)
            table = table_for(ent, parsed.filters.get(ent))
            tid = plan.new_input(table)
            out_cols, out_attrs = [], []
            for rent, col in required:
                if rent != ent:
                    raise JobError(f"Cannot found the required attr: {rent}.{col} in children's output")
                ci = column_index(ent[0], col)
                ty = SCHEMA[ent[0]][ci][1]
                out_cols.append((ent, col, ty))
                out_attrs.append((ci, abi.TYPE_IDS[ty]))
            return plan.new_scan_node(tid, out_attrs), out_cols
        raise JobError(f"Not supported node type: {t}")

    plan.root, cols = rec(explain, parsed.output_attrs)
    return plan, cols
