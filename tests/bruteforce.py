"""Pure-Python brute-force evaluator of the contest join semantics, for small
inputs only.  It is written from the documented behaviour of the reference
(not from the oracle's code) and pins the oracle on random plans:

* build side = the child with fewer rows, ties to the left
  (join_setup.h:108-136); nested loop when the build side has < 8 rows
  (execute.cpp:236-238);
* NULL never equals anything;
* non-root join results are value_t intermediates: an INT32 cell holds
  INT32_MIN for NULL (intermediate.h:40-63), so a key read from an
  intermediate is skipped on probe and nested-loop build when it equals
  INT32_MIN (hash_join.h:120-127, nested_loop.h:85) but a hash build over an
  intermediate inserts it (hashtable.h:357);
* root INT32 outputs equal to INT32_MIN come out NULL (page_builders.h:123);
* INT64 / FP64 / VARCHAR values are carried unchanged.
"""
from __future__ import annotations

I32_MIN = -(2 ** 31)


def evaluate(plan_nodes, tables, root):
    """plan_nodes: list of dicts like cjtest.Plan.nodes; tables: list of
    (types, rows).  Returns the result rows (tuples) in any order."""

    def run(idx, is_root):
        n = plan_nodes[idx]
        if n["kind"] == 1:  # scan
            types, rows = tables[n["table"]]
            cols = [a for a, _ in n["attrs"]]
            out_types = [types[a] for a in cols]
            return False, out_types, [tuple(r[a] for a in cols) for r in rows]
        l_inter, l_types, l_rows = run(n["left"], False)
        r_inter, r_types, r_rows = run(n["right"], False)
        build_left = len(l_rows) <= len(r_rows)
        if build_left:
            b_inter, b_rows, battr = l_inter, l_rows, n["left_attr"]
            p_inter, p_rows, pattr = r_inter, r_rows, n["right_attr"]
        else:
            b_inter, b_rows, battr = r_inter, r_rows, n["right_attr"]
            p_inter, p_rows, pattr = l_inter, l_rows, n["left_attr"]
        nested = len(b_rows) < 8

        def bkey(row):
            k = row[battr]
            if b_inter:
                if nested and k == I32_MIN:
                    return None
                return k  # hash build over an intermediate keeps INT32_MIN
            return k

        def pkey(row):
            k = row[pattr]
            if p_inter and k == I32_MIN:
                return None
            return k

        pairs = []
        for pr in p_rows:
            pk = pkey(pr)
            if pk is None:
                continue
            for br in b_rows:
                bk = bkey(br)
                if bk is not None and bk == pk:
                    pairs.append((br, pr) if build_left else (pr, br))
        # pairs are (left_row, right_row)
        types = l_types + r_types
        out_idx = [a for a, _ in n["attrs"]]
        out_types = [types[a] for a in out_idx]
        out = []
        for lr, rr in pairs:
            full = lr + rr
            row = []
            for a in out_idx:
                v = full[a]
                t = types[a]
                if t == "INT32":
                    if is_root:
                        v = None if (v is None or v == I32_MIN) else v
                    else:
                        v = I32_MIN if v is None else v
                row.append(v)
            out.append(tuple(row))
        return True, out_types, out

    _, _, rows = run(root, True)
    return rows
