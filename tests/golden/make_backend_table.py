"""Writes backend_functions.json: the names, result kinds and argument counts of the
reference's FFI table (`backend_functions`, traceon/backend/__init__.py:287-328) and the
data symbols its loader reads with in_dll (:58-66).  The reference's loader binds every
name of that table with getattr (:341-342), so a library that replaces its backend must
export all of them.  Run in a checkout that has the reference at REF (read as text, not
imported)."""
import ast
import json
import os
import sys

REF = os.environ.get("TRACEON_REF", "/root/reference/traceon/backend/__init__.py")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "backend_functions.json")


def main():
    tree = ast.parse(open(REF).read())
    table = None
    data = []
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and any(isinstance(t, ast.Name) and t.id == "backend_functions" for t in node.targets):
            table = node.value
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) and node.func.attr == "in_dll":
            if len(node.args) == 2 and isinstance(node.args[1], ast.Constant):
                data.append(node.args[1].value)
    assert isinstance(table, ast.Dict)
    entries = []
    for k, v in zip(table.keys, table.values):
        res = v.elts[0]
        kind = "void" if isinstance(res, ast.Constant) and res.value is None else ast.unparse(res)
        entries.append({"name": k.value, "result": kind, "nargs": len(v.elts) - 1})
    json.dump({"source": "traceon/backend/__init__.py:287-328 (backend_functions), :58-66 (in_dll)",
               "functions": entries, "data": sorted(set(data))}, open(OUT, "w"), indent=1)
    print(f"wrote {OUT}: {len(entries)} functions, {len(set(data))} data symbols")


if __name__ == "__main__":
    sys.exit(main())
