"""bench.py's command line: every --workload choice dispatches to a function
that exists (the script is parsed, not run: its legs need a GPU)."""
import ast
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def test_every_workload_has_a_leg():
    src = open(os.path.join(HERE, "..", "bench.py")).read()
    tree = ast.parse(src)
    defined = {n.name for n in tree.body if isinstance(n, ast.FunctionDef)}
    choices = None
    for node in ast.walk(tree):
        if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "add_argument" and node.args and \
                isinstance(node.args[0], ast.Constant) and node.args[0].value == "--workload":
            for kw in node.keywords:
                if kw.arg == "choices":
                    choices = [e.value for e in kw.value.elts]
    assert choices and "classify" in choices
    called = {n.func.id for n in ast.walk(tree) if isinstance(n, ast.Call) and isinstance(n.func, ast.Name)}
    for w in choices:
        if w == "classify":
            continue  # the default line, in main()
        assert f"bench_{w}" in defined, f"--workload {w} has no bench_{w}()"
        assert f"bench_{w}" in called
