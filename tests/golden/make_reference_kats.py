"""Extract the reference's known-answer vectors for the perception path into
tests/golden/reference_kats.json.

Reads abm/projects/cooperative_signaling/cs_agent/tests/test_cs_supcalc.py
from a reference checkout as text (the module is parsed with ``ast``, never
imported or run) and keeps the inputs and expected outputs of the cases that
exercise functions this repository implements:

* test_projection_field (:143-158): projection_field inputs and the field;
* test_calculate_closed_angle (:161-169): vector pairs and closed angles;
* test_random_walk (:10-34): the uniform draws of numpy's legacy MT19937
  seeded 42 (random_walk(desired_vel, exp_theta_min, exp_theta_max)).

Run in a checkout that holds the reference:
    python tests/golden/make_reference_kats.py /root/reference
"""
import ast
import json
import os
import sys

SRC = "abm/projects/cooperative_signaling/cs_agent/tests/test_cs_supcalc.py"


def _value(node):
    """Literal value of a parametrize entry; np.pi and np.array(...) spelled out."""
    if isinstance(node, ast.Attribute) and node.attr == "pi":
        return {"pi": 1}
    if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Div, ast.Mult)):
        left, right = _value(node.left), _value(node.right)
        if isinstance(left, dict) and "pi" in left:
            f = right if isinstance(node.op, ast.Mult) else 1 / right
            return {"pi": left["pi"] * f}
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        v = _value(node.operand)
        return {"pi": -v["pi"]} if isinstance(v, dict) else -v
    if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "array":
        return _value(node.args[0])
    if isinstance(node, (ast.Tuple, ast.List)):
        return [_value(e) for e in node.elts]
    return ast.literal_eval(node)


def main(ref_root):
    tree = ast.parse(open(os.path.join(ref_root, SRC)).read())
    out = {"source": SRC, "projection_field": [], "closed_angle": [], "random_walk_seed42": []}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef):
            continue
        for dec in fn.decorator_list:
            if isinstance(dec, ast.Call) and getattr(dec.func, "attr", "") == "parametrize":
                names = [n.strip() for n in ast.literal_eval(dec.args[0]).split(",")]
                for case in dec.args[1].elts:
                    row = dict(zip(names, _value(case)))
                    if fn.name == "test_projection_field":
                        out["projection_field"].append(row)
                    elif fn.name == "test_calculate_closed_angle":
                        out["closed_angle"].append(row)
        if fn.name == "test_random_walk":
            # assert dtheta == <float> in order, with the (min, max) of each call
            calls, vals = [], []
            for node in ast.walk(fn):
                if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "random_walk":
                    kw = {k.arg: _value(k.value) for k in node.keywords}
                    calls.append([kw.get("exp_theta_min"), kw.get("exp_theta_max")])
                if (isinstance(node, ast.Compare) and isinstance(node.left, ast.Name)
                        and node.left.id == "dtheta"):
                    vals.append(_value(node.comparators[0]))
            # the second call takes exp_theta_min from a patched module value
            patched = [n for n in ast.walk(fn) if isinstance(n, ast.Call)
                       and getattr(n.func, "attr", "") == "patch"]
            mins = [_value(p.args[1]) for p in patched if "exp_theta_min" in ast.literal_eval(p.args[0])]
            for c in calls:
                if c[0] is None and mins:
                    c[0] = mins.pop(0)
            out["random_walk_seed42"] = [{"lo": c[0], "hi": c[1], "dtheta": v} for c, v in zip(calls, vals)]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
