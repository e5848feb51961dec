"""Write fading_tables.json: the tap tables of the standard fading parameterisations.

Data: 3GPP TR 38.901 Table 7.7.2-1..5 (TDL-A..E, normalised delays, powers in
dB, K-factors) and COST 259 (typical urban, rural area, hilly terrain), as the
reference tabulates them in ``hermespy/channel/fading/tdl.py:101-369`` and
``hermespy/channel/fading/cost259.py:94-202``.  The reference files are parsed
with ``ast`` (literal arrays and scalars only; nothing is imported or run).

    python tests/golden/make_fading_tables.py /root/reference

The output is the product's table file, hermespy_amd/data/fading_tables.json.
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path


def _literal_array(node):
    """np.array([...]) or 1e-6 * np.array([...]) -> (scale, list)."""

    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
        scale = ast.literal_eval(node.left)
        _, vals = _literal_array(node.right)
        return scale, vals
    if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "array":
        return 1.0, ast.literal_eval(node.args[0])
    return None, None


def _branches(tree, enum_name):
    out = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.If) and isinstance(node.test, ast.Compare):
            cmp = node.test.comparators[0]
            if isinstance(cmp, ast.Attribute) and getattr(cmp.value, "id", "") == enum_name:
                out[cmp.attr] = node.body
    return out


def extract(root: Path) -> dict:
    tables = {"TDL": {}, "COST259": {}}
    tdl = ast.parse((root / "hermespy/channel/fading/tdl.py").read_text())
    for name, body in _branches(tdl, "TDLType").items():
        entry = {"rice_factor_first": 0.0, "los_doppler_frequency": None}
        for st in body:
            if isinstance(st, ast.Assign):
                tgt = st.targets[0]
                if isinstance(tgt, ast.Name) and tgt.id in ("normalized_delays", "power_db"):
                    entry[tgt.id] = _literal_array(st.value)[1]
                elif isinstance(tgt, ast.Subscript) and getattr(tgt.value, "id", "") == "rice_factors":
                    entry["rice_factor_first"] = float(ast.literal_eval(st.value))
                elif isinstance(tgt, ast.Name) and tgt.id == "los_doppler_frequency":
                    entry["los_doppler_frequency"] = float(ast.literal_eval(st.value))
        entry["source"] = f"hermespy/channel/fading/tdl.py:{body[0].lineno}"
        tables["TDL"][name] = entry
    cost = ast.parse((root / "hermespy/channel/fading/cost259.py").read_text())
    for name, body in _branches(cost, "Cost259Type").items():
        entry = {"los_first": False}
        for st in body:
            if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name):
                nm = st.targets[0].id
                if nm == "delays":
                    scale, vals = _literal_array(st.value)
                    entry["delays_s"] = [scale * v for v in vals]
                elif nm == "power_db":
                    entry["power_db"] = _literal_array(st.value)[1]
                elif nm == "rice_factors" and "inf" in ast.dump(st.value):
                    entry["los_first"] = True
                elif nm == "los_angle":
                    entry["los_angle_cos"] = 0.7
        entry["source"] = f"hermespy/channel/fading/cost259.py:{body[0].lineno}"
        tables["COST259"][name] = entry
    return tables


if __name__ == "__main__":
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    t = extract(root)
    out = Path(__file__).resolve().parents[2] / "hermespy_amd" / "data" / "fading_tables.json"
    out.write_text(json.dumps(t, indent=1))
    print({k: {m: len(v.get("power_db", [])) for m, v in d.items()} for k, d in t.items()})
