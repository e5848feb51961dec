"""Extract the reference's PSK/QAM/PAM known-answer vectors into psk_qam_kat.json.

Source: ``tests/unit_tests/modem/tools/test_psk_qam_mapping.py:80-144`` of the
reference (the ``test_symbols_*`` methods).  The file is parsed with ``ast``
(nothing of it is imported or executed); for each method the literal bit list,
the literal expected symbol list, its ``np.sqrt(k)`` normalisation and the
``PskQamMapping(order, is_complex=...)`` arguments are written out as data.

Run from the repository root where the reference is mounted:
    python tests/golden/make_psk_qam_kat.py /root/reference
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

REL = "tests/unit_tests/modem/tools/test_psk_qam_mapping.py"


def _complex_list(node: ast.AST) -> list:
    """Literal list of ints / complex expressions (a + bj, a - bj) -> [[re, im], ...]."""

    out = []
    for el in node.elts:
        v = complex(ast.literal_eval(el))
        out.append([v.real, v.imag])
    return out


def extract(root: Path) -> list[dict]:
    tree = ast.parse((root / REL).read_text())
    cases = []
    for fn in ast.walk(tree):
        if not (isinstance(fn, ast.FunctionDef) and fn.name.startswith("test_symbols_")) or fn.name.endswith("provided"):
            continue
        case = {"name": fn.name, "source": f"{REL}:{fn.lineno}", "is_complex": True}
        for st in ast.walk(fn):
            if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name):
                name, val = st.targets[0].id, st.value
                if name == "bits":
                    case["bits"] = ast.literal_eval(val.args[0])
                elif name.endswith("_symbols") and name != "symbols":
                    if isinstance(val, ast.BinOp):  # np.array([...]) / np.sqrt(k)
                        case["norm"] = ast.literal_eval(val.right.args[0])
                        case["symbols"] = _complex_list(val.left.args[0])
                    else:
                        case["norm"] = 1
                        case["symbols"] = _complex_list(val.args[0])
                elif name == "psk_qam_mapping":
                    case["order"] = ast.literal_eval(val.args[0])
                    for kw in val.keywords:
                        if kw.arg == "is_complex":
                            case["is_complex"] = ast.literal_eval(kw.value)
        cases.append(case)
    return cases


if __name__ == "__main__":
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    cases = extract(root)
    out = Path(__file__).with_name("psk_qam_kat.json")
    out.write_text(json.dumps(cases, indent=1))
    print(f"wrote {len(cases)} cases to {out}")
