"""Write tests/golden/lgl_literals.json: the decimal Lobatto node and weight
tables of the reference (main/basis.py:9-57) for the node counts whose
entries are plain decimal literals (7..10 nodes).

    python tests/golden/make_lgl_fixture.py [path/to/main/basis.py]

The reference file is read as text: the numbers are taken from the literal
tokens of the ``lgl_nodes`` / ``lgl_weights`` dictionaries (kept as strings,
so that every printed digit survives); nothing of the reference is imported
or evaluated.  Entries written as expressions (``np.sqrt(...)``, fractions;
node counts 1..6) are skipped -- the tests check those against their closed
forms.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/main/basis.py"
NUMBER = re.compile(r"^-?\d+(\.\d+)?$")


def literal_tables(text, name):
    """{count: [decimal strings]} of the dictionary ``name = {...}`` in text,
    for the keys whose entries are all plain decimal literals."""
    start = text.index(name + " = {")
    body = text[text.index("{", start) + 1:]
    depth, end = 1, 0
    for end, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            break
    body = body[:end]
    out = {}
    for m in re.finditer(r"(\d+)\s*:\s*\[(.*?)\]", body, flags=re.S):
        tokens = [t.strip() for t in m.group(2).replace("\n", " ").split(",") if t.strip()]
        if tokens and all(NUMBER.match(t) for t in tokens):
            out[int(m.group(1))] = tokens
    return out


def main(src):
    with open(src) as fh:
        text = fh.read()
    nodes = literal_tables(text, "lgl_nodes")
    weights = literal_tables(text, "lgl_weights")
    keys = sorted(k for k in nodes if k in weights and k >= 7)
    out = {"source": "main/basis.py:9-57 (decimal literals)",
           "nodes": {str(k): nodes[k] for k in keys},
           "weights": {str(k): weights[k] for k in keys}}
    path = os.path.join(HERE, "lgl_literals.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"node counts {keys} -> {path}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC)
