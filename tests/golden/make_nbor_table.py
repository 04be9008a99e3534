"""Write tests/golden/nbor_table.json: the neighbour rule of the reference,
read from the text of src/fmm/nbor.cpp (its 26 direction cases, spelled out
branch by branch) and src/fmm/fmm.h (the order of FMM::Dir).

For every direction and branch index 0-7 of getNeighborGeqSize
(nbor.cpp:10-483) the table holds what the case does:
  ["sibling", k]             return base->branches[k]
  ["parent", E, k, "leaf"]   ask the parent along E; a Leaf answer is
                             returned, a Stem answer gives its child k
  ["parent", E, k, "drop"]   the same, but a Leaf answer is dropped (it is
                             found along the lower-dimensional direction E)
and for getNeighborsLeqSize (nbor.cpp:490-644) the children of a stem
neighbour pushed for each direction, in order.

The case bodies are read with a few patterns that cover every case of the
file; conditions and index expressions (integers, branchIdx, + - == < >= ||)
are evaluated for each branch index after checking that they contain
nothing else.  Run from the repository root; the reference is read from
/root/reference."""
import json
import os
import re

REF = "/root/reference/src/fmm"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nbor_table.json")

SAFE = re.compile(r"^[\sbranchIdx0-9+\-=<>|()]*$")


def ev(expr, b):
    expr = expr.strip()
    if not SAFE.match(expr):
        raise ValueError(f"unexpected expression {expr!r}")
    py = expr.replace("||", " or ")
    return eval(py, {"__builtins__": {}}, {"branchIdx": b})


def dir_order():
    txt = open(os.path.join(REF, "fmm.h")).read()
    m = re.search(r"enum class Dir\s*\{([^}]*)\}", txt)
    return [t.strip() for t in m.group(1).split(",") if t.strip()]


def split_cases(body):
    parts = re.split(r"\bcase\s+([A-Z]+)\s*:", body)
    return {parts[i]: parts[i + 1] for i in range(1, len(parts), 2)}


def geq_table(names):
    txt = open(os.path.join(REF, "nbor.cpp")).read()
    a = txt.index("getNeighborGeqSize(const Dir dir) const")
    b = txt.index("getNeighborsLeqSize(")
    cases = split_cases(txt[a:b])
    table = {}
    for name in names:
        body = " ".join(cases[name].split())
        rows = []
        for br in range(8):
            rows.append(geq_case(name, body, br))
        table[name] = rows
    return table


def geq_case(name, body, b):
    # statements in order; the first that applies decides
    pos = 0
    # 1. sibling returns: if (C) return base->branches[E];
    for m in re.finditer(r"if \(([^{};]*?)\) return base->branches\[([^\]]+)\];", body):
        if ev(m.group(1), b):
            return ["sibling", ev(m.group(2), b)]
    # 2. full exit inside a block: if (C) { nbor = base->getNeighborGeqSize(dir); ... return nbor->branches[E]; }
    m = re.search(r"if \(([^{};]*?)\) \{ nbor = base->getNeighborGeqSize\(dir\);.*?return nbor->branches\[([^\]]+)\]; \}",
                  body)
    if m and ev(m.group(1), b):
        return ["parent", name, ev(m.group(2), b), "leaf"]
    # 3. one-axis directions: unconditional full exit
    m = re.search(r"\} ?nbor = base->getNeighborGeqSize\(dir\);|; nbor = base->getNeighborGeqSize\(dir\);", body)
    if m and "?" not in body and "switch" not in body:
        r = re.search(r"if \(nbor->isNodeType<Leaf>\(\)\) return nbor; return nbor->branches\[([^\]]+)\];", body)
        return ["parent", name, ev(r.group(1), b), "leaf"]
    # 4. two-axis directions: nbor = (C) ? base->get...(A) : base->get...(B); ... return (C2) ? nbor->branches[E1] : nbor->branches[E2];
    m = re.search(r"nbor = \(([^?]*?)\) \? base->getNeighborGeqSize\(([A-Z]+)\) : base->getNeighborGeqSize\(([A-Z]+)\);",
                  body)
    if m:
        e = m.group(2) if ev(m.group(1), b) else m.group(3)
        r = re.search(r"if \(nbor->isNodeType<Leaf>\(\)\) return nullptr; return \(([^?]*?)\) \? "
                      r"nbor->branches\[([^\]]+)\] : nbor->branches\[([^\]]+)\];", body)
        k = ev(r.group(2), b) if ev(r.group(1), b) else ev(r.group(3), b)
        return ["parent", e, k, "drop"]
    # 5. three-axis directions: a switch over branchIdx picks the direction
    m = re.search(r"switch \(branchIdx\) \{(.*?)\} \} \(\);", body)
    if m:
        sw = dict((int(c), d) for c, d in re.findall(r"case (\d+): return base->getNeighborGeqSize\(([A-Z]+)\);",
                                                      m.group(1)))
        r = re.search(r"if \(nbor->isNodeType<Leaf>\(\)\) return nullptr; return nbor->branches\[([^\]]+)\];", body)
        return ["parent", sw[b], ev(r.group(1), b), "drop"]
    raise ValueError(f"case {name}: branch {b} not matched")


def leq_table(names):
    txt = open(os.path.join(REF, "nbor.cpp")).read()
    a = txt.index("getNeighborsLeqSize(")
    cases = split_cases(txt[a:])
    out = {}
    for name in names:
        out[name] = [int(k) for k in re.findall(r"queue\.push\(nbor->branches\[(\d+)\]\)", cases[name])]
    return out


def main():
    names = dir_order()
    assert len(names) == 26
    data = {"source": "src/fmm/nbor.cpp (getNeighborGeqSize, getNeighborsLeqSize), src/fmm/fmm.h (FMM::Dir)",
            "dirs": names, "geq": geq_table(names), "leq": leq_table(names)}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
