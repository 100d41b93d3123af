"""Extract the golden vectors and fixed test matrices held in the reference's
own Julia tests into tests/golden/fixtures.json (data only: inputs and
expected outputs). Run where the reference's sources are present; the JSON is
committed, and the tests read only the JSON, never the reference.

    python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def nums(text):
    """Numbers of a Julia vector/matrix literal, in reading order."""
    return [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?", text)]


def after(text, anchor, start=0):
    i = text.index(anchor, start)
    return i + len(anchor)


def bracket(text, pos):
    """The [...] literal starting at or after pos (balanced)."""
    i = text.index("[", pos)
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                return text[i + 1:j], j + 1
    raise ValueError("unbalanced literal")


def ints(v):
    return [int(x) for x in v]


fx = {}

# --- test_graph.jl: Figure 3.1.3 (mgrap001) and its unsymmetric variant
# (mgrap003/mgrap004) ---
g = read("test/test_graph.jl")
p = after(g, "module mgrap001")
I, p = bracket(g, after(g, "I =", p))
J, p = bracket(g, after(g, "J =", p))
xadj, p = bracket(g, after(g, "graph.xadj ==", p))
adj, p = bracket(g, after(g, "graph.adj ==", p))
fx["fig313"] = {"source": "test/test_graph.jl mgrap001",
                "I": ints(nums(I)), "J": ints(nums(J)),
                "xadj": ints(nums(xadj)), "adj": ints(nums(adj))}
p = after(g, "module mgrap004")
I, p = bracket(g, after(g, "I =", p))
J, p = bracket(g, after(g, "J =", p))
xadj, p = bracket(g, after(g, "graph.xadj ==", p))
adj, p = bracket(g, after(g, "graph.adj ==", p))
fx["fig313_unsym"] = {"source": "test/test_graph.jl mgrap004",
                      "I": ints(nums(I)), "J": ints(nums(J)),
                      "xadj": ints(nums(xadj)), "adj": ints(nums(adj))}

# --- test_sparse_method.jl: tridiagonal n = 11 (msprs003-006) ---
t = read("test/test_sparse_method.jl")
p = after(t, "module msprs003")
rperm, p = bracket(t, after(t, "o.rperm == vec(", p))
rinvp, p = bracket(t, after(t, "o.rinvp == vec(", p))
p = after(t, "module msprs004")
xlnz, p = bracket(t, after(t, "s.slvr.xlnz ==", p))
xunz, p = bracket(t, after(t, "s.slvr.xunz ==", p))
xlindx, p = bracket(t, after(t, "s.slvr.xlindx ==", p))
lindx, p = bracket(t, after(t, "s.slvr.lindx ==", p))
p = after(t, "module msprs005")
unz0, p = bracket(t, after(t, "s.slvr.unz  == vec(T", p))
lnz0, p = bracket(t, after(t, "s.slvr.lnz  == vec(T", p))
p = after(t, "module msprs006")
lnz1, p = bracket(t, after(t, "s_lnz  = T", p))
unz1, p = bracket(t, after(t, "s_unz = T", p))
fx["tridiag11"] = {
    "source": "test/test_sparse_method.jl msprs003-006 (4 on the diagonal, -1 off)",
    "n": 11,
    "rperm": ints(nums(rperm)), "rinvp": ints(nums(rinvp)),
    "xlnz": ints(nums(xlnz)), "xunz": ints(nums(xunz)),
    "xlindx": ints(nums(xlindx)), "lindx": ints(nums(lindx)),
    "lnz_inmatrix": nums(lnz0), "unz_inmatrix": nums(unz0),
    "lnz_factor": nums(lnz1), "unz_factor": nums(unz1),
    "factor_rtol": 1e-6,
}

# --- msprs010: fixed 31 x 31 matrix, rhs = 1:31 ---
p = after(t, "module msprs010")
p = after(t, "spm = sparse(", p)
I31, p = bracket(t, p)
J31, p = bracket(t, p)
V31, p = bracket(t, p)
fx["m31"] = {"source": "test/test_sparse_method.jl msprs010", "n": 31,
             "I": ints(nums(I31)), "J": ints(nums(J31)), "V": nums(V31),
             "rhs": list(range(1, 32)), "rtol": 1e-6}


# --- fixed 4 x 4 matrices (test_small.jl, test_structunsymm.jl) ---
def dense_after(text, anchor, start):
    q = after(text, anchor, start)
    body, q = bracket(text, q)
    rows = [r for r in body.split(";") if r.strip()]
    return [nums(r) for r in rows], q


s = read("test/test_small.jl")
small = []
q = after(s, "module m_inconsistency_data_structure_symm")
A, q = dense_after(s, "A = ", q)
small.append({"source": "test/test_small.jl m_inconsistency_data_structure_symm",
              "A": A, "b_index": 1})
q = after(s, "module m_inconsistency_data_structure_unsymm")
A, q = dense_after(s, "A = ", q)
small.append({"source": "test/test_small.jl m_inconsistency_data_structure_unsymm",
              "A": A, "b_index": 1})
u = read("test/test_structunsymm.jl")
q = 0
for k in range(2):
    A, q = dense_after(u, "A=", q)
    small.append({"source": f"test/test_structunsymm.jl matrix {k + 1}", "A": A,
                  "b_index": 1})
fx["small4"] = small

with open(OUT, "w") as f:
    json.dump(fx, f, indent=1)
print("wrote", OUT)
