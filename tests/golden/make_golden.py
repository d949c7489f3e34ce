"""Extract the golden numbers the reference's own unit tests hold for the
sparse-matrix path into tests/golden/sparsematrix_golden.json.

Sources (data only, read as text from the reference checkout):
  tests/ref/sm_transpose.out  -- nnz of set_as_transpose_of (keeps explicit
                                 zeros) and set_as_transpose_of2 (drops them)
                                 for the inputs built by tests/sm_transpose.cc
  tests/ref/boost_test0.out   -- adjacency (non-zero pattern) of a 7x7 matrix
                                 and of its transpose, from tests/boost_test0.cc

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import re

REF = "/root/reference/tests/ref"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sparsematrix_golden.json")


def parse_sm_transpose(text):
    cases = []
    blocks = re.split(r"^== ", text, flags=re.M)[1:]
    for b in blocks:
        lines = b.strip().splitlines()
        name = lines[0].strip()
        nnz = [int(m.group(1)) for m in re.finditer(r"nnz (\d+)", b)]
        shapes = re.findall(r"\[ (\d+) x (\d+) \]", b)
        cases.append({"name": name, "input_shape": [int(x) for x in shapes[0]],
                      "nnz_first": nnz[0], "nnz_second": nnz[1]})
    return cases


def parse_boost_test0(text):
    lines = text.strip().splitlines()
    adj_b_row1 = [int(x) for x in lines[0].split()]
    edges = []
    i = 2
    while not lines[i].startswith("pg"):
        m = re.match(r"(\d+):(\d+) -- (\S+)", lines[i])
        edges.append([int(m.group(1)), int(m.group(2)), float(m.group(3))])
        i += 1
    i += 1
    graph = {}
    while "-->" in lines[i]:
        a, b = lines[i].split("-->")
        graph[int(a)] = [int(x) for x in b.split()]
        i += 1
    degree = {}
    for ln in lines[i:]:
        a, b = ln.split()
        degree[int(a)] = int(b)
    return {"B_row1_adjacent": adj_b_row1, "B_row1_edges": edges, "A_nonzero_adjacency": graph,
            "A_out_degree": degree}


def main():
    with open(os.path.join(REF, "sm_transpose.out")) as f:
        smt = parse_sm_transpose(f.read())
    with open(os.path.join(REF, "boost_test0.out")) as f:
        bt0 = parse_boost_test0(f.read())
    data = {"sm_transpose": smt, "boost_test0": bt0,
            "source": ["tests/ref/sm_transpose.out", "tests/ref/boost_test0.out"]}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
