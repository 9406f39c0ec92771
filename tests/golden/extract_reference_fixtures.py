"""Extract the QP data sets that the reference commits as C source into JSON
fixtures under tests/golden/.

Run once in the build container (the reference tree is only mounted there):

    python tests/golden/extract_reference_fixtures.py /root/reference

Sources (committed C data of the reference; values only, no code):
  examples/problems/{hs21,hs35,qptest,cvxqp2_s,qpcblend,largeqp,primalc1}.c
  tests/large_qp/large_qp_data.c   (objective 0.106081, checked at rel 1e-4 by
                                     tests/large_qp/test_large_qp.cpp:33-37)

Each fixture holds n, m, P (upper-triangular CSC), q, A (CSC), l, u, the
objective value when the file states one, and the settings initialiser as a
list of numbers in declaration order.
"""
import json
import os
import re
import sys

ARRAY_RE = re.compile(r"^(OSQPInt|OSQPFloat)\s+(\w+)\[(\d+)\]\s*=\s*\{(.*?)\};", re.S | re.M)
SCALAR_RE = re.compile(r"^(OSQPInt|OSQPFloat)\s+(\w+)\s*=\s*([^;{]+);", re.M)
SETTINGS_RE = re.compile(r"^OSQPSettings\s+(\w+)\s*=\s*\{(.*?)\};", re.S | re.M)


def _num(tok):
    tok = tok.strip()
    tok = re.sub(r"\(OSQPFloat\)", "", tok).strip()
    if tok in ("OSQP_DIRECT_SOLVER",):
        return 1
    if tok in ("OSQP_INDIRECT_SOLVER",):
        return 2
    if tok in ("OSQP_DIAGONAL_PRECONDITIONER",):
        return 1
    if tok in ("OSQP_NO_PRECONDITIONER",):
        return 0
    if re.fullmatch(r"-?\d+", tok):
        return int(tok)
    return float(tok)


def parse_c_data(path, prefix):
    text = open(path).read()
    arrays = {}
    for typ, name, n, body in ARRAY_RE.findall(text):
        vals = [_num(t) for t in body.split(",") if t.strip()]
        assert len(vals) == int(n), (name, len(vals), n)
        if typ == "OSQPFloat":
            vals = [float(v) for v in vals]
        arrays[name] = vals
    scalars = {}
    for typ, name, val in SCALAR_RE.findall(text):
        scalars[name] = _num(val)
    settings = None
    m = SETTINGS_RE.search(text)
    if m:
        settings = [_num(t) for t in m.group(2).split(",") if t.strip()]
    d = f"{prefix}_data_"
    out = {
        "name": prefix,
        "source": os.path.relpath(path, sys.argv[1] if len(sys.argv) > 1 else "/root/reference"),
        "n": scalars[d + "n"],
        "m": scalars[d + "m"],
        "P": {"p": arrays[d + "P_csc_p"], "i": arrays[d + "P_csc_i"], "x": arrays[d + "P_csc_x"]},
        "A": {"p": arrays[d + "A_csc_p"], "i": arrays[d + "A_csc_i"], "x": arrays[d + "A_csc_x"]},
        "q": arrays[d + "q_val"],
        "l": arrays[d + "l_val"],
        "u": arrays[d + "u_val"],
        "obj_val": scalars.get(prefix + "_obj_val"),
        "settings_initializer": settings,
    }
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    jobs = [(os.path.join(ref, "examples", "problems", f"{p}.c"), p)
            for p in ("hs21", "hs35", "qptest", "cvxqp2_s", "qpcblend", "largeqp", "primalc1")]
    jobs.append((os.path.join(ref, "tests", "large_qp", "large_qp_data.c"), "prob1"))
    for path, prefix in jobs:
        fx = parse_c_data(path, prefix)
        name = "large_qp" if prefix == "prob1" else prefix
        with open(os.path.join(here, f"{name}.json"), "w") as f:
            json.dump(fx, f)
        print(f"{name}: n={fx['n']} m={fx['m']} nnzP={len(fx['P']['x'])} nnzA={len(fx['A']['x'])} obj={fx['obj_val']}")


if __name__ == "__main__":
    main()
