"""The batched engine's matrix layout, checked on the CPU (no GPU needed).

tests/layout_check.hip builds each problem's stage image with the library's own
host code and replays, lane by lane and with a bounds check on every index,
the setup kernel's derivation of the row groups, slices and compact long rows
and the product loops that read them; see its header for the list of checks.
This test writes the problems (MPC, portfolios on both sides of the sliced
buckets' threshold, a random QP, a lasso and a matrix mixing every row-length
class), builds the harness against the library's objects and runs it; a run
with one slice length corrupted must fail, so the checks are not vacuous.  The
harness also builds every image in an image arena over ordinary memory, sized
to hold about half of them, and checks the arena's images against heap-built
ones, the fall-back to the heap, and concurrent take / give from 8 threads.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from osqp_amd import build as B
from osqp_amd import problems as pr

from .conftest import ROOT


def _problems():
    out = [pr.mpc(N, seed=N) for N in (10, 25, 50)]
    out += [pr.portfolio(n, seed=n) for n in (500, 2000, 3500, 5000)]
    out.append(pr.random_qp(300, 3000, 0.02, 0.01, seed=3))
    out.append(pr.lasso(60, 600, 0.15, seed=4))
    # every row-length class: empty, short, medium, 64 wide (slices wider
    # than one load round), long; enough entries for the sliced buckets
    rng = np.random.default_rng(5)
    rows, cols = 2400, 2800
    lens = rng.choice([0, 1, 3, 7, 8, 9, 15, 16, 31, 33, 63, 64, 65, 90], rows)
    ri = np.repeat(np.arange(rows), lens)
    ci = np.concatenate([np.sort(rng.choice(cols, L, replace=False)) for L in lens])
    A = sp.csc_matrix((rng.standard_normal(ri.size), (ri, ci)), shape=(rows, cols))
    out.append(dict(P=sp.csc_matrix(sp.eye(cols)), A=A, q=np.zeros(cols), l=-np.ones(rows), u=np.ones(rows)))
    return out


def _write(path, probs):
    with open(path, "wb") as f:
        f.write(struct.pack("i", len(probs)))
        for p in probs:
            P = sp.csc_matrix(sp.triu(p["P"]))
            A = sp.csc_matrix(p["A"])
            P.sort_indices()
            A.sort_indices()
            f.write(struct.pack("ii", P.shape[0], A.shape[0]))
            for M in (P, A):
                f.write(struct.pack("i", M.nnz))
                f.write(M.indptr.astype(np.int32).tobytes())
                f.write(M.indices.astype(np.int32).tobytes())
                f.write(M.data.astype(np.float64).tobytes())


def test_layout_replay(tmp_path):
    objs = [os.path.join(B.BUILD, f) for f in sorted(os.listdir(B.BUILD))] if os.path.isdir(B.BUILD) else []
    objs = [o for o in objs if o.endswith(".o") and "serve_batch" not in o]
    if not os.path.exists(B.HIPCC) or not objs:
        pytest.skip("needs hipcc and the library's objects (python -m osqp_amd.build)")
    probs = tmp_path / "probs.bin"
    _write(probs, _problems())
    obj, exe = tmp_path / "layout_check.o", tmp_path / "layout_check"
    subprocess.run([B.HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-Wno-unused-result",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "osqp_amd", "csrc"),
                    "-c", os.path.join(ROOT, "tests", "layout_check.hip"), "-o", str(obj)],
                   check=True, capture_output=True)
    subprocess.run([B.HIPCC, "--offload-arch=gfx950", str(obj)] + objs +
                   ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe), str(probs)], capture_output=True, text=True)
    assert r.returncode == 0 and "all checks passed" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    # the problems exercised slices, slices wider than one load round, and long rows
    counts = [ln for ln in r.stdout.splitlines() if ln.startswith("sliced matrices")][0]
    sliced, wide, longs = (int(v.split()[-1]) for v in counts.split(","))
    assert sliced > 0 and wide > 0 and longs > 0, counts
    bad = subprocess.run([str(exe), str(probs), "plant"], capture_output=True, text=True)
    assert bad.returncode == 1 and "FAILURES" in bad.stdout, bad.stdout[-2000:]
    shutil.rmtree(tmp_path, ignore_errors=True)
