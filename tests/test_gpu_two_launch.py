"""The two-launch Jacobi-PCG iteration (direction formed in the SpMV's
gathers from interleaved (z, p) pairs, then the update), selected by
FFH_PCG_LAUNCHES=2 for the whole process: a child process runs FP64 and
mixed solves of the reduced P2 Poisson system on a 10^3 x 6 box (graph
replay on and off) and the parent checks them against the oracle's
Jacobi-PCG (iterations within one, solutions within 1e-8 relative: PCG stops
at a relative residual of 1e-10)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import fortfem_amd as ff
import oracle as O
v, t = O.unit_box_mesh(10)
a, g, G = O.lagrange_stiffness_csc(v, t, 2, 6)
cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
red, free, rb = O.constrained_reduction(a, cons, np.zeros(G), np.ones(G))
m = ff.Matrix.upload(ff.CSC(red.n_rows, red.n_cols, red.col_ptr, red.row_idx, red.val))
out = {"launches": [ff.pcg_launches_per_iteration(red.n_rows, p) for p in ("fp64", "fp32")], "runs": []}
for prec in ("fp64", "mixed"):
    for graph in (False, True):
        x = np.zeros(red.n_rows)
        opts = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10,
                                 max_iterations=5000, breakdown="relative", precision=prec,
                                 use_graph=graph)
        st = m.solve(rb, x, opts)
        out["runs"].append({"prec": prec, "graph": graph, "conv": bool(st.converged),
                            "iters": int(st.iterations), "x": x.tolist()})
print(json.dumps(out))
"""


def test_two_launch_form_matches_oracle():
    env = dict(os.environ, FFH_PCG_LAUNCHES="2")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["launches"] == [2, 2]
    v, t = O.unit_box_mesh(10)
    a, g, G = O.lagrange_stiffness_csc(v, t, 2, 6)
    cons, _ = O.lagrange_dirichlet(v, t, 2, g, G)
    red, free, rb = O.constrained_reduction(a, cons, np.zeros(G), np.ones(G))
    xo, sto = O.solve_sparse(red, rb, method="pcg", preconditioner="jacobi", tolerance=1e-10,
                             max_iterations=5000, breakdown="relative")
    for run in out["runs"]:
        x = np.array(run["x"])
        assert run["conv"], run["prec"]
        if run["prec"] == "fp64":
            assert abs(run["iters"] - sto.iterations) <= 1, (run["iters"], sto.iterations)
        assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo), (run["prec"], run["graph"])
