"""GPU tests of the served (continuously batched) solve path.

* every served solution is within 1e-8 (relative) of the
  oracle's FP64 Jacobi-PCG solution of the same reduced system;
* a request's result is bit-for-bit the same alone and in a mix (each system's
  reductions run over its own row chunks in a fixed order);
* admission respects the device-memory budget and rejects requests that can
  never fit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ff():
    import fortfem_amd

    return fortfem_amd


def tangential_mask(k, t):
    from test_gpu_nedelec_parity import _tangential_mask

    return _tangential_mask(k, None, t)


def oracle_system(ff, family, k, N, coeff):
    """The reduced FP64 system a served request solves (zero Dirichlet data)."""
    v, t = O.unit_box_mesh(N)
    if family == ff.capi.FAMILY_H1:
        q = 2 * k + 4
        a, g, G = O.lagrange_stiffness_csc(v, t, k, q, cell_kappa=coeff)
        cons, _ = O.lagrange_dirichlet(v, t, k, g, G)
    else:
        K = coeff[:, None, None] * np.eye(3)[None]
        a = O.nedelec_csc(k, v, t, beta=1.0, cell_tensor=K)
        G = a.n_rows
        cons = tangential_mask(k, t).astype(np.int32)
    red, free, _ = O.constrained_reduction(a, cons, np.zeros(G), np.zeros(G))
    return v, t, red


CASES = [(1, 1, 5, 6), (1, 2, 3, 2), (2, 1, 3, 7), (2, 2, 2, 3), (1, 1, 3, 1)]  # family, order, N, nrhs


def make_requests(ff, server, seed=0):
    rng = np.random.default_rng(seed)
    reqs = []
    for key, (fam, k, N, nrhs) in enumerate(CASES):
        v, t, red = oracle_system(ff, fam, k, N, np.ones(6 * N ** 3))
        n, _ = server.prepare(key, fam, k, v, t)
        assert n == red.n_rows
        coeff = np.exp(rng.normal(size=t.shape[0]))
        b = rng.uniform(-1, 1, size=(nrhs, n))
        mass = 0.0 if fam == ff.capi.FAMILY_H1 else 1.0
        reqs.append(dict(key=key, fam=fam, k=k, N=N, coeff=coeff, b=b, mass=mass, n=n))
    return reqs


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_served_solutions_match_oracle(ff, policy):
    srv = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    srv.set_precision(policy)
    reqs = make_requests(ff, srv)
    tickets = []
    for r in reqs:
        tk, nrhs, n = srv.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"], tol=1e-10)
        r["ticket"] = tk
        tickets.append((tk, nrhs, n))
    res = srv.run_until_done(tickets, max_steps=20000)
    for r in reqs:
        x, it, conv, rel = res[r["ticket"]]
        assert conv.all() and (rel <= 1e-10).all(), (r["fam"], r["k"], rel)
        _, _, red = oracle_system(ff, r["fam"], r["k"], r["N"], r["coeff"])
        for j in range(r["b"].shape[0]):
            xo, st = O.solve_sparse(red, r["b"][j], method="pcg", preconditioner="jacobi", tolerance=1e-13,
                                    max_iterations=100000, breakdown="relative")
            err = np.linalg.norm(x[j] - xo) / np.linalg.norm(xo)
            assert err <= 1e-8, (r["fam"], r["k"], j, err)


def test_alone_equals_in_mix(ff):
    alone = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    reqs = make_requests(ff, alone, seed=3)
    target = reqs[1]
    tk, nrhs, n = alone.submit(target["key"], target["fam"], target["k"], target["coeff"], target["b"],
                               mass=target["mass"])
    xa = alone.run_until_done([(tk, nrhs, n)], max_steps=20000)[tk][0]
    mix = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    reqs2 = make_requests(ff, mix, seed=3)
    order = [reqs2[3], reqs2[0], reqs2[1], reqs2[2]]
    tickets = []
    for i, r in enumerate(order):
        tk2 = mix.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"])
        tickets.append(tk2)
        if i == 1:
            mix.step()  # the target joins while others are running
    out = mix.run_until_done(tickets, max_steps=20000)
    xb = out[tickets[2][0]][0]
    assert np.array_equal(xa, xb)


def test_graph_replay_matches_direct_launches(ff):
    """Each step's chunk replayed from a HIP graph (re-captured when a request
    joins or leaves) gives bit-for-bit the results of launching the kernels
    one by one: same solutions, iteration counts and residuals, with the
    host's own synchronisation and polling between the replays."""
    outs = []
    for graph in (False, True):
        srv = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
        srv.use_graph(graph)
        reqs = make_requests(ff, srv, seed=5)
        tickets = []
        for i, r in enumerate(reqs):
            tickets.append(srv.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"]))
            if i in (1, 3):
                srv.step()  # joins while others run: the graph is captured again
                srv.step()
        outs.append(srv.run_until_done(tickets, max_steps=20000))
    a, b = outs
    assert a.keys() == b.keys()
    for t in a:
        xa, ia, ca, ra = a[t]
        xb, ib, cb, rb = b[t]
        assert np.array_equal(xa, xb) and np.array_equal(ia, ib)
        assert np.array_equal(ca, cb) and np.array_equal(ra, rb)
        assert ca.all()


def test_product_forms_match_plain(ff):
    """The batched SpMV with each item's column tile of p staged in LDS, and
    the row-split ones (2, 4 or 8 lanes per row, each lane a share of the
    right-hand sides, p.Ap products handed back to one lane per row through
    LDS), add every row's products in the order of the plain SELL-64 product
    and reduce them in its order, so servers forced to each form (wherever
    it applies), one forced to the plain product and one left to the
    autotuner return bit-for-bit the same solutions, iteration counts and
    residuals.  The autotuner records one choice per system type (pattern,
    right-hand sides, iteration precision).  CASES holds R = 1, 2, 4, 6 and
    8 (7 padded), so every split width meets a system it applies to."""
    outs, stats = [], []
    modes = (0, 1, 2, 4, 8, -1)
    for mode in modes:
        srv = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
        srv.set_tiling(mode)
        reqs = make_requests(ff, srv, seed=7)
        tickets = [srv.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"]) for r in reqs]
        outs.append(srv.run_until_done(tickets, max_steps=20000))
        stats.append(srv.product_forms())
        assert srv.tiling_stats() == stats[-1][:2]
    for other in outs[1:]:
        assert outs[0].keys() == other.keys()
        for t in outs[0]:
            for u, w in zip(outs[0][t], other[t]):
                assert np.array_equal(u, w)
    by = dict(zip(modes, stats))
    assert by[0][1:] == (0, 0)  # plain: neither tiled nor split
    assert by[1][1] >= 1 and by[1][2] == 0  # forced tiles fit somewhere
    assert by[2][2] >= 3 and by[4][2] >= 2 and by[8][2] >= 1  # R = 2, 4, 8 / R = 4, 8 / R = 8
    n, t, l = by[-1]
    assert n == len(CASES) and t + l <= n


def test_tight_tolerance_fp32_iteration(ff):
    """At a relative tolerance of 1e-13 the FP32 iterations' recurrence
    residual can reach the tolerance while the FP64 residual of x has not:
    such a right-hand side stops until the chunk's FP64 refresh, which then
    restarts it from D r (its direction missed the last update), and a
    request whose FP64 residual stalls moves to FP64 iterations.  Every
    right-hand side still ends converged on the FP64 residual, within 1e-10
    of the oracle's FP64 solution."""
    srv = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    srv.set_precision(0)
    reqs = make_requests(ff, srv, seed=3)
    tickets = []
    for r in reqs:
        tk, nrhs, n = srv.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"], tol=1e-13)
        r["ticket"] = tk
        tickets.append((tk, nrhs, n))
    res = srv.run_until_done(tickets, max_steps=50000)
    for r in reqs:
        x, it, conv, rel = res[r["ticket"]]
        assert conv.all() and (rel <= 1e-13).all(), (r["fam"], r["k"], rel)
        _, _, red = oracle_system(ff, r["fam"], r["k"], r["N"], r["coeff"])
        for j in range(r["b"].shape[0]):
            xo, st = O.solve_sparse(red, r["b"][j], method="pcg", preconditioner="jacobi", tolerance=1e-14,
                                    max_iterations=100000, breakdown="relative")
            err = np.linalg.norm(x[j] - xo) / np.linalg.norm(xo)
            assert err <= 1e-10, (r["fam"], r["k"], j, err)


def test_admission_by_memory(ff):
    v, t = O.unit_box_mesh(4)
    srv = ff.Server(budget_bytes=1 << 30, max_systems=8, chunk=8)
    n, nnz = srv.prepare(0, ff.capi.FAMILY_H1, 1, v, t)
    srv.submit(0, ff.capi.FAMILY_H1, 1, np.ones(t.shape[0]), np.ones((4, n)))
    srv.step()
    one = srv.stats()["used_bytes"]  # footprint of one admitted request
    assert one > nnz * 12 + n * 4 * 8
    small = ff.Server(budget_bytes=int(one * 1.5), max_systems=8, chunk=8)
    small.prepare(0, ff.capi.FAMILY_H1, 1, v, t)
    rng = np.random.default_rng(1)
    tickets = [small.submit(0, ff.capi.FAMILY_H1, 1, np.ones(t.shape[0]), rng.uniform(-1, 1, (4, n)))
               for _ in range(3)]
    seen = 0
    done = {}
    for _ in range(10000):
        a, q, f = small.step()
        assert a <= 1
        seen = max(seen, a)
        for tk in tickets:
            if tk[0] not in done:
                r = small.poll(*tk)
                if r is not None:
                    done[tk[0]] = r
        if len(done) == 3:
            break
    assert len(done) == 3 and seen == 1
    tiny = ff.Server(budget_bytes=1024, max_systems=8, chunk=8)
    tiny.prepare(0, ff.capi.FAMILY_H1, 1, v, t)
    with pytest.raises(ff.FFHError) as e:
        tiny.submit(0, ff.capi.FAMILY_H1, 1, np.ones(t.shape[0]), np.ones(n))
    assert e.value.code == ff.capi.FFH_ERR_ADMISSION


def test_wait_returns_the_request_result(ff):
    """ffh_server_wait steps the whole batch until the ticket finishes and
    returns what poll would; the result equals the request solved alone."""
    alone = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    reqs = make_requests(ff, alone, seed=5)
    target = reqs[2]
    tk = alone.submit(target["key"], target["fam"], target["k"], target["coeff"], target["b"],
                      mass=target["mass"])
    xa = alone.run_until_done([tk], max_steps=20000)[tk[0]][0]
    srv = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    reqs2 = make_requests(ff, srv, seed=5)
    tickets = [srv.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"]) for r in reqs2]
    assert srv.wait(*tickets[2], max_steps=0) is None  # nothing run yet
    x, it, conv, rel = srv.wait(*tickets[2], max_steps=20000)
    assert conv.all() and np.array_equal(x, xa)
    for tk2 in tickets[:2] + tickets[3:]:
        assert srv.wait(*tk2, max_steps=20000) is not None


def test_non_finite_request_fails_alone(ff):
    """A request whose right-hand side (or coefficient) holds a NaN or an
    infinity is refused at submission with -1 (FFH_ERR_INVALID, the
    reference C API's invalid-input code, include/fortfem.h:22-23) naming the
    offending entry; the server keeps serving, and the other requests'
    results are bit for bit what they are alone."""
    alone = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    reqs = make_requests(ff, alone, seed=7)
    ref = {}
    for i in (0, 2):
        r = reqs[i]
        tk = alone.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"])
        ref[i] = alone.run_until_done([tk], max_steps=20000)[tk[0]][0]
    mix = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    reqs2 = make_requests(ff, mix, seed=7)
    t0 = mix.submit(reqs2[0]["key"], reqs2[0]["fam"], reqs2[0]["k"], reqs2[0]["coeff"], reqs2[0]["b"],
                    mass=reqs2[0]["mass"])
    mix.step()
    bad = reqs2[1]
    b = bad["b"].copy()
    b[0, 3] = np.nan
    with pytest.raises(ff.FFHError) as e:
        mix.submit(bad["key"], bad["fam"], bad["k"], bad["coeff"], b, mass=bad["mass"])
    assert e.value.code == ff.capi.FFH_ERR_INVALID and "non-finite" in e.value.message
    c = bad["coeff"].copy()
    c[5] = np.inf
    with pytest.raises(ff.FFHError) as e:
        mix.submit(bad["key"], bad["fam"], bad["k"], c, bad["b"], mass=bad["mass"])
    assert e.value.code == ff.capi.FFH_ERR_INVALID
    t2 = mix.submit(reqs2[2]["key"], reqs2[2]["fam"], reqs2[2]["k"], reqs2[2]["coeff"], reqs2[2]["b"],
                    mass=reqs2[2]["mass"])
    out = mix.run_until_done([t0, t2], max_steps=20000)
    assert np.array_equal(out[t0[0]][0], ref[0])
    assert np.array_equal(out[t2[0]][0], ref[2])


@pytest.mark.parametrize("policy", [0, 1])
def test_iteration_cap_reports_true_residual(ff, policy):
    """A right-hand side stopped by its iteration cap in mid-chunk (cap 13,
    chunks of 8) reports iterations = 13, converged = false, and the relative
    residual ||b - A x|| / ||b|| of the x it returns (the FP64 refresh after
    the cap measures it), checked against the oracle's matrix."""
    srv = ff.Server(budget_bytes=1 << 30, max_systems=16, chunk=8)
    srv.set_precision(policy)
    reqs = make_requests(ff, srv, seed=9)
    r = reqs[0]
    tk = srv.submit(r["key"], r["fam"], r["k"], r["coeff"], r["b"], mass=r["mass"], tol=1e-14,
                    max_iterations=13)
    x, it, conv, rel = srv.run_until_done([tk], max_steps=100)[tk[0]]
    assert (it == 13).all() and not conv.any()
    _, _, red = oracle_system(ff, r["fam"], r["k"], r["N"], r["coeff"])
    for j in range(r["b"].shape[0]):
        res = np.linalg.norm(r["b"][j] - red.matvec(x[j])) / np.linalg.norm(r["b"][j])
        assert abs(rel[j] - res) <= 1e-9 * res, (j, rel[j], res)
