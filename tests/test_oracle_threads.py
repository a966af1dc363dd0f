"""The oracle's multi-threaded stages (the all-cores CPU baseline) against
its serial restatement on the same inputs: FEM triplets, the converted
BCOO, the MAS tables, Galerkin bank matrices and inverses bit-identical; the
vertex gradient and the PCG iterate to round-off (the threaded SpMV sums a
row's blocks in another order and the dots are blocked), with the same
iteration count."""
import numpy as np
import pytest


def _step(O, sc, pairs, threads):
    with O.threads(threads):
        shape = O.kinetic_gradient(sc.x, sc.x_tilde, sc.mass)
        grad, vals, rows, cols = O.fem_assemble(sc.x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt, sc.btype, 0,
                                                shape.copy())
        mv, mr, mc = O.mass_triplets(sc.mass)
        kr, kc, kv = O.convert(np.concatenate([rows, mr]), np.concatenate([cols, mc]), np.concatenate([vals, mv]))
        ns, nn, nl = O.vertex_neighbors(sc.n, sc.tets)
        pm, rm = O.bank_map(sc.part, sc.bank)
        mas = O.Mas(sc.n, pm, rm, ns, nn, nl, sc.bank)
        mas.build(kr, kc, kv, pairs=pairs, fp32_store=False)
        b = O.fem_rhs(np.zeros_like(grad), grad, sc.btype).reshape(-1)
        banks = [mas.bank_matrix(i) for i in (0, mas.total // sc.bank - 1)]
        out = {}
        for prec in (1, 2):
            out[prec] = O.pcg(kr, kc, kv, sc.n, b, precond=prec, mas=mas, tol=1e-8, rule=1, max_iter=100000)
        # both thread counts on the same right-hand side: the preconditioner
        # applied once (PCG budget of two iterations)
        x2 = O.pcg(kr, kc, kv, sc.n, np.linspace(-1, 1, 3 * sc.n), precond=2, mas=mas, tol=1e-30, rule=0,
                   max_iter=2)[0]
        return dict(grad=grad, vals=vals, rows=rows, cols=cols, bcoo=(kr, kc, kv), tables=mas.tables(),
                    banks=banks, pcg=out, x2=x2)


@pytest.mark.parametrize("name", ["cube10", "twobox8"])
@pytest.mark.parametrize("threads", [3, 8])
def test_threaded_oracle_matches_serial(oracle, name, threads):
    from sgipc import scene
    if name == "cube10":
        sc, pairs = scene.cube_scene(10, seed=1), None
    else:
        sc = scene.two_box_scene(8, seed=2)
        pt, ee = scene.facing_contact_candidates(sc)
        pairs = oracle.classify(pt, ee, sc.x, sc.x_rest, sc.dhat)[0]
    a = _step(oracle, sc, pairs, 1)
    b = _step(oracle, sc, pairs, threads)
    assert oracle.get_threads() == 1
    assert np.abs(a["grad"] - b["grad"]).max() <= 1e-14 * np.abs(a["grad"]).max()
    for k in ("vals", "rows", "cols"):
        assert np.array_equal(a[k], b[k]), k
    for u, v in zip(a["bcoo"], b["bcoo"]):
        assert np.array_equal(u, v)
    for k in a["tables"]:
        assert np.array_equal(a["tables"][k], b["tables"][k]), k
    for (A1, I1), (A2, I2) in zip(a["banks"], b["banks"]):
        assert np.array_equal(A1, A2) and np.array_equal(I1, I2)
    for prec in (1, 2):
        x1, it1, r1 = a["pcg"][prec]
        x2, it2, r2 = b["pcg"][prec]
        assert it1 == it2
        assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x1)
        assert r2 <= 1e-8
    assert np.linalg.norm(a["x2"] - b["x2"]) <= 1e-12 * np.linalg.norm(a["x2"])
