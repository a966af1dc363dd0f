"""GPU parity of contact classification, IPC barrier assembly and the
contact Newton step (MAS with contact connectivity) against the oracle."""
import numpy as np
import pytest

from conftest import dense_from_bcoo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def twobox8():
    from sgipc import scene
    sc = scene.two_box_scene(8, seed=2)
    pt, ee = scene.facing_contact_candidates(sc)
    return sc, pt, ee


def oracle_contact_system(O, sc, pairs, mi, counts):
    fb, bv, br, bc = O.barrier_assemble(pairs, mi, counts, sc.x, sc.x_rest, sc.dhat, sc.kappa, sc.btype)
    shape = O.kinetic_gradient(sc.x, sc.x_tilde, sc.mass)
    grad, vals, rows, cols = O.fem_assemble(sc.x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt, sc.btype, 0,
                                            shape.copy())
    mv, mr, mc = O.mass_triplets(sc.mass)
    R = np.concatenate([br, rows, mr]); Cc = np.concatenate([bc, cols, mc]); V = np.concatenate([bv, vals, mv])
    b = O.fem_rhs(fb.reshape(-1, 3), grad, sc.btype).reshape(-1)
    kr, kc, kv = O.convert(R, Cc, V)
    return dict(fb=fb, trip=(R, Cc, V), b=b, bcoo=(kr, kc, kv))


def test_classify_bitexact(torch_cuda, oracle, twobox8):
    from sgipc.system import NewtonSystem
    sc, pt, ee = twobox8
    print("case: classify twobox8 candidates", len(pt), len(ee), flush=True)
    sysm = NewtonSystem(sc, "cuda:0")
    npair = sysm.setup_contacts(pt, ee)
    po, mo, co = oracle.classify(pt, ee, sc.x, sc.x_rest, sc.dhat)
    assert npair == len(po) and npair > 100
    assert np.array_equal(sysm.pairs.cpu().numpy().reshape(-1, 4), po)
    assert np.array_equal(sysm.mat_index.cpu().numpy(), mo)
    assert np.array_equal(sysm.counts, co)


def test_barrier_assembly_matches_oracle(torch_cuda, oracle, twobox8):
    from sgipc.system import NewtonSystem
    sc, pt, ee = twobox8
    print("case: barrier assembly twobox8", flush=True)
    sysm = NewtonSystem(sc, "cuda:0")
    sysm.setup_contacts(pt, ee)
    sysm.assemble()
    pairs = sysm.pairs.cpu().numpy().reshape(-1, 4)
    ref = oracle_contact_system(oracle, sc, pairs, sysm.mat_index.cpu().numpy(), sysm.counts)
    R, Cc, V = sysm.triplets()
    Ro, Co, Vo = ref["trip"]
    assert np.array_equal(R, Ro) and np.array_equal(Cc, Co)
    nct = sysm.n_contact_triplets
    sc_b = np.abs(Vo[:nct]).max()
    assert np.abs(V[:nct] - Vo[:nct]).max() <= 1e-10 * sc_b
    fb = sysm.fb.cpu().numpy()
    assert np.abs(fb - ref["fb"]).max() <= 1e-10 * np.abs(ref["fb"]).max()
    b = sysm.b.cpu().numpy()
    assert np.abs(b - ref["b"]).max() <= 1e-10 * np.abs(ref["b"]).max()


def test_contact_newton_step_matches_oracle(torch_cuda, oracle, twobox8):
    from sgipc.system import NewtonSystem
    sc, pt, ee = twobox8
    print("case: contact newton step twobox8", flush=True)
    sysm = NewtonSystem(sc, "cuda:0")
    sysm.setup_contacts(pt, ee)
    sysm.assemble(); K = sysm.convert(); sysm.build_precond()
    pairs = sysm.pairs.cpu().numpy().reshape(-1, 4)
    ref = oracle_contact_system(oracle, sc, pairs, sysm.mat_index.cpu().numpy(), sysm.counts)
    kr, kc, kv = ref["bcoo"]
    gr, gc, gv = sysm.bcoo()
    assert K == len(kr) and np.array_equal(gr, kr) and np.array_equal(gc, kc)
    assert np.abs(gv - kv).max() <= 1e-10 * np.abs(kv).max()
    mas = oracle.Mas(sc.n, sysm.part_map, sysm.real_map, sysm.ns, sysm.nn, sysm.nl, sc.bank)
    mas.build(kr, kc, kv, pairs=pairs)
    t, to = sysm.mas_tables(), mas.tables()
    assert np.array_equal(t["level_size"], to["level_size"])
    assert np.array_equal(t["coarse_space"], to["coarse_space"])
    assert np.array_equal(t["fine_mask"], to["fine_mask"])
    st = sysm.solve(tol=1e-4, stop_rule=0)
    xo, ito, _ = oracle.pcg(kr, kc, kv, sc.n, ref["b"], precond=2, mas=mas, tol=1e-4, rule=0)
    assert st.iterations == ito
    x = sysm.dx.cpu().numpy()
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    st = sysm.solve(tol=1e-6, stop_rule=1, max_iter=20000)
    A = dense_from_bcoo(kr, kc, kv, sc.n)
    x = sysm.dx.cpu().numpy()
    assert st.converged and np.linalg.norm(A @ x - ref["b"]) <= 1.5e-6 * np.linalg.norm(ref["b"])


def test_convert_plan_reuse_is_bitexact(torch_cuda, oracle, twobox8):
    """Convert keeps the sort of the FEM + mass segment between calls; the
    BCOO must equal the oracle's full sort bit for bit when the plan is
    built, reused, and reused after the contact prefix changed length."""
    from sgipc.system import NewtonSystem
    sc, pt, ee = twobox8
    sysm = NewtonSystem(sc, "cuda:0")
    for step, (p, e) in enumerate([(pt, ee), (pt, ee), (pt[: len(pt) // 2], ee[: len(ee) // 3]), (pt, ee)]):
        print("case: convert plan step %d (%d PT, %d EE candidates)" % (step, len(p), len(e)), flush=True)
        sysm.setup_contacts(p, e)
        sysm.assemble()
        R, Cc, V = sysm.triplets()
        K = sysm.convert()
        kr, kc, kv = oracle.convert(R, Cc, V)
        gr, gc, gv = sysm.bcoo()
        assert K == len(kr)
        assert np.array_equal(gr, kr) and np.array_equal(gc, kc)
        assert np.array_equal(gv, kv)
