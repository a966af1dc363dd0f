"""GPU parity of the MAS preconditioner with 32-vertex banks (96x96 dense
bank inverses, the reference's BANKSIZE 32 build, MASPreconditioner.cu
__inverse6_P96x96 :630-726) against the CPU oracle, through the C ABI.
Scene: the seeded 10^3-cell cube partitioned into 2x4x4 vertex blocks."""
import numpy as np
import pytest

from conftest import oracle_system, dense_from_bcoo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _system(sc, precond):
    from sgipc.system import NewtonSystem
    s = NewtonSystem(sc, "cuda:0", precond=precond)
    s.assemble(); s.convert(); s.build_precond()
    return s


def _oracle_mas(oracle, sc, sys_, bcoo):
    mas = oracle.Mas(sc.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, sc.bank)
    mas.build(*bcoo)
    return mas


def test_bank32_tables_bitexact(torch_cuda, oracle, cube10_b32):
    from sgipc import capi
    sc = cube10_b32
    print("case: bank32 tables cube10", flush=True)
    ref = oracle_system(oracle, sc)
    sys_ = _system(sc, capi.PRECOND_MAS)
    mas = _oracle_mas(oracle, sc, sys_, ref["bcoo"])
    t, to = sys_.mas_tables(), mas.tables()
    assert t["levelnum"] == mas.levelnum >= 3
    assert np.array_equal(t["level_size"], to["level_size"])
    assert np.array_equal(t["coarse_space"], to["coarse_space"])
    assert np.array_equal(t["coarse_table"][:, :mas.levelnum - 1], to["coarse_table"][:, :mas.levelnum - 1])
    assert np.array_equal(t["fine_mask"], to["fine_mask"])


@pytest.mark.parametrize("fp32", [False, True])
def test_bank32_apply_and_inverse(torch_cuda, oracle, cube10_b32, fp32):
    from sgipc import capi
    sc = cube10_b32
    print("case: bank32 apply fp32=%s" % fp32, flush=True)
    ref = oracle_system(oracle, sc)
    sys_ = _system(sc, capi.PRECOND_MAS_FP32 if fp32 else capi.PRECOND_MAS)
    mas = _oracle_mas(oracle, sc, sys_, ref["bcoo"])
    r = np.random.default_rng(11).normal(size=3 * sc.n)
    z, zo = sys_.precond_apply(r), mas.apply(r, fp32=False)
    tol = 1e-5 if fp32 else 1e-11          # FP32 inverses and FP32 apply vs the FP64 oracle
    assert np.abs(z - zo).max() <= tol * np.abs(zo).max()
    for b in (0, 7):
        inv = sys_.bank_inverse(b)          # 96 x 96, from the FP64 Gauss-Jordan
        A, Ainv = mas.bank_matrix(b)
        assert inv.shape == (96, 96)
        assert np.abs(inv - Ainv).max() <= 1e-10 * np.abs(Ainv).max()
        assert np.abs(inv @ A - np.eye(96)).max() <= 1e-9


def test_bank32_pcg_matches_oracle(torch_cuda, oracle, cube10_b32):
    from sgipc import capi
    sc = cube10_b32
    print("case: bank32 pcg", flush=True)
    ref = oracle_system(oracle, sc)
    sys_ = _system(sc, capi.PRECOND_MAS)
    kr, kc, kv = ref["bcoo"]
    mas = _oracle_mas(oracle, sc, sys_, ref["bcoo"])
    st = sys_.solve(tol=1e-4, stop_rule=0)
    xo, ito, _ = oracle.pcg(kr, kc, kv, sc.n, ref["b"], precond=2, mas=mas, tol=1e-4, rule=0)
    x = sys_.dx.cpu().numpy()
    assert st.iterations == ito
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    A = dense_from_bcoo(kr, kc, kv, sc.n)
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        sys2 = _system(sc, capi.PRECOND_MAS_FP32 if prec == capi.PRECISION_MIXED else capi.PRECOND_MAS)
        st = sys2.solve(tol=1e-6, stop_rule=1, precision=prec, max_iter=5000)
        x = sys2.dx.cpu().numpy()
        assert st.converged and st.rel_residual <= 1e-6
        assert np.linalg.norm(A @ x - ref["b"]) <= 1.5e-6 * np.linalg.norm(ref["b"])
