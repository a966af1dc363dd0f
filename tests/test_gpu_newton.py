"""GPU parity of the HIP Newton-step path against the CPU oracle, through the
C ABI (libsgipc.so).  Inputs are the seeded synthetic cube scenes."""
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


def make_system(sc, precond=None):
    from sgipc import capi
    from sgipc.system import NewtonSystem
    return NewtonSystem(sc, "cuda:0", precond=capi.PRECOND_MAS if precond is None else precond)


@pytest.mark.parametrize("name", ["cube4", "cube10"])
def test_fem_assembly_matches_oracle(torch_cuda, oracle, request, name):
    sc = request.getfixturevalue(name)
    print("case: fem assembly", name, "n", sc.n, "T", sc.T, flush=True)
    ref = oracle_system(oracle, sc)
    sys_ = make_system(sc)
    sys_.assemble()
    R, Cc, V = sys_.triplets()
    Ro, Co, Vo = ref["trip"]
    assert np.array_equal(R, Ro) and np.array_equal(Cc, Co)          # bit-exact indices
    scale = np.abs(Vo).max()
    assert np.abs(V - Vo).max() <= 1e-12 * scale                      # FP64 values
    g = sys_.shape.cpu().numpy().reshape(-1, 3)
    assert np.abs(g - ref["shape"]).max() <= 1e-12 * np.abs(ref["shape"]).max()
    b = sys_.b.cpu().numpy()
    assert np.abs(b - ref["b"]).max() <= 1e-12 * np.abs(ref["b"]).max()


@pytest.mark.parametrize("name", ["cube4", "cube10"])
def test_convert_matches_oracle(torch_cuda, oracle, request, name):
    sc = request.getfixturevalue(name)
    print("case: convert", name, flush=True)
    ref = oracle_system(oracle, sc)
    sys_ = make_system(sc)
    sys_.assemble()
    K = sys_.convert()
    kr, kc, kv = sys_.bcoo()
    orow, ocol, oval = ref["bcoo"]
    assert K == len(orow)
    assert np.array_equal(kr, orow) and np.array_equal(kc, ocol)     # bit-exact BCOO index set and order
    assert np.abs(kv - oval).max() <= 1e-12 * np.abs(oval).max()


def test_convert_on_oracle_triplets_is_bitexact(torch_cuda, oracle, cube4):
    """Fed the oracle's own triplets, the GPU convert reproduces the oracle's
    BCOO bit for bit (same per-block summation order)."""
    import torch
    print("case: convert bit-exact on oracle triplets", flush=True)
    ref = oracle_system(oracle, cube4)
    sys_ = make_system(cube4)
    R, Cc, V = ref["trip"]
    nt = len(R)
    sys_.tr[:nt] = torch.tensor(R, dtype=torch.int32, device=sys_.dev)
    sys_.tc[:nt] = torch.tensor(Cc, dtype=torch.int32, device=sys_.dev)
    sys_.tv[:9 * nt] = torch.tensor(V.reshape(-1), dtype=torch.float64, device=sys_.dev)
    sys_.nt = nt
    sys_.convert()
    kr, kc, kv = sys_.bcoo()
    orow, ocol, oval = ref["bcoo"]
    assert np.array_equal(kr, orow) and np.array_equal(kc, ocol)
    assert np.array_equal(kv, oval)


def test_spmv_matches_dense(torch_cuda, oracle, cube10):
    print("case: spmv cube10", flush=True)
    sys_ = make_system(cube10)
    sys_.assemble(); sys_.convert()
    kr, kc, kv = sys_.bcoo()
    A = dense_from_bcoo(kr, kc, kv, cube10.n)
    rng = np.random.default_rng(3)
    x = rng.normal(size=3 * cube10.n); y0 = rng.normal(size=3 * cube10.n)
    y = sys_.spmv(x, a=0.7, b=-0.3, y=y0)
    ref = 0.7 * A @ x - 0.3 * y0
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_mas_tables_bitexact(torch_cuda, oracle, cube10):
    print("case: mas tables cube10", flush=True)
    ref = oracle_system(oracle, cube10)
    sys_ = make_system(cube10)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    t = sys_.mas_tables()
    kr, kc, kv = ref["bcoo"]
    mas = oracle.Mas(cube10.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, cube10.bank)
    tot = mas.build(kr, kc, kv)
    to = mas.tables()
    assert t["levelnum"] == mas.levelnum and t["total"] == tot
    assert np.array_equal(t["level_size"], to["level_size"])
    assert np.array_equal(t["coarse_space"], to["coarse_space"])
    assert np.array_equal(t["coarse_table"][:, :mas.levelnum - 1], to["coarse_table"][:, :mas.levelnum - 1])
    assert np.array_equal(t["fine_mask"], to["fine_mask"])
    assert np.array_equal(t["prefix0"], to["prefix0"])


@pytest.mark.parametrize("fp32", [False, True])
def test_mas_apply_matches_oracle(torch_cuda, oracle, cube10, fp32):
    from sgipc import capi
    print("case: mas apply fp32=%s" % fp32, flush=True)
    ref = oracle_system(oracle, cube10)
    sys_ = make_system(cube10, capi.PRECOND_MAS_FP32 if fp32 else capi.PRECOND_MAS)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    kr, kc, kv = ref["bcoo"]
    mas = oracle.Mas(cube10.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, cube10.bank)
    mas.build(kr, kc, kv)
    r = np.random.default_rng(5).normal(size=3 * cube10.n)
    z = sys_.precond_apply(r)
    zo = mas.apply(r, fp32=False)
    tol = 1e-5 if fp32 else 1e-11
    assert np.abs(z - zo).max() <= tol * np.abs(zo).max()
    b0 = sys_.bank_inverse(0)
    _, Ainv = mas.bank_matrix(0)
    assert np.abs(b0 - Ainv).max() <= 1e-10 * np.abs(Ainv).max()


@pytest.mark.parametrize("precond", [0, 1, 2])
def test_pcg_matches_oracle(torch_cuda, oracle, cube10, precond):
    print("case: pcg precond=%d" % precond, flush=True)
    ref = oracle_system(oracle, cube10)
    sys_ = make_system(cube10, precond)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    kr, kc, kv = ref["bcoo"]
    mas = None
    if precond == 2:
        mas = oracle.Mas(cube10.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, cube10.bank)
        mas.build(kr, kc, kv)
    # reference stopping rule: identical iteration count, direction within 1e-10
    st = sys_.solve(tol=1e-4, stop_rule=0)
    xo, ito, _ = oracle.pcg(kr, kc, kv, cube10.n, ref["b"], precond=precond, mas=mas, tol=1e-4, rule=0)
    x = sys_.dx.cpu().numpy()
    assert st.iterations == ito
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    # true relative residual 1e-6 against a direct solve
    st = sys_.solve(tol=1e-6, stop_rule=1, max_iter=5000)
    x = sys_.dx.cpu().numpy()
    A = dense_from_bcoo(kr, kc, kv, cube10.n)
    xs = np.linalg.solve(A, ref["b"])
    assert st.converged and st.rel_residual <= 1.5e-6
    assert np.linalg.norm(A @ x - ref["b"]) <= 1.5e-6 * np.linalg.norm(ref["b"])
    assert np.linalg.norm(x - xs) <= 1e-3 * np.linalg.norm(xs)


@pytest.mark.parametrize("budget", [6, 7, 1, 2])
def test_pcg_iteration_budget_matches_oracle(torch_cuda, oracle, cube10, budget):
    # an unreachable tolerance: both stop at the budget with the same iterate
    # (graphs hold two iterations: odd budgets stop inside one; a budget of
    # one runs no iteration)
    print("case: pcg budget max_iter=%d" % budget, flush=True)
    ref = oracle_system(oracle, cube10)
    sys_ = make_system(cube10)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    kr, kc, kv = ref["bcoo"]
    mas = oracle.Mas(cube10.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, cube10.bank)
    mas.build(kr, kc, kv)
    st = sys_.solve(tol=1e-30, stop_rule=0, max_iter=budget)
    xo, ito, _ = oracle.pcg(kr, kc, kv, cube10.n, ref["b"], precond=2, mas=mas, tol=1e-30, rule=0, max_iter=budget)
    assert st.iterations == ito and not st.converged
    x = sys_.dx.cpu().numpy()
    assert np.linalg.norm(x - xo) <= 1e-10 * max(np.linalg.norm(xo), 1e-300)


def test_pcg_graph_replays_match_eager(torch_cuda, cube10):
    # captured iteration graphs (profiling off) against the eager launch
    # sequence (profiling on), replayed over several solves with other
    # device work in between; results must be bit-identical
    import torch
    from sgipc import capi
    print("case: pcg graph replay vs eager", flush=True)
    sys_ = make_system(cube10)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    b0 = sys_.b.clone()
    xs = []
    for rep in range(3):
        sys_.b.copy_(b0 * (1.0 + rep))
        st = sys_.solve(tol=1e-6, stop_rule=1, max_iter=5000)
        xs.append((st.iterations, sys_.dx.clone()))
        torch.cuda.synchronize()
        _ = (sys_.b * 2.0).sum().item()  # unrelated eager work between replays
    sys_.ctx.call("sgipc_profile_enable", 1)
    try:
        for rep in range(3):
            sys_.b.copy_(b0 * (1.0 + rep))
            st = sys_.solve(tol=1e-6, stop_rule=1, max_iter=5000)
            it, xg = xs[rep]
            assert st.iterations == it
            assert torch.equal(sys_.dx, xg)
    finally:
        sys_.ctx.call("sgipc_profile_enable", 0)
    sys_.b.copy_(b0)


@pytest.mark.parametrize("mode", ["mas_fp32", "mixed_mas", "mixed_diag", "mixed_none"])
def test_reduced_precision_solves(torch_cuda, oracle, cube10, mode):
    # FP32 preconditioner storage and the mixed FP32/FP64 solve reach the
    # same FP64 true relative residual as the FP64 solve
    from sgipc import capi
    print("case: reduced precision %s" % mode, flush=True)
    ref = oracle_system(oracle, cube10)
    kind = {"mas_fp32": capi.PRECOND_MAS_FP32, "mixed_mas": capi.PRECOND_MAS, "mixed_diag": capi.PRECOND_DIAG,
            "mixed_none": capi.PRECOND_NONE}[mode]
    prec = capi.PRECISION_FP64 if mode == "mas_fp32" else capi.PRECISION_MIXED
    sys_ = make_system(cube10, kind)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    kr, kc, kv = ref["bcoo"]
    A = dense_from_bcoo(kr, kc, kv, cube10.n)
    xs = np.linalg.solve(A, ref["b"])
    st64 = sys_.solve(tol=1e-6, stop_rule=1, max_iter=20000)
    st = sys_.solve(tol=1e-6, stop_rule=1, max_iter=20000, precision=prec)
    x = sys_.dx.cpu().numpy()
    assert st.converged and st.rel_residual <= 1e-6
    assert np.linalg.norm(A @ x - ref["b"]) <= 1.5e-6 * np.linalg.norm(ref["b"])
    assert np.linalg.norm(x - xs) <= 1e-3 * np.linalg.norm(xs)
    # reduced precision costs at most a modest number of extra iterations
    assert st.iterations <= 1.5 * st64.iterations + 10
    if prec == capi.PRECISION_MIXED:
        assert st.refinements >= 1


def test_convert_long_rows_bitexact(torch_cuda, oracle, cube4):
    """Rows of every length class (one wavefront up to 256 triplets, one
    workgroup's bitonic sort up to 4096, and the device radix-sort path above
    that, which the rows of affine bodies reach), at and either side of each
    class boundary, convert bit for bit as the oracle does."""
    import torch
    print("case: convert long rows", flush=True)
    rng = np.random.default_rng(21)
    n = cube4.n
    lens = {0: 9000, 1: 5000, 2: 3000, 3: 200, 4: 257, 5: 4096, 6: 256}
    R, Cc = [], []
    for r, L in lens.items():
        R.append(np.full(L, r)); Cc.append(rng.integers(r, n, L))
    for r in range(len(lens), n):
        L = int(rng.integers(0, 12))
        R.append(np.full(L, r)); Cc.append(rng.integers(r, n, L))
    R = np.concatenate(R).astype(np.int32); Cc = np.concatenate(Cc).astype(np.int32)
    perm = rng.permutation(len(R))                     # rows interleaved in the store
    R, Cc = R[perm], Cc[perm]
    V = rng.normal(size=(len(R), 9))
    sys_ = make_system(cube4)
    nt = len(R)
    sys_._set_capacity(nt)
    sys_.tr[:nt] = torch.tensor(R, dtype=torch.int32, device=sys_.dev)
    sys_.tc[:nt] = torch.tensor(Cc, dtype=torch.int32, device=sys_.dev)
    sys_.tv[:9 * nt] = torch.tensor(V.reshape(-1), dtype=torch.float64, device=sys_.dev)
    sys_.nt = nt
    K = sys_.convert()
    kr, kc, kv = sys_.bcoo()
    orow, ocol, oval = oracle.convert(R, Cc, V)
    assert K == len(orow)
    assert np.array_equal(kr, orow) and np.array_equal(kc, ocol)
    assert np.array_equal(kv, oval)


@pytest.mark.parametrize("mode", ["fp64", "fp32_apply", "mixed", "diag"])
def test_zero_rhs_returns_zero(torch_cuda, cube10, mode):
    """A zero right-hand side stops the loop before its first iteration (the
    device sees rz0 = 0 itself): x = 0, no iterations, converged, and a
    following solve of a real right-hand side is unaffected."""
    from sgipc import capi
    print("case: zero rhs", mode, flush=True)
    kind = {"fp64": capi.PRECOND_MAS, "fp32_apply": capi.PRECOND_MAS_FP32, "mixed": capi.PRECOND_MAS,
            "diag": capi.PRECOND_DIAG}[mode]
    prec = capi.PRECISION_MIXED if mode == "mixed" else capi.PRECISION_FP64
    sys_ = make_system(cube10, kind)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    b0 = sys_.b.clone()
    ref = sys_.solve(tol=1e-6, stop_rule=1, max_iter=5000, precision=prec)
    x_ref = sys_.dx.clone()
    sys_.dx.fill_(1.0)
    sys_.b.zero_()
    st = sys_.solve(tol=1e-6, stop_rule=1, max_iter=5000, precision=prec)
    assert st.iterations == 0 and st.converged and st.rz0 == 0.0
    assert not torch_cuda.any(sys_.dx != 0)
    sys_.b.copy_(b0)
    st2 = sys_.solve(tol=1e-6, stop_rule=1, max_iter=5000, precision=prec)
    assert st2.converged and st2.iterations == ref.iterations
    assert torch_cuda.equal(sys_.dx, x_ref)
