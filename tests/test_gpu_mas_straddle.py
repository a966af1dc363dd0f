"""MAS restriction fused into the PCG update, on a partition whose level-0
banks split into many components of varying count.

Level-1 nodes are the components of the level-0 banks, numbered bank by
bank, so a level-0 bank whose components run over a multiple of the bank
size has coarse nodes in two level-1 banks (it "straddles" them) and is
visited by the two workgroups that own those level-1 banks.  Here every
level-0 bank holds 10-16 vertices drawn at random from the mesh, so most
vertices are singleton components, straddling banks are everywhere and
there are more level-1 banks than the fused update launches workgroups for
them (its grid-stride loop runs).  PCG iterates after a fixed budget depend
on the restriction of each updated residual, so they are compared with the
oracle's PCG (its own MAS apply of the same residual) at 1e-10.
"""
import numpy as np
import pytest

from conftest import oracle_system

pytestmark = pytest.mark.gpu

# workgroups the fused x/r update launches for level-1 banks (one partial
# slot each; the library's partial-sum width)
FUSED_UPDATE_WORKGROUPS = 1024


def scrambled_scene(cells=26, seed=11, min_size=10, bank=16):
    """cube_scene with its vertices regrouped into parts of min_size..bank
    randomly drawn vertices (numbering part-contiguous again)."""
    from sgipc import scene
    sc = scene.cube_scene(cells, seed=1)
    rng = np.random.default_rng(seed)
    n = sc.n
    order = rng.permutation(n)
    sizes = []
    tot = 0
    while tot < n:
        s = int(rng.integers(min_size, bank + 1))
        s = min(s, n - tot)
        sizes.append(s)
        tot += s
    part_of = np.empty(n, np.int32)
    part_of[order] = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    perm, inv, part_sorted = scene.renumber_by_part(part_of)
    tets = inv[sc.tets.astype(np.int64)].astype(np.uint32)
    return scene.Scene("scrambled%d" % cells, sc.x_rest[perm], sc.x[perm], sc.x_tilde[perm], tets,
                       part_sorted.astype(np.int32), sc.btype[perm], sc.DmInv, sc.vol, sc.mass[perm], sc.mu,
                       sc.lam, sc.dt, bank=bank, meta=dict(sc.meta, scrambled_seed=seed))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def scrambled():
    return scrambled_scene()


@pytest.fixture(scope="module")
def scrambled_system(torch_cuda, oracle, scrambled):
    from sgipc import capi
    from sgipc.system import NewtonSystem
    ref = oracle_system(oracle, scrambled)
    sys_ = NewtonSystem(scrambled, "cuda:0", precond=capi.PRECOND_MAS)
    sys_.assemble(); sys_.convert(); sys_.build_precond()
    kr, kc, kv = ref["bcoo"]
    mas = oracle.Mas(scrambled.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, scrambled.bank)
    mas.build(kr, kc, kv)
    return sys_, ref, mas


def test_scrambled_partition_exercises_straddling_banks(scrambled_system):
    sys_, _, _ = scrambled_system
    t = sys_.mas_tables()
    B = sys_.sc.bank
    ls = t["level_size"]
    n1 = int(ls[1, 0])
    nb1 = (n1 + B - 1) // B
    print("case: scrambled partition levels", t["levelnum"], "level-1 nodes", n1, "level-1 banks", nb1, flush=True)
    assert nb1 > FUSED_UPDATE_WORKGROUPS           # grid-stride loop over level-1 banks runs
    # level-0 banks whose coarse nodes fall in two level-1 banks
    p0 = t["prefix0"].astype(np.int64)
    psum = np.concatenate([[0], np.cumsum(p0)[:-1]])
    nz = p0 > 0
    straddle = ((psum[nz] // B) != ((psum[nz] + p0[nz] - 1) // B)).sum()
    assert straddle > 100


@pytest.mark.parametrize("budget", [2, 3, 7])
def test_pcg_budget_matches_oracle_on_straddling_banks(scrambled_system, oracle, budget):
    sys_, ref, mas = scrambled_system
    print("case: scrambled pcg budget", budget, flush=True)
    kr, kc, kv = ref["bcoo"]
    st = sys_.solve(tol=1e-30, stop_rule=0, max_iter=budget)
    xo, ito, _ = oracle.pcg(kr, kc, kv, sys_.sc.n, ref["b"], precond=2, mas=mas, tol=1e-30, rule=0,
                            max_iter=budget)
    assert st.iterations == ito
    x = sys_.dx.cpu().numpy()
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


def test_solves_repeat_bitwise_on_straddling_banks(scrambled_system):
    from sgipc import capi
    sys_, _, _ = scrambled_system
    print("case: scrambled repeat", flush=True)
    for prec in (capi.PRECISION_FP64, capi.PRECISION_MIXED):
        xs = []
        for _ in range(3):
            sys_.solve(tol=1e-6, stop_rule=capi.STOP_TRUE_REL, precision=prec, max_iter=2000)
            xs.append(sys_.dx.cpu().numpy().copy())
        assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])
