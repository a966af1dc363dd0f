"""The ordered (deterministic) MAS Galerkin projection: the same
preconditioner as the oracle and bitwise identical from build to build and
solve to solve, for 16- and 32-vertex banks and a contact system."""
import numpy as np
import pytest

from conftest import oracle_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _det_system(sc, pt=None, ee=None):
    from sgipc import capi
    from sgipc.system import NewtonSystem
    s = NewtonSystem(sc, "cuda:0", precond=capi.PRECOND_MAS)
    s.ctx.call("sgipc_set_deterministic", 1)
    if pt is not None:
        s.setup_contacts(pt, ee)
    s.assemble(); s.convert()
    return s


@pytest.mark.parametrize("name", ["cube10", "cube10_b32"])
def test_ordered_projection_matches_oracle_and_repeats(torch_cuda, oracle, request, name):
    sc = request.getfixturevalue(name)
    print("case: ordered projection", name, flush=True)
    ref = oracle_system(oracle, sc)
    s = _det_system(sc)
    s.build_precond()
    mas = oracle.Mas(sc.n, s.part_map, s.real_map, s.ns, s.nn, s.nl, sc.bank)
    mas.build(*ref["bcoo"])
    r = np.random.default_rng(2).normal(size=3 * sc.n)
    z1 = s.precond_apply(r)
    zo = mas.apply(r)
    assert np.abs(z1 - zo).max() <= 1e-11 * np.abs(zo).max()
    inv1 = [s.bank_inverse(b) for b in (0, 3)]
    for _ in range(2):
        s.build_precond()
        assert np.array_equal(s.precond_apply(r), z1)                   # bitwise, build after build
        assert all(np.array_equal(s.bank_inverse(b), i) for b, i in zip((0, 3), inv1))


def test_ordered_projection_contact_solves_repeat(torch_cuda):
    from sgipc import scene, capi
    print("case: ordered projection two_box 8 solves", flush=True)
    sc = scene.two_box_scene(8, seed=2)
    pt, ee = scene.facing_contact_candidates(sc)
    xs = []
    for _ in range(2):
        s = _det_system(sc, pt, ee)
        s.build_precond()
        st = s.solve(tol=1e-6, stop_rule=capi.STOP_TRUE_REL)
        assert st.converged
        xs.append((st.iterations, s.dx.cpu().numpy()))
    assert xs[0][0] == xs[1][0] and np.array_equal(xs[0][1], xs[1][1])
    # and the ordered and atomic projections agree to round-off
    s2 = _det_system(sc, pt, ee)
    s2.ctx.call("sgipc_set_deterministic", 0)
    s2.build_precond()
    st = s2.solve(tol=1e-6, stop_rule=capi.STOP_TRUE_REL)
    assert abs(st.iterations - xs[0][0]) <= 1
    assert np.linalg.norm(s2.dx.cpu().numpy() - xs[0][1]) <= 1e-5 * np.linalg.norm(xs[0][1])


def _top_banks(ls, bank=16):
    """Banks of levels >= 3, as the banked projection's top levels hold them
    (level_size rows: count, first node)."""
    L = len(ls) - 1
    return sum((ls[l + 1][1] - ls[l][1]) // bank for l in range(3, L))


@pytest.mark.parametrize("cells,contacts", [(40, True), (50, False)])
def test_banked_projection_matches_ordered(torch_cuda, cells, contacts):
    """The default (banked) Galerkin projection -- register sums for the
    blocks inside the workgroup's level-2 node, LDS windows, LDS top levels
    in the two-workgroups-per-CU or the twelve-bank form -- against the
    ordered projection, through the MAS apply, to round-off: two boxes of
    40^3 cells with contacts (top levels in six banks) and of 50^3 cells
    (more than six: the larger form)."""
    from sgipc import scene
    print("case: banked vs ordered projection two_box %d contacts=%s" % (cells, contacts), flush=True)
    sc = scene.two_box_scene(cells, seed=2)
    pt, ee = scene.facing_contact_candidates(sc) if contacts else (None, None)
    s = _det_system(sc, pt, ee)
    s.build_precond()
    tops = _top_banks(s.mas_tables()["level_size"])
    assert (tops <= 6) if cells == 40 else (6 < tops <= 12), tops
    r = np.random.default_rng(5).normal(size=3 * sc.n)
    z_ord = s.precond_apply(r)
    s.ctx.call("sgipc_set_deterministic", 0)
    s.build_precond()
    z_bank = s.precond_apply(r)
    assert np.abs(z_bank - z_ord).max() <= 1e-10 * np.abs(z_ord).max()
