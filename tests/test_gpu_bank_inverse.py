"""MAS bank inverses on the matrix cores (blocked Gauss-Jordan over 16x16
tiles on MFMA) against the oracle's scalar FP64 Gauss-Jordan
(__inverse6_P96x96, StiffGIPC/MASPreconditioner.cu:630-726), for 16- and
32-vertex banks, with and without contact, through the C ABI."""
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


def _built(sc, precond, mode, pairs=None):
    from sgipc.system import NewtonSystem
    s = NewtonSystem(sc, "cuda:0", precond=precond)
    s.ctx.call("sgipc_set_bank_inverse", mode)
    if pairs is not None:
        s.setup_contacts(*pairs)
    s.assemble(); s.convert(); s.build_precond()
    return s


def _oracle_mas(oracle, sc, sys_):
    """the oracle's MAS over the GPU's converted system and contact pairs"""
    kr, kc, kv = sys_.bcoo()
    mas = oracle.Mas(sc.n, sys_.part_map, sys_.real_map, sys_.ns, sys_.nn, sys_.nl, sc.bank)
    pairs = sys_.pairs.cpu().numpy().reshape(-1, 4)[:sys_.n_pairs] if sys_.n_pairs else None
    mas.build(kr, kc, kv, pairs=pairs)
    return mas


def _with_unit_pads(A):
    """the bank matrix as inverted: a zero diagonal entry (an empty slot) reads as 1"""
    A = A.copy()
    d = np.diag(A).copy()
    d[d == 0.0] = 1.0
    np.fill_diagonal(A, d)
    return A


def _nbanks(sys_):
    import ctypes as C
    L, tot = C.c_int(0), C.c_int(0)
    sys_.ctx.call("sgipc_mas_info", C.byref(L), C.byref(tot), None)
    return tot.value // sys_.sc.bank


@pytest.mark.parametrize("name", ["cube10", "cube10_b32"])
def test_mfma64_inverse_matches_scalar_gauss_jordan(torch_cuda, oracle, request, name):
    from sgipc import capi
    sc = request.getfixturevalue(name)
    print("case: MFMA64 bank inverses", name, flush=True)
    s = _built(sc, capi.PRECOND_MAS, capi.BANK_INVERSE_MFMA64)
    mas = _oracle_mas(oracle, sc, s)
    nb = _nbanks(s)
    worst, worst_bwd = 0.0, 0.0
    for b in range(nb):
        A, Ai = mas.bank_matrix(b)
        A = _with_unit_pads(A)
        X = s.bank_inverse(b)
        worst = max(worst, np.abs(X - Ai).max() / np.abs(Ai).max())
        # backward error of the GPU inverse against the bank matrix itself
        worst_bwd = max(worst_bwd, np.abs(A @ X - np.eye(len(A))).max() / (np.abs(A).max() * np.abs(X).max()))
        # the FP32 store is the rounding of the FP64 inverse
        assert np.array_equal(s.bank_inverse_f32(b), X.astype(np.float32))
    print("  banks %d, max |X - X_oracle| / max|X| %.2e, backward error %.2e" % (nb, worst, worst_bwd), flush=True)
    assert worst <= 1e-10
    assert worst_bwd <= 1e-12


@pytest.mark.parametrize("name", ["cube10", "cube10_b32"])
def test_mfma32_refined_inverse(torch_cuda, oracle, request, name):
    """FP32 elimination + one refinement step whose residual I - A X is
    accumulated in FP64: the FP32 inverses agree with the FP64 inverse to
    FP32 rounding (a few units in the last place of the largest entry)."""
    from sgipc import capi
    sc = request.getfixturevalue(name)
    print("case: MFMA32 refined bank inverses", name, flush=True)
    s = _built(sc, capi.PRECOND_MAS_FP32, capi.BANK_INVERSE_MFMA32_REFINED)
    mas = _oracle_mas(oracle, sc, s)
    nb = _nbanks(s)
    worst = 0.0
    for b in range(nb):
        _, Ai = mas.bank_matrix(b)
        X = s.bank_inverse_f32(b).astype(np.float64)
        worst = max(worst, np.abs(X - Ai).max() / np.abs(Ai).max())
    print("  banks %d, max |X32 - X_oracle| / max|X| %.2e" % (nb, worst), flush=True)
    assert worst <= 4 * 2.0 ** -24
    # an FP64 MAS build after the mode switch still makes FP64 inverses
    s.build_precond(capi.PRECOND_MAS)
    _, Ai = mas.bank_matrix(0)
    assert np.abs(s.bank_inverse(0) - Ai).max() <= 1e-10 * np.abs(Ai).max()
    # and an FP32-only build refuses FP64 reads
    s.build_precond(capi.PRECOND_MAS_FP32)
    with pytest.raises(capi.SgipcError):
        s.bank_inverse(0)


def test_modes_on_contact_system(torch_cuda, oracle):
    """Two boxes in contact (8^3 cells): the three inversion modes give
    inverses within their precision of the oracle's, and PCG solves to the
    same true residual with the same iteration counts up to one."""
    from sgipc import capi, scene
    sc = scene.two_box_scene(8, seed=2)
    pairs = scene.facing_contact_candidates(sc)
    print("case: bank inverse modes, two_box 8", flush=True)
    its = {}
    for mode, kind in ((capi.BANK_INVERSE_VALU64, capi.PRECOND_MAS),
                       (capi.BANK_INVERSE_MFMA64, capi.PRECOND_MAS),
                       (capi.BANK_INVERSE_MFMA32_REFINED, capi.PRECOND_MAS_FP32)):
        s = _built(sc, kind, mode, pairs)
        mas = _oracle_mas(oracle, sc, s)
        for b in (0, _nbanks(s) // 2, _nbanks(s) - 1):
            _, Ai = mas.bank_matrix(b)
            if kind == capi.PRECOND_MAS:
                err = np.abs(s.bank_inverse(b) - Ai).max() / np.abs(Ai).max()
                assert err <= 1e-10, (mode, b, err)
            else:
                err = np.abs(s.bank_inverse_f32(b).astype(np.float64) - Ai).max() / np.abs(Ai).max()
                assert err <= 4 * 2.0 ** -24, (mode, b, err)
        st = s.solve(tol=1e-6, stop_rule=capi.STOP_TRUE_REL,
                     precision=capi.PRECISION_MIXED if kind == capi.PRECOND_MAS_FP32 else capi.PRECISION_FP64)
        assert st.converged and st.rel_residual <= 1e-6
        its[mode] = st.iterations
        print("  mode %d: %d iterations, residual %.2e" % (mode, st.iterations, st.rel_residual), flush=True)
    assert abs(its[capi.BANK_INVERSE_MFMA64] - its[capi.BANK_INVERSE_VALU64]) <= 1


def _half_bound(Ai):
    """Entrywise bound on D S D - inv for the half storage: S = D^-1 inv D^-1
    rounded to FP16 (half an ulp: 2^-11 relative for normal values, 2^-25
    absolute below 2^-14), times the FP32 scales d_i d_j, plus FP32 rounding
    of the rebuilt product, plus the 1e-10 (of the largest entry) by which
    the GPU's FP64 inverse, the one packed, may differ from the oracle's."""
    d = np.sqrt(np.where(np.diag(Ai) > 0, np.diag(Ai), 1.0))
    S = Ai / np.outer(d, d)
    return (np.outer(d, d) * (2.0 ** -11 * np.abs(S) + 2.0 ** -25) * (1 + 1e-6) + 4 * 2.0 ** -24 * np.abs(Ai)
            + 1e-10 * np.abs(Ai).max())


@pytest.mark.parametrize("mode", ["mfma64", "valu64", "mfma32_refined"])
def test_half_storage_rebuilds_the_inverse(torch_cuda, oracle, cube10, mode):
    """SGIPC_BANK_STORE_F16: D S D from the half storage is the oracle's FP64
    inverse up to FP16 rounding of S, entry by entry; the diagonal (S_ii = 1,
    exact in FP16) to FP32 rounding; a build back in FP32 storage drops it.
    MFMA64 writes the half form in its own store phase; the other modes pack
    it in a separate pass (from the FP64 inverse, or from the refined FP32
    one, whose own error of up to 4 units of FP32 of the largest entry the
    bound then adds)."""
    from sgipc import capi
    inv_mode, kind = {"mfma64": (capi.BANK_INVERSE_MFMA64, capi.PRECOND_MAS),
                      "valu64": (capi.BANK_INVERSE_VALU64, capi.PRECOND_MAS),
                      "mfma32_refined": (capi.BANK_INVERSE_MFMA32_REFINED, capi.PRECOND_MAS_FP32)}[mode]
    print("case: half-storage bank inverses, cube10,", mode, flush=True)
    s = _built(cube10, kind, inv_mode)
    with pytest.raises(capi.SgipcError):
        s.bank_inverse_f16(0)
    s.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F16)
    s.build_precond()
    mas = _oracle_mas(oracle, cube10, s)
    nb = _nbanks(s)
    worst = 0.0
    for b in range(nb):
        _, Ai = mas.bank_matrix(b)
        X = s.bank_inverse_f16(b).astype(np.float64)
        err = np.abs(X - Ai)
        extra = 8 * 2.0 ** -24 * np.abs(Ai).max() if kind == capi.PRECOND_MAS_FP32 else 0.0
        bound = _half_bound(Ai) * (1 + 2.0 ** -10 if extra else 1) + extra
        assert np.all(err <= bound), (b, np.max(err / np.maximum(bound, 1e-300)))
        assert np.allclose(np.diag(X), np.diag(Ai), rtol=4 * 2.0 ** -24, atol=1e-10 * np.abs(Ai).max() + extra)
        worst = max(worst, err.max() / np.abs(Ai).max())
    print("  banks %d, max |DSD - X_oracle| / max|X| %.2e" % (nb, worst), flush=True)
    assert worst > 0.0  # the half storage is in use, not a copy of the FP32 one
    with pytest.raises(capi.SgipcError):
        s.ctx.call("sgipc_set_bank_storage", 7)
    s.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F32)
    s.build_precond()
    with pytest.raises(capi.SgipcError):
        s.bank_inverse_f16(0)


def test_half_storage_mixed_solve(torch_cuda):
    """Two boxes in contact (8^3 cells), mixed precision (FP32 inner PCG,
    FP64 refinement) to a 1e-6 true relative residual: with the half storage
    the solve converges to the same residual bar, its iteration count within
    a few of the FP32 storage's, and its answer agrees with the FP64 solve's
    to the accuracy the stop test allows."""
    from sgipc import capi, scene
    sc = scene.two_box_scene(8, seed=2)
    pairs = scene.facing_contact_candidates(sc)
    print("case: half-storage mixed solve, two_box 8", flush=True)
    s = _built(sc, capi.PRECOND_MAS, capi.BANK_INVERSE_MFMA64, pairs)
    kw = dict(tol=1e-6, stop_rule=capi.STOP_TRUE_REL)
    st64 = s.solve(precision=capi.PRECISION_FP64, **kw)
    x64 = s.dx.double().cpu().numpy().copy()
    res = {}
    for store in (capi.BANK_STORE_F32, capi.BANK_STORE_F16):
        s.ctx.call("sgipc_set_bank_storage", store)
        s.build_precond()
        st = s.solve(precision=capi.PRECISION_MIXED, **kw)
        assert st.converged and st.rel_residual <= 1e-6, (store, st.rel_residual)
        x = s.dx.double().cpu().numpy()
        rel = np.linalg.norm(x - x64) / np.linalg.norm(x64)
        res[store] = (st.iterations, rel)
        print("  store %d: %d iterations (FP64 %d), residual %.2e, |x - x64|/|x64| %.2e"
              % (store, st.iterations, st64.iterations, st.rel_residual, rel), flush=True)
    assert abs(res[capi.BANK_STORE_F16][0] - res[capi.BANK_STORE_F32][0]) <= 3
    assert res[capi.BANK_STORE_F16][1] <= max(10 * res[capi.BANK_STORE_F32][1], 1e-5)
