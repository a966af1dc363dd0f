"""GPU FP64 line-search energies (kinetic, SNK1 elastic, IPC barrier) against
the oracle and closed forms, their determinism, and their consistency with
the assembled gradient (central finite differences of the GPU energy)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def twobox8_sys(torch_cuda):
    from sgipc import scene
    from sgipc.system import NewtonSystem
    sc = scene.two_box_scene(8, seed=2)
    pt, ee = scene.facing_contact_candidates(sc)
    sysm = NewtonSystem(sc, "cuda:0")
    sysm.setup_contacts(pt, ee)
    return sc, sysm


def test_energies_match_oracle(torch_cuda, oracle, twobox8_sys):
    sc, sysm = twobox8_sys
    print("case: energies twobox8", flush=True)
    e = sysm.energies()
    ek = 0.5 * np.sum(sc.mass[:, None] * (sc.x - sc.x_tilde) ** 2)
    assert abs(e["kinetic"] - ek) <= 1e-12 * abs(ek)
    ef = sum(oracle.snk1_energy(sc.x[sc.tets[t].astype(int)].reshape(-1), sc.DmInv[t], sc.vol[t], sc.mu[t], sc.lam[t])
             for t in range(sc.T)) * sc.dt ** 2
    assert abs(e["elastic"] - ef) <= 1e-11 * abs(ef)
    pairs = sysm.pairs.cpu().numpy().reshape(-1, 4)
    eb = oracle.barrier_energy(pairs, sc.x, sc.x_rest, sc.dhat, sc.kappa)
    assert eb > 0 and abs(e["barrier"] - eb) <= 1e-11 * abs(eb)
    # fixed-order reductions: bitwise repeatable
    assert sysm.energies() == e


def test_energy_gradient_consistency(torch_cuda, twobox8_sys):
    """b (the assembled gradient, zero on pinned vertices) equals the
    directional derivative of the GPU total energy along free directions."""
    import torch
    sc, sysm = twobox8_sys
    print("case: energy FD vs assembled gradient", flush=True)
    sysm.assemble()
    b = sysm.b.cpu().numpy()
    x0 = sysm.x.clone()
    rng = np.random.default_rng(11)
    free = np.repeat(sc.btype == 0, 3)
    for trial in range(3):
        v = rng.normal(size=3 * sc.n) * free
        v /= np.linalg.norm(v)
        h = 1e-4 * np.sqrt(sc.dhat)  # small against the contact gap
        vt = torch.tensor(v, dtype=torch.float64, device=sysm.dev)
        sysm.x.copy_(x0 + h * vt)
        ep = sum(sysm.energies().values())
        sysm.x.copy_(x0 - h * vt)
        em = sum(sysm.energies().values())
        sysm.x.copy_(x0)
        fd = (ep - em) / (2 * h)
        assert abs(fd - b @ v) <= 1e-5 * max(abs(b @ v), 1e-3 * np.linalg.norm(b))
