#!/usr/bin/env python3
"""Newton-step benchmark on one MI355X.

Workload (the contact configuration): two soft boxes of 40^3 cells (5 tets per
cell, 68,921 vertices each) stacked with a gap of 0.5 dhat, E=1e6, nu=0.4,
kappa=1e4, dt=0.01, seeded velocities in [-0.5, 0.5].  One step = one Newton
iteration's hot path on a fixed iterate: kinetic + FEM (+ barrier) gradient
and Hessian triplets, mass diagonal, triplet -> BCOO, MAS build, MAS-PCG to a
1e-6 true relative residual.  Inputs are resident on the GPU before timing.

By default both forms run on the same system: the mixed FP32/FP64 solve
(reported as `value`) and the plain FP64 solve (`fp64`, with the ratio).
Prints one JSON line (metric: Newton solves/s) with the roofline of the
dominant kernel (the PCG SpMV) measured live with HIP events, and the CPU
baseline (the oracle's FP64 restatement of the same step, one thread).
"""
import argparse
import json
import subprocess
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "stiff-gipc_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--bank", type=int, choices=[16, 32], default=16, help="MAS subdomain size (vertices)")
    ap.add_argument("--precision", choices=["fp64", "mixed", "fp32apply", "both"], default="both",
                    help="both: every form of the workload (contact: fp64 + mixed; hybrid: fp64, fp32apply, mixed)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-half-store", action="store_true",
                    help="skip the side record of the mixed form with FP16 bank-inverse storage")
    ap.add_argument("--bank-store", choices=["f32", "f16"], default="f32",
                    help="batched: storage of the MAS bank inverses the FP32 bank solves of the launch sequence "
                         "read (f16: D S D with S in FP16, opt-in)")
    ap.add_argument("--workload", choices=["contact", "cantilever_adjoint", "batched", "hybrid"], default="contact",
                    help="contact: the Newton-step benchmark (default); cantilever_adjoint: the differentiable "
                         "rollout (60x12x12-cell cantilever over a ground barrier, 20 steps) with its adjoint "
                         "gradient checked against central differences; batched: 256 independent boxes of "
                         "uneven size, one batched Newton step against 256 sequential plain FP64 steps; hybrid: "
                         "64 affine cubes (kappa 1e11) on a 100x20x100-cell soft slab, one coupled Newton step, "
                         "mixed vs FP64")
    ap.add_argument("--scenes", type=int, default=256, help="batched: number of scenes")
    ap.add_argument("--scene-vertices", type=int, nargs=2, default=[500, 20000], metavar=("MIN", "MAX"),
                    help="batched: range of the log-uniform per-scene vertex counts")
    ap.add_argument("--batch-form", choices=["auto", "sequence", "persistent"], default="auto",
                    help="batched: the autotuner's choice, the graph-captured launch sequence, or the persistent "
                         "launch (a workgroup per scene)")
    ap.add_argument("--fd-dirs", type=int, default=32, help="cantilever_adjoint: FD directions")
    ap.add_argument("--rollout-steps", type=int, default=20, help="cantilever_adjoint: time steps")
    ap.add_argument("--newton-tol", type=float, default=1e-13,
                    help="cantilever_adjoint: Newton stop on |dx|_inf (and round-off stall below 1e3 x this)")
    ap.add_argument("--pcg-tol", type=float, default=1e-12, help="cantilever_adjoint: PCG true rel. residual")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    # a plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE of its
    # dispatches calibrate the counters against exact byte counts
    import torch
    n = mib * (1 << 20) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8}))


def spmv_alg_bytes(K, nb, fp32):
    # the PCG SpMV kernel (direction update fused): upper BCOO, 9 values
    # (72 B FP64 / 36 B FP32) + 8 B (row, col) per block; per block row z
    # read once, p and Ap each read and written once.  (The FP32 copy is
    # stored padded to 48 B per block; the padding is not counted.)
    vb = 36.0 if fp32 else 72.0
    xb = 12.0 if fp32 else 24.0
    return (vb + 8.0) * K + 5.0 * xb * nb


def broad_phase_timing(sc, sysm, reps=5):
    """LBVH broad phase + GPU narrow phase on the benchmark iterate (outside
    the Newton-step metric, as collision detection is): median wall ms over
    `reps`, candidate and pair counts."""
    import ctypes as C
    import numpy as np
    import torch
    from sgipc import scene
    from sgipc.system import _p
    faces, edges, surf = scene.surface_primitives(sc.tets)
    dev = dict(dtype=torch.int32, device="cuda:0")
    f = torch.tensor(faces.astype(np.int32).reshape(-1), **dev)
    e = torch.tensor(edges.astype(np.int32).reshape(-1), **dev)
    v = torch.tensor(surf.astype(np.int32), **dev)
    cap = 4 * (len(surf) + len(edges)) + 1024
    pt = torch.zeros(4 * cap, **dev); ee = torch.zeros(4 * cap, **dev)
    pairs = torch.zeros(8 * cap, **dev); mi = torch.zeros(2 * cap, **dev)
    npt, nee, npair = C.c_int(), C.c_int(), C.c_int()
    counts = np.zeros(3, np.int32)
    ts = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sysm.ctx.call("sgipc_broad_phase", len(surf), _p(v), len(faces), _p(f), len(edges), _p(e), _p(sysm.x),
                      _p(sysm.btype), None, float(sc.dhat), _p(pt), cap, C.byref(npt), _p(ee), cap, C.byref(nee))
        sysm.ctx.call("sgipc_classify_contacts", npt.value, _p(pt), nee.value, _p(ee), _p(sysm.x), _p(sysm.x_rest),
                      float(sc.dhat), _p(pairs), _p(mi), 2 * cap, C.byref(npair), counts.ctypes.data)
        torch.cuda.synchronize()
        ts.append(1e3 * (time.perf_counter() - t0))
    ts = sorted(ts[1:])
    out = {"ms": round(ts[len(ts) // 2], 3), "surface_vertices": int(len(surf)), "faces": int(len(faces)),
           "edges": int(len(edges)), "pt_candidates": npt.value, "ee_candidates": nee.value,
           "pairs": npair.value, "note": "LBVH broad phase + narrow phase; the Newton step uses the grid-cell "
           "candidate pairs (a subset of these)"}
    # line-search step bound along the Newton direction of this iterate
    sysm.newton_step(tol=1e-6, stop_rule=1)
    a = C.c_double(1.0)
    tc = []
    ccap = 8 * cap
    pt = torch.zeros(4 * ccap, **dev); ee = torch.zeros(4 * ccap, **dev)
    cap = ccap
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sysm.ctx.call("sgipc_broad_phase_ccd", len(surf), _p(v), len(faces), _p(f), len(edges), _p(e), _p(sysm.x),
                      _p(sysm.dx), 1.0, _p(sysm.btype), None, float(sc.dhat), _p(pt), cap, C.byref(npt), _p(ee), cap,
                      C.byref(nee))
        sysm.ctx.call("sgipc_ccd_step_bound", npt.value, _p(pt), nee.value, _p(ee), _p(sysm.x), _p(sysm.dx), 0.8,
                      C.byref(a))
        torch.cuda.synchronize()
        tc.append(1e3 * (time.perf_counter() - t0))
    tc = sorted(tc[1:])
    out["ccd"] = {"ms": round(tc[len(tc) // 2], 3), "pt_candidates": npt.value, "ee_candidates": nee.value,
                  "step_bound": a.value, "note": "swept-box broad phase + additive CCD along the Newton direction, "
                  "slackness 0.8"}
    return out


def cpu_baseline(sc, pairs, mat_index, counts, min_secs=10.0, max_steps=5, threads=1):
    """The oracle's FP64 C restatement of the same Newton step on `threads`
    host threads (FEM assembly, convert, MAS build and PCG split over
    std::thread-style pthreads; 1 = the serial restatement), repeated until
    at least min_secs of CPU work (bounded by max_steps)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    with O.threads(threads):
        return _cpu_baseline(O, sc, pairs, mat_index, counts, min_secs, max_steps)


def _cpu_baseline(O, sc, pairs, mat_index, counts, min_secs, max_steps):
    import numpy as np
    ns, nn, nl = O.vertex_neighbors(sc.n, sc.tets)  # topology setup: one-time, untimed (as on the GPU)
    pm, rm = O.bank_map(sc.part, sc.bank)
    total, steps, it = 0.0, 0, 0
    while steps < max_steps and (steps == 0 or total < min_secs):
        t0 = time.perf_counter()
        if pairs is not None:
            fb, bv, br, bc = O.barrier_assemble(pairs, mat_index, counts, sc.x, sc.x_rest, sc.dhat, sc.kappa,
                                                sc.btype)
        else:
            fb = np.zeros(3 * sc.n); bv = np.zeros((0, 9)); br = np.zeros(0, np.int32); bc = np.zeros(0, np.int32)
        shape = O.kinetic_gradient(sc.x, sc.x_tilde, sc.mass)
        grad, vals, rows, cols = O.fem_assemble(sc.x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt, sc.btype, 0,
                                                shape.copy())
        mv, mr, mc = O.mass_triplets(sc.mass)
        kr, kc, kv = O.convert(np.concatenate([br, rows, mr]), np.concatenate([bc, cols, mc]),
                               np.concatenate([bv, vals, mv]))
        b = O.fem_rhs(fb.reshape(-1, 3), grad, sc.btype).reshape(-1)
        mas = O.Mas(sc.n, pm, rm, ns, nn, nl, sc.bank)
        mas.build(kr, kc, kv, pairs=pairs, fp32_store=False)
        x, it, rr = O.pcg(kr, kc, kv, sc.n, b, precond=2, mas=mas, tol=1e-6, rule=1, max_iter=100000)
        total += time.perf_counter() - t0
        steps += 1
        del mas
    return total, steps, it


def oracle_ground_step(O, sc, topo, tol):
    """One Newton step of a scene over the ground plane with the oracle's FP64
    C restatement (ground detection + barrier, kinetic + FEM, mass, convert,
    MAS build, MAS-PCG to `tol` true rel. residual); returns PCG iterations."""
    import numpy as np
    m = sc.meta
    nrm, off = m["ground_normal"], float(m["ground_offset"])
    gp = O.ground_detect(m["surf"], sc.x, nrm, off, sc.dhat)
    fb, gv, gr, gc = O.ground_assemble(gp, sc.x, nrm, off, sc.dhat, sc.kappa, sc.btype)
    shape = O.kinetic_gradient(sc.x, sc.x_tilde, sc.mass)
    grad, vals, rows, cols = O.fem_assemble(sc.x, sc.tets, sc.DmInv, sc.vol, sc.mu, sc.lam, sc.dt, sc.btype, 0,
                                            shape.copy())
    mv, mr, mc = O.mass_triplets(sc.mass)
    kr, kc, kv = O.convert(np.concatenate([gr, rows, mr]), np.concatenate([gc, cols, mc]),
                           np.concatenate([gv, vals, mv]))
    b = O.fem_rhs(fb.reshape(-1, 3), grad, sc.btype).reshape(-1)
    ns, nn, nl, pm, rm = topo
    mas = O.Mas(sc.n, pm, rm, ns, nn, nl, sc.bank)
    mas.build(kr, kc, kv, fp32_store=False)
    x, it, rr = O.pcg(kr, kc, kv, sc.n, b, precond=2, mas=mas, tol=tol, rule=1, max_iter=100000)
    del mas
    return it


def oracle_hybrid_step(O, A, hs, x_all, x_rest, btype, pairs, mat_index, counts, tol=1e-6):
    """One coupled Newton step of the hybrid scene on the CPU: the oracle's C
    restatement for the barrier, FEM, mass, convert and MAS, its numpy
    affine-body restatement for the J^T H J mapping, body terms and the 12x12
    block-Jacobi, and PCG (true rel. residual) in numpy around them.  Returns
    the PCG iteration count."""
    import numpy as np
    fem = hs.fem
    nb, nav, nf = hs.nb, hs.nav, fem.n
    nbr = 4 * nb + nf
    body_of = np.repeat(np.arange(nb), np.diff(hs.vstart))
    fb, bv, br, bc = O.barrier_assemble(pairs, mat_index, counts, x_all, x_rest, fem.dhat, fem.kappa, btype)
    R, Cc, V = A.map_contacts(br, bc, bv.reshape(-1, 3, 3).transpose(0, 2, 1), nav, nb, body_of, hs.xbar, hs.fixed)
    xf = x_all.reshape(-1, 3)[nav:]
    shape = O.kinetic_gradient(xf, fem.x_tilde, fem.mass)
    grad, vals, rows, cols = O.fem_assemble(xf, fem.tets, fem.DmInv, fem.vol, fem.mu, fem.lam, fem.dt, fem.btype,
                                            4 * nb, shape.copy())
    mv, mr, mc = O.mass_triplets(fem.mass, 4 * nb)
    G, H = A.body_system(hs.q, hs.q_tilde, hs.mass12, hs.volume, hs.fixed, hs.kappa_abd, fem.dt)
    hr, hc, hv = A.body_triplets(H)
    col9 = lambda m: np.ascontiguousarray(m.transpose(0, 2, 1).reshape(-1, 9))
    kr, kc, kv = O.convert(np.concatenate([R, rows, mr, hr]).astype(np.int32),
                           np.concatenate([Cc, cols, mc, hc]).astype(np.int32),
                           np.concatenate([col9(V), vals, mv, col9(hv)]))
    gq = G + A.contact_gradient(fb.reshape(-1, 3)[:nav], body_of, hs.xbar, nb, hs.fixed)
    bf = O.fem_rhs(fb.reshape(-1, 3)[nav:], grad, fem.btype).reshape(-1)
    b = np.concatenate([gq.reshape(-1), bf])
    ns, nn, nl = O.vertex_neighbors(nf, fem.tets)
    pm, rm = O.bank_map(fem.part, fem.bank)
    mas = O.Mas(nf, pm, rm, ns, nn, nl, fem.bank)
    mas.build(kr, kc, kv, offset=4 * nb, pairs=pairs, node_offset=nav, fp32_store=False)
    # body diagonal blocks from the BCOO, inverted
    D = np.zeros((nb, 12, 12))
    for k in np.nonzero(kr < 4 * nb)[0]:
        r, c = int(kr[k]), int(kc[k])
        if c // 4 != r // 4:
            continue
        Bk = kv[k].reshape(3, 3).T
        bb, i, j = r // 4, r % 4, c % 4
        D[bb, 3 * i:3 * i + 3, 3 * j:3 * j + 3] = Bk
        if i != j:
            D[bb, 3 * j:3 * j + 3, 3 * i:3 * i + 3] = Bk.T
    Dinv = np.linalg.inv(D)
    n12 = 12 * nb

    def precond(r):
        return np.concatenate([np.einsum("bij,bj->bi", Dinv, r[:n12].reshape(nb, 12)).reshape(-1), mas.apply(r[n12:])])
    x = np.zeros(3 * nbr)
    r = b.copy()
    bn = np.linalg.norm(b)
    z = precond(r)
    p = z.copy()
    rz = r @ z
    it = 0
    while np.linalg.norm(r) > tol * bn and it < 20000:
        Ap = O.spmv_sym(kr, kc, kv, nbr, p)
        a = rz / (p @ Ap)
        x += a * p
        r -= a * Ap
        z = precond(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    del mas
    return it


def oracle_topology(O, sc):
    ns, nn, nl = O.vertex_neighbors(sc.n, sc.tets)
    pm, rm = O.bank_map(sc.part, sc.bank)
    return ns, nn, nl, pm, rm


def run_mode(sysm, precision, steps, warmup):
    """Warm-up, then exactly `steps` timed Newton steps; then two steps with
    the library's HIP-event timing for the kernel figures."""
    import ctypes as C
    import torch
    from sgipc import capi
    solve_kw = dict(tol=1e-6, stop_rule=capi.STOP_TRUE_REL, precision=precision, max_iter=100000)
    for _ in range(warmup):
        sysm.newton_step(**solve_kw)
    torch.cuda.synchronize()
    iters, pcg_s, rel, per = 0, 0.0, [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ta = time.perf_counter()
        sysm.assemble()
        sysm.convert()
        sysm.build_precond()
        ts = time.perf_counter()
        st = sysm.solve(**solve_kw)
        pcg_s += time.perf_counter() - ts
        iters += st.iterations
        rel.append(st.rel_residual)
        torch.cuda.synchronize()
        per.append(time.perf_counter() - ta)
    secs = time.perf_counter() - t0
    per.sort()
    dx = sysm.dx.detach().clone()

    ctx = sysm.ctx
    ctx.call("sgipc_profile_enable", 1)
    for slot in range(7):
        ctx.call("sgipc_profile_read", slot, None, None, 1)
    for _ in range(2):
        sysm.newton_step(**solve_kw)
    calls = C.c_longlong(0); ms = C.c_double(0.0)
    ctx.call("sgipc_profile_read", capi.PROF_SPMV, C.byref(calls), C.byref(ms), 0)
    phase = {}
    for name, slot in (("fem_kernel", capi.PROF_FEM), ("barrier_kernel", capi.PROF_BARRIER),
                       ("convert", capi.PROF_CONVERT), ("precond_build", capi.PROF_PRECOND_BUILD),
                       ("solve_ungraphed", capi.PROF_SOLVE), ("mas_bank_solve_kernel_per_launch", capi.PROF_MAS_SOLVE)):
        c2 = C.c_longlong(0); m2 = C.c_double(0.0)
        ctx.call("sgipc_profile_read", slot, C.byref(c2), C.byref(m2), 0)
        phase[name] = round(m2.value / max(c2.value, 1), 4)
    ctx.call("sgipc_profile_enable", 0)
    return dict(value=steps / secs, ms_per_step=1e3 * secs / steps, iters=iters / steps,
                pcg_iters_per_s=iters / pcg_s, rel_max=max(rel), phase_ms=phase,
                spmv_ms=ms.value / max(calls.value, 1), spmv_calls=calls.value,
                ms_median=1e3 * per[len(per) // 2], ms_p90=1e3 * per[max(0, -(-9 * len(per) // 10) - 1)], dx=dx)


def reference_rule(sysm, steps):
    """FP64 Newton steps with the reference's stop rule (|r.z| <= 1e-4 r0.z0)."""
    import torch
    from sgipc import capi
    kw = dict(tol=1e-4, stop_rule=capi.STOP_REF_RZ, precision=capi.PRECISION_FP64)
    sysm.newton_step(**kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its, rel = 0, 0.0
    for _ in range(steps):
        st = sysm.newton_step(**kw)
        its += st.iterations
        rel = max(rel, st.rel_residual)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    return {"value": round(steps / secs, 4), "unit": "Newton solves/s", "ms_per_step": round(1e3 * secs / steps, 4),
            "pcg_iters_per_step": its / steps, "true_rel_residual_max": rel,
            "stop": "reference rule |r.z| <= 1e-4 r0.z0, FP64"}


def host_cpu():
    """The host's CPU: model, logical CPUs, sockets and physical cores (from
    /proc/cpuinfo; lscpu where present), and the CPUs this process may use."""
    info = {"model": None, "logical_cpus": os.cpu_count(), "sockets": None, "physical_cores": None}
    try:
        phys, cores_of = set(), set()
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and info["model"] is None:
                    info["model"] = v
                elif k == "physical id":
                    cur["p"] = v
                    phys.add(v)
                elif k == "core id":
                    cur["c"] = v
                elif not k and cur:
                    if "p" in cur and "c" in cur:
                        cores_of.add((cur["p"], cur["c"]))
                    cur = {}
        if "p" in cur and "c" in cur:
            cores_of.add((cur["p"], cur["c"]))
        info["sockets"] = len(phys) or None
        info["physical_cores"] = len(cores_of) or None
    except OSError:
        pass
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["usable_cpus"] = os.cpu_count()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket" and info.get("sockets"):
                info["physical_cores"] = int(v) * info["sockets"]
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def baseline_threads():
    """Threads for the all-cores CPU baseline: the CPUs this process may use,
    at most 16 (the GPU box's CPU share for one GPU)."""
    h = host_cpu()
    return max(1, min(16, h.get("usable_cpus") or os.cpu_count() or 1))


def mas_banks(sysm, bank):
    import ctypes as C
    L = C.c_int(0); tot = C.c_int(0)
    sysm.ctx.call("sgipc_mas_info", C.byref(L), C.byref(tot), None)
    return tot.value // bank


def kernel_rooflines(sysm, r, fp32_mas):
    """Achieved GB/s of the assembly and MAS kernels from the library's
    HIP-event timing and their algorithmic bytes per unit."""
    out = {}
    sc = sysm.sc
    fem_b = 1104.0 * sc.T
    if r["phase_ms"]["fem_kernel"] > 0:
        out["k_fem_snk1"] = {"alg_bytes": fem_b, "avg_ms": r["phase_ms"]["fem_kernel"],
                             "achieved_GBs": round(fem_b / (r["phase_ms"]["fem_kernel"] * 1e-3) / 1e9, 1),
                             "unit_bytes": "1104 B per tet"}
    c = [int(v) for v in sysm.counts]
    bar_b = sum(n * (116.0 + nblk * 80.0 + 96.0) for n, nblk in ((c[0], 4), (c[1], 9), (c[2], 16)))
    if r["phase_ms"]["barrier_kernel"] > 0 and bar_b > 0:
        out["k_barrier"] = {"alg_bytes": bar_b, "avg_ms": r["phase_ms"]["barrier_kernel"],
                            "achieved_GBs": round(bar_b / (r["phase_ms"]["barrier_kernel"] * 1e-3) / 1e9, 1),
                            "unit_bytes": "116 B read + 80 B per block + 96 B gradient per pair"}
    nbk = mas_banks(sysm, sc.bank)
    D, w = 3 * sc.bank, (4.0 if fp32_mas else 8.0)
    mas_b = nbk * (D * (D + 1) / 2 * w + 2 * D * w)
    if r["phase_ms"]["mas_bank_solve_kernel_per_launch"] > 0:
        out["k_bank_solve"] = {"alg_bytes": mas_b, "avg_ms": r["phase_ms"]["mas_bank_solve_kernel_per_launch"],
                               "achieved_GBs": round(mas_b / (r["phase_ms"]["mas_bank_solve_kernel_per_launch"] * 1e-3)
                                                     / 1e9, 1),
                               "unit_bytes": "packed %dx%d inverse + R and Z per bank, %s"
                                             % (D, D, "FP32" if fp32_mas else "FP64"),
                               "banks": nbk}
    return out


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (run with --gpus 1)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "cantilever_adjoint":
        cantilever_adjoint(a)
        return
    if a.workload == "batched":
        batched(a)
        return
    if a.workload == "hybrid":
        hybrid(a)
        return
    import torch
    from sgipc import scene, capi
    from sgipc.system import NewtonSystem

    sc = scene.two_box_scene(a.cells, bank=a.bank)
    pt, ee = scene.facing_contact_candidates(sc)
    sysm = NewtonSystem(sc, "cuda:0", precond=capi.PRECOND_MAS)
    sysm.setup_contacts(pt, ee)  # narrow phase once per Newton iterate, outside the timed hot path

    try:
        bp = broad_phase_timing(sc, sysm)
    except Exception as e:  # informational line only; the Newton-step metric does not depend on it
        bp = {"error": str(e)}
    modes = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    res = {}
    for m in modes:
        sysm.precond = capi.PRECOND_MAS_FP32 if m == "fp32apply" else capi.PRECOND_MAS
        res[m] = run_mode(sysm, capi.PRECISION_MIXED if m == "mixed" else capi.PRECISION_FP64, a.steps, a.warmup)
    sysm.precond = capi.PRECOND_MAS
    main_mode = "mixed" if "mixed" in res else modes[0]
    r = res[main_mode]
    # the opt-in half storage of the bank inverses (FP16 S, FP32 row
    # scales) for the mixed form: a side record; the headline keeps FP32
    half = None
    if "mixed" in res and not a.no_half_store:
        sysm.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F16)
        try:
            half = run_mode(sysm, capi.PRECISION_MIXED, a.steps, a.warmup)
        finally:
            sysm.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F32)
    ref_rule = reference_rule(sysm, 5)

    # roofline of the dominant kernel (PCG SpMV) of the reported mode
    fp32 = main_mode == "mixed"
    alg = spmv_alg_bytes(sysm.K, sysm.n, fp32)
    achieved = alg / (r["spmv_ms"] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if os.path.exists(tf):
        try:
            with open(tf) as f:
                traffic = json.load(f)[main_mode]["hbm_bytes_per_launch"]
        except (OSError, ValueError, KeyError, TypeError):
            traffic = None

    cpu = None
    if not a.no_cpu_baseline:
        cp = sysm.pairs.cpu().numpy().reshape(-1, 4) if sysm.n_pairs else None
        cm = sysm.mat_index.cpu().numpy() if sysm.n_pairs else None
        nt = baseline_threads()
        print("CPU baseline (oracle): %d threads ..." % nt, file=sys.stderr, flush=True)
        csecs, csteps, cit = cpu_baseline(sc, cp, cm, sysm.counts, min_secs=8.0, threads=nt)
        print("CPU baseline (oracle): 1 thread ...", file=sys.stderr, flush=True)
        ssecs, ssteps, sit = cpu_baseline(sc, cp, cm, sysm.counts, min_secs=5.0, max_steps=2, threads=1)
        cpu = {"value": round(csteps / csecs, 5), "unit": "Newton solves/s", "cores": nt, "kind": "port",
               "sample": "%d full Newton step(s) of the same workload (barrier + FEM assembly, convert, MAS build, "
                         "MAS-PCG to 1e-6 true rel. residual: %d iterations), oracle FP64 C restatement on %d "
                         "threads, %.1f s" % (csteps, cit, nt, csecs),
               "single_thread": {"value": round(ssteps / ssecs, 5), "unit": "Newton solves/s", "cores": 1,
                                 "sample": "%d step(s), %d iterations, %.1f s" % (ssteps, sit, ssecs)},
               "label": "CPU restatement of the reference algorithm (baseline only)", "host": host_cpu()}

    g = os.environ.get("SGIPC_SPMV_G", "8")
    g = g if g in ("4", "8", "16") else "8"   # lanes per block row, as the library reads it
    full32 = os.environ.get("SGIPC_SPMV_FULL32", "1") != "0"   # the library's default for mixed solves
    if fp32 and full32:
        kname = "k_spmv_pap_full<%s, float> (PCG SpMV + direction update, row-contiguous FP32 copy)" % g
    elif fp32:
        kname = "k_spmv_pap<%s, float, float> (PCG SpMV + direction update, FP32 upper BCOO)" % g
    else:
        kname = "k_spmv_pap<%s, double, double> (PCG SpMV + direction update, FP64)" % g
    out = {
        "metric": "Newton solves/sec (assembly + BCOO convert + MAS build + MAS-PCG to 1e-6 true rel. residual)",
        "value": round(r["value"], 4),
        "unit": "Newton solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_mode == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {
            "workload": "two stacked soft boxes, 2 x %d^3 cells, 5 tets/cell, %d vertices, %d tets, gap 0.5 dhat"
                        % (a.cells, sc.n, sc.T),
            "contact_pairs": int(sysm.n_pairs),
            "upper_blocks": int(sysm.K),
            "precision": main_mode,
            "mixed_inner_tol_floor": float(os.environ.get("SGIPC_INNER_TOL", "1e-6")),
            "preconditioner": "MAS, %d-vertex banks, %d levels" % (a.bank, sysm.mas_levels()),
            "stop": "true relative residual 1e-6",
            "parallelism": "single GPU",
        },
        "ms_per_step_median": round(r["ms_median"], 4),
        "ms_per_step_p90": round(r["ms_p90"], 4),
        "pcg_iters_per_step": r["iters"],
        "pcg_iters_per_s": round(r["pcg_iters_per_s"], 2),
        "rel_residual_max": r["rel_max"],
        "phase_ms": r["phase_ms"],
        "phase_note": "HIP-event averages per call over two extra Newton steps run with event timing on (not the "
                      "timed steps): fem_kernel, barrier_kernel, convert, precond_build per step; solve_ungraphed "
                      "is the whole PCG solve in that timing mode, where the iteration graphs are not used (so it "
                      "exceeds the graphed solve inside ms_per_step); mas_bank_solve_kernel_per_launch is one bank "
                      "solve launch (several per iteration); the phases do not add up to ms_per_step",
        "kernel_rooflines": kernel_rooflines(sysm, r, main_mode == "mixed"),
        "reference_rule": ref_rule,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("stored counter measurement, profiles/spmv_traffic.json (rocprofv3 --pmc "
                                        "FETCH_SIZE / WRITE_SIZE runs of this kernel, scaled by a device copy); "
                                        "not measured in this run") if traffic is not None else None,
                     "avg_ms": round(r["spmv_ms"], 5), "launches": r["spmv_calls"], "alg_bytes": alg,
                     "cache_note": "the SpMV's matrix copy (about 76 MB FP32 / 68 MB FP64) fits MI355X's 256 MB "
                                   "Infinity Cache; FETCH_SIZE counts L2 misses, which it may serve instead of HBM"},
        "cpu_baseline": cpu,
        "broad_phase": bp,
    }
    if main_mode == "mixed" and "fp64" in res:
        f = res["fp64"]
        out["fp64"] = {"value": round(f["value"], 4), "ms_per_step": round(f["ms_per_step"], 4),
                       "pcg_iters_per_step": f["iters"], "pcg_iters_per_s": round(f["pcg_iters_per_s"], 2),
                       "rel_residual_max": f["rel_max"], "phase_ms": f["phase_ms"],
                       "spmv_avg_ms": round(f["spmv_ms"], 5),
                       "spmv_achieved_GBs": round(spmv_alg_bytes(sysm.K, sysm.n, False) / (f["spmv_ms"] * 1e-3) / 1e9,
                                                  2)}
        out["mixed_over_fp64"] = round(r["value"] / f["value"], 4)
        out["fp64"]["ms_per_step_median"] = round(f["ms_median"], 4)
        out["fp64"]["ms_per_step_p90"] = round(f["ms_p90"], 4)
        out["fp64"]["kernel_rooflines"] = kernel_rooflines(sysm, f, False)
        d64 = f["dx"]
        out["mixed_dx_rel_diff_vs_fp64"] = float((r["dx"] - d64).norm() / d64.norm())
    if half is not None:
        out["mixed_half_bank_store"] = {
            "value": round(half["value"], 4), "ms_per_step": round(half["ms_per_step"], 4),
            "ms_per_step_median": round(half["ms_median"], 4), "pcg_iters_per_step": half["iters"],
            "rel_residual_max": half["rel_max"], "phase_ms": half["phase_ms"],
            "over_mixed": round(half["value"] / r["value"], 4),
            "note": ("mixed form with sgipc_set_bank_storage(SGIPC_BANK_STORE_F16): the FP32 bank solves read "
                     "each 48x48 inverse as D S D, S in FP16 and D in FP32 (2544 instead of 4704 bytes a bank); "
                     "same FP64 true-residual stop; opt-in, not the headline value")}
        if "fp64" in res:
            d64 = res["fp64"]["dx"]
            out["mixed_half_bank_store"]["dx_rel_diff_vs_fp64"] = float((half["dx"] - d64).norm() / d64.norm())
    print(json.dumps(out))


HYBRID_INNER_TOL = 1e-4  # mixed form of the hybrid workload: floor of the inner relative tolerance


def hybrid(a):
    """Hybrid stiff/soft workload: 64 affine cubes (8 vertices / 5 tets, 12
    DOF, kappa_abd 1e11) on a seeded jittered 4x4x4 lattice of columns over a
    soft SNK1 slab of 100x20x100 cells (214,221 vertices, 1.0M tets), E=1e5,
    nu=0.45, dt=0.005, every layer 0.5 dhat above the one below.  One step =
    one coupled Newton iteration's hot path on a fixed iterate: barrier blocks
    in vertex space mapped to DOF space (J^T H J), FEM + mass + body
    triplets, convert, MAS over the slab rows + 12x12 block-Jacobi on the
    body rows, PCG to a 1e-6 true relative residual; mixed vs FP64."""
    import torch
    from sgipc import scene, capi
    from sgipc.hybrid import HybridSystem
    hs = scene.hybrid_scene()
    s = HybridSystem(hs, "cuda:0")
    s.detect()   # LBVH broad phase + narrow phase once per iterate, outside the timed hot path
    # fp64: FP64 PCG + FP64 MAS; fp32apply: FP64 PCG with the MAS stored and
    # applied in FP32; mixed: FP32 inner PCG with FP64 refinement
    forms = {"fp64": (capi.PRECISION_FP64, capi.PRECOND_MAS),
             "fp32apply": (capi.PRECISION_FP64, capi.PRECOND_MAS_FP32),
             "mixed": (capi.PRECISION_MIXED, capi.PRECOND_MAS)}
    modes = list(forms) if a.precision == "both" else [a.precision]
    # FP32 inner solves of this stiff system stall early (body rows ~1e8
    # stiffer than the slab rows), so its mixed form restarts from the FP64
    # residual after every 1e-4 reduction instead of solving each inner
    # system to the final tolerance
    s.ctx.call("sgipc_set_inner_tol", HYBRID_INNER_TOL)
    res = {}
    for m in modes:
        s.precond = forms[m][1]
        res[m] = run_mode(s, forms[m][0], a.steps, a.warmup)
    s.precond = capi.PRECOND_MAS
    main_mode = "fp32apply" if "fp32apply" in res else modes[0]
    r = res[main_mode]
    fp32 = main_mode == "mixed"
    alg = spmv_alg_bytes(s.K, s.nbr, fp32)
    achieved = alg / (r["spmv_ms"] * 1e-3) / 1e9
    g = os.environ.get("SGIPC_SPMV_G", "8")
    g = g if g in ("4", "8", "16") else "8"
    full32 = os.environ.get("SGIPC_SPMV_FULL32", "1") != "0"
    kname = (("k_spmv_pap_full<%s, float>" if full32 else "k_spmv_pap<%s, float, float>") % g) if fp32 else \
        ("k_spmv_pap<%s, double, double>" % g)
    cpu = None
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        import abd_ref as A
        print("hybrid: CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        x_all = s.x.cpu().numpy(); x_rest = s.x_rest.cpu().numpy(); bt = s.btype.cpu().numpy()
        pairs = s.pairs.cpu().numpy().reshape(-1, 4); mi = s.mat_index.cpu().numpy()
        t0 = time.perf_counter()
        cit = oracle_hybrid_step(O, A, hs, x_all, x_rest, bt, pairs, mi, s.counts)
        csecs = time.perf_counter() - t0
        cpu = {"value": round(1.0 / csecs, 5), "unit": "Newton solves/s", "cores": 1, "kind": "port",
               "sample": "1 full coupled Newton step of the same workload (barrier, J^T H J mapping, FEM, mass, "
                         "body terms, convert, MAS + 12x12 block-Jacobi, FP64 PCG to 1e-6 true rel. residual: %d "
                         "iterations): the oracle's C restatement with its numpy affine-body restatement and a numpy "
                         "PCG loop, 1 thread, %.1f s" % (cit, csecs)}
    s.assemble()            # the mapped blocks, before a convert overwrites the store with the BCOO
    m, nb4 = s.n_mapped, 4 * s.nb
    rb, cb = s.tr[:m] < nb4, s.tc[:m] < nb4
    cls = {"body_body": int((rb & cb).sum()), "body_fem": int((rb ^ cb).sum()), "fem_fem": int((~rb & ~cb).sum())}
    out = {
        "metric": "Newton solves/sec (hybrid: assembly + J^T H J mapping + BCOO convert + MAS/ABD build + PCG to "
                  "1e-6 true rel. residual)",
        "value": round(r["value"], 4), "unit": "Newton solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(r["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if main_mode == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": "hybrid: %d affine cubes on a 100x20x100-cell slab, %d slab vertices, %d tets, "
                               "%d block rows" % (s.nb, s.nf, s.T, s.nbr),
                   "contact_pairs": int(s.n_pairs), "mapped_blocks": cls, "upper_blocks": int(s.K),
                   "precision": main_mode + (" (FP64 PCG, MAS stored and applied in FP32)"
                                             if main_mode == "fp32apply" else ""),
                   "mixed_inner_tol_floor": HYBRID_INNER_TOL,
                   "preconditioner": "MAS (16-vertex banks) on slab rows + 12x12 "
                   "block-Jacobi on body rows", "stop": "true relative residual 1e-6",
                   "parallelism": "single GPU"},
        "pcg_iters_per_step": r["iters"], "pcg_iters_per_s": round(r["pcg_iters_per_s"], 2),
        "rel_residual_max": r["rel_max"], "phase_ms": r["phase_ms"],
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_ms": round(r["spmv_ms"], 5), "launches": r["spmv_calls"], "alg_bytes": alg},
        "cpu_baseline": cpu,
    }
    for m in res:
        if m == main_mode:
            continue
        f = res[m]
        out[m] = {"value": round(f["value"], 4), "ms_per_step": round(f["ms_per_step"], 4),
                  "pcg_iters_per_step": f["iters"], "pcg_iters_per_s": round(f["pcg_iters_per_s"], 2),
                  "rel_residual_max": f["rel_max"], "phase_ms": f["phase_ms"], "spmv_avg_ms": round(f["spmv_ms"], 5)}
    if "fp64" in res and main_mode != "fp64":
        out["%s_over_fp64" % main_mode] = round(r["value"] / res["fp64"]["value"], 4)
    if "mixed" in res and "fp64" in res:
        out["mixed_over_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 4)
    if "fp64" in res:
        d64 = res["fp64"]["dx"]
        out["dx_rel_diff_vs_fp64"] = {m: float((res[m]["dx"] - d64).norm() / d64.norm()) for m in res if m != "fp64"}
    out["ms_per_step_median"] = round(r["ms_median"], 4)
    out["ms_per_step_p90"] = round(r["ms_p90"], 4)
    print(json.dumps(out))


def cantilever_adjoint(a):
    """Differentiable workload: per-tet ln E ~ N(ln 1e5, 0.2^2) on a
    60x12x12-cell cantilever clamped at x = 0 falling onto the ground
    barrier, `rollout_steps` implicit FP64 steps (Newton to round-off,
    MAS-PCG), L = |x_T - target|^2; gradient in (ln E, v0) by the
    checkpointed adjoint (every 5 steps), checked by central differences
    (h = 1e-6) along `fd_dirs` seeded unit directions, and in every direction
    by Ridders' extrapolation of central differences (the step chosen per
    direction)."""
    import torch
    from sgipc import scene
    from sgipc.sim import GroundSim, loss_and_gradient, fd_check
    sc = scene.cantilever_scene(60, 12, 12, seed=4)
    sim = GroundSim(sc, "cuda:0", newton_tol=a.newton_tol, stall_tol=1e3 * a.newton_tol, pcg_tol=a.pcg_tol,
                    adj_tol=a.pcg_tol)
    m = sc.meta
    dev = "cuda:0"
    logE = torch.tensor(m["log_E"], dtype=torch.float64, device=dev)
    v0 = torch.tensor(m["v0"], dtype=torch.float64, device=dev)
    target = torch.tensor(m["target"], dtype=torch.float64, device=dev)
    steps = a.rollout_steps
    print("cantilever_adjoint: n %d, T %d, %d steps; forward ..." % (sc.n, sc.T, steps), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xT, _ = sim.rollout(logE, v0, steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    fwd_stats = dict(sim.stats)
    contacts = sim.n_close
    print("  forward %.2f s, %s; forward + adjoint ..." % (t_fwd, fwd_stats), file=sys.stderr, flush=True)
    for k in sim.stats:
        sim.stats[k] = 0
    t0 = time.perf_counter()
    L, gE, gv = loss_and_gradient(sim, logE, v0, target, steps, ckpt=5)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    grad_stats = dict(sim.stats)
    print("  forward + adjoint %.2f s, %s; %d FD directions ..." % (t_grad, grad_stats, a.fd_dirs),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    def progress(i, fd, ad, rel):
        print("  direction %d: fd %.9e adjoint %.9e rel %.2e" % (i, fd, ad, rel), file=sys.stderr, flush=True)
    res, _ = fd_check(sim, logE, v0, target, steps, ckpt=5, n_dir=a.fd_dirs, h=1e-6, seed=0,
                      progress=progress, ridders_above=0) if a.fd_dirs else ([], L)
    t_fd = time.perf_counter() - t0
    rels = [r[2] for r in res]
    gnorm = float(torch.sqrt((gE * gE).sum() + (gv * gv).sum()))
    cpu = None
    if not a.no_cpu_baseline:
        # the oracle (FP64 C restatement, one thread): the first Newton
        # iteration of step 1 (ground barrier, SNK1 with the per-tet E, MAS-PCG
        # to the same tolerance), repeated for about 10 s
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        print("  CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        topo = oracle_topology(O, sc)
        n_it, secs, pits = 0, 0.0, 0
        while n_it < 5 and (n_it == 0 or secs < 10.0):
            t0 = time.perf_counter()
            pits = oracle_ground_step(O, sc, topo, a.pcg_tol)
            secs += time.perf_counter() - t0
            n_it += 1
        cpu = {"value": round(n_it / secs, 4), "unit": "Newton iterations/s", "cores": 1, "kind": "port",
               "sample": "%d x the first Newton iteration of step 1 (ground barrier + SNK1 assembly, convert, MAS "
                         "build, MAS-PCG to %g true rel. residual: %d iterations), oracle C restatement, 1 thread, "
                         "%.1f s" % (n_it, a.pcg_tol, pits, secs),
               "gpu_forward_newton_iterations_per_s": round(fwd_stats["newton"] / t_fwd, 3)}
    out = {
        "metric": "differentiable rollout: adjoint gradient vs central FD (max rel. error) and wall time",
        "value": max(r[4] for r in res) if res else None,
        "unit": "max relative error over FD directions (Ridders-extrapolated central differences)",
        "max_rel_err_central_h1e-6": max(rels) if rels else None,
        "higher_is_better": False,
        "n_gpus": 1,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cantilever 60x12x12 cells (%d vertices, %d tets) clamped at x=0 over a ground "
                               "barrier, %d implicit steps, dt=%g, ln E ~ N(ln 1e5, 0.2^2) per tet" %
                               (sc.n, sc.T, steps, sc.dt),
                   "checkpoint_every": 5, "fd_h": 1e-6, "fd_directions": a.fd_dirs,
                   "newton_tol": a.newton_tol, "pcg_tol": a.pcg_tol,
                   "parameters": "per-tet ln E (%d) and free initial velocities" % sc.T},
        "loss": L,
        "grad_norm_logE": float(gE.norm()),
        "grad_norm_v0": float(gv.norm()),
        "max_abs_err_over_grad_norm": max(abs(r[0] - r[1]) for r in res) / gnorm if res else None,
        "directions_within_1e-4": sum(1 for r in rels if r <= 1e-4),
        "extrapolated_fd": "every direction also by Ridders' extrapolation of central differences from h = 1e-4 "
                           "(steps halved, Neville tableau in h^2, the entry with the smallest error estimate; "
                           "the step is chosen per direction)",
        "max_rel_err_extrapolated": max(r[4] for r in res) if res else None,
        "directions_within_1e-4_extrapolated": sum(1 for r in res if r[4] <= 1e-4),
        "fd_formula": "central, (x+ - x-) . (x+ + x- - 2 target) / 2h (no cancellation of O(L) terms)",
        "forward_s": round(t_fwd, 3),
        "forward_plus_adjoint_s": round(t_grad, 3),
        "fd_s": round(t_fd, 3),
        "forward_stats": fwd_stats,
        "forward_plus_adjoint_stats": grad_stats,
        "ground_contacts_at_T": contacts,
        "fd": [dict(zip(("fd", "adjoint", "rel", "fd_extrapolated", "rel_extrapolated", "error_estimate",
                         "smallest_h"), r)) for r in res],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def batched(a):
    """Batched workload: `scenes` free boxes on a ground plane, vertex counts
    log-uniform in [500, 20000], side ratios in [1, 4], E log-uniform in
    [1e4, 1e7], nu in [0.3, 0.48], dt = 0.01.  One step = one Newton step of
    every scene (ground detection, assembly, convert, MAS build, PCG to a
    1e-6 true relative residual per scene): batched (one union system, FP32
    inner solves with per-scene FP64 refinement, schedule from the
    autotuner) against the scenes one after another in plain FP64."""
    import numpy as np
    import torch
    from sgipc import scene, capi
    from sgipc.batch import BatchedNewton, SequentialNewton
    print("batched: generating %d scenes ..." % a.scenes, file=sys.stderr, flush=True)
    scs = scene.batch_scenes(a.scenes, n_min=a.scene_vertices[0], n_max=a.scene_vertices[1])
    bn = BatchedNewton(scs, "cuda:0", precond=capi.PRECOND_MAS_FP32)
    if a.bank_store == "f16":
        bn.sys.ctx.call("sgipc_set_bank_storage", capi.BANK_STORE_F16)
    print("batched: %d vertices, %d tets; warm-up + autotune ..." % (bn.sc.n, bn.sc.T), file=sys.stderr, flush=True)
    bn.newton_step(precision=capi.PRECISION_MIXED)
    tuned = bn.autotune(capi.PRECISION_MIXED)
    tuned_form = bn.mode()
    # the solve alone (same assembled system) in each form at the tuned lanes per row
    forms = {}
    for name, pers, threads in (("sequence", False, 1024), ("persistent_512", True, 512),
                                ("persistent_1024", True, 1024)):
        bn.set_mode(pers, threads)
        bn.solve(precision=capi.PRECISION_MIXED)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            fst = bn.solve(precision=capi.PRECISION_MIXED)
        torch.cuda.synchronize()
        forms[name] = {"solve_ms": round(1e3 * (time.perf_counter() - t0) / 3, 3),
                       "all_converged": all(fst[i].converged for i in range(bn.S)),
                       "pcg_iters_max": max(fst[i].iterations for i in range(bn.S))}
        print("  %s: %.2f ms per batched solve" % (name, forms[name]["solve_ms"]), file=sys.stderr, flush=True)
    if a.batch_form == "auto":
        bn.set_mode(*tuned_form)
    else:
        bn.set_mode(a.batch_form == "persistent", tuned_form[1])
    form = bn.mode()
    for _ in range(max(a.warmup - 1, 0)):
        bn.newton_step(precision=capi.PRECISION_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st = bn.newton_step(precision=capi.PRECISION_MIXED)
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / a.steps
    its = np.array([st[i].iterations for i in range(bn.S)])
    rel = max(st[i].rel_residual for i in range(bn.S))
    conv = all(st[i].converged for i in range(bn.S))
    print("  batched %.2f ms per step; sequential FP64 ..." % (1e3 * t_b), file=sys.stderr, flush=True)
    seq = SequentialNewton(scs, "cuda:0")
    seq.newton_step()
    torch.cuda.synchronize()
    reps = max(1, min(a.steps, 3))
    t0 = time.perf_counter()
    for _ in range(reps):
        sres = seq.newton_step()
    torch.cuda.synchronize()
    t_s = (time.perf_counter() - t0) / reps
    sits = np.array([r[0] for r in sres])
    cpu = None
    if not a.no_cpu_baseline:
        # the oracle (FP64 C restatement, one thread) on the first scenes in
        # order, until about 10 s of CPU work; topology setup untimed
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        print("  CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        done, secs, nverts = 0, 0.0, 0
        for sc_i in scs:
            topo = oracle_topology(O, sc_i)
            t0 = time.perf_counter()
            oracle_ground_step(O, sc_i, topo, 1e-6)
            secs += time.perf_counter() - t0
            done += 1
            nverts += sc_i.n
            if secs >= 10.0:
                break
        cpu = {"value": round(done / secs, 4), "unit": "scene Newton solves/s", "cores": 1, "kind": "port",
               "sample": "the first %d of the %d scenes (%d vertices), one FP64 Newton step each (ground barrier + "
                         "FEM assembly, convert, MAS build, MAS-PCG to 1e-6 true rel. residual), oracle C "
                         "restatement, 1 thread, %.1f s" % (done, len(scs), nverts, secs)}
    out = {
        "metric": "batched Newton steps: scenes per second (each scene to 1e-6 true rel. residual)",
        "value": round(bn.S / t_b, 3),
        "unit": "scene Newton solves/s",
        "higher_is_better": True,
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * t_b, 3),
        "dtype": "f32/f64",
        "data": "synthetic",
        "config": {"workload": "%d free boxes on a ground plane, %d vertices, %d tets in total (per scene %d-%d "
                               "vertices), E log-U[1e4,1e7], nu U[0.3,0.48], dt=0.01" %
                               (bn.S, bn.sc.n, bn.sc.T, min(s.n for s in scs), max(s.n for s in scs)),
                   "precision": "FP32 inner PCG, per-scene FP64 refinement", "preconditioner": "MAS, 16-vertex banks",
                   "bank_inverse_storage": ("FP16 S with FP32 row scales (opt-in; read by the launch sequence, the "
                                            "persistent form reads FP32)") if a.bank_store == "f16" else "FP32",
                   "schedule": {"spmv_lanes_per_row": tuned[0], "iterations_per_graph": tuned[1],
                                "form": ("persistent launch, %d threads per workgroup" % form[1]) if form[0]
                                else "graph-captured launch sequence",
                                "autotuned_form": "persistent" if tuned_form[0] else "sequence"}},
        "batched_solve_by_form": forms,
        "pcg_iters_per_scene": {"min": int(its.min()), "median": float(np.median(its)), "max": int(its.max())},
        "rel_residual_max": rel,
        "all_converged": bool(conv),
        "upper_blocks": int(bn.sys.K),
        "spmv_alg_bytes_all_active": int((2 * bn.sys.K - bn.sc.n) * 40 + 24 * bn.sc.n),
        "spmv_alg_bytes_note": "k_bspmv_full per launch with every scene active: 40 B per entry of both triangles "
                               "(nine FP32 block values and the column; 2K - n entries), p read and Ap written per "
                               "block row",
        "sequential_fp64": {"value": round(bn.S / t_s, 3), "ms_per_pass": round(1e3 * t_s, 3),
                            "pcg_iters_per_scene": {"min": int(sits.min()), "median": float(np.median(sits)),
                                                    "max": int(sits.max())},
                            "rel_residual_max": max(r[1] for r in sres)},
        "batched_over_sequential": round(t_s / t_b, 3),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
