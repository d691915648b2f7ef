#!/usr/bin/env python
"""Benchmark: isolated build_fock + hf_energy at nk1 = nk2 = 96, nb = 16.

Workload (one "step"): one Fock build in the reference's API semantics on
seeded synthetic inputs resident in HBM -- layer-resolved real V (layer =
band mod 2, U0 = 5, q0 = 0.3, d = 0.5), Haar-random Hermitian density, all
256 (i, j) exchange channels, outputs Sigma, F and the HF energy:

    exchange kernel   P, VR -> Sigma          (fused 2D FFT convolution)
    epilogue kernel   h, Sigma, P -> F, E     (herm + energy trace, FP64 sums)

``value`` = Fock builds per second on one GPU.  ``roofline`` is for the
exchange kernel (the dominant one): algorithmic bytes per launch
c * Nk * (2 nb^2 + nb^2) (P and VR read, Sigma written) over its average
HIP-event duration inside the timed region.  ``cpu_baseline`` times the CPU
oracle (NumPy/SciPy FP64 restatement of the reference algorithm) on the same
inputs for a bounded number of builds.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = ("converged HF solves/s @96²k,nb=16, mixed vs FP64 (|ΔE|<1e-8); "
          "Fock-build GB/s")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nk", type=int, default=96)
    ap.add_argument("--nb", type=int, default=16)
    ap.add_argument("--dtype", choices=["c128", "c64"], default="c128")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in exchange-kernel variant instead of timing all of them first")
    ap.add_argument("--extra", action="store_true", help="also time the packed and complex64 forms")
    ap.add_argument("--solve", action="store_true",
                    help="also time converged CG solves (FP64 and mixed) at the same mesh")
    ap.add_argument("--solve-tol", type=float, default=1e-9)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of the exchange kernel's byte count "
                         "(hfx_lincomb out = X), to scale FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def cpu_baseline(inp, budget_s):
    """Oracle build_fock + hf_energy on the host cores, bounded to ~budget_s."""
    import oracle as O
    from oracle import hf_oracle as OH
    ncores = len(os.sched_getaffinity(0))
    OH.FFT_WORKERS = ncores
    K = O.OracleKernel(inp["weights"], inp["hamiltonian"], inp["coulomb_q"], inp["T"])
    a = K.as_args()
    P = inp["P"]

    def one():
        S, H, F = O.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=K.w2d,
                               include_exchange=True, include_hartree=False,
                               exchange_hermitian_channel_packing=False, contact_g=a["contact_g"],
                               contact_Oi=a["contact_Oi"], contact_Oj=a["contact_Oj"])
        return O.hf_energy(P, h=a["h"], Sigma=S, H=H, weights_b=K.weights_b), F

    one()
    n = 0
    t0 = time.perf_counter()
    while True:
        E, F = one()
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="Fock builds/s", cores=ncores, kind="port",
                sample=f"{n} FP64 build_fock+hf_energy calls of the NumPy/SciPy oracle at "
                       f"{P.shape[0]}x{P.shape[1]}, nb={P.shape[-1]} (scipy.fft workers={ncores})"), E, F


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; use --gpus 1"}))
        sys.exit(2)
    import torch
    import jax_hf_amd as J
    from jax_hf_amd import _lib as L
    from jax_hf_amd import synthetic as syn

    dev = torch.device("cuda:0")
    nk, nb = args.nk, args.nb
    if args.copy_calibration:
        n = nk * nk * nb * nb * 3 // 2          # complex128 values: read n*16 B, write n*16 B
        X = torch.ones(n, dtype=torch.complex128, device=dev)
        Y = torch.empty_like(X)
        for _ in range(5):
            L.call("hfx_lincomb", n, 2, 1.0, 0.0, L.ptr(X), 0.0, 0.0, None, L.ptr(Y), L.stream(dev))
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 16, "copy_calibration_bytes_written": n * 16}))
        return
    inp = syn.config1(nk=nk, nb=nb, seed=1)
    cdt = torch.complex128 if args.dtype == "c128" else torch.complex64
    npd = np.complex128 if args.dtype == "c128" else np.complex64
    G = J.HartreeFockKernel(inp["weights"], inp["hamiltonian"].astype(npd), inp["coulomb_q"], inp["T"],
                            device=dev)
    eng = G.fock_engine(herm=False)        # reference semantics: all nb^2 channels
    P = torch.from_numpy(np.ascontiguousarray(inp["P"], dtype=npd)).to(dev)
    F = torch.empty_like(P)
    Sig = torch.empty_like(P)
    E = eng.E
    st = L.stream(dev)
    csz = 16 if args.dtype == "c128" else 8
    Nk = nk * nk
    bytes_exch = csz * Nk * (3 * nb * nb)               # P + VR in, Sigma out
    bytes_build = csz * Nk * (4 * nb * nb + nb * nb)    # + h in, F out

    variant_ms = None
    if not args.no_autotune:
        best, variant_ms = eng.plan.autotune(P, eng.VR, Sig, reps=10)
    kname = {1: "exch_ct_kernel"}.get(eng.plan.exchange_variant, "exch_v2_kernel")

    def step(evs=None):
        if evs is not None:
            evs[0].record()
        eng.selfenergy(P, Sig, refP=None, rho=False)
        if evs is not None:
            evs[1].record()
        L.call("hfx_fock_epilogue", eng.plan.handle, L.ptr(eng.h), L.ptr(Sig), None, None, L.ptr(P),
               L.ptr(eng.w2d), L.ptr(F), None, L.ptr(eng.plan.scratch), L.ptr(E), 0.0, st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    exch_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    ms_step = dt * 1e3 / args.steps
    achieved = bytes_exch / (exch_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded tight-binding h, layer-resolved V, Haar density)",
        "config": {"workload": f"build_fock+hf_energy nk={nk}x{nk} nb={nb} layer-resolved V, "
                               f"all {nb * nb} channels, Sigma+F+E outputs",
                   "nk": nk, "nb": nb, "parallelism": "1 GPU"},
        "fock_build_GBps": bytes_build / (ms_step * 1e-3) / 1e9,
        "exchange_variant": L.Plan.VARIANTS[eng.plan.exchange_variant],
        "exchange_variants_ms": variant_ms,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": bytes_exch, "avg_ms": exch_ms},
        "energy": float(E),
    }
    if args.extra:
        out["extra"] = extra_forms(J, L, syn, inp, dev, nk, nb, args)
    if args.solve:
        out["solves"] = solve_forms(J, syn, inp, dev, nk, nb, args)
    if not args.no_cpu_baseline:
        cb, E_cpu, F_cpu = cpu_baseline(inp, args.cpu_seconds)
        out["cpu_baseline"] = cb
        if args.dtype == "c128":
            Fg = F.cpu().numpy()
            out["parity"] = {"E_abs_diff": abs(float(E) - E_cpu),
                             "F_max_rel": float(np.abs(Fg - F_cpu).max() / np.abs(F_cpu).max())}
    print(json.dumps(out))


def extra_forms(J, L, syn, inp, dev, nk, nb, args):
    """Packed (upper-triangle) and complex64 variants, same timing method."""
    import torch
    res = {}
    for name, npd, herm in (("c128_packed", np.complex128, True), ("c64_full", np.complex64, False),
                            ("c64_packed", np.complex64, True)):
        G = J.HartreeFockKernel(inp["weights"], inp["hamiltonian"].astype(npd), inp["coulomb_q"],
                                inp["T"], device=dev)
        eng = G.fock_engine(herm=herm)
        P = torch.from_numpy(np.ascontiguousarray(inp["P"], dtype=npd)).to(dev)
        F = torch.empty_like(P)
        for _ in range(10):
            eng.build(P, F, energy=True, refP=None)
        torch.cuda.synchronize()
        n = max(20, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(n):
            eng.build(P, F, energy=True, refP=None)
        torch.cuda.synchronize()
        res[name] = {"ms_per_build": (time.perf_counter() - t0) * 1e3 / n}
    return res


def solve_forms(J, syn, inp, dev, nk, nb, args):
    """Converged CG solves from P0 = 0 at half filling, FP64 vs mixed precision.

    Same h / layer-resolved V as the Fock-build workload, T = 0.1,
    n_electrons = nb/2 per k-point, tol_E = args.solve_tol (windowed test).
    |dE| of the mixed solve is against the FP64 GPU solve at the same tolerance.
    """
    import torch
    res = {}
    G = J.HartreeFockKernel(inp["weights"], inp["hamiltonian"], inp["coulomb_q"], 0.1, device=dev)
    P0 = np.zeros_like(inp["hamiltonian"])
    ne = nb / 2.0
    E_ref = None
    for prec in ("fp64", "mixed"):
        cfg = J.SolverConfig(max_iter=500, tol_E=args.solve_tol, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = J.solve(G, P0, ne, config=cfg)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        E = float(r.energy)
        if E_ref is None:
            E_ref = E
        res[prec] = {"solves_per_s": 1.0 / dt, "seconds": dt, "n_iter": r.n_iter, "converged": r.converged,
                     "energy": E, "abs_dE_vs_fp64": abs(E - E_ref)}
        if "phase_iters" in r.history:
            hE = r.history["E"].cpu().numpy()
            res[prec]["phase_iters"] = [int(x) for x in r.history["phase_iters"]]
            res[prec]["E_head"] = [float(x) for x in hE[:12]]
            res[prec]["E_tail"] = [float(x) for x in hE[max(r.n_iter - 6, 0):r.n_iter]]
    res["speedup_mixed_vs_fp64"] = res["fp64"]["seconds"] / res["mixed"]["seconds"]
    return res


if __name__ == "__main__":
    main()
