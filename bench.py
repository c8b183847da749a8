#!/usr/bin/env python3
"""Benchmark of the MI355X toroidal ERI engine on one GPU.

One step = one complete two-electron integral build of the workload (t-tables,
pair data, Schwarz and primitive-Schwarz passes, unique-quartet list, quartet
kernel, scatter into the SAO layout ERI(nfuc,nBas,nBas,nBas) resident in HBM),
i.e. the work of the reference's ERI_integral_toroidal
(src/Toroidal/ERI_toroidal.f90:36) for the Tori1D H64 chain (320 functions).

The start-up fast_quad self-check runs first, as in the reference
(src/calculation_type/Tori1D.f90:88): on the H chains it finds the 32-point
rule 1.9e-9 away from the 64-point table and switches the run to the table,
so both the GPU steps and the CPU baseline use the 64-point table.

Prints one JSON line.  ``value`` = nBas^4/8-equivalent unique quartets
delivered per second (the unit of the workload's "1.3e9 unique quartets");
the evaluated shell quartets per second are reported beside it.
``roofline`` prices the dominant kernel, the row MFMA kernel k_rows_mfma
(I_d(t_k) = sum_n F_n T_d[n][k] on v_mfma_f64_16x16x4f64): its algorithmic
FP64 flops (counted by the kernel itself, per row 2 sum_k min(nm, t_nmax(k))
+ 10 (nm+1)) over its duration, measured by HIP events around its launch on
the engine's stream, against MI355X's FP64 matrix peak.  ``cpu_baseline`` times the
reference's own Fortran ERI driver (oracle/_ref, OpenMP) on this host's cores.

Beside the timed ERI steps the line reports the energy path on the same
workload: RHF (k-space, J/K from the resident ERI) and k-point MP2 wall times;
``learned`` repeats ERI + RHF + MP2 with the learned per-quartet orders (build
time, energy differences to the plain path, speed-ups); ``df_mp2`` runs the
density-fitted path (41 even-tempered auxiliary functions per H): df_init,
DF-RHF, DF-kMP2, and the DF correlation energy on the exact orbitals against the
exact k-MP2 (and in FP32 on the matrix cores); ``learned_direct`` runs the
integral-direct ERI + RHF + MP2 of the Tori2D 8x8 H2 lattice (or, with
--direct-workload h2_4x4x4_d, of the Tori3D 4x4x4 H2 + d lattice) plain, plain on
the converged (128-point) quadrature, and learned (per-quartet orders,
incremental Fock builds with density screening).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def _traffic_from_profiles(workload: str):
    """HBM bytes per quartet-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", f"k_rows_mfma_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(sysm, quad, seconds: float):
    """The reference's own ERI driver (Fortran + OpenMP, oracle/_ref) on a
    bounded sample: repeated full builds of the same workload for ~seconds,
    on the same quadrature rule as the GPU (quad = (fast, nquad, t, w))."""
    fast, nquad, tq, wq = quad
    try:
        from oracle import ref
    except Exception:
        ref = None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    flags = "-O2"
    if ref is not None and ref.fast_available():
        # the reference built with its own optimisation flags (its Makefile:37)
        ref.use_fast_build()
        flags = "-O3 -funroll-loops -ftree-vectorize -ffast-math -march=x86-64-v3"
    if ref is not None and ref.available():
        r = ref.RefSystem(sysm, nq=nquad, fast_quad=fast)
        n, t0 = 0, time.perf_counter()
        while True:
            r.eri_full()
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds or el / n * (n + 1) > 3 * seconds:
                break
        nb = sysm.nbas
        return {"value": n * nb ** 4 / 8 / el, "unit": "unique quartets/s (nBas^4/8 per build)",
                "cores": threads, "kind": "reference",
                "sample": f"{n} full ERI builds of {sysm.kind} {sysm.nbas}-function workload, {nquad}-point "
                          f"{'fast' if fast else 'table'} rule "
                          f"(reference Fortran ERI_integral_toroidal, amdflang {flags}, OpenMP, {el:.1f} s)",
                "seconds_per_build": el / n}
    from oracle import orc
    from william_ci_amd.basis import normalize
    o = orc.OracleSystem(sysm, normalize(sysm), nq=nquad, nodes=None if fast else (tq, wq))
    n, t0 = 0, time.perf_counter()
    while True:
        o.eri_full()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    nb = sysm.nbas
    return {"value": n * nb ** 4 / 8 / el, "unit": "unique quartets/s (nBas^4/8 per build)",
            "cores": threads, "kind": "port",
            "sample": f"{n} full ERI builds (C restatement of the reference driver, OpenMP, {el:.1f} s)",
            "seconds_per_build": el / n}


def _learned_leg(args, s, coef, eng, S, T, V, res, mp2, plain_total, ms_step_plain):
    """The same ERI + RHF + MP2 with learned per-quartet orders (the committed
    network): build time, energies against the plain path, speed-up."""
    from william_ci_amd import learned, scf
    model = learned.Model.load()
    model.install(eng)
    for _ in range(args.warmup):
        eng.refresh_basis(coef)
        learned.eri_build_learned(eng, model, to_host=False)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.refresh_basis(coef)
        _, counts = learned.eri_build_learned(eng, model, to_host=False)
    step = (time.perf_counter() - t0) / args.steps
    tm = eng.times()
    t1 = time.perf_counter()
    rl = scf.rhf_sao_gpu(s, eng, S, T, V)
    t2 = time.perf_counter()
    ml = eng.kmp2(rl.mo_coef, rl.mo_energies, rl.occupied)
    t3 = time.perf_counter()
    total = step + (t2 - t1) + (t3 - t2)
    nord = len(model.orders)
    # learned skipping (bound head at 1e-13): build time, quartets skipped, energy
    learned.set_skip(eng, 1e-13)
    for _ in range(args.warmup):
        eng.refresh_basis(coef)
        learned.eri_build_learned(eng, model, to_host=False)
    t4 = time.perf_counter()
    for _ in range(args.steps):
        eng.refresh_basis(coef)
        _, counts_sk = learned.eri_build_learned(eng, model, to_host=False)
    step_sk = (time.perf_counter() - t4) / args.steps
    rs = scf.rhf_sao_gpu(s, eng, S, T, V)
    learned.set_skip(eng, 0.0)
    # spot check of a learned build on the host copy of its tensor
    eng.refresh_basis(coef)
    E_l, _ = learned.eri_build_learned(eng, model, to_host=True)
    spot = learned.spot_check(eng, E_l, n=256)
    del E_l
    # the plain path at the learned path's accuracy: the 128-point table (the
    # converged quadrature the network reproduces), ERI + RHF + MP2 timed alike
    quad = eng.quadrature()
    eng.set_quadrature(False, 128)
    for _ in range(max(1, args.warmup)):
        eng.refresh_basis(coef)
        eng.eri_build(to_host=False)
    t5 = time.perf_counter()
    for _ in range(args.steps):
        eng.refresh_basis(coef)
        eng.eri_build(to_host=False)
    step128 = (time.perf_counter() - t5) / args.steps
    t6 = time.perf_counter()
    r128 = scf.rhf_sao_gpu(s, eng, S, T, V)
    t7 = time.perf_counter()
    m128 = eng.kmp2(r128.mo_coef, r128.mo_energies, r128.occupied)
    t8 = time.perf_counter()
    total128 = step128 + (t7 - t6) + (t8 - t7)
    eng.set_quadrature(bool(quad[0]), int(quad[1]))
    eng.refresh_basis(coef)
    return {"orders": list(model.orders), "margin": model.margin,
            "shell_quartets_per_rule": {"plain": int(counts[nord]),
                                        **{str(o): int(c) for o, c in zip(model.orders, counts[:nord])}},
            "spot_check_max_abs": spot, "spot_check_rule": "40-point mapped Gauss-Legendre (converged)",
            "skipping": {"threshold": 1e-13, "schwarz_cap": 1e-10, "margin_b": 1.5, "skipped_shell_quartets": int(counts_sk[nord + 1]),
                         "ms_per_step": step_sk * 1e3, "rhf_energy": rs.energy, "dE_rhf": rs.energy - res.energy},
            "ms_per_step": step * 1e3, "step_ms": {k: round(v, 3) for k, v in tm.items()},
            "speedup_eri": ms_step_plain / step,
            "rhf_energy": rl.energy, "rhf_iterations": rl.iterations, "mp2_ec": ml["ec"],
            "dE_rhf": rl.energy - res.energy, "dE_total": (rl.energy + ml["ec"]) - (res.energy + mp2["ec"]),
            "eri_rhf_mp2_wall_s": total, "speedup_eri_rhf_mp2": plain_total / total,
            "plain_128": {"ms_per_step": step128 * 1e3, "rhf_energy": r128.energy, "mp2_ec": m128["ec"],
                          "eri_rhf_mp2_wall_s": total128},
            "speedup_eri_vs_plain_128": step128 / step, "speedup_eri_rhf_mp2_vs_plain_128": total128 / total,
            "dE_total_vs_plain_128": (rl.energy + ml["ec"]) - (r128.energy + m128["ec"])}


DIRECT_WORKLOADS = {"h2_8x8": "Tori2D 8x8 H2", "h2_4x4x4_d": "Tori3D 4x4x4 H2 + d"}


def _learned_direct_leg(workload: str = "h2_8x8"):
    """Integral-direct ERI + RHF + MP2 on a lattice of H2 (default the Tori2D
    8x8 lattice, 640 functions; --direct-workload h2_4x4x4_d: the Tori3D 4x4x4
    lattice with d functions, 1408 functions, a few minutes), the path where the
    integrals are recomputed every SCF iteration: direct RHF then the direct
    k-MP2 (no stored tensor), three ways:
      plain      the reference's rule after its start-up self-check (the 64-point
                 table here), every quartet at every iteration (direct_JK);
      plain_128  the same on the 128-point table (the converged quadrature the
                 network is trained to reproduce: the plain path at the learned
                 path's accuracy);
      learned    per-quartet learned orders and the incremental Fock build with
                 density screening at 1e-14 (bound: the Schwarz product);
    and learned_bound_screen: the learned path with the network's integral bound
    in the density screening."""
    from william_ci_amd import learned, scf
    from william_ci_amd import system as tsys
    from william_ci_amd.engine import Engine, normalized_coef
    s = tsys.workload(workload)
    eng = Engine(s, normalized_coef(s))
    sc = eng.fast_quad_self_check(64)
    S, T, V = eng.one_electron()
    model = learned.Model.load()
    model.install(eng)

    def run(setup, source, label):
        # (a progress line per path on stderr: the 3D lattice's paths take minutes)
        print(f"bench: direct leg {workload}: {label}", file=sys.stderr, flush=True)
        setup()
        t0 = time.perf_counter()
        r = scf.rhf_sao_gpu(s, eng, S, T, V, source=source)
        t1 = time.perf_counter()
        m = eng.kmp2_direct(r.mo_coef, r.mo_energies, r.occupied)
        t2 = time.perf_counter()
        info = eng.direct_screen_info()
        return {"rhf_energy": r.energy, "iterations": r.iterations, "rhf_s": t1 - t0, "mp2_ec": m["ec"],
                "mp2_s": t2 - t1, "e_total": r.energy + m["ec"], "wall_s": t2 - t0,
                "quartets_evaluated": info["kept_total"], "quartets_listed": info["listed_total"]}
    counts = {}

    def learned_setup():
        counts["c"] = learned.direct_init_learned(eng, model)
        eng.set_direct_screen(1e-14, 50, False)
    def learned_bound_setup():
        learned.direct_init_learned(eng, model)
        eng.set_direct_screen(1e-14, 50, True)
    if workload == "h2_8x8":
        run(lambda: eng.direct_init(), "direct", "warm-up")                    # first-call buffers
        plain = run(lambda: eng.direct_init(), "direct", "plain")
        lrn = run(learned_setup, "direct_incremental", "learned")
    else:
        # (minutes per path: no warm-up; the learned path runs first and carries the
        # first-call allocations)
        lrn = run(learned_setup, "direct_incremental", "learned")
        plain = run(lambda: eng.direct_init(), "direct", "plain")
    # the same learned path with the network's integral bound in the density screening
    # (learned Schwarz-style skipping of the incremental builds)
    lrn_b = run(learned_bound_setup, "direct_incremental", "learned, learned bound")
    quad = eng.quadrature()
    eng.set_quadrature(False, 128)
    p128 = run(lambda: eng.direct_init(), "direct", "plain, 128-point table")
    eng.close()
    c = counts["c"]
    nord = len(model.orders)
    total = int(c.sum())
    return {"workload": f"{workload} ({DIRECT_WORKLOADS[workload]}, {s.nbas} functions)",
            "self_check": sc, "plain_rule": f"{quad[1]}-point {'fast' if quad[0] else 'table'}",
            "plain": plain, "plain_128": p128, "learned": lrn, "learned_bound_screen": lrn_b,
            "learned_orders": {"plain_rule": int(c[nord]), **{str(o): int(v) for o, v in zip(model.orders, c[:nord])}},
            "fallback_fraction": int(c[nord]) / max(total, 1),
            "speedup_vs_plain": plain["wall_s"] / lrn["wall_s"],
            "speedup_vs_plain_128": p128["wall_s"] / lrn["wall_s"],
            "dE_total_vs_plain": lrn["e_total"] - plain["e_total"],
            "dE_total_vs_plain_128": lrn["e_total"] - p128["e_total"],
            "dE_total_plain_vs_plain_128": plain["e_total"] - p128["e_total"],
            "speedup_vs_plain_128_learned_bound": p128["wall_s"] / lrn_b["wall_s"],
            "dE_total_vs_plain_128_learned_bound": lrn_b["e_total"] - p128["e_total"]}


def _df_leg(s, eng, S, T, V, res, mp2):
    """Density-fitted MP2 on the same chain (keyword density_fitting with the
    41-function even-tempered AuxBasis): df_init, DF-RHF, DF-kMP2; and DF-kMP2
    on the exact-integral orbitals against the exact k-MP2."""
    from william_ci_amd import df, scf
    aux = df.aux_list(s, df.read_aux(df.AUX_H_ETB41))
    t0 = time.perf_counter()
    d = df.DF(eng, aux)
    t1 = time.perf_counter()
    njk, tjk = 0, 0.0

    def jk(P):
        nonlocal njk, tjk
        ta = time.perf_counter()
        out = d.jk(P)
        tjk += time.perf_counter() - ta
        njk += 1
        return out
    rd = scf.rhf_sao(s, S, T, V, jk)
    t2 = time.perf_counter()
    md = d.kmp2(rd.mo_coef, rd.mo_energies, rd.occupied)
    t3 = time.perf_counter()
    on_exact = d.kmp2(res.mo_coef, res.mo_energies, res.occupied)
    d.kmp2(res.mo_coef, res.mo_energies, res.occupied, fp32=True)      # first call: the FP32 path's buffers
    fp32 = d.kmp2(res.mo_coef, res.mo_energies, res.occupied, fp32=True)
    info, times = d.info(), d.times()
    d.close()
    return {"aux": "AuxBasis_H_etb41 (s 8: 0.1-40, p 5: 0.2-10, d 3: 0.3-3; 41 per H)",
            "nau": info["nau"], "naux": info["naux"], "dropped": info["dropped"],
            "df_init_s": t1 - t0, "df_init_ms": {k: round(v, 3) for k, v in times.items()},
            "rhf_energy": rd.energy, "rhf_iterations": rd.iterations, "rhf_wall_s": t2 - t1,
            "df_jk_ms": tjk / max(njk, 1) * 1e3,
            "kmp2_ec": md["ec"], "kmp2_gpu_ms": md["gpu_ms"], "kmp2_wall_s": t3 - t2,
            "e_total": rd.energy + md["ec"], "df_init_rhf_mp2_wall_s": t3 - t0,
            "ec_on_exact_orbitals": on_exact["ec"], "ec_exact": mp2["ec"],
            "fit_error_ec": on_exact["ec"] - mp2["ec"],
            "fp32": {"kind": "FP32 L, (ov|ov) as FP32 matrix-core GEMMs, compensated FP32 energy sums",
                     "ec": fp32["ec"], "gpu_ms": fp32["gpu_ms"], "ec_fp64": on_exact["ec"], "gpu_ms_fp64": on_exact["gpu_ms"],
                     "dE_vs_fp64": fp32["ec"] - on_exact["ec"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="h64_chain")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-self-check", action="store_true",
                    help="keep the 32-point fast rule (skip the start-up fast_quad self-check)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-order leg")
    ap.add_argument("--no-df", action="store_true", help="skip the density-fitted MP2 leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the learned integral-direct leg")
    ap.add_argument("--direct-workload", default="h2_8x8", choices=sorted(DIRECT_WORKLOADS),
                    help="lattice of the integral-direct leg (h2_4x4x4_d: a few minutes)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = (1 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the bra-pair task list over GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    from william_ci_amd import system as tsys
    from william_ci_amd.engine import normalized_coef
    from william_ci_amd.engine import Engine

    s = tsys.workload(args.workload)
    coef = normalized_coef(s)                   # normalize_AO_toroidal on the GPU
    eng = Engine(s, coef)
    # start-up self-check of the fast rule (the reference runs it before any
    # integral, src/calculation_type/Tori1D.f90:88): it keeps the 32-point rule
    # or switches the whole run to the 64-point table
    t_sc = time.perf_counter()
    selfcheck = None if args.no_self_check else eng.fast_quad_self_check(64)
    t_sc = time.perf_counter() - t_sc
    quad = eng.quadrature()
    for _ in range(args.warmup):
        eng.refresh_basis(coef)
        eng.eri_build(to_host=False)
    kms, steps = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        eng.refresh_basis(coef)               # every step rebuilds tables and pair data from the AO list
        eng.eri_build(to_host=False)          # synchronous: returns after the GPU work completed
        steps.append(time.perf_counter() - ts)
        kms.append(eng.times()["rows_kernel_ms"])
    el = time.perf_counter() - t0
    st = eng.stats()
    tm = eng.times()

    # the energy path on the same workload: one-electron integrals + RHF (Bloch
    # basis, J/K from the resident ERI on the GPU); the ERI build is the last step's
    from william_ci_amd import scf
    t_one = time.perf_counter()
    S, T, V = eng.one_electron()
    t_one = time.perf_counter() - t_one
    # untimed: first-call costs (column lists of the resident tensor, work buffers)
    warm = scf.rhf_sao_gpu(s, eng, S, T, V)
    eng.kmp2(warm.mo_coef, warm.mo_energies, warm.occupied)
    t_rhf = time.perf_counter()
    res = scf.rhf_sao_gpu(s, eng, S, T, V)                               # the whole iteration on the GPU
    rhf_wall = time.perf_counter() - t_rhf
    # untimed cross-check: the host loop (numpy algebra around the GPU J/K)
    t_host = time.perf_counter()
    res_host = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P))
    host_wall = time.perf_counter() - t_host
    # MP2 on the RHF orbitals (SAO run: k-point MP2 from the resident ERI,
    # main.f90:823-847 with transform_ERI_AO_to_kMO + kMP2_calc)
    t_mp2 = time.perf_counter()
    mp2 = eng.kmp2(res.mo_coef, res.mo_energies, res.occupied)
    mp2_wall = time.perf_counter() - t_mp2
    plain_total = ms_step_plain = el / args.steps
    plain_total += rhf_wall + mp2_wall
    learned_part = None if args.no_learned else _learned_leg(args, s, coef, eng, S, T, V, res, mp2, plain_total,
                                                             ms_step_plain)
    df_part = None if args.no_df else _df_leg(s, eng, S, T, V, res, mp2)
    eng.close()
    direct_part = None if args.no_direct else _learned_direct_leg(args.direct_workload)

    nb = s.nbas
    unique = nb ** 4 / 8
    ms_per_step = el / args.steps * 1e3
    kernel_ms = sum(kms) / len(kms)
    achieved = st["row_flops"] / (kernel_ms * 1e-3) / 1e12
    traffic = _traffic_from_profiles(args.workload)
    line = {
        "metric": "Clifford ERI quartets/s (FP64 ERI build)",
        "value": unique * args.steps / el,
        "unit": "unique quartets/s (nBas^4/8 per build)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (deterministic H chain, no external data)",
        "config": {"workload": args.workload, "kind": s.kind, "nbas": nb, "nfuc": s.nfuc,
                   "cells": s.n_cell, "L": s.L[0], "nquad": quad[1],
                   "quadrature": ("fast_quad %d-point mapped Gauss-Legendre" % quad[1]) if quad[0]
                   else "%d-point table (fast_quad self-check: above 1e-10)" % quad[1],
                   "layout": "SAO ERI(nfuc,nBas^3)", "parallelism": "single GPU"},
        "self_check": None if selfcheck is None else dict(selfcheck, seconds=t_sc),
        "shell_quartets_per_s": st["shell_quartets"] * args.steps / el,
        "evaluated": {"shell_quartets": st["shell_quartets"], "prim_quartets": st["prim_quartets"],
                      "rows": st["rows"]},
        "step_ms": {k: round(v, 3) for k, v in tm.items()},
        "rhf": {"energy": res.energy, "converged": res.converged, "iterations": res.iterations,
                "orbital_error": res.orbital_error, "wall_s": rhf_wall, "one_electron_s": t_one,
                "jk_s": res.times["jk_s"], "eri_plus_rhf_wall_s": rhf_wall + ms_per_step / 1e3,
                "loop": "tcg_rhf_sao (J/K, lattice transforms, per-k Jacobi, DIIS, density on the GPU)",
                "loop_ms": res.times["gpu_loop_s"] * 1e3,
                "host_loop_check": {"energy": res_host.energy, "dE": res.energy - res_host.energy,
                                    "iterations": res_host.iterations, "wall_s": host_wall}},
        "mp2": {"kind": "k-point MP2 (transform_ERI_AO_to_kMO + kMP2_calc)", "ec": mp2["ec"],
                "e_total": res.energy + mp2["ec"], "zero_denominators": mp2["zero_denominators"],
                "gpu_ms": mp2["gpu_ms"], "wall_s": mp2_wall,
                "eri_rhf_mp2_wall_s": ms_per_step / 1e3 + rhf_wall + mp2_wall},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "k_rows_mfma", "kernel_ms": kernel_ms, "flops_per_launch": st["row_flops"],
                     "all_quartet_flops": st["flops"]},
    }
    line["learned"] = learned_part
    line["learned_direct"] = direct_part
    line["df_mp2"] = df_part
    if not args.no_cpu_baseline:
        cb = cpu_baseline(s, quad, args.cpu_seconds)
        cb["gpu_speedup"] = line["value"] / cb["value"]
        line["cpu_baseline"] = cb
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
