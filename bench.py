"""Benchmark: Kohn-Sham SCF steps/s and H|psi> GB/s on one MI355X.

Workload (default): 32 H2O at 1.0 g/cm^3 in an 18.63-bohr cubic box, GTH-PBE,
70 Ry (100^3 grid, 31,995 half-sphere plane waves), 128 doubly occupied
bands; mixed precision (FP32 FFT/H|psi>, FP64 overlap accumulation, Cholesky,
energies).  One "step" is one iteration of the reference's wfn_opt loop
(density -> V_KS -> energies -> H|psi> -> Lagrange -> PCG -> ortho) on
device-resident wavefunctions.  Synthetic data: seeded water lattice and
seeded random initial coefficients.

Prints one JSON line: the steps/s figure with the step time, the H|psi>
roofline (algorithmic bytes per launch over the HIP-event launch time,
against 8 TB/s), the per-region and per-GEMM times, the FP64 comparison,
the 64 H2O figures and the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCF steps/sec + H|ψ⟩ GB/s on 32/64 H2O; mixed vs FP64 GPU at 1e-6 Ha/atom"
HBM_PEAK = 8000.0  # GB/s, MI355X HBM3E
HPSI_KERNELS = ("rows_inv", "cols_inv", "z_vpsi", "cols_fwd", "rows_fwd_hpsi")
CONFIGS = {
    "h2o_32_pbe": "32 H2O PBE 70 Ry, 18.63 bohr, 128 bands, 100^3",
    "h2o_64_pbe": "64 H2O PBE 80 Ry, 23.47 bohr, 256 bands, 144^3",
    "h2o_128_pbe": "128 H2O PBE 80 Ry, 29.57 bohr, 512 bands, 180^3",
    "h2o_32_ref": "32 H2O of the reference's recorded run: MT O/H, Teter93, 80 Ry, 18.61 bohr, 128 bands, 108^3",
}
# The reference's recorded run of that system: SCF step times (ms) in the
# last column of tests/32_water/wf_opt/SP_ENERGY.dat, 48 MPI ranks.
REF32_S_PER_STEP = 0.597
REF32_NOTE = "tests/32_water/wf_opt/SP_ENERGY.dat: ~597 ms per SCF step on 48 MPI ranks (its step-time column)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hpsi_h2o32_pbe_mixed_pmc_traffic.md")


def pmc_traffic(config, precision):
    """HBM bytes per launch measured with rocprofv3 PMC counters (FETCH_SIZE,
    WRITE_SIZE, calibrated against a plain copy) in a separate profiling run
    of this benchmark; read from the committed summary.  None for workloads
    that have no such measurement."""
    if config != "h2o_32_pbe" or precision != "mixed" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    text = open(TRAFFIC_PROFILE).read()
    i = text.find("```json")
    if i < 0:
        return None
    j = text.find("```", i + 7)
    return json.loads(text[i + 7:j])


def hpsi_bytes(pw, nb, prec):
    """Algorithmic bytes of one local H|psi> application (all bands):
    (nb/2) * (36 N_r + 48 N_pw) in FP32, doubled in FP64."""
    s = 1 if prec == "mixed" else 2
    return (nb / 2.0) * (36.0 * pw.nr + 48.0 * pw.ngpw) * s


def kernel_bytes(ctx, pw, nb, prec):
    """Algorithmic bytes per launch of each H|psi> pass (all pairs in one launch):
    the sphere coefficients are FP64 (16 B), the row/plane intermediates are in
    the FFT precision."""
    cb = 8 if prec == "mixed" else 16
    rb = cb // 2
    n1, n2, n3 = pw.n
    geo = ctx.sphere_geometry()
    npair = (nb + 1) // 2
    pg = 4 if prec == "mixed" else 2
    plane = geo["nplane"] * n1 * n2 * cb
    rows = geo["nslot"] * n1 * cb
    return {
        "rows_inv": npair * (2 * pw.ngpw * 16 + rows),
        "cols_inv": npair * (rows + plane),
        "z_vpsi": npair * (2 * plane) + ((npair + pg - 1) // pg) * pw.nr * rb,
        "cols_fwd": npair * (plane + rows),
        "rows_fwd_hpsi": npair * (rows + 4 * pw.ngpw * 16),
    }


def run_gpu(sysobj, prec, steps, warmup, c0, profile_steps=3, tfor_steps=0, ortho="cholesky"):
    import torch

    from amdkiit_amd.context import KSContext

    s = sysobj
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    ctx.set_ortho_method(ortho)
    c = ctx.ortho(c0)
    ctx.wfn_load(c, s.occ)
    it = 0
    for _ in range(warmup):
        it += 1
        ctx.scf_step(it, s.xc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = None
    for _ in range(steps):
        it += 1
        e = ctx.scf_step(it, s.xc)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # per-kernel HIP-event timing on the library stream, separate pass
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(profile_steps):
        it += 1
        ctx.scf_step(it, s.xc)
    ctx.profile(False)
    prof = {}
    for name in ctx.profile_names():
        tot, cnt = ctx.profile_query(name)
        prof[name] = (tot, cnt)
    if tfor_steps:
        # forces-on steps (the TFOR branch: E_SR + POTFOR + NLFOR inside the step)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(tfor_steps):
            it += 1
            ctx.scf_step_tfor(it, s.xc, True)
        torch.cuda.synchronize()
        prof["_tfor_ms_per_step"] = ((time.perf_counter() - t1) / tfor_steps * 1e3, tfor_steps)
    return ctx, dt, e, prof, it


def copy_ceiling(nbytes=1 << 30, reps=10):
    """Measured device-to-device copy rate (read + write bytes / time)."""
    import torch

    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (time.perf_counter() - t0) / 1e9
    del x, y
    return gbs


MFMA_PEAK = {"f64": 78.6, "f32": 157.3}  # TFLOP/s, MI355X dense


def gemm_table(prof, pw, nb, nproj, prec, nsteps):
    """Achieved TFLOP/s of the band-space GEMMs of one SCF step (real form,
    K = 2 N_pw reals; updates run over the padded leading dimension)."""
    ldc = (pw.ngpw + 63) // 64 * 64
    f32 = "f32" if prec == "mixed" else "f64"
    shapes = {  # region: (M, N, K, arithmetic)
        "fnl": (nproj, nb, 2 * pw.ngpw, f32),
        "fnonloc": (nproj, nb, 2 * ldc, f32),
        "lagrange_ovl": (nb, nb, 2 * pw.ngpw, f32),
        "lagrange_upd": (nb, nb, 2 * ldc, f32),
        "ortho_ovl": (nb, nb, 2 * pw.ngpw, "f64"),
        "ortho_trmm": (nb, nb, 2 * ldc, "f64"),
    }
    # the orthonormalisation computes only the tiles the product needs: the
    # 64 x 64 overlap tiles on and above the diagonal (S is symmetric) and,
    # in C U^-1, the 32-row x 16-column MFMA tiles not below U^-1's diagonal
    tn = (nb + 63) // 64
    pairs = done = 0
    for n0 in range(0, nb, 128):
        for mb in range(0, (nb + 31) // 32 * 32, 32):
            for t in range(8):
                if n0 + 16 * t < nb:
                    pairs += 1
                    done += t >= min(8, max(0, int((mb - n0) / 16)))
    executed = {"ortho_ovl": tn * (tn + 1) / 2 / tn ** 2, "ortho_trmm": done / pairs}
    out = {}
    for k, (M, N, K, ar) in shapes.items():
        if k not in prof or not prof[k][1] or M == 0:
            continue
        ms = prof[k][0] / prof[k][1]
        tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        out[k] = {"M": M, "N": N, "K": K, "arith": ar, "avg_ms": ms, "TFLOPs": tf,
                  "peak": MFMA_PEAK[ar], "frac": tf / MFMA_PEAK[ar]}
        if k in executed:  # TFLOPs above counts the full product
            out[k]["executed_fraction"] = executed[k]
            out[k]["frac_executed"] = tf * executed[k] / MFMA_PEAK[ar]
    return out


def cpu_baseline(s, c0, seconds_hint, cores=None):
    """FP64 numpy/scipy restatement of the same SCF step on the host cores
    (all of them, or `cores` threads for FFT and BLAS)."""
    from threadpoolctl import threadpool_limits

    from oracle.ks_oracle import KSOracle

    ncores = cores or os.cpu_count() or 1
    with threadpool_limits(limits=ncores):
        orc = KSOracle(s.pw, s.tables, s.sp, s.tau, s.xc, workers=ncores)
        c = c0.copy()
        t0 = time.perf_counter()
        c, _ = orc.scf_step(c, s.occ, 1)  # warm-up (plans, caches)
        t1 = time.perf_counter()
        n = 0
        while True:
            c, _ = orc.scf_step(c, s.occ, 2 + n)
            n += 1
            if time.perf_counter() - t1 > seconds_hint or n >= 3:
                break
        dt = (time.perf_counter() - t1) / n
    return {"value": 1.0 / dt, "unit": "SCF steps/s", "cores": ncores, "kind": "port",
            "sample": f"{n} SCF step(s) of the same workload after 1 warm-up step, FP64 numpy/scipy "
                      f"(scipy.fft workers={ncores}, BLAS threads={ncores})",
            "s_per_step": dt, "warmup_s": t1 - t0}


def measure_secondary(cfg_name, sysobj, steps, warmup, seed, bf16=False):
    """Mixed and FP64 steps/s of another workload from one seeded start (and
    with bf16, the BF16 recipe with FP64 Loewdin orthonormalisation)."""
    s2 = sysobj
    rng2 = np.random.default_rng(seed)
    c02 = (rng2.random((s2.nbands, s2.pw.ngpw)) - 0.5 + 1j * (rng2.random((s2.nbands, s2.pw.ngpw)) - 0.5)) / (1.0 + s2.pw.g2_pw)
    c02[:, 0] = c02[:, 0].real
    cx, d2, e2, p2, _ = run_gpu(s2, "mixed", steps, warmup, c02, profile_steps=2)
    hp2 = sum(p2[k][0] / p2[k][1] for k in HPSI_KERNELS if k in p2 and p2[k][1])
    hb2 = hpsi_bytes(s2.pw, s2.nbands, "mixed")
    g2 = gemm_table(p2, s2.pw, s2.nbands, cx.nproj, "mixed", 2)
    cx.close()
    c64, d64, e64b, p64, _ = run_gpu(s2, "fp64", steps, warmup, c02, profile_steps=1)
    c64.close()
    rb = None
    if bf16:
        cb, db, eb, pb, _ = run_gpu(s2, "bf16", steps, warmup, c02, profile_steps=1, ortho="lowdin")
        cb.close()
        rb = {"recipe": "BF16 projector/Lagrange products with one refinement term, FP64 Loewdin",
              "steps_per_s": steps / db, "ms_per_step": db / steps * 1e3, "speedup_vs_fp64": d64 / db,
              "etot_diff_per_atom_vs_fp64": abs(eb["etot"] - e64b["etot"]) / s2.nat,
              "ms_by_region": {k: v[0] for k, v in sorted(pb.items(), key=lambda x: -x[1][0])},
              "note": "same seeded start and step count; the Loewdin path differs from the Cholesky one "
                      "step by step and meets it at convergence"}
    return {"workload": cfg_name, "steps": steps, "grid": list(s2.pw.n), "ngpw": s2.pw.ngpw, "bands": s2.nbands,
            "bf16_lowdin": rb,
            "mixed_steps_per_s": steps / d2, "mixed_ms_per_step": d2 / steps * 1e3,
            "fp64_steps_per_s": steps / d64, "fp64_ms_per_step": d64 / steps * 1e3, "mixed_speedup": d64 / d2,
            "hpsi_ms": hp2, "hpsi_GBps": hb2 / (hp2 * 1e-3) / 1e9,
            "hpsi_frac_hbm": hb2 / (hp2 * 1e-3) / 1e9 / HBM_PEAK,
            "etot_diff_per_atom": abs(e2["etot"] - e64b["etot"]) / s2.nat,
            "mixed_ms_by_region": {k: v[0] / 2 for k, v in sorted(p2.items(), key=lambda x: -x[1][0])},
            "mixed_gemm_mfma": g2,
            "fp64_ms_by_region": {k: v[0] for k, v in sorted(p64.items(), key=lambda x: -x[1][0])}}


def run_md(cfg, prec, nsteps, dt_fs=0.5, temperature=300.0, gcut=1e-6):
    """Born-Oppenheimer NVE MD (velocity Verlet) driven by the device SCF and
    the adjoint forces; E_cons drift in Ha/atom/ps."""
    from amdkiit_amd import drivers as D
    from amdkiit_amd import systems
    from amdkiit_amd.context import KSContext

    s = systems.CONFIGS[cfg]()
    rng = np.random.default_rng(7)
    c0 = (rng.random((s.nbands, s.pw.ngpw)) - 0.5 + 1j * (rng.random((s.nbands, s.pw.ngpw)) - 0.5)) / (1.0 + s.pw.g2_pw)
    c0[:, 0] = c0[:, 0].real
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c = ctx.ortho(c0)
    prov = D.GPUForceProvider(ctx, s, c, gcut=gcut)
    m = np.array([D.MASS_AMU[s.species[i].symbol] for i in s.sp]) * D.AMU_AU
    v = D.maxwell_boltzmann(m, temperature, seed=2024)
    t0 = time.perf_counter()
    h, drift, _, _ = D.bomd(prov, s.tau.copy(), m, v, dt_fs=dt_fs, nsteps=nsteps,
                            log=lambda i, e, ek, ec: log(f"[bench]   md {i} E_cons {ec:.8f}") if i % 25 == 0 else None)
    wall = time.perf_counter() - t0
    ctx.close()
    return {"config": cfg, "precision": prec, "steps": nsteps, "dt_fs": dt_fs, "T0_K": temperature,
            "scf_gcut": gcut, "drift_Ha_per_atom_ps": drift,
            "max_dev_Econs_Ha_per_atom": float(np.abs(h[:, 2] - h[0, 2]).max() / s.nat),
            "scf_steps_first": prov.nscf[0], "scf_steps_per_md_step": float(np.mean(prov.nscf[1:])),
            "wall_s": wall, "epot": h[:, 0].tolist()}


def run_lbfgs(cfg, prec, nsteps, gcut=3e-5):
    """SCF to GEMAX < gcut, then L-BFGS on the GPU adjoint forces."""
    from amdkiit_amd import drivers as D
    from amdkiit_amd import systems
    from amdkiit_amd.context import KSContext

    s = systems.CONFIGS[cfg]()
    rng = np.random.default_rng(9)
    c0 = (rng.random((s.nbands, s.pw.ngpw)) - 0.5 + 1j * (rng.random((s.nbands, s.pw.ngpw)) - 0.5)) / (1.0 + s.pw.g2_pw)
    c0[:, 0] = c0[:, 0].real
    ctx = KSContext(s.pw, s.nbands, prec)
    ctx.set_species(s.tables)
    ctx.set_positions(s.sp, s.tau)
    c = ctx.ortho(c0)
    prov = D.GPUForceProvider(ctx, s, c, gcut=gcut, maxit=6000)
    t0 = time.perf_counter()
    _, hist, _ = D.lbfgs(prov, s.tau.copy(), gtol=1e-4, maxiter=nsteps,
                         log=lambda i, e, fm: log(f"[bench]   lbfgs {i} E {e:.8f} max|g| {fm:.2e}"))
    wall = time.perf_counter() - t0
    ctx.close()
    return {"config": cfg, "precision": prec, "lbfgs_steps": len(hist) - 1, "scf_gcut": gcut,
            "energy_force_history": hist, "scf_steps_total": int(np.sum(prov.nscf)),
            "scf_steps_first": prov.nscf[0], "evaluations": len(prov.nscf), "wall_s": wall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="h2o_32_pbe", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64", "bf16"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 64 H2O secondary measurement")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--md-steps", type=int, default=0, help="also run BO-MD (both precisions) on --md-config")
    ap.add_argument("--md-config", default="h2o_64_pbe")
    ap.add_argument("--lbfgs-steps", type=int, default=0, help="also run L-BFGS (mixed) on --lbfgs-config")
    ap.add_argument("--lbfgs-config", default="h2o_128_pbe")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 512 MiB device-to-device copy (PMC byte-counter calibration)")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = 512 << 20
        x = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 5}))
        return
    if a.gpus != 1:
        print(f"bench.py: this implementation runs on one GPU only (asked for --gpus {a.gpus}); "
              "splitting the bands over GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)

    from amdkiit_amd import systems
    from amdkiit_amd.pwgrid import PWSetup  # noqa: F401

    log(f"[bench] building {a.config}")
    s = systems.CONFIGS[a.config]()
    rng = np.random.default_rng(2024)
    g2 = s.pw.g2_pw
    c0 = (rng.random((s.nbands, s.pw.ngpw)) - 0.5 + 1j * (rng.random((s.nbands, s.pw.ngpw)) - 0.5)) / (1.0 + g2)
    c0[:, 0] = c0[:, 0].real
    log(f"[bench] n={s.pw.n} ngpw={s.pw.ngpw} nb={s.nbands} nat={s.nat}")

    ctx, dt, e, prof, it = run_gpu(s, a.precision, a.steps, a.warmup, c0, tfor_steps=min(a.steps, 10))
    tfor_ms = prof.pop("_tfor_ms_per_step")[0]
    ms = dt / a.steps * 1e3
    value = a.steps / dt
    log(f"[bench] {a.precision}: {ms:.3f} ms/step  E={e['etot']:.8f}  gemax={e['gemax']:.3e}; "
        f"forces-on step {tfor_ms:.3f} ms")
    copy_gbs = copy_ceiling()
    log(f"[bench] device copy ceiling {copy_gbs:.0f} GB/s")

    kb = kernel_bytes(ctx, s.pw, s.nbands, a.precision)
    hp_ms = sum(prof[k][0] / prof[k][1] for k in HPSI_KERNELS if k in prof and prof[k][1])
    hb = hpsi_bytes(s.pw, s.nbands, a.precision)
    hp_gbs = hb / (hp_ms * 1e-3) / 1e9
    per_kernel = {}
    for k in HPSI_KERNELS:
        if k in prof and prof[k][1]:
            avg = prof[k][0] / prof[k][1]
            per_kernel[k] = {"avg_ms": avg, "alg_bytes": kb[k], "GBps": kb[k] / (avg * 1e-3) / 1e9}
    step_breakdown = {k: v[0] / 3 for k, v in sorted(prof.items(), key=lambda x: -x[1][0])}
    dominant = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"])
    gemms = gemm_table(prof, s.pw, s.nbands, ctx.nproj, a.precision, 3)
    traffic = pmc_traffic(a.config, a.precision)
    if traffic:
        for k in per_kernel:
            tb = traffic["per_launch_bytes"].get("k_" + k + "_r")
            per_kernel[k]["hbm_bytes_pmc"] = tb

    fp64 = None
    ctx.close()
    if not a.no_fp64 and a.precision == "mixed":
        ctx64, dt64, e64, prof64, _ = run_gpu(s, "fp64", a.steps, a.warmup, c0, profile_steps=1)
        fp64 = {"steps_per_s": a.steps / dt64, "ms_per_step": dt64 / a.steps * 1e3,
                "mixed_speedup": value / (a.steps / dt64),
                "etot_fp64": e64["etot"],
                "etot_diff_per_atom": abs(e["etot"] - e64["etot"]) / s.nat,
                "note": "same seeded start and step count (1 optimisation) in both precisions"}
        ctx64.close()
        log(f"[bench] fp64: {fp64['ms_per_step']:.3f} ms/step; mixed speedup {fp64['mixed_speedup']:.2f}x; "
            f"|dE|/atom={fp64['etot_diff_per_atom']:.2e}")

    secondary = sec128 = ref32 = None
    if not a.no_secondary and a.config == "h2o_32_pbe":
        log("[bench] secondary: 64 H2O")
        secondary = measure_secondary(CONFIGS["h2o_64_pbe"], systems.CONFIGS["h2o_64_pbe"](),
                                      max(5, a.steps // 2), a.warmup, 2025, bf16=True)
        log(f"[bench] 64 H2O: mixed {secondary['mixed_ms_per_step']:.2f} ms/step, fp64 speedup "
            f"{secondary['mixed_speedup']:.2f}x, H|psi> {secondary['hpsi_GBps']:.0f} GB/s")
        log("[bench] secondary: 128 H2O")
        sec128 = measure_secondary(CONFIGS["h2o_128_pbe"], systems.CONFIGS["h2o_128_pbe"](), 5, 2, 2026)
        log(f"[bench] 128 H2O: mixed {sec128['mixed_ms_per_step']:.2f} ms/step, fp64 speedup "
            f"{sec128['mixed_speedup']:.2f}x, H|psi> {sec128['hpsi_GBps']:.0f} GB/s")
        log("[bench] secondary: the reference's recorded 32 H2O run (MT/Teter93, 80 Ry, 108^3)")
        ref32 = measure_secondary(CONFIGS["h2o_32_ref"], systems.reference_water(32), 10, a.warmup, 2027)
        ref32["reference_s_per_step"] = REF32_S_PER_STEP
        ref32["reference_run"] = REF32_NOTE
        ref32["vs_reference_run"] = REF32_S_PER_STEP / (ref32["mixed_ms_per_step"] * 1e-3)
        log(f"[bench] ref 32 H2O: mixed {ref32['mixed_ms_per_step']:.3f} ms/step, fp64 "
            f"{ref32['fp64_ms_per_step']:.3f} ms/step (recorded reference run {REF32_S_PER_STEP} s/step)")

    md = None
    if a.md_steps > 0:
        md = {}
        for prec in ("mixed", "fp64"):
            log(f"[bench] BO-MD {a.md_config} {prec} {a.md_steps} steps ...")
            md[prec] = run_md(a.md_config, prec, a.md_steps)
            log(f"[bench]   drift {md[prec]['drift_Ha_per_atom_ps']:.2e} Ha/atom/ps, "
                f"{md[prec]['wall_s'] / a.md_steps * 1e3:.1f} ms/MD step")
        ep = np.array(md["mixed"]["epot"]) - np.array(md["fp64"]["epot"])
        md["max_abs_dEpot_mixed_vs_fp64_Ha_per_atom"] = float(np.abs(ep).max() / len(systems.CONFIGS[a.md_config]().tau))
    lb = None
    if a.lbfgs_steps > 0:
        log(f"[bench] L-BFGS {a.lbfgs_config} {a.lbfgs_steps} steps ...")
        lb = run_lbfgs(a.lbfgs_config, "mixed", a.lbfgs_steps)
        log(f"[bench]   E {lb['energy_force_history'][0][0]:.6f} -> {lb['energy_force_history'][-1][0]:.6f}, "
            f"{lb['scf_steps_total']} SCF steps in {lb['wall_s']:.1f} s")

    cpu = None
    if not a.no_cpu:
        log("[bench] cpu baseline (FP64 port) ...")
        cpu = cpu_baseline(s, c0, a.cpu_seconds)
        log(f"[bench] cpu: {cpu['s_per_step']:.2f} s/step on {cpu['cores']} threads")
        one = cpu_baseline(s, c0, 1.0, cores=1)
        log(f"[bench] cpu: {one['s_per_step']:.2f} s/step on 1 thread")
        cpu["one_core"] = {k: one[k] for k in ("value", "unit", "cores", "kind", "sample", "s_per_step")}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "SCF steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32", "fp64": "f64", "bf16": "bf16+f32"}[a.precision],
        "data": "synthetic (seeded water lattice + seeded random coefficients)",
        "config": {"workload": CONFIGS[a.config], "precision": a.precision, "grid": list(s.pw.n),
                   "ngpw": s.pw.ngpw, "bands": s.nbands, "atoms": s.nat, "xc": s.xc},
        "hpsi": {"ms": hp_ms, "alg_bytes": hb, "GBps": hp_gbs, "frac_hbm": hp_gbs / HBM_PEAK,
                 "frac_copy_ceiling": hp_gbs / copy_gbs, "launches": list(HPSI_KERNELS)},
        "forces_on_ms_per_step": tfor_ms,
        "forces_off_ms_per_step": ms,
        "hbm_copy_ceiling_GBps": copy_gbs,
        "roofline": {"bound": "hbm", "kernel": "hpsi_local (" + "+".join(HPSI_KERNELS) + ")",
                     "achieved": hp_gbs, "peak": HBM_PEAK, "unit": "GB/s", "frac": hp_gbs / HBM_PEAK,
                     "traffic": traffic["hpsi_local_bytes"] if traffic else None,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
                     "dominant_launch": dominant, "per_launch": per_kernel},
        "step_ms_by_region": step_breakdown,
        "gemm_mfma": gemms,
        "energy": {"etot": e["etot"], "gemax": e["gemax"], "nelec": e["nelec"], "after_steps": it},
        "fp64_gpu": fp64,
        "secondary_64h2o": secondary,
        "secondary_128h2o": sec128,
        "secondary_reference_32h2o": ref32,
        "cpu_baseline": cpu,
    }
    if md:
        out["bomd"] = md
    if lb:
        out["lbfgs"] = lb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
