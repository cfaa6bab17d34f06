#!/usr/bin/env python3
"""Benchmark: RI J+K Fock builds/sec, mixed precision vs FP64, on one MI355X.

Workload (large real system): B[n_aux=4000][n=1000][n=1000] symmetric, N(0,1) exp(-|mu-nu|/50),
generated on the device and resident in HBM in fp64, fp32 and bf16 copies; n_occ = 200 orthonormal
orbitals from a seeded QR; a sequence of 20 further orbital sets, each rotated from the previous
one by an angle that shrinks 10x per step (so |D_k - D_{k-1}| shrinks 10x per step).

One step = one SCF-like sequence of 22 Fock builds:
  mixed : full FP64 build at D_0, 20 incremental builds F_k = F_{k-1} + G(D_k - D_{k-1}) with bf16
          operands on the matrix cores (fp32 accumulation), final full FP64 refinement build at D_20;
  fp64  : the same 22 builds as full FP64 builds (reported beside it as fp64_builds_per_s).
value = mixed builds/sec = 22 * steps / time. |dE| and max|dF| against the FP64 build at D_20 are
reported before and after the refinement. All inputs are resident in HBM before the timed region.

cpu_baseline: the oracle's FP64 restatement of the reference CPU algorithm (two GEMV passes for J,
OpenMP static loop over the aux index with per-thread K partials, serial BLAS inside) timed on the
host on a bounded sample of the same workload (the first n_aux_sample aux rows), scaled to n_aux.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_AUX, N, N_OCC, N_DELTA = 4000, 1000, 200, 20
SEED = 20261203 + 1
PEAK = {"hbm": 8000.0, "f64": 78.6, "f32": 157.3, "bf16": 2500.0}  # GB/s, TFLOP/s (dense)


def progress(msg):
    """One line on stderr per finished step of the run (the JSON result goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def legs(a, torch, rr, ctx):
    out = {}
    for key, on, fn in (("batch_uneven_molecules", a.batch_molecules > 0, lambda: batch_leg(torch, rr, ctx, a.batch_molecules)),
                        ("sigma_trial_densities", a.sigma_densities > 0, lambda: sigma_leg(torch, rr, ctx, a.sigma_densities)),
                        ("relativistic", bool(a.relativistic), lambda: relativistic_leg(torch, rr, ctx, cpu=a.cpu_sample > 0)),
                        ("other_rows", bool(a.rows), lambda: rows_leg(torch, rr, ctx))):
        if not on:
            out[key] = None
            continue
        t0 = time.time()
        progress(f"{key}: start")
        out[key] = fn()
        progress(f"{key}: done in {time.time() - t0:.1f} s")
    return out


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--naux", type=int, default=N_AUX)
    p.add_argument("--n", type=int, default=N)
    p.add_argument("--nocc", type=int, default=N_OCC)
    p.add_argument("--delta-precision", default="bf16", choices=["bf16", "f32", "f64"])
    p.add_argument("--cpu-sample", type=int, default=1500, help="aux rows timed on the CPU (0: skip)")
    p.add_argument("--no-fp64-leg", action="store_true")
    p.add_argument("--batch-molecules", type=int, default=256,
                   help="molecules in the uneven-batch leg (0: skip); n ~ U[20,200], n_aux = 4n, n_occ = n/5")
    p.add_argument("--sigma-densities", type=int, default=64,
                   help="trial densities in the shared-tensor leg (0: skip); n = 800 (100 occ + 700 virt), n_aux = 3200")
    p.add_argument("--relativistic", type=int, default=1,
                   help="1: time the X2C (n=400, n_aux=2400) and 4c RKB (nL=300, nS=700, n_aux=4200) builds")
    p.add_argument("--rows", type=int, default=1,
                   help="1: time the other entry points (RKB from a density, Cholesky builds, MO transforms, "
                        "functional-stage products, metric) at mid-size shapes")
    p.add_argument("--copy-probe-gib", type=float, default=0.0,
                   help="only run a device copy of this many GiB (calibrates HBM byte counters) and exit")
    return p.parse_args()


def orbital_sequence(n, nocc, nsteps, seed):
    from rerdmft_amd import synthetic as syn
    g = syn.rng(seed)
    q, _ = np.linalg.qr(g.standard_normal((n, n)))
    cs = [np.ascontiguousarray(q[:, :nocc])]
    angle = 0.1
    basis = q.copy()
    for k in range(nsteps):
        c = syn.rotate_occ(cs[-1], basis, angle, seed + 1 + k)
        # keep an orthonormal completion of the new occupied space for the next rotation
        proj = basis[:, nocc:] - c @ (c.T @ basis[:, nocc:])
        v, _ = np.linalg.qr(proj)
        basis = np.concatenate([c, v[:, : n - nocc]], axis=1)
        cs.append(c)
        angle *= 0.1
    return cs


def copy_probe(torch, gib):
    """A plain device-to-device copy of known size: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it
    gives the factor between the counters and the bytes actually moved."""
    nbytes = int(gib * (1 << 30)) // 16 * 16
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}))
    return 0


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (two rocprofv3 --pmc
    runs, FETCH_SIZE and WRITE_SIZE, each scaled by the copy-probe calibration), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_large_real.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel + ":":
            return float(parts[1])
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (the path would shard the "
              "aux index over GPUs with one all-reduce of J and K per build).", file=sys.stderr)
        return 2
    import torch
    if a.copy_probe_gib > 0:
        return copy_probe(torch, a.copy_probe_gib)
    import rerdmft_amd as rr

    dev = torch.device("cuda:0")
    n, naux, nocc = a.n, a.naux, a.nocc
    ctx = rr.Context(0)
    t0 = time.time()
    ctx.generate(0, naux, n, 1.0, 50.0, SEED, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    cs = orbital_sequence(n, nocc, N_DELTA, SEED + 7)
    c_dev = [torch.tensor(c, device=dev) for c in cs]
    l_dev = [np.sqrt(2.0) * c for c in c_dev]
    d_dev = [2.0 * (c @ c.T) for c in c_dev]
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    dprec = {"bf16": rr.BF16, "f32": rr.FP32, "f64": rr.FP64}[a.delta_precision]

    J = torch.empty((n, n), dtype=torch.float64, device=dev)
    K = torch.empty((n, n), dtype=torch.float64, device=dev)
    dJ = torch.empty_like(J)
    dK = torch.empty_like(K)

    def full_build(k, prec=rr.FP64):
        # D_k = L_k L_k^T: the Coulomb contraction is fused into the exchange half-transform
        ctx.ri_jk_factor(0, None, l_dev[k], precision=prec, j_out=J, k_out=K)
        return J - 0.5 * K  # two-electron part of F (h = 0)

    def mixed_sequence(prec=None):
        prec = dprec if prec is None else prec
        f = full_build(0)
        for k in range(1, N_DELTA + 1):
            ctx.ri_jk_delta(0, c_dev[k], c_dev[k - 1], precision=prec, dj_out=dJ, dk_out=dK)
            f = f + (dJ - 0.5 * dK)
        pre = f
        f = full_build(N_DELTA)
        return pre, f

    def fp64_sequence():
        f = None
        for k in range(N_DELTA + 2):
            f = full_build(min(k, N_DELTA))
        return f

    builds = N_DELTA + 2
    for _ in range(a.warmup):
        mixed_sequence()
    torch.cuda.synchronize()
    ctx.timing(enable=True, reset=True)
    t_start = time.perf_counter()
    for _ in range(a.steps):
        pre, f_mixed = mixed_sequence()
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t_start
    phases_p = ctx.timing(enable=False, reset=True, by_precision=True)
    phases = {}
    for (pn, ph), (ms_, c_) in phases_p.items():
        phases.setdefault(ph, [0.0, 0])
        phases[ph][0] += ms_
        phases[ph][1] += c_

    # accuracy at D_20 against the FP64 build
    f_ref = full_build(N_DELTA).clone()
    d20 = d_dev[N_DELTA]
    de_pre = 0.5 * float(torch.sum(d20 * (pre - f_ref)))
    de_fin = 0.5 * float(torch.sum(d20 * (f_mixed - f_ref)))
    df_pre = float(torch.max(torch.abs(pre - f_ref)))
    df_fin = float(torch.max(torch.abs(f_mixed - f_ref)))
    e_ref = 0.5 * float(torch.sum(d20 * f_ref))

    fp64_rate = None
    fp64_phases = None
    if not a.no_fp64_leg:
        fp64_sequence()
        torch.cuda.synchronize()
        ctx.timing(enable=True, reset=True)
        t1 = time.perf_counter()
        fp64_sequence()
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t1
        fp64_phases = ctx.timing(enable=False, reset=True, by_precision=True)
        fp64_rate = builds / t_fp64

    rate = builds * a.steps / t_mixed

    # Every scheme of incremental builds: rate of the same sequence, and the accuracy of its
    # accumulated Fock matrix at D_20 before the final FP64 build (|dE| = |1/2 tr(D (F - F_64))|).
    schemes = {}
    for pname, prec in (("bf16", rr.BF16), ("f32", rr.FP32)):
        mixed_sequence(prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pre_s, _ = mixed_sequence(prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        schemes[pname] = {"builds_per_s": builds / dt,
                          "dE_before_refinement": 0.5 * float(torch.sum(d20 * (pre_s - f_ref))),
                          "max_dF_before_refinement": float(torch.max(torch.abs(pre_s - f_ref)))}
    progress("incremental schemes timed")

    # The reference's density-taking build recovers the occupied orbitals from D (FP64, n x n):
    # time it on the device (rr_nonrel_fock_ri at D_20, phase "recovery").
    eye = torch.eye(n, dtype=torch.float64, device=dev)
    hz = torch.zeros((n, n), dtype=torch.float64, device=dev)
    fo = torch.empty_like(hz)
    ctx.nonrel_fock_ri(0, hz, eye, eye, d_dev[N_DELTA], out=fo)
    torch.cuda.synchronize()
    ctx.timing(enable=True, reset=True)
    for _ in range(3):
        ctx.nonrel_fock_ri(0, hz, eye, eye, d_dev[N_DELTA], out=fo)
    torch.cuda.synchronize()
    rec_ph = ctx.timing(enable=False, reset=True)
    gpu_recovery_ms = rec_ph["recovery"][0] / max(1, rec_ph["recovery"][1])
    progress("device recovery timed")

    # Roofline of the dominant kernel of the mixed sequence: the (phase, precision) with the largest
    # device time, from HIP events recorded on the library stream around each launch.
    rp = (nocc + 7) // 8 * 8
    prec_bytes = {"f64": 8, "f32": 4, "bf16": 2}
    n_full = 2 * a.steps
    n_delta = N_DELTA * a.steps

    def algorithmic(phase, pname, launches_full, launches_delta):
        """(work per launch, unit, bound) for a phase; full builds use rank n_occ, incremental 2 n_occ."""
        nb = prec_bytes[pname]
        if phase == "exchange_half":
            # flops 2 n_aux n^2 r; HBM bytes: B read once + W written once. The binding roofline is
            # the one with the larger time at peak (BF16 sits just past the ridge: HBM-bound).
            r = nocc * (2 if launches_delta else 1)
            flops = 2.0 * naux * n * n * r
            nbytes = (naux * n * n + naux * n * r) * nb
            if nbytes / (PEAK["hbm"] * 1e9) > flops / (PEAK[pname] * 1e12):
                return nbytes, "GB/s", "hbm"
            return flops, "TFLOP/s", "mfma"
        if phase == "exchange_rank_k":
            # full build: symmetric K, n(n+1)/2 unique elements; incremental: G (non-symmetric) n^2
            if launches_delta:
                return 2.0 * naux * n * n * nocc, "TFLOP/s", "mfma"
            return 1.0 * naux * n * (n + 1) * nocc, "TFLOP/s", "mfma"
        if phase == "coulomb":
            # symmetric tensor: its unique upper triangle is read once
            return 0.5 * naux * n * (n + 1) * nb, "GB/s", "hbm"
        if phase == "gamma":
            return 1.0 * naux * n * nocc * (2 if launches_delta else 1) * nb, "GB/s", "hbm"
        return None, None, None

    dom_key, (dom_ms, dom_cnt) = max(((k, v) for k, v in phases_p.items() if v[1] > 0), key=lambda kv: kv[1][0])
    dom_phase_prec = dom_key[0]
    dom_phase = dom_key[1]
    is_delta = dom_phase_prec != "f64"  # low-precision launches are the incremental builds
    work, unit, bound = algorithmic(dom_phase, dom_phase_prec, 0 if is_delta else n_full, n_delta if is_delta else 0)
    avg_ms = dom_ms / dom_cnt
    if unit == "TFLOP/s":
        achieved = work / (avg_ms * 1e-3) / 1e12
        peak = PEAK[dom_phase_prec]
    else:
        achieved = work / (avg_ms * 1e-3) / 1e9
        peak = PEAK["hbm"]
    roof = {"kernel": f"{dom_phase}[{dom_phase_prec}]", "bound": bound, "achieved": achieved, "peak": peak,
            "unit": unit, "frac": achieved / peak, "traffic": read_traffic(f"{dom_phase}[{dom_phase_prec}]"), "avg_launch_ms": avg_ms,
            "algorithmic_per_launch": work, "launches": dom_cnt}

    # per-kernel rooflines of the plain FP64 build (for the record)
    fp64_kernels = {}
    if fp64_phases:
        for k2, (ms2, c2) in fp64_phases.items():
            if c2 == 0 or k2[0] != "f64":
                continue
            avg = ms2 / c2
            w2, u2, _ = algorithmic(k2[1], "f64", 1, 0)
            if w2 is None:
                fp64_kernels[k2[1]] = {"avg_ms": avg}
            elif u2 == "TFLOP/s":
                fp64_kernels[k2[1]] = {"avg_ms": avg, "TFLOP/s": w2 / (avg * 1e-3) / 1e12,
                                       "frac": w2 / (avg * 1e-3) / 1e12 / PEAK["f64"]}
            else:
                fp64_kernels[k2[1]] = {"avg_ms": avg, "GB/s": w2 / (avg * 1e-3) / 1e9,
                                       "frac": w2 / (avg * 1e-3) / 1e9 / PEAK["hbm"]}

    progress(f"headline sequence timed ({rate:.1f} builds/s)")
    # The Coulomb stream alone (J only from a density; in the FP64 builds it overlaps the rank-k
    # update on a second stream, so its in-build time above is shared): unique upper triangle of B
    # read once, device time of the stream kernel from the library's phase events.
    coul = {}
    for pname, prec_c, nb in (("f64", rr.FP64, 8), ("bf16", rr.BF16, 2)):
        ctx.ri_jk_factor(0, d_dev[0], None, precision=prec_c, want_k=False, j_out=J)
        ctx.timing(enable=True, reset=True)
        for _ in range(5):
            ctx.ri_jk_factor(0, d_dev[0], None, precision=prec_c, want_k=False, j_out=J)
        torch.cuda.synchronize()
        ph = ctx.timing(enable=False, reset=True)
        if "coulomb" in ph and ph["coulomb"][1]:
            ms_c = ph["coulomb"][0] / ph["coulomb"][1]
            gbs = 0.5 * naux * n * (n + 1) * nb / (ms_c * 1e-3) / 1e9
            coul[pname] = {"avg_ms": ms_c, "GB/s": gbs, "frac": gbs / PEAK["hbm"]}
    cpu = None
    if a.cpu_sample > 0:
        cpu = cpu_baseline(n, nocc, min(a.cpu_sample, naux), naux, cs[0])
        progress("CPU baseline timed")

    out = {
        "metric": "RI J+K Fock builds/sec (mixed vs FP64) at |dE|<=1e-8 Ha; % HBM/MFMA peak",
        "value": rate,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_mixed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{a.delta_precision}+f64",
        "data": "synthetic",
        "config": {"workload": "real RI J/K, B[4000][1000][1000] sym N(0,1)exp(-|mu-nu|/50), n_occ=200, "
                               "20 incremental builds (|dD| /10 per step) + FP64 refinement",
                   "n_aux": naux, "n": n, "n_occ": nocc, "builds_per_step": builds,
                   "delta_precision": a.delta_precision},
        "fp64_builds_per_s": fp64_rate,
        "speedup_vs_fp64": (rate / fp64_rate) if fp64_rate else None,
        "incremental_schemes": {k: {**v, "speedup_vs_fp64": (v["builds_per_s"] / fp64_rate) if fp64_rate else None}
                                for k, v in schemes.items()},
        "gpu_recovery_ms": gpu_recovery_ms,
        "value_with_recovery": builds * a.steps / (t_mixed + 2 * a.steps * gpu_recovery_ms * 1e-3),
        "fp64_builds_per_s_with_recovery": (1.0 / (1.0 / fp64_rate + gpu_recovery_ms * 1e-3)) if fp64_rate else None,
        "scf_check": scf_check(rr, ctx),
        "dE_after_refinement": de_fin,
        "dF_after_refinement": df_fin,
        "dE_before_refinement": de_pre,
        "dF_before_refinement": df_pre,
        "E_two_electron": e_ref,
        "roofline": roof,
        "fp64_build_kernels": fp64_kernels,
        "coulomb_standalone": coul,
        "phase_ms_mixed": {f"{k[1]}[{k[0]}]": round(v[0], 3) for k, v in phases_p.items() if v[1]},
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    out.update(legs(a, torch, rr, ctx))
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def scf_check(rr, ctx):
    """Mixed-precision SCF (incremental BF16 builds on FP64 full builds, FP64 final iterations) against
    the all-FP64 SCF through nonRelFockMatrix's RI path: water dimer in the s-type model basis."""
    from rerdmft_amd import guess, molecule, scf
    w = molecule.WATER_GEOMETRY
    mol = molecule.Molecule(np.concatenate([molecule.WATER_CHARGES] * 2),
                            np.concatenate([w, w + np.array([0.0, 0.0, 5.6])]), molecule.WATER_BASIS,
                            molecule.WATER_AUX)
    tol = dict(diis_size=5, energy_tolerance=1e-11, density_tolerance=1e-8, max_iterations=80)
    ref, ints, _ = guess.run_scf(ctx, mol, slot=12, copies=rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16, **tol)
    full, delta = scf.gpu_builds(ctx, 12, rr.BF16)
    r = scf.nonrel_scf_incremental(full, delta, ints["h"], ints["X"], ints["S"], mol.n_electrons, ints["e_nuc"],
                                   **tol)
    ctx.release(12)
    return {"system": f"water dimer, s-type model basis (n={mol.n}, n_aux={mol.n_aux})",
            "fp64_scf_builds": ref.iterations, "mixed_full_builds": r.full_builds, "mixed_bf16_builds": r.delta_builds,
            "E_fp64": ref.total_energy, "dE_mixed_vs_fp64": r.total_energy - ref.total_energy,
            "converged": bool(ref.converged and r.converged)}


def batch_leg(torch, rr, ctx, nmol, reps=5):
    """Uneven batch: nmol molecules, n ~ U[20,200], n_aux = 4n, n_occ = n/5; J/K of D = L L^T for every
    molecule in one HIP-graph launch (FP64). The first call runs the batch eagerly and captures it."""
    g = np.random.default_rng(20261203 + 5)
    ns = [int(x) for x in g.integers(20, 201, nmol)]
    dev = torch.device("cuda:0")
    tens = []
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261208)
    for n in ns:
        t = torch.randn((4 * n, n, n), device=dev, dtype=torch.float64, generator=gen)
        tens.append(0.5 * (t + t.transpose(1, 2)))
    batch = ctx.batch_create(tens, rr.COPY_FP64 | rr.COPY_FP32)
    del tens
    ls = [torch.randn((n, max(1, n // 5)), device=dev, dtype=torch.float64, generator=gen) for n in ns]
    js = [torch.empty((n, n), device=dev, dtype=torch.float64) for n in ns]
    ks = [torch.empty((n, n), device=dev, dtype=torch.float64) for n in ns]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.batch_jk_factor(batch, ls, outs=(js, ks))  # eager run + capture
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    ctx.batch_jk_factor(batch, ls, outs=(js, ks))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _, _, rep = ctx.batch_jk_factor(batch, ls, outs=(js, ks))
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    flops = sum(2.0 * (4 * n) * n * n * max(1, n // 5) + (4 * n) * n * (n + 1) * max(1, n // 5) for n in ns)
    # FP32-MFMA mode (the bulk builds of a mixed SCF, FP64 for the final ones): the same batch from
    # the FP32 copies, its own graph; errors against the FP64 results above
    j64 = [j.clone() for j in js]
    k64 = [k.clone() for k in ks]
    js32 = [torch.empty_like(j) for j in js]
    ks32 = [torch.empty_like(k) for k in ks]
    ctx.batch_jk_factor(batch, ls, precision=rr.FP32, outs=(js32, ks32))
    ctx.batch_jk_factor(batch, ls, precision=rr.FP32, outs=(js32, ks32))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.batch_jk_factor(batch, ls, precision=rr.FP32, outs=(js32, ks32))
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / reps
    dj = max(float(torch.max(torch.abs(a - b)) / torch.max(torch.abs(b))) for a, b in zip(js32, j64))
    dk = max(float(torch.max(torch.abs(a - b)) / torch.max(torch.abs(b))) for a, b in zip(ks32, k64))
    ctx.batch_destroy(batch)
    return {"molecules": nmol, "n_range": [min(ns), max(ns)], "dtype": "f64",
            "builds_per_s_graph": nmol / t_graph, "ms_per_batch_graph": 1e3 * t_graph,
            "ms_first_call_eager_plus_capture": 1e3 * t_first, "replayed": bool(rep),
            "exchange_TFLOP_per_s": flops / t_graph / 1e12,
            "f32": {"builds_per_s_graph": nmol / t32, "ms_per_batch_graph": 1e3 * t32,
                    "exchange_TFLOP_per_s": flops / t32 / 1e12,
                    "max_rel_dJ_vs_f64": dj, "max_rel_dK_vs_f64": dk}}


def sigma_leg(torch, rr, ctx, nd, reps=2, slot=14):
    """Shared-tensor batch: nd orbital-Hessian trial densities D_k = C_occ T_k C_virt^T (T_k N(0,1),
    n_occ = 100, n_virt = 700, n = 800, n_aux = 3200) built in one graph-captured call per precision;
    J_k and K_k of every density. Inputs and outputs stay in HBM."""
    n, no, naux = 800, 100, 3200
    dev = torch.device("cuda:0")
    ctx.generate(slot, naux, n, 1.0, 40.0, SEED + 7, rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16)
    gen = torch.Generator(device=dev)
    gen.manual_seed(20261209)
    q, _ = torch.linalg.qr(torch.randn((n, n), device=dev, dtype=torch.float64, generator=gen))
    c_occ, c_virt = q[:, :no].contiguous(), q[:, no:].contiguous()
    t = torch.randn((nd, no, n - no), device=dev, dtype=torch.float64, generator=gen)
    rfacs = torch.matmul(c_virt.unsqueeze(0), t.transpose(1, 2)).contiguous()  # R_k = C_virt T_k^T
    js = torch.empty((nd, n, n), device=dev, dtype=torch.float64)
    ks = torch.empty((nd, n, n), device=dev, dtype=torch.float64)
    out = {"densities": nd, "n": n, "n_occ": no, "n_aux": naux}
    ref = None
    for pname, prec in (("f64", rr.FP64), ("f32", rr.FP32), ("bf16", rr.BF16)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.ri_jk_multi(slot, c_occ, rfacs, precision=prec, j_out=js, k_out=ks)  # eager run + capture
        torch.cuda.synchronize()
        t_first = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(reps):
            _, _, rep = ctx.ri_jk_multi(slot, c_occ, rfacs, precision=prec, j_out=js, k_out=ks)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res = {"builds_per_s": nd / dt, "ms_per_call": 1e3 * dt, "ms_first_call_eager_plus_capture": 1e3 * t_first,
               "replayed": bool(rep)}
        if ref is None:
            ref = (js.clone(), ks.clone())
        else:
            res["max_rel_dJ_vs_f64"] = float((js - ref[0]).abs().max() / ref[0].abs().max())
            res["max_rel_dK_vs_f64"] = float((ks - ref[1]).abs().max() / ref[1].abs().max())
        out[pname] = res
    # exchange flops per density: half-transform of R_k and the rank-k product (B L once per call)
    out["exchange_TFLOP_per_density"] = 2 * 2.0 * naux * n * n * no / 1e12
    ctx.release(slot)
    return out


def _timed(torch, fn, reps):
    fn()  # warm-up (first call of a shape: workspace sizing, tile autotuning)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def _phases(torch, ctx, fn):
    """Device time per phase (ms) of one call of fn."""
    ctx.timing(enable=True, reset=True)
    fn()
    torch.cuda.synchronize()
    ph = ctx.timing(enable=False, reset=True)
    return {k: round(v[0], 3) for k, v in ph.items() if v[1]}


def _phase_rates(ph, g1, rk):
    """Achieved TFLOP/s of the half-transform and the rank-k update from their device times (ms)."""
    out = {}
    for key, fl in (("exchange_half", g1), ("exchange_rank_k", rk)):
        if ph.get(key):
            out[key] = round(fl / ph[key] / 1e9, 1)
    return out


def relativistic_cpu(x2c_shape, rkb_shape):
    """The oracle's X2C and RKB builds timed on the host over sampled aux rows (same sizes otherwise):
    X2C at two sample sizes, extrapolated linearly to the full aux count (the occupied recovery is a
    per-build constant); RKB (no recovery) scaled by n_aux / sample."""
    from oracle import oracle as orc
    from rerdmft_amd import synthetic as syn
    orc.lib()
    out = {}
    nl, naux, npairs = x2c_shape
    g = np.random.default_rng(SEED + 30)
    q = syn.kramers_pair_occ(nl, npairs, SEED + 31)
    p = q @ q.conj().T
    eye = np.eye(2 * nl, dtype=np.complex128)
    h = np.zeros((2 * nl, 2 * nl), dtype=np.complex128)
    times = {}
    for ns in (24, 96):
        b = g.standard_normal((ns, nl * nl))
        orc.x2c_fock_ri(b[:2], eye, eye, h, p)  # warm-up
        t0 = time.perf_counter()
        orc.x2c_fock_ri(b, eye, eye, h, p)
        times[ns] = time.perf_counter() - t0
    per_row = (times[96] - times[24]) / 72
    t_full = times[24] + per_row * (naux - 24)
    out["x2c"] = {"value": 1.0 / t_full, "unit": "Fock builds/s", "cores": orc.max_threads(), "kind": "port",
                  "sample": f"x2cFockMatrix over 24 and 96 of {naux} aux rows ({times[24]:.2f} s, {times[96]:.2f} s), "
                            f"extrapolated linearly to {naux}"}
    nl, ns_, naux, npairs = rkb_shape
    n = 2 * nl + 2 * ns_
    nsamp = 48
    bll = g.standard_normal((nsamp, nl * nl))
    bss = g.standard_normal((nsamp, ns_ * ns_))
    c, _ = np.linalg.qr(g.standard_normal((n, 2 * npairs)) + 1j * g.standard_normal((n, 2 * npairs)))
    d = c @ c.conj().T
    occ = np.ones(2 * npairs)
    orc.ri_fock_two_electron_direct(nl, ns_, bll[:2], bss[:2], d, c, occ)  # warm-up
    t0 = time.perf_counter()
    orc.ri_fock_two_electron_direct(nl, ns_, bll, bss, d, c, occ)
    dt = time.perf_counter() - t0
    out["rkb"] = {"value": nsamp / (dt * naux), "unit": "Fock builds/s", "cores": orc.max_threads(), "kind": "port",
                  "sample": f"riFockTwoElectronDirect over {nsamp} of {naux} aux rows ({dt:.2f} s), scaled by {naux}/{nsamp}"}
    return out


def rows_leg(torch, rr, ctx, reps=3):
    """Per-call device time of the remaining entry points at mid-size shapes: inputs and outputs are
    device tensors (no host transfer inside the timed calls), FP64, one warm-up call each. A row whose
    call raises is reported with its error instead of a time."""
    from rerdmft_amd import synthetic as syn
    dev = torch.device("cuda:0")
    D = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    out = {}

    def row(name, fn, flop=None, **info):
        progress(f"row {name}")
        try:
            ms = 1e3 * _timed(torch, fn, reps)
            r = {"ms_per_call": round(ms, 3), **info}
            if flop:
                r["TFLOP_per_s"] = round(flop / ms / 1e9, 1)
            out[name] = r
        except Exception as e:  # reported, not fatal: the other rows still run
            out[name] = {"error": f"{type(e).__name__}: {e}"[:300], **info}

    # a4: RKB from a density (recovery in the RKB basis, V transforms, 4-flavor build)
    nl, ns, naux, nocc = 150, 350, 1200, 120
    try:
        ctx.generate(5, naux, nl, 0.4, 20.0, SEED + 31, rr.COPY_FP64)
        ctx.generate(6, naux, ns, 0.2, 40.0, SEED + 32, rr.COPY_FP64)
        n_ukb, n_rkb = 2 * nl + 2 * ns, 4 * nl
        g = syn.rng(SEED + 33)
        v = (g.standard_normal((n_ukb, n_rkb)) + 1j * g.standard_normal((n_ukb, n_rkb))) * 0.3
        sr = v.conj().T @ v + 0.5 * np.eye(n_rkb)
        w, u = np.linalg.eigh(sr)
        x = u / np.sqrt(w)[None, :]
        c = x @ syn.unitary(n_rkb, SEED + 34)[:, :nocc]
        pd = (c * np.linspace(1.0, 0.6, nocc)) @ c.conj().T
        h = syn.hermitian(n_rkb, SEED + 35)
        args = [D(a) for a in (h, v, sr, x, pd)]
        fo = torch.empty((n_rkb, n_rkb), dtype=torch.complex128, device=dev)
        row("a4_rkb_fock_ri", lambda: ctx.rkb_fock_ri(5, 6, nl, ns, *args, out=fo),
            shape=f"nL={nl} nS={ns} n_aux={naux} n_occ={nocc}")
    except Exception as e:
        out["a4_rkb_fock_ri"] = {"error": f"{type(e).__name__}: {e}"[:300]}
    for sl in (5, 6):
        try:
            ctx.release(sl)
        except Exception:
            pass

    # a5 / a15 / a8 / a10 on one set of n x n vectors
    n, nchol, nmo = 500, 1600, 250
    try:
        ctx.generate(7, nchol, n, 1.0, 50.0, SEED + 41, rr.COPY_FP64)
        g = syn.rng(SEED + 42)
        h = D(syn.symmetric(n, SEED + 43))
        c0 = np.linalg.qr(g.standard_normal((n, 100)))[0]
        pdn = D(2.0 * c0 @ c0.T)
        fo = torch.empty((n, n), dtype=torch.float64, device=dev)
        row("a5_nonrel_fock_cholesky", lambda: ctx.nonrel_fock_cholesky(7, h, pdn, out=fo), flop=4.0 * n ** 3 * nchol,
            shape=f"n={n} n_chol={nchol}")
        cx = np.linalg.qr(g.standard_normal((2 * n, 100)) + 1j * g.standard_normal((2 * n, 100)))[0]
        hx, px = D(syn.hermitian(2 * n, SEED + 44)), D(cx @ cx.conj().T)
        fx = torch.empty((2 * n, 2 * n), dtype=torch.complex128, device=dev)
        row("a5_x2c_fock_cholesky", lambda: ctx.x2c_fock_cholesky(7, hx, px, out=fx), shape=f"n={n} n_chol={nchol}")
        # factor form from the same densities (occupied-set recovery, then the RI exchange pipeline with
        # N_chol in place of n_aux); the densities are idempotent in the plain metric: S = X = I
        eye1, eye2 = D(np.eye(n)), D(np.eye(2 * n, dtype=np.complex128))
        row("a5_nonrel_fock_cholesky_factor", lambda: ctx.nonrel_fock_ri(7, h, eye1, eye1, pdn, out=fo),
            shape=f"n={n} n_chol={nchol} n_occ=100 (recovered)")
        row("a5_x2c_fock_cholesky_factor", lambda: ctx.x2c_fock_ri(7, hx, eye2, eye2, px, out=fx),
            shape=f"n={n} n_chol={nchol} n_occ=100 spinors (recovered)")
        cm = D(np.linalg.qr(g.standard_normal((n, nmo)))[0])
        mo = torch.empty((nchol, nmo, nmo), dtype=torch.float64, device=dev)
        mo_flop = nchol * (2.0 * n * n * nmo + 2.0 * n * nmo * nmo)
        row("a15_cholesky_to_mo_spatial", lambda: ctx.cholesky_to_mo_spatial(7, cm, out=mo), flop=mo_flop,
            shape=f"n={n} n_chol={nchol} n_mo={nmo}")
        row("a8_build_ri_nonrel_spin_mo_eri", lambda: ctx.build_ri_nonrel_spin_mo_eri(7, cm, out=mo), flop=mo_flop,
            shape=f"n={n} n_aux={nchol} n_mo={nmo}")
        occ = D(g.uniform(0.05, 0.95, 2 * n))
        h2, x2 = D(g.standard_normal((2 * n, 2 * n))), D(g.standard_normal((2 * n, 2 * n)))
        # spin-orbital coefficients, n x 2n (RiNonRelFullOptFock.cpp:53-98): alpha and beta columns
        q = np.linalg.qr(g.standard_normal((n, n)))[0]
        cfull = D(np.concatenate([q, q], axis=1))
        fm = torch.empty((2 * n, 2 * n), dtype=torch.float64, device=dev)
        row("a10_ri_mo_fock_nonrel", lambda: ctx.ri_mo_fock_nonrel(7, h, cfull, occ, h2, x2, out=fm),
            shape=f"n={n} n_aux={nchol} (MO tensor, h_MO and the generalized Fock)")
        del mo
    except Exception as e:
        out["a5_a8_a10_a15_setup"] = {"error": f"{type(e).__name__}: {e}"[:300]}
    try:
        ctx.release(7)
    except Exception:
        pass

    # a5 RKB: density form (K_fg = sum_L B^f P_fg B^g) and factor form (occupied set recovered in the
    # RKB metric, K = sum_L W_L W_L^H) on the same Hermitian vectors and SCF-like density
    try:
        nl, nchol, nocc = 120, 1000, 100
        n4 = 4 * nl
        g = np.random.default_rng(SEED + 50)
        large = g.standard_normal((nchol, nl, nl))
        large = 0.5 * (large + large.transpose(0, 2, 1))
        herm = lambda: (lambda a: 0.5 * (a + a.conj().transpose(0, 2, 1)))(  # noqa: E731
            g.standard_normal((nchol, nl, nl)) + 1j * g.standard_normal((nchol, nl, nl)))
        small0, small1 = herm(), herm()
        cross = g.standard_normal((nchol, nl, nl)) + 0j
        ctx.rkb_cholesky_upload(0, large, small0, small1, cross, cross)
        del large, small0, small1, cross
        a = g.standard_normal((n4, n4)) + 1j * g.standard_normal((n4, n4))
        s = a @ a.conj().T / n4 + np.eye(n4)
        w, v = np.linalg.eigh(s)
        x = v / np.sqrt(w)
        cq = x @ np.linalg.qr(g.standard_normal((n4, nocc)) + 1j * g.standard_normal((n4, nocc)))[0]
        hr, pr, sr, xr = D(syn.hermitian(n4, SEED + 51)), D(cq @ cq.conj().T), D(s), D(x)
        fr = torch.empty((n4, n4), dtype=torch.complex128, device=dev)
        flop_d = 16 * 2 * 8.0 * nl ** 3 * nchol
        row("a5_rkb_fock_cholesky", lambda: ctx.rkb_fock_cholesky(0, hr, pr, out=fr), flop=flop_d,
            shape=f"nl={nl} (n_rkb={n4}) n_chol={nchol}")
        row("a5_rkb_fock_cholesky_factor", lambda: ctx.rkb_fock_cholesky_factor(0, hr, sr, xr, pr, out=fr),
            shape=f"nl={nl} (n_rkb={n4}) n_chol={nchol} n_occ={nocc} (recovered)")
    except Exception as e:
        out["a5_rkb_setup"] = {"error": f"{type(e).__name__}: {e}"[:300]}
    try:
        ctx.rkb_cholesky_release(0)
    except Exception:
        pass

    # a9 / a13 on a device MO tensor
    n, naux = 200, 800
    try:
        g = syn.rng(SEED + 51)
        b = torch.randn((naux, n, n), dtype=torch.float64, device=dev)
        b = 0.5 * (b + b.transpose(1, 2))
        hm = D(syn.symmetric(n, SEED + 52))
        occ = D(g.uniform(0.05, 0.95, n))
        h2 = g.standard_normal((n, n))
        x2 = g.standard_normal((n, n))
        h2d, x2d = D(h2), D(x2)
        fo = torch.empty((n, n), dtype=torch.float64, device=dev)
        row("a9_ri_hx_fock_gemm", lambda: ctx.ri_hx_fock_gemm(b, hm, occ, h2d, x2d, out=fo), shape=f"n={n} n_aux={naux}")
        pairs = [(p_, q_) for p_ in range(n) for q_ in range(p_)]
        vv = D(g.standard_normal(len(pairs)))
        h2s, x2s = D(0.5 * (h2 + h2.T)), D(0.5 * (x2 + x2.T))
        row("a13_jk_only_sigma_hessian_vector_ri", lambda: ctx.jk_only_sigma_hessian_vector_ri(hm, b, occ, h2s, x2s, pairs, vv),
            shape=f"n={n} n_aux={naux} pairs={len(pairs)} (toolkit a11 and pair contractions a12 inside)")
        del b
    except Exception as e:
        out["a9_a13_setup"] = {"error": f"{type(e).__name__}: {e}"[:300]}

    # a7: rotation of a complex MO tensor, B'_P = U^H B_P U
    nt, naux = 400, 600
    try:
        bz = torch.randn((naux, nt, nt), dtype=torch.complex128, device=dev)
        uz = D(syn.unitary(nt, SEED + 55))
        bo = torch.empty_like(bz)
        row("a7_ri_mo_eri_rotated", lambda: ctx.ri_mo_eri_rotated(bz, uz, out=bo), flop=naux * 2 * 8.0 * nt ** 3,
            shape=f"n_tot={nt} n_aux={naux} (complex)")
        del bz, bo
    except Exception as e:
        out["a7_setup"] = {"error": f"{type(e).__name__}: {e}"[:300]}

    # a14: metric orthogonalization and the three-centre contraction
    n, naux_raw = 200, 2000
    try:
        g = syn.rng(SEED + 61)
        q = np.linalg.qr(g.standard_normal((naux_raw, naux_raw)))[0]
        metric = D((q * np.geomspace(1.0, 1e-6, naux_raw)) @ q.T)
        row("a14_canonical_orthogonalize", lambda: ctx.canonical_orthogonalize(metric, 1e-10), shape=f"n_aux={naux_raw}")
        xk = D(ctx.canonical_orthogonalize(metric, 1e-10))
        raw = torch.randn((naux_raw, n * n), dtype=torch.float64, device=dev)
        bo = torch.empty((xk.shape[1], n * n), dtype=torch.float64, device=dev)
        row("a14_ri_three_center_tensor", lambda: ctx.ri_three_center_tensor(raw, xk, out=bo),
            flop=2.0 * naux_raw * xk.shape[1] * n * n, shape=f"n={n} n_aux={naux_raw}")
        del raw, bo
    except Exception as e:
        out["a14_setup"] = {"error": f"{type(e).__name__}: {e}"[:300]}
    torch.cuda.empty_cache()
    return out


def relativistic_leg(torch, rr, ctx, reps=2, cpu=False):
    """X2C 2-spinor Fock build (n = 400 spatial, 800 spinors, n_aux = 2400, 80 Kramers pairs) and its
    AO->MO transform to 800 x 800 per P; 4c RKB riFockTwoElectronDirect (nL = 300, nS = 700, one shared
    n_aux = 4200 with the LL block non-zero on its first 1800 rows, 100 Kramers pairs). The AO tensors
    are real, as in the reference's RI sources; tensors are generated in HBM, and densities, orbitals
    and outputs are device arrays. "achieved_TFLOP_per_s" divides the flops the kernels perform by the
    build time, and per phase by that phase's device time (real flops; a complex multiply-add is 8):
      half-transform  2 n_aux_f fs_f^2 (2 n_occ) per flavor f (real B_P times the Re / Im rows of the
                      occupied coefficients; only the aux rows where the flavor's tensor is non-zero)
      rank-k update   8 n_occ sum_P m_P (m_P + 1) / 2, m_P = the rows of the flavors live at P
                      (lower triangle of the Hermitian K over all flavor blocks at once)
      X2C MO          n_aux (8 n^2 n_mo + 16 n_mo^2 n)   (real legs per spin, complex legs per spin)"""
    from rerdmft_amd import synthetic as syn
    dev = torch.device("cuda:0")
    cp = rr.COPY_FP64 | rr.COPY_FP32 | rr.COPY_BF16
    names = (("f64", rr.FP64), ("f32", rr.FP32), ("bf16", rr.BF16))
    out = {}
    # X2C
    nl, naux, npairs, slot = 400, 2400, 80, 13
    ctx.generate(slot, naux, nl, 1.0, 40.0, 20261203 + 2, cp)
    q = syn.kramers_pair_occ(nl, npairs, 20261203 + 12)
    eye = torch.eye(2 * nl, dtype=torch.complex128, device=dev)
    c = torch.from_numpy(q).to(dev)
    p = c @ c.conj().T
    h = torch.zeros((2 * nl, 2 * nl), dtype=torch.complex128, device=dev)
    f = torch.empty_like(h)
    torch.cuda.synchronize()
    x2c = {"n_spatial": nl, "n_aux": naux, "n_occ": 2 * npairs}
    no = 2 * npairs
    g1 = 2 * (2.0 * naux * nl * nl * 2 * no)
    rk = 8.0 * no * naux * (2 * nl) * (2 * nl + 1) / 2
    ref = None
    for pname, prec in names:
        dt = _timed(torch, lambda: ctx.x2c_fock_ri(slot, h, eye, eye, p, precision=prec, out=f), reps)
        ph = _phases(torch, ctx, lambda: ctx.x2c_fock_ri(slot, h, eye, eye, p, precision=prec, out=f))
        r = {"ms_per_build": 1e3 * dt, "builds_per_s": 1.0 / dt, "achieved_TFLOP_per_s": (g1 + rk) / dt / 1e12,
             "phase_ms": ph, "phase_TFLOP_per_s": _phase_rates(ph, g1, rk)}
        if ref is None:
            ref = f.clone()
        else:
            r["max_rel_dF_vs_f64"] = float((f - ref).abs().max() / ref.abs().max())
        x2c[pname] = r
    nt = 2 * nl
    g = torch.Generator(device=dev)
    g.manual_seed(20261203 + 22)
    u, _ = torch.linalg.qr(torch.randn((nt, nt), dtype=torch.complex128, device=dev, generator=g))
    mo_flops = naux * (8.0 * nl * nl * nt + 16.0 * nt * nt * nl)
    for pname, prec, odt in (("f64", rr.FP64, torch.complex128), ("f32", rr.FP32, torch.complex64)):
        mo = torch.empty((naux, nt, nt), dtype=odt, device=dev)
        torch.cuda.synchronize()
        dt = _timed(torch, lambda: ctx.build_ri_mo_eri(slot, slot, nl, 0, u, precision=prec, out_precision=prec,
                                                       out=mo), 1)
        x2c["mo_transform_" + pname] = {"ms": 1e3 * dt, "achieved_TFLOP_per_s": mo_flops / dt / 1e12}
        del mo
        torch.cuda.empty_cache()
    ctx.release(slot)
    out["x2c"] = x2c
    # 4c RKB
    nl, ns, naux, nll, npairs = 300, 700, 4200, 1800, 100
    sl, ss = 10, 11
    gl = torch.Generator(device=dev)
    gl.manual_seed(20261203 + 3)
    idx = torch.arange(nl, device=dev, dtype=torch.float64)
    dec = torch.exp(-(idx[:, None] - idx[None, :]).abs() / 30.0)
    bll = torch.zeros((naux, nl, nl), dtype=torch.float64, device=dev)
    z = torch.randn((nll, nl, nl), dtype=torch.float64, device=dev, generator=gl)
    bll[:nll] = 0.5 * (z + z.transpose(1, 2)) * dec
    del z
    torch.cuda.synchronize()
    ctx.upload(sl, bll, cp)
    del bll
    ctx.generate(ss, naux, ns, 1.0, 70.0, 20261203 + 13, cp)
    n = 2 * nl + 2 * ns
    no = 2 * npairs
    cq, _ = torch.linalg.qr(torch.randn((n, no), dtype=torch.complex128, device=dev, generator=gl))
    dens = cq @ cq.conj().T
    occ = np.ones(no)
    gout = torch.empty((n, n), dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    rkb = {"n_large": nl, "n_small": ns, "n_aux": naux, "n_aux_large_nonzero": nll, "n_occ": no}
    g1 = 2 * (2.0 * nll * nl * nl * 2 * no) + 2 * (2.0 * naux * ns * ns * 2 * no)
    rk = 8.0 * no * (nll * (2 * nl + 2 * ns) * (2 * nl + 2 * ns + 1) / 2 + (naux - nll) * (2 * ns) * (2 * ns + 1) / 2)
    ref = None
    for pname, prec in names:
        dt = _timed(torch, lambda: ctx.ri_fock_two_electron_direct(sl, ss, nl, ns, dens, cq, occ, precision=prec,
                                                                     out=gout), reps)
        ph = _phases(torch, ctx, lambda: ctx.ri_fock_two_electron_direct(sl, ss, nl, ns, dens, cq, occ,
                                                                          precision=prec, out=gout))
        r = {"ms_per_build": 1e3 * dt, "builds_per_s": 1.0 / dt, "achieved_TFLOP_per_s": (g1 + rk) / dt / 1e12,
             "phase_ms": ph, "phase_TFLOP_per_s": _phase_rates(ph, g1, rk)}
        if ref is None:
            ref = gout.clone()
        else:
            r["max_rel_dG_vs_f64"] = float((gout - ref).abs().max() / ref.abs().max())
        rkb[pname] = r
    ctx.release(sl)
    ctx.release(ss)
    torch.cuda.empty_cache()
    out["rkb"] = rkb
    if cpu:
        out["cpu_baseline"] = relativistic_cpu((400, 2400, 80), (nl, ns, naux, npairs))
    return out


def cpu_baseline(n, nocc, nsample, naux, c0):
    """Oracle timed on the host: J + K over `nsample` aux rows of the same distribution, scaled."""
    from oracle import oracle as orc
    from rerdmft_amd import synthetic as syn
    orc.lib()
    b = np.empty((nsample, n * n))
    g = syn.rng(SEED)
    dec = syn.decay_matrix(n, 50.0).ravel()
    for p in range(nsample):
        z = g.standard_normal((n, n))
        z = np.triu(z) + np.triu(z, 1).T
        b[p] = z.ravel() * dec
    d = 2.0 * c0 @ c0.T
    l = np.sqrt(2.0) * c0
    orc.ri_jk_factor(b[:8], d, l)  # warm-up (thread pool, MKL dispatch)
    t0 = time.perf_counter()
    orc.ri_jk_factor(b, d, l)
    dt = time.perf_counter() - t0
    per_build = dt * naux / nsample
    # the reference additionally recovers C_occ from D each build (FP64 eigensolve, n x n)
    x = np.eye(n)
    t1 = time.perf_counter()
    orc.nonrel_fock_ri(b[:1], x, x, np.zeros((n, n)), d)
    t_rec = time.perf_counter() - t1
    return {"value": 1.0 / per_build, "unit": "Fock builds/s", "cores": orc.max_threads(), "kind": "port",
            "sample": f"J+K over the first {nsample} of {naux} aux rows at n={n}, n_occ={nocc} "
                      f"({dt:.2f} s), scaled by {naux}/{nsample}; occupied recovery excluded",
            "value_with_recovery": 1.0 / (per_build + t_rec), "recovery_s": t_rec,
            "blas": "mkl" if orc.has_mkl() else "portable"}


if __name__ == "__main__":
    sys.exit(main())
