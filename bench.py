"""Benchmark of the Debye F(Q) + gradient hot path on one MI355X.

Workload: Au FCC sphere of radius 30 A (6,699 atoms, a = 4.08 A) with seeded
sigma = 0.05 A Gaussian jitter, float32 positions, Q = 0 ... 24.98 A^-1 in
steps of 0.02 (1,250 bins).  One step = one F(Q) + grad F(Q) evaluation
(pyiid_grad_fq with F output) with inputs resident in HBM and outputs written
to device memory.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Debye F(Q)+grad pair·Q evals/s; HMC wall-s to target Rw "
          "(surrogate vs exact)")
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (published)
FLOP_PER_PAIR_Q = 21       # F + grad, sin/cos counted as 1 FLOP, FMA as 2
QBIN, NQ, RADIUS, LATTICE, SIGMA, SEED = 0.02, 1250, 30.0, 4.08, 0.05, 0


def workload():
    from pyiid_amd import structures
    from pyiid_amd.formfactors import element_table
    pos = structures.jitter(structures.fcc_sphere(RADIUS, LATTICE), SIGMA, SEED)
    pos = pos.astype(np.float32)
    table, index, _ = element_table(np.full(len(pos), 79), QBIN, NQ)
    return pos, table, index


def read_traffic(label="au6699_q1250", kernel="fused_grad"):
    """HBM bytes per launch of the F + gradient kernel (`kernel`: fused_grad,
    the ordered one, or sym_grad) on workload `label` from the committed
    rocprof counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.txt" % (kernel, label))
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


GRAD_KERNEL = "auto"  # --grad-kernel


def grad_kernel_setup(ctx):
    """Restrict the F + gradient path to the kernel chosen by --grad-kernel."""
    from pyiid_amd import _lib
    ctx.set_grad_kernel({"auto": _lib.GRAD_AUTO, "ordered": _lib.GRAD_ORDERED,
                         "symmetric": _lib.GRAD_SYMMETRIC}[GRAD_KERNEL])


def grad_kernel_name(ctx):
    """(profile name, description) of the kernel the last grad_fq ran."""
    sym, bpw, group = ctx.grad_shape()
    if sym:
        return "sym_grad", ("sym_grad_kernel (F + grad, unordered pairs), %d bins per wave, "
                            "%d tiles per item" % (bpw, group))
    return "fused_grad", ("fused_grad_kernel (F + grad, ordered pairs), %d bins per wave, "
                          "%d j ranges per tile" % (bpw, group))


def host_cpu():
    """The host cores this process may use and the CPU model: the affinity
    mask, capped by the cgroup's CPU quota when one is set (a job limited to
    16 CPUs of a 128-core host gains nothing from more threads)."""
    import math
    import subprocess
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, period = fh.read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    cores = affinity if quota is None else max(1, min(affinity, int(math.floor(quota))))
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "cgroup_quota_cpus": quota, "threads_used": cores}


def cpu_baseline(pos, table, index, target_s=15.0):
    """The oracle's C restatement of the reference 'flat' CPU algorithm
    (float32 pair terms, threaded over the flat pair list like the
    reference's process pool) on every core this process may use, built with
    -march=native on this host.  The C2 sample is a strided subset of the Q
    bins (the cost is linear in bins).  C1 is timed in full, and so is exact
    HMC at C1, each beside the device's time for the same work."""
    import tempfile
    from oracle import native
    host = host_cpu()
    threads = host["threads_used"]
    build_dir = tempfile.mkdtemp(prefix="pyiid_cpu_")
    lib = native.build_native(build_dir)
    per_atom = table[index]
    n = len(pos)
    pairs = n * (n - 1) // 2
    # per-bin cost from 8 strided bins (one bin would mostly time the threads' start-up)
    probe = np.linspace(0, NQ - 1, 8).round().astype(np.int32)
    t0 = time.perf_counter()
    native.flat_f32(pos, per_atom, QBIN, q_index=probe, nthreads=threads, lib=lib)
    dt = (time.perf_counter() - t0) / len(probe)
    nsel = int(max(2, min(NQ, target_s / max(dt, 1e-3))))
    qi = np.unique(np.linspace(0, NQ - 1, nsel).round().astype(np.int32))
    t0 = time.perf_counter()
    native.flat_f32(pos, per_atom, QBIN, q_index=qi, nthreads=threads, lib=lib)
    dt = time.perf_counter() - t0
    out = {
        "value": pairs * len(qi) / dt,
        "unit": "pair·Q evals/s",
        "cores": threads,
        "kind": "port",
        "sample": "same structure, F+grad over all %d pairs at %d of the %d Q bins "
                  "(evenly strided), %.1f s" % (pairs, len(qi), NQ, dt),
        "host": host,
        "build": "gcc -O3 -march=native" if lib else "gcc -O3 -march=x86-64-v2 (native build failed)",
    }
    out["c1"] = c1_compare(lib, threads)
    return out


C1_EXP = {"qmin": 0.0, "qmax": 25.0, "qbin": 0.1, "rmin": 0.0, "rmax": 40.0, "rstep": 0.01}


def _c1_cpu_energy_forces(lib, e, pq, per_pdf, gobs, conv, transform="gemm", workers=1):
    """conv Rw and -dE/dq of the 55-atom particle on the host.
    * F + grad F: the reference's float32 flat algorithm (C restatement, one
      thread: 1,485 pairs).
    * G(r) and grad G(r): ``transform`` "gemm" applies the dense matrix T of
      the F -> G map (the oracle's transform_matrix, equal to the FFT path) as
      one BLAS product over the 165 gradient rows; "fft" runs the reference's
      FFT path row by row, batched on ``workers`` threads.
    * Rw and its gradient."""
    from oracle import native
    from oracle import scatter as S
    r = S.r_grid(e)
    nq = per_pdf.shape[1]
    tmat = S.transform_matrix(e["rstep"], pq, r, e["qmin"], nq) if transform == "gemm" else None

    def run(x):
        fp, gfp = native.flat_f32(x, per_pdf, pq, nthreads=1, lib=lib)
        gfp = gfp.astype(np.float64)
        gfp[:, :, :int(np.ceil(e["qmin"] / pq))] = 0.0
        if tmat is not None:
            g = tmat @ fp.astype(np.float64)
            gg = (gfp.reshape(-1, nq) @ tmat.T).reshape(len(x), 3, len(r))
        else:
            g = S.get_pdf_at_qmin(fp.astype(np.float64), e["rstep"], pq, r, e["qmin"])
            gg = S.grad_pdf_rows(gfp, e["rstep"], pq, r, e["qmin"], workers=workers)
        rw = S.get_rw(gobs, g)[0]
        # the reference's forces are half of -dE/dq (its -1/2 dF/dq convention)
        return conv * rw, 2.0 * conv * S.wrap_grad_rw(gg, g, gobs)
    return run


def c1_compare(lib, threads, n_prop=10, n_leap=10, eps=2e-3, beta=50.0, conv=100.0):
    """The 55-atom Au cuboctahedron (a = 4.08 A, sigma 0.05 A jitter, seed 0;
    Q 0-25 step 0.1, r 0-40 step 0.01) in full on the host and on the device:
    one evaluation of F(Q), grad F, G(r), grad G and Rw against the ideal
    particle's G(r); and exact HMC (10 velocity-Verlet steps of 2e-3 per
    proposal, ASE masses, energy 100 Rw, beta 50), seconds per proposal."""
    from oracle import native
    from oracle import scatter as S
    from pyiid_amd import structures
    from pyiid_amd.formfactors import MASSES, element_table
    ideal = structures.cuboctahedron(4.08)
    pos = structures.jitter(ideal, 0.05, 0)
    n = len(pos)
    e, pq = S.make_experiment(dict(C1_EXP))
    nq = S.qmax_bin(e["qmax"], e["qbin"])
    npq = S.qmax_bin(e["qmax"], pq)
    zeros = np.zeros(n, int)
    per_q = element_table(np.full(n, 79), e["qbin"], nq)[0][zeros]
    per_p = element_table(np.full(n, 79), pq, npq)[0][zeros]
    r = S.r_grid(e)
    f_ideal, _ = native.flat_f32(ideal, per_p, pq, grad=False, lib=lib)
    gobs = S.get_pdf_at_qmin(f_ideal.astype(np.float64), e["rstep"], pq, r, e["qmin"])
    energy_forces = _c1_cpu_energy_forces(lib, e, pq, per_p, gobs, conv)
    fft_energy_forces = _c1_cpu_energy_forces(lib, e, pq, per_p, gobs, conv, "fft", threads)

    def cpu_eval():
        native.flat_f32(pos, per_q, e["qbin"], nthreads=1, lib=lib)     # F(Q), grad F
        return energy_forces(pos)                                       # G, grad G, Rw

    def cpu_eval_fft():
        native.flat_f32(pos, per_q, e["qbin"], nthreads=1, lib=lib)
        return fft_energy_forces(pos)

    def timed(fn, min_s=2.0):
        fn()
        k, t0 = 0, time.perf_counter()
        while True:
            fn()
            k += 1
            dt = time.perf_counter() - t0
            if dt >= min_s:
                return dt / k, k

    cpu_s, cpu_k = timed(cpu_eval)
    cpu_fft_s, _ = timed(cpu_eval_fft)

    # exact HMC: momenta ~ N(0, m / beta), velocity Verlet with the force
    # carried between steps, Metropolis on exp(beta (H0 - H1))
    m = np.array([MASSES[79]] * n)
    rs = np.random.RandomState(0)
    x = np.asarray(pos, np.float32).astype(np.float64)
    ex, fx = energy_forces(x)
    acc = 0
    t0 = time.perf_counter()
    for _ in range(n_prop):
        p = rs.normal(0.0, 1.0, (n, 3)) * np.sqrt(m / beta)[:, None]
        h0 = ex + np.sum(p * p / (2 * m[:, None]))
        y, f, q = x.copy(), fx.copy(), p.copy()
        for _ in range(n_leap):
            q += 0.5 * eps * f
            y += eps * q / m[:, None]
            ey, f = energy_forces(y)
            q += 0.5 * eps * f
        h1 = ey + np.sum(q * q / (2 * m[:, None]))
        if rs.random_sample() < np.exp(min(0.0, beta * (h0 - h1))):
            x, ex, fx, acc = y, ey, f, acc + 1
    cpu_hmc_s = (time.perf_counter() - t0) / n_prop
    dev = c1_device(n_prop, n_leap, eps, beta, conv)
    return {
        "workload": "55-atom Au cuboctahedron (a=4.08 A), sigma=0.05 A seeded jitter; "
                    "Q 0-25 step 0.1, r 0-40 step 0.01; Rw against the ideal particle",
        "eval": "F(Q), grad F, G(r), grad G, Rw",
        "cpu_s_per_eval": cpu_s, "cpu_evals": cpu_k,
        "cpu_transform": "dense T (R x Q) applied by one BLAS product over the gradient rows",
        "cpu_s_per_eval_fft_rows": cpu_fft_s,
        "cpu_threads": "1 for the pair sums (1,485 pairs); BLAS / %d FFT workers for the "
                       "transform" % threads,
        "gpu_s_per_eval": dev["s_per_eval"], "gpu_evals": dev["evals"],
        "gpu_eval_path": "ElasticScatter getters (host arrays, PCIe round trips included)",
        "hmc": {"leapfrog_steps": n_leap, "step": eps, "beta": beta, "conv": conv,
                "masses": "ASE (Au 196.97 amu)",
                "cpu_s_per_proposal": cpu_hmc_s, "cpu_proposals": n_prop,
                "cpu_acceptance": acc / n_prop,
                "gpu_s_per_proposal": dev["s_per_proposal"], "gpu_proposals": dev["proposals"],
                "gpu_acceptance": dev["acceptance"]},
    }


def c1_device(n_prop=10, n_leap=10, eps=2e-3, beta=50.0, conv=100.0, min_s=2.0):
    """The device half of the 55-atom cuboctahedron comparison: one evaluation
    of F(Q), grad F, G(r), grad G and Rw through the ElasticScatter getters
    (host arrays), and exact HMC on the device (4 n_prop proposals after a
    captured warm-up), seconds per evaluation and per proposal."""
    from pyiid_amd import Atoms, structures
    from pyiid_amd.calc import wrap_rw
    from pyiid_amd.elasticscatter import ElasticScatter
    from pyiid_amd.sim import DeviceHMC
    ideal = structures.cuboctahedron(4.08)
    pos = structures.jitter(ideal, 0.05, 0)
    n = len(pos)
    scat = ElasticScatter(dict(C1_EXP), seed=0)
    atoms = Atoms("Au%d" % n, pos)
    g_ideal = scat.get_pdf(Atoms("Au%d" % n, ideal))

    def gpu_eval():
        scat.get_fq(atoms)
        scat.get_grad_fq(atoms)
        g = scat.get_pdf(atoms)
        scat.get_grad_pdf(atoms)
        return wrap_rw(g, g_ideal)

    gpu_eval()
    k, t0 = 0, time.perf_counter()
    while True:
        gpu_eval()
        k += 1
        if time.perf_counter() - t0 >= min_s:
            break
    gpu_s = (time.perf_counter() - t0) / k
    hmc = DeviceHMC(scat, [(pos, np.full(n, 79))], g_ideal[None, :], conv=conv)
    hmc.run(2, n_leapfrog=n_leap, step=eps, beta=beta, seed=1)      # capture
    _, a_dev, t = hmc.run(4 * n_prop, n_leapfrog=n_leap, step=eps, beta=beta, seed=1,
                          times=True)
    hmc.close()
    return {"s_per_eval": gpu_s, "evals": k, "s_per_proposal": float(t[-1]) / 1e3 / (4 * n_prop),
            "proposals": 4 * n_prop, "acceptance": float(a_dev.mean())}


def headline_20k(steps=3, warmup=1):
    """F(Q) + grad F(Q) on the two 20k-atom particles: a 20,113-atom Au FCC
    sphere (r = 43.3 A, 1,250 bins of 0.02) and the 20,000-atom Pt/Ni random
    alloy (50/50 seeded, nearest FCC sites at a = 3.724 A, 1,500 bins of
    0.02), device-resident inputs and outputs.  Per workload: wall time per
    step, the fused kernel's average launch (HIP events) and its algorithmic
    fraction of FP32 peak."""
    import torch
    from pyiid_amd import _lib, structures
    from pyiid_amd.formfactors import element_table
    out = {}
    cases = [
        ("au20113_q1250", lambda: structures.fcc_sphere(43.3, LATTICE), lambda n: np.full(n, 79),
         1250),
        ("ptni20000_q1500", lambda: structures.fcc_nearest(20000, a=3.724),
         lambda n: structures.random_alloy(n, 78, 28, seed=SEED), 1500),
    ]
    for label, sites, numbers, nq in cases:
        pos = structures.jitter(sites(), SIGMA, SEED).astype(np.float32)
        n = len(pos)
        table, index, _ = element_table(numbers(n), QBIN, nq)
        ctx = _lib.Context(0)
        ctx.set_positions(pos)
        ctx.set_elements(index)
        ctx.set_scatter(_lib.SLOT_FQ, table)
        grad_kernel_setup(ctx)
        g = torch.empty((n, 3, nq), dtype=torch.float32, device="cuda")
        f = torch.empty(nq, dtype=torch.float32, device="cuda")

        def step():
            ctx.grad_fq(_lib.SLOT_FQ, QBIN, n, nq, out=g.data_ptr(), fq_out=f.data_ptr(),
                        flags=_lib.PYIID_DEVICE)

        for _ in range(warmup):
            step()
        ctx.synchronize()
        ctx.enable_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / steps
        ms, k = ctx.kernel_time(_lib.KERNEL_FUSED)
        ctx.enable_timing(False)
        kname, kdesc = grad_kernel_name(ctx)
        ctx.close()
        pairs = n * (n - 1) // 2
        avg_s = ms / max(1, k) / 1e3
        achieved = FLOP_PER_PAIR_Q * pairs * nq / avg_s / 1e12
        out[label] = {"atoms": n, "q_bins": nq, "steps": steps, "ms_per_step": wall * 1e3,
                      "pair_q_per_s": pairs * nq / wall, "kernel": kdesc,
                      "fused_avg_launch_ms": avg_s * 1e3, "achieved_tflops": achieved,
                      "frac": achieved / PEAK_FP32_TFLOPS, "traffic": read_traffic(label, kname)}
        del g, f
    return out


def _first_at_or_below(rw, target):
    hit = np.nonzero(rw <= target)[0]
    return int(hit[0]) if hit.size else None


def hmc_comparison(n_exact=40, n_sur=80, seeds=(0, 1, 2), train_every=10, energy_weight=None,
                   n_leap=2, eps=3.0, sur_n_leap=None, sur_eps=None):
    """Wall-seconds to a target Rw: exact-force HMC against the learned
    integrator on 64 chains of 5,000-atom Au particles refined against the
    G(r) of independently rattled copies.  ASE masses (Au 196.97 amu),
    energy 100 Rw, beta 1000, n_leap velocity-Verlet steps of eps per
    proposal for both samplers (sur_n_leap / sur_eps for the learned
    sampler when given).  The default 2 steps of 3.0 is the exact sampler's
    fastest trajectory to the target among those it accepts at least 90% of
    the time (a sweep over 1-10 steps of 0.56-5.6; larger steps reach the
    target sooner at lower acceptance).  Per sampler seed: the exact run's plateau is its mean Rw over the last
    10 of its n_exact proposals, the target that plateau + 0.005, and each
    sampler's time to target the HIP-event time at which the chains' mean
    Rw first reaches it.  The learned integrator trains every train_every
    proposals and falls back to exact forces for a block when the learned
    block before it, or the block's own first three learned proposals,
    accepted under 30% (the time includes its training steps and its exact
    proposals).  Both samplers start from the same
    structures."""
    from pyiid_amd import workloads
    from pyiid_amd.elasticscatter import ElasticScatter
    from pyiid_amd.sim import DeviceHMC, SurrogateHMC
    conv, beta = 100.0, 1000.0
    sur_n_leap = n_leap if sur_n_leap is None else sur_n_leap
    sur_eps = eps if sur_eps is None else sur_eps
    starts, targs, numbers = workloads.build("large", 0)
    scat = ElasticScatter(dict(workloads.EXP), seed=0)
    gt = workloads.target_pdfs(scat, targs, numbers)
    structs = list(zip(starts, numbers))
    runs = []
    for seed in seeds:
        ex = DeviceHMC(scat, structs, gt, conv=conv)
        e, a, t = ex.run(n_exact, n_leapfrog=n_leap, step=eps, beta=beta, seed=seed, times=True)
        ex.close()
        rw_e = e.mean(axis=1) / conv
        plateau = float(rw_e[-10:].mean())
        drift = float(rw_e[-20:-10].mean() - plateau)
        target = plateau + 0.005
        k_e = _first_at_or_below(rw_e, target)
        kw = {} if energy_weight is None else {"energy_weight": energy_weight}
        sur = SurrogateHMC(scat, structs, gt, conv=conv, net_seed=seed, **kw)
        out = sur.run(n_sur, n_leapfrog=sur_n_leap, step=sur_eps, beta=beta, seed=seed,
                      train_every=train_every, forget=0.5, ridge=1e-6, min_r2=0.5,
                      min_acceptance=0.3)
        sur.close()
        rw_s = out["energies"].mean(axis=1) / conv
        ts = out["times_ms"]
        k_s = _first_at_or_below(rw_s, target)
        t_e = float(t[k_e]) / 1e3
        t_s = float(ts[k_s]) / 1e3 if k_s is not None else None
        learned = out["exact_mode"] == 0
        runs.append({
            "seed": seed, "plateau_rw": plateau, "plateau_drift_last_20": drift,
            "target_rw": target,
            "exact": {"proposals": n_exact, "s_per_proposal": float(t[-1]) / 1e3 / n_exact,
                      "rw_start": float(rw_e[0]), "rw_final": float(rw_e[-1]),
                      "acceptance": float(a.mean()), "proposals_to_target": k_e + 1,
                      "wall_s_to_target": t_e},
            "learned": {"proposals": n_sur, "rw_final": float(rw_s[-1]),
                        "learned_force_proposals": int(learned.sum()),
                        "acceptance_learned": float(out["accepted"][learned].mean())
                        if learned.any() else None,
                        "acceptance_exact_blocks": float(out["accepted"][~learned].mean())
                        if (~learned).any() else None,
                        "s_per_proposal": float(ts[-1]) / 1e3 / n_sur,
                        "guard_rejections": int(out["violations"].sum()),
                        "proposals_to_target": (k_s + 1) if k_s is not None else None,
                        "wall_s_to_target": t_s},
            "speedup_to_target": (t_e / t_s) if t_s else None,
        })
    sp = [r["speedup_to_target"] for r in runs if r["speedup_to_target"]]
    return {
        "workload": "64 Au chains x 5000 atoms (nearest FCC sites, a=4.08 A), start sigma=0.05 A, "
                    "targets from sigma=0.1 A rattled copies; Q 0-25 step 0.1, r 0-40 step 0.01",
        "leapfrog_steps": n_leap, "step": eps, "learned_leapfrog_steps": sur_n_leap,
        "learned_step": sur_eps, "beta": beta, "conv": conv,
        "masses": "ASE (Au 196.97 amu)", "seeds": list(seeds), "runs": runs,
        "targets_reached_by_learned": len(sp), "speedup_to_target_mean": (sum(sp) / len(sp)) if sp else None,
    }


def batched_hmc(n_prop=3, seed=0):
    """Seconds per proposal for 256 uneven chains (100-3000 atoms), exact and
    learned integrator, each proposal one captured HIP graph."""
    from pyiid_amd import workloads
    from pyiid_amd.elasticscatter import ElasticScatter
    from pyiid_amd.sim import SurrogateHMC
    starts, targs, numbers = workloads.build("batched", seed)
    scat = ElasticScatter(dict(workloads.EXP), seed=0)
    gt = workloads.target_pdfs(scat, targs, numbers)
    h = SurrogateHMC(scat, list(zip(starts, numbers)), gt, conv=100.0, masses="unit")
    kw = dict(n_leapfrog=10, step=2e-3, beta=50.0, seed=seed)
    h.run_exact(1, **kw)                       # capture
    _, _, t = h.run_exact(n_prop, times=True, **kw)
    h.run(2, train_every=0, **kw)               # fit + capture
    out = h.run(4 * n_prop, train_every=0, **kw)
    h.close()
    return {"chains": 256, "atoms": int(sum(len(s) for s in starts)),
            "exact_s_per_proposal": float(t[-1]) / 1e3 / n_prop,
            "surrogate_s_per_proposal": float(out["times_ms"][-1]) / 1e3 / (4 * n_prop)}


def kernel_suite(reps=3):
    """Each hot kernel on its own workload, one warm-up and `reps` timed
    launches (HIP events on the context stream), for profiling runs:
    F-only and pair stage on the 6,699-atom sphere (1,250 bins; all 22.4 M
    pairs), F-only and fused F + gradient on a 20,113-atom Au sphere (1,250 bins) and on
    the 20,000-atom Pt/Ni alloy (1,500 bins), the transform GEMM on its grad-G rows (3 N x 330 -> 4,000 r
    points), contracted forces, the F-only kernel of an energy call (with the
    steps it ran in each recurrence form) and learned-integrator forces on 64
    chains of 5,000 atoms."""
    import torch
    from pyiid_amd import _lib, workloads
    from pyiid_amd.elasticscatter import ElasticScatter
    from pyiid_amd.sim import SurrogateHMC
    dev = _lib.PYIID_DEVICE
    out = {}
    pos, table, index = workload()
    n = len(pos)
    pairs = n * (n - 1) // 2

    def timed(ctx, kind, fn, forms=None):
        fn()
        ctx.synchronize()
        ctx.enable_timing(True)
        for _ in range(reps):
            fn()
        ms, cnt = ctx.kernel_time(kind)
        if forms is not None:
            three, reinsch = ctx.fq_form_counts()
            forms.update({"three_term_steps": three // reps, "reinsch_steps": reinsch // reps,
                          "reinsch_share": reinsch / max(1, three + reinsch)})
        ctx.enable_timing(False)
        return ms / max(1, cnt)

    ctx = _lib.Context(0)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    fq = torch.empty(NQ, dtype=torch.float32, device="cuda")
    ms = timed(ctx, _lib.KERNEL_FQ, lambda: ctx.fq(_lib.SLOT_FQ, QBIN, NQ, out=fq.data_ptr(),
                                                   flags=dev))
    out["fq_kernel"] = {"ms": ms, "pair_q_per_s": pairs * NQ / (ms * 1e-3)}
    # both Debye kernels on a 20k-atom particle (Au FCC sphere r = 43.3 A)
    from pyiid_amd import structures
    from pyiid_amd.formfactors import element_table
    pos20 = structures.jitter(structures.fcc_sphere(43.3, LATTICE), SIGMA, SEED).astype(np.float32)
    t20, i20, _ = element_table(np.full(len(pos20), 79), QBIN, NQ)
    c20 = _lib.Context(0)
    c20.set_positions(pos20)
    c20.set_elements(i20)
    c20.set_scatter(_lib.SLOT_FQ, t20)
    grad_kernel_setup(c20)
    n20 = len(pos20)
    p20 = n20 * (n20 - 1) // 2
    g20 = torch.empty((n20, 3, NQ), dtype=torch.float32, device="cuda")
    f20 = torch.empty(NQ, dtype=torch.float32, device="cuda")
    ms = timed(c20, _lib.KERNEL_FQ, lambda: c20.fq(_lib.SLOT_FQ, QBIN, NQ, out=f20.data_ptr(),
                                                  flags=dev))
    out["fq_kernel_20k"] = {"ms": ms, "atoms": n20, "pair_q_per_s": p20 * NQ / (ms * 1e-3)}
    ms = timed(c20, _lib.KERNEL_FUSED, lambda: _lib.check(c20.lib.pyiid_grad_fq(
        c20.handle, _lib.SLOT_FQ, ctypes.c_double(QBIN), ctypes.c_void_p(g20.data_ptr()),
        ctypes.c_void_p(f20.data_ptr()), dev)))
    out["fused_grad_kernel_20k"] = {"ms": ms, "atoms": n20, "kernel": grad_kernel_name(c20)[1],
                                    "pair_q_per_s": p20 * NQ / (ms * 1e-3),
                                    "tflops_alg": FLOP_PER_PAIR_Q * p20 * NQ / (ms * 1e-3) / 1e12}
    del g20, f20
    c20.close()
    # the 20,000-atom Pt/Ni random alloy (50/50, Cromer-Mann per element), 1,500 bins
    posa = structures.jitter(structures.fcc_nearest(20000, a=3.724), SIGMA, SEED).astype(np.float32)
    nums = structures.random_alloy(20000, 78, 28, seed=SEED)
    ta, ia, _ = element_table(nums, QBIN, 1500)
    ca = _lib.Context(0)
    ca.set_positions(posa)
    ca.set_elements(ia)
    ca.set_scatter(_lib.SLOT_FQ, ta)
    grad_kernel_setup(ca)
    pa = 20000 * 19999 // 2
    ga = torch.empty((20000, 3, 1500), dtype=torch.float32, device="cuda")
    fa = torch.empty(1500, dtype=torch.float32, device="cuda")
    ms = timed(ca, _lib.KERNEL_FQ, lambda: ca.fq(_lib.SLOT_FQ, QBIN, 1500, out=fa.data_ptr(),
                                                 flags=dev))
    out["fq_kernel_alloy20k"] = {"ms": ms, "pair_q_per_s": pa * 1500 / (ms * 1e-3)}
    ms = timed(ca, _lib.KERNEL_FUSED, lambda: _lib.check(ca.lib.pyiid_grad_fq(
        ca.handle, _lib.SLOT_FQ, ctypes.c_double(QBIN), ctypes.c_void_p(ga.data_ptr()),
        ctypes.c_void_p(fa.data_ptr()), dev)))
    out["fused_grad_kernel_alloy20k"] = {"ms": ms, "kernel": grad_kernel_name(ca)[1],
                                         "pair_q_per_s": pa * 1500 / (ms * 1e-3),
                                         "tflops_alg": FLOP_PER_PAIR_Q * pa * 1500 / (ms * 1e-3) / 1e12}
    del ga, fa
    ca.close()
    d = torch.empty((pairs, 3), dtype=torch.float32, device="cuda")
    r = torch.empty(pairs, dtype=torch.float32, device="cuda")
    ms = timed(ctx, _lib.KERNEL_PAIRS, lambda: _lib.check(ctx.lib.pyiid_pair_geometry(
        ctx.handle, 0, pairs, d.data_ptr(), r.data_ptr(), dev)))
    out["pair_geometry_kernel"] = {"ms": ms, "bytes": 16 * pairs + 32 * n,
                                   "gb_per_s": (16 * pairs + 32 * n) / (ms * 1e-3) / 1e9}
    del d, r
    scat = ElasticScatter(dict(workloads.EXP))
    nqp = int(np.floor(scat.exp["qmax"] / scat.pdf_qbin))
    rgrid = scat.get_r()
    ctx.set_transform(scat.exp["rstep"], scat.pdf_qbin, rgrid, scat.exp["qmin"], nqp)
    gq = torch.randn((n, 3, nqp), dtype=torch.float32, device="cuda")
    gg = torch.empty((n, 3, len(rgrid)), dtype=torch.float64, device="cuda")
    ms = timed(ctx, _lib.KERNEL_GEMM, lambda: ctx.grad_pdf(gq, len(rgrid), out=gg, flags=dev))
    flop = 2.0 * 3 * n * nqp * len(rgrid)
    out["gemm_nt_kernel"] = {"ms": ms, "rows": 3 * n, "k": nqp, "n": len(rgrid),
                             "tflops": flop / (ms * 1e-3) / 1e12}
    del gq, gg
    ctx.close()

    starts, targs, numbers = workloads.build("large", 0)
    gt = workloads.target_pdfs(scat, targs, numbers)
    h = SurrogateHMC(scat, list(zip(starts, numbers)), gt, conv=100.0, masses="unit")
    nat = int(h.offsets[-1])
    ms = timed(h.ctx, _lib.KERNEL_PAIR_FORCE, lambda: h.ctx.pdf_energy_forces(
        _lib.SLOT_PDF, scat.pdf_qbin, gt, _lib.POT_RW, 100.0, nat))
    kpairs = sum(len(x) * (len(x) - 1) // 2 for x in starts)
    out["pair_force_kernel"] = {"ms": ms, "pair_q_per_s": kpairs * nqp / (ms * 1e-3)}
    # the F-only kernel of the same energy evaluation, with the (wave, j atom)
    # steps it ran in each recurrence form
    forms = {}
    ms = timed(h.ctx, _lib.KERNEL_FQ, lambda: h.ctx.pdf_energy_forces(
        _lib.SLOT_PDF, scat.pdf_qbin, gt, _lib.POT_RW, 100.0, nat, forces=False), forms)
    out["fq_kernel_large"] = dict({"ms": ms, "chains": 64, "atoms_per_chain": 5000,
                                   "q_bins": nqp, "pair_q_per_s": kpairs * nqp / (ms * 1e-3)},
                                  **forms)
    h.train()
    h.forces()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        h.ctx.sur_forces(nat)
    out["surrogate_forces_call"] = {"ms": (time.perf_counter() - t0) / reps * 1e3,
                                    "atoms": nat}
    h.close()
    # the same chains as 50/50 Pt/Ni random alloys (weighted contracted forces)
    alloy = [structures.random_alloy(len(x), 78, 28, seed=k) for k, x in enumerate(starts)]
    gta = workloads.target_pdfs(scat, targs, alloy)
    ta, ia, _ = element_table(np.concatenate(alloy), scat.pdf_qbin, nqp)
    cpa = _lib.Context(0)
    cpa.set_batch(starts)
    cpa.set_elements(ia)
    cpa.set_scatter(_lib.SLOT_PDF, ta)
    cpa.set_transform(scat.exp["rstep"], scat.pdf_qbin, rgrid, scat.exp["qmin"], nqp)
    ms = timed(cpa, _lib.KERNEL_PAIR_FORCE, lambda: cpa.pdf_energy_forces(
        _lib.SLOT_PDF, scat.pdf_qbin, gta, _lib.POT_RW, 100.0, nat))
    out["pair_force_kernel_alloy"] = {"ms": ms, "pair_q_per_s": kpairs * nqp / (ms * 1e-3)}
    cpa.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-suite", action="store_true",
                    help="only run every hot kernel on its own workload (for profiling)")
    ap.add_argument("--only-20k", action="store_true",
                    help="only run the two 20k-atom F + grad workloads (for profiling)")
    ap.add_argument("--skip-20k", action="store_true",
                    help="leave the 20k-atom workloads out of the JSON line (for profiling)")
    ap.add_argument("--only-batched", action="store_true",
                    help="only run the 256-chain sampler timing (for tracing its graphs)")
    ap.add_argument("--only-hmc", action="store_true",
                    help="only run the exact / learned-integrator sampler comparison")
    ap.add_argument("--only-c1-device", action="store_true",
                    help="only the device half of the 55-atom cuboctahedron comparison "
                         "(evaluation through the getters and exact HMC)")
    ap.add_argument("--hmc-proposals", type=int, default=40,
                    help="exact-HMC proposals per seed of the sampler comparison (0 skips it)")
    ap.add_argument("--grad-kernel", choices=("auto", "ordered", "symmetric"), default="auto",
                    help="F + gradient kernel of the timed steps and the 20k workloads "
                         "(default: the autotuner's choice)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (the known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    global GRAD_KERNEL
    GRAD_KERNEL = args.grad_kernel
    if args.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("copied 3 x %d bytes" % (4 * src.numel()))
        return
    if args.kernel_suite:
        print(json.dumps(kernel_suite()))
        return
    if args.only_20k:
        print(json.dumps(headline_20k()))
        return
    if args.only_batched:
        print(json.dumps(batched_hmc()))
        return
    if args.only_hmc:
        print(json.dumps(hmc_comparison()))
        return
    if args.only_c1_device:
        print(json.dumps(c1_device()))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from pyiid_amd import _lib

    pos, table, index = workload()
    n = len(pos)
    pairs = n * (n - 1) // 2
    ctx = _lib.Context(0)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    grad_kernel_setup(ctx)
    grad = torch.empty((n, 3, NQ), dtype=torch.float32, device="cuda")
    fq = torch.empty(NQ, dtype=torch.float32, device="cuda")
    flags = _lib.PYIID_DEVICE

    def step():
        ctx.grad_fq(_lib.SLOT_FQ, QBIN, n, NQ, out=grad.data_ptr(), fq_out=fq.data_ptr(),
                    flags=flags)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    ctx.enable_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fused_ms, fused_n = ctx.kernel_time(_lib.KERNEL_FUSED)
    ctx.enable_timing(False)
    kname, kdesc = grad_kernel_name(ctx)
    elapsed = t1 - t0
    value = pairs * NQ * args.steps / elapsed
    avg_s = fused_ms / max(1, fused_n) / 1e3
    achieved = FLOP_PER_PAIR_Q * pairs * NQ / avg_s / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair·Q evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Au FCC sphere r=30 A (a=4.08 A, 6699 atoms), sigma=0.05 A seeded "
                        "jitter (seed 0), float32 positions, Q 0-24.98 A^-1 step 0.02 "
                        "(1250 bins); F(Q)+grad F(Q) per step",
            "n_atoms": n,
            "pairs": pairs,
            "q_bins": NQ,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kdesc,
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": read_traffic("au6699_q1250", kname),
            "avg_launch_ms": avg_s * 1e3,
            "launches": fused_n,
            "flop_per_unit": FLOP_PER_PAIR_Q,
        },
    }
    ctx.close()
    if not args.skip_20k:
        out["headline_20k"] = headline_20k()
    if args.hmc_proposals > 0:
        out["hmc"] = hmc_comparison(args.hmc_proposals, 2 * args.hmc_proposals)
        out["batched_hmc"] = batched_hmc()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos, table, index)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
