#!/usr/bin/env python3
"""Headline benchmark: HMC on a 32^4 lattice (D = 4) on one MI355X.

Workload (one "step" = one HMC trajectory, everything resident in HBM):
  beta = 6.0, cold start, 40 leapfrog steps of dt = 0.025 with the
  su(3)-exponential link update, links and E field stored in FP32, FP64 for
  the Hamiltonian / plaquette sums, the accept decision and the
  reunitarisation that follows every trajectory.  Momenta N(0, 0.5/beta) are
  drawn on the device each trajectory.

value = link updates/s = links * MD steps * trajectories / wall time.
The same trajectories are run with FP64 storage for the traj/s comparison at
the same seeds.  roofline = algorithmic HBM bytes of the fused kick+drift
kernel (52 reals per link: read U and E, write E and U) over its average
duration from HIP events recorded around every launch in the timed region.
cpu_baseline = the oracle (C restatement of the reference's rayon
SymplecticEulerRayon step, f64, OpenMP over all host threads) on a bounded
sample of the same lattice.

Other workloads (--workload), one JSON line each, same keys:
  batched  64 independent chains, 16 each of 8^4, 12^4, 16^4, 20^4, beta
           5.6..6.2, FP32 links with FP64 dH, hot start per chain seed, 20
           leapfrog steps dt = 0.01 per trajectory, one launch per MD step,
           trajectories replayed from a captured HIP graph, autotuned shape.
  large    48^3 x 96 FP32-mixed HMC, beta = 6.2, cold start, trajectories of
           50 leapfrog steps dt = 0.02, plaquette after every trajectory.
  mh3d     D = 3 128^3 FP64 checkerboard Metropolis sweeps (spread 0.1,
           beta = 6.0), proposals and accept draws generated on the device.
  sweeps   D = 4 32^4 FP64 quenched update: one HeatBathSweep, normalize_link_matrices,
           four OverrelaxationSweepRotation per step (checkerboard, device streams).
  stats    statistics::mean_and_variance (two-pass) of 2^27 f64 values resident
           in device memory.
  small    8^4 FP64: plaquette, HMC trajectories of 20 steps dt = 0.01
           (reference kick-drift-kick scheme, additive drift) and Metropolis
           sweeps (spread 0.1) with reunitarisation every 10 sweeps.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "lattice-qcd-rs_amd"))

METRIC = ("link updates/s + % HBM peak; FP32-mixed vs FP64 HMC traj/s at matched ΔH, 32^4")
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, AMD specification
REALS_KICK_DRIFT = 52   # read U (18) + E (8), write E (8) + U (18) per link
REALS_KICK = 34         # read U (18) + E (8), write E (8)
# with two-row link reads (f32, exp drift, links in SU(3)) U is read as 12 reals
REALS_KICK_DRIFT_R12 = 46
REALS_KICK_R12 = 28
# VALU instructions per site of the fused f32 kernel (k_md_site_xcd<4, float, float, 1, R12>),
# counted in its gfx950 assembly (hipcc -O3 --cuda-device-only -S), and the issue rate of
# the vector ALU: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles per SIMD
VALU_PER_SITE = {True: 7171, False: 6902}
VALU_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4
REALS_PLAQ = 18
REALS_EENERGY = 8
REALS_REUNIT = 36
REALS_MOMENTA = 8


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=30, help="timed trajectories")
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--extent", type=int, default=32)
    p.add_argument("--md-steps", type=int, default=40)
    p.add_argument("--dt", type=float, default=0.025)
    p.add_argument("--beta", type=float, default=6.0)
    p.add_argument("--f64-steps", type=int, default=10, help="trajectories of the FP64 leg")
    p.add_argument("--cpu-md-steps", type=int, default=15,
                   help="MD steps of the CPU baseline sample (0 = skip)")
    p.add_argument("--seed", type=int, default=0x457893F44AB067F0)
    p.add_argument("--copy-only", action="store_true",
                   help="only run the 1 GiB streaming copy (byte-counter calibration)")
    p.add_argument("--workload", default="hmc32",
                   choices=["hmc32", "batched", "large", "mh3d", "small", "fields", "state", "stats",
                            "sweeps"])
    p.add_argument("--precision-trajectories", type=int, default=3,
                   help="trajectories of the FP32-mixed against FP64 error table (0 = skip)")
    p.add_argument("--gauss-checks", type=int, default=4000,
                   help="limit of the separately timed Gauss projection (0 = skip it)")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="target length of the CPU baseline sample of the other workloads")
    return p.parse_args()


def stream_copy_gbs(lqcd, nbytes=1 << 30, iters=3):
    import ctypes
    out = ctypes.c_double()
    ctx = lqcd.Context.default()
    lqcd._raise(lqcd._ffi.lib().lqcd_stream_copy(ctx.handle, nbytes, iters, ctypes.byref(out)))
    return out.value


def run_leg(lqcd, precision, args, steps, warmup, profile):
    F = lqcd._ffi
    L = args.extent
    st = lqcd.LatticeStateDefault.new_cold(1.0, args.beta, L, 4, precision)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    hmc = lqcd.HybridMonteCarloDiagnostic(args.dt, args.md_steps, integ, args.seed, project_gauss=False)
    ctx = st._dev.ctx

    def trajectory():
        hmc.next_element(st)
        st.normalize_link_matrices()
        return hmc.last_stats

    for _ in range(warmup):
        trajectory()
    kinds = [F.KERNEL_KICK_DRIFT, F.KERNEL_KICK, F.KERNEL_PLAQUETTE, F.KERNEL_EFIELD_ENERGY,
             F.KERNEL_REUNIT, F.KERNEL_MOMENTA]
    ctx.profile_enable(sum(1 << k for k in kinds) if profile else 0)
    ctx.synchronize()
    t0 = time.perf_counter()
    dh, acc = [], 0
    for _ in range(steps):
        s = trajectory()
        dh.append(s.delta_h)
        acc += s.accepted
    ctx.synchronize()
    t1 = time.perf_counter()
    kern = {}
    if profile:
        for k in kinds:
            kern[k] = ctx.profile_read(k)
        ctx.profile_enable(0)
    plaq = st.average_trace_plaquette().real / 3.0
    nlink = st.lattice().number_of_canonical_links_space()
    return {
        "md_variant": st.md_variant(),
        "r12": st.md_reconstruct_last(),
        "seconds": t1 - t0,
        "traj_per_s": steps / (t1 - t0),
        "mean_abs_dh": float(sum(abs(x) for x in dh) / len(dh)),
        "acceptance": acc / steps,
        "plaquette": plaq,
        "nlink": nlink,
        "kernels": kern,
        "dh": dh,
    }


def measured_traffic(variant, r12):
    """HBM bytes per kick+drift launch (fetched + written) from the committed
    rocprofv3 counter summary, when it covers the selected MD variant and
    link-read mode."""
    path = os.path.join(ROOT, "profiles", "kick_drift_32x4_f32_pmc.md")
    key = f"traffic_bytes_per_launch variant={variant}{' r12' if r12 else ''}:"
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(args):
    import oracle as O
    n_md = args.cpu_md_steps
    L = args.extent
    ol = O.OracleLattice(4, L)
    u = ol.hot_start(args.seed)
    e = ol.random_efield(args.seed + 1, 0.5 / args.beta)
    threads = O.get_threads()
    t0 = time.perf_counter()
    ol.simulate_symplectic_n(u, e, args.dt, n_md, O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
    t = time.perf_counter() - t0
    return {
        "value": ol.nlink * n_md / t,
        "unit": "link updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{n_md} SymplecticEulerRayon::integrate_symplectic steps (kick-drift-kick, "
                   f"additive drift, f64) of the oracle on the {L}^4 lattice, {threads} OpenMP "
                   f"threads, {t:.1f} s"),
    }


def precision_table(lqcd, args, n_traj=3):
    """FP32-mixed against FP64 on the same trajectories: from the cold start,
    momenta n of the same device stream, the full MD trajectory of the
    headline (40 leapfrog steps, exp drift) in both storage precisions; the
    per-observable differences of the proposals (every trajectory of this
    workload is rejected, so the kept states would compare trivially)."""
    rows = {lqcd.F32: [], lqcd.F64: []}
    for prec in (lqcd.F32, lqcd.F64):
        for t in range(n_traj):
            st = lqcd.LatticeStateDefault.new_cold(1.0, args.beta, args.extent, 4, prec)
            ef = lqcd.LatticeStateEFSyncDefault(st)
            ef.reset_e_field(args.seed, 10_000 + t, 0.5 / args.beta)
            h0 = ef.hamiltonian_total()
            ef.simulate_symplectic_n(lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG), args.dt,
                                     args.md_steps)
            h1 = ef.hamiltonian_total()
            rows[prec].append({"h_old": h0, "h_new": h1, "dh": h1 - h0,
                               "plaquette": st.average_trace_plaquette().real / 3.0,
                               "e_energy": ef.hamiltonian_efield()})
    out = {}
    for key in ("h_old", "h_new", "dh", "plaquette", "e_energy"):
        diffs = [abs(a[key] - b[key]) for a, b in zip(rows[lqcd.F32], rows[lqcd.F64])]
        scale = max(abs(b[key]) for b in rows[lqcd.F64])
        out[key] = {"max_abs_err": max(diffs), "rel_to_f64": max(diffs) / scale if scale else None,
                    "f64": [b[key] for b in rows[lqcd.F64]]}
    out["trajectories"] = n_traj
    return out


def gauss_projection_cost(lqcd, args):
    """One project_to_gauss (field.rs:1265-1294) of fresh momenta on the
    headline lattice after a trajectory: the refresh step of the reference's
    HMC, which the timed trajectories leave out."""
    st = lqcd.LatticeStateDefault.new_cold(1.0, args.beta, args.extent, 4, lqcd.F32)
    hmc = lqcd.HybridMonteCarloDiagnostic(args.dt, args.md_steps,
                                          lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG),
                                          args.seed, project_gauss=False)
    hmc.next_element(st)
    st.normalize_link_matrices()
    ef = lqcd.LatticeStateEFSyncDefault(st)
    ef.reset_e_field(args.seed, 1000, 0.5 / args.beta)
    st._dev.ctx.synchronize()
    t0 = time.perf_counter()
    try:
        checks = ef.project_to_gauss(args.gauss_checks)
        out = {"converged": True, "checks": checks}
    except lqcd.StateInitializationError as exc:
        out = {"converged": False, "message": str(exc)}
    out["seconds"] = round(time.perf_counter() - t0, 4)
    return out


def contract_line(args, metric, value, unit, steps, seconds, dtype, config, roofline, cpu, **extra):
    line = {
        "metric": metric,
        "value": round(value, 1),
        "unit": unit,
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * seconds / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    return line


def kernel_rate(ctx, kind, bytes_per_launch):
    n, ms = ctx.profile_read(kind)
    if n == 0:
        return None
    avg_s = ms / n / 1e3
    gbs = bytes_per_launch / avg_s / 1e9
    return {"launches": n, "avg_us": round(avg_s * 1e6, 2), "GB_s": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": int(bytes_per_launch)}


def roofline_from(rate, kernel):
    return {"bound": "hbm", "achieved": rate["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": rate["hbm_frac"], "traffic": None, "kernel": kernel,
            "bytes_per_launch": rate["bytes_per_launch"]}


def cpu_md_sample(extents, beta, dt, seconds, seed):
    """The oracle's kick-drift-kick f64 MD steps (SymplecticEulerRayon) on a
    lattice of the given extents, as many steps as fit in about `seconds`."""
    import oracle as O
    ol = O.OracleLattice(len(extents), extents)
    u = ol.hot_start(seed)
    e = ol.random_efield(seed + 1, 0.5 / beta)
    t0 = time.perf_counter()
    u, e = ol.simulate_symplectic_n(u, e, dt, 1, O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
    one = time.perf_counter() - t0
    n = max(1, int(seconds / max(one, 1e-6)) - 1)
    t0 = time.perf_counter()
    ol.simulate_symplectic_n(u, e, dt, n, O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
    t = time.perf_counter() - t0
    threads = O.get_threads()
    return {"value": ol.nlink * n / t, "unit": "link updates/s", "cores": threads, "kind": "port",
            "sample": (f"{n} kick-drift-kick f64 MD steps (SymplecticEulerRayon restatement) of the "
                       f"oracle on a {'x'.join(map(str, extents))} lattice, {threads} OpenMP threads, "
                       f"{t:.1f} s")}


def bench_batched(args, lqcd):
    F = lqcd._ffi
    sizes, n_steps, dt = [8, 12, 16, 20], 20, 0.01
    specs = [lqcd.ChainSpec([sizes[k // 16]] * 4, 5.6 + 0.6 * k / 63, args.seed + k) for k in range(64)]
    b = lqcd.BatchedChains(specs, 4, lqcd.F32)
    ctx = b.ctx
    b.hot_start()
    config = b.autotune(lqcd.DRIFT_EXP)
    nlink = sum(b.number_of_links(i) for i in range(len(specs)))
    # eager pass with a timing event around every launch
    b.set_graph(False)
    b.hmc(dt, n_steps, 1, reunitarize=True)
    ctx.profile_enable((1 << F.KERNEL_KICK_DRIFT) | (1 << F.KERNEL_KICK) | (1 << F.KERNEL_PLAQUETTE)
                       | (1 << F.KERNEL_MOMENTA))
    t0 = time.perf_counter()
    b.hmc(dt, n_steps, 2, reunitarize=True)
    eager_s = (time.perf_counter() - t0) / 2
    r12 = b.md_reconstruct_last()
    kd = kernel_rate(ctx, F.KERNEL_KICK_DRIFT, nlink * (REALS_KICK_DRIFT_R12 if r12 else REALS_KICK_DRIFT) * 4)
    kick = kernel_rate(ctx, F.KERNEL_KICK, nlink * (REALS_KICK_R12 if r12 else REALS_KICK) * 4)
    sums = kernel_rate(ctx, F.KERNEL_PLAQUETTE, nlink * (REALS_PLAQ + REALS_EENERGY) * 4)
    mom = kernel_rate(ctx, F.KERNEL_MOMENTA, nlink * REALS_MOMENTA * 4)
    ctx.profile_enable(0)
    # graph replays
    b.set_graph(True)
    b.hmc(dt, n_steps, max(args.warmup, 1), reunitarize=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    stats = b.hmc(dt, n_steps, args.steps, reunitarize=True)
    seconds = time.perf_counter() - t0
    value = nlink * n_steps * args.steps / seconds
    groups = {}
    for g, L in enumerate(sizes):
        s = stats[:, 16 * g:16 * (g + 1)]
        groups[f"{L}^4"] = {"acceptance": round(float(s["accepted"].mean()), 4),
                            "mean_abs_dH": float(abs(s["delta_h"]).mean()),
                            "plaquette_last": float(s["plaquette"][-1].mean())}
    shape = {0: "64", 1: "128", 2: "256"}[config % 3] + (" per-site" if (config // 3) % 2 else " per-link") \
        + (" persistent" if config >= 6 else " full-grid")
    cfg = {"workload": ("64 chains (16 each of 8^4, 12^4, 16^4, 20^4), beta 5.6..6.2, FP32 links/E with "
                        "FP64 dH, hot start per chain seed, 20 leapfrog steps dt=0.01 per trajectory, "
                        "exp link update, reunitarisation after every trajectory, trajectories replayed "
                        "from a captured HIP graph"),
           "chains": 64, "sites": b.total_sites(), "links": nlink, "md_steps": n_steps, "dt": dt,
           "launch_shape": {"config": config, "shape": shape}, "two_row_link_reads": r12}
    cpu = cpu_md_sample([16] * 4, 5.9, dt, args.cpu_seconds, args.seed) if args.cpu_md_steps > 0 else None
    return contract_line(
        args, "link updates/s, 64 batched chains", value, "link updates/s", args.steps, seconds, "f32",
        cfg, roofline_from(kd, "k_batch_md<4,float,EXP> (fused force + kick + exp drift, all chains"
                               + (", links read as two rows)" if r12 else ")")), cpu,
        chain_traj_per_s=round(64 * args.steps / seconds, 2),
        graph_traj_ms=round(1e3 * seconds / args.steps, 3), eager_traj_ms=round(1e3 * eager_s, 3),
        kernels={"kick_drift": kd, "kick": kick, "sums": sums, "momenta": mom}, by_size=groups)


def hmc_leg(lqcd, precision, extents, beta, dt, md_steps, steps, warmup, seed, cold=True):
    F = lqcd._ffi
    d = len(extents)
    if cold:
        st = lqcd.LatticeStateDefault.new_cold(1.0, beta, extents, d, precision)
    else:
        st = lqcd.LatticeStateDefault.new_determinist(1.0, beta, extents, seed, d, precision)
    hmc = lqcd.HybridMonteCarloDiagnostic(dt, md_steps, lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP,
                                                                               lqcd.SCHEME_LEAPFROG), seed,
                                          project_gauss=False)
    ctx = st._dev.ctx
    for _ in range(warmup):
        hmc.next_element(st)
        st.normalize_link_matrices()
    ctx.profile_enable(1 << F.KERNEL_KICK_DRIFT)
    ctx.synchronize()
    plaq, dh, acc = [], [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        hmc.next_element(st)
        st.normalize_link_matrices()
        plaq.append(st.average_trace_plaquette().real / 3.0)
        dh.append(hmc.last_stats.delta_h)
        acc += hmc.last_stats.accepted
    ctx.synchronize()
    seconds = time.perf_counter() - t0
    nlink = st.lattice().number_of_canonical_links_space()
    elem = 4 if precision == lqcd.F32 else 8
    r12 = st.md_reconstruct_last()
    kd = kernel_rate(ctx, F.KERNEL_KICK_DRIFT,
                     nlink * (REALS_KICK_DRIFT_R12 if r12 else REALS_KICK_DRIFT) * elem)
    ctx.profile_enable(0)
    return {"seconds": seconds, "nlink": nlink, "plaquette": plaq, "dh": dh, "acceptance": acc / steps,
            "kick_drift": kd, "md_variant": st.md_variant(), "r12": r12}


def bench_large(args, lqcd):
    ext, beta, dt, md = [48, 48, 48, 96], 6.2, 0.02, 50
    leg = hmc_leg(lqcd, lqcd.F32, ext, beta, dt, md, args.steps, args.warmup, args.seed)
    value = leg["nlink"] * md * args.steps / leg["seconds"]
    cfg = {"workload": ("HMC 48^3x96 D=4 beta=6.2, cold start, 50 leapfrog steps dt=0.02 per trajectory, "
                        "exp link update, FP32 links/E with FP64 dH/plaquette/accept/reunitarisation, "
                        "plaquette after every trajectory"),
           "lattice": ext, "links": leg["nlink"], "md_steps": md, "dt": dt, "beta": beta,
           "md_variant": leg["md_variant"], "two_row_link_reads": leg["r12"]}
    cpu = cpu_md_sample([24, 24, 24, 48], beta, dt, args.cpu_seconds, args.seed) if args.cpu_md_steps > 0 else None
    kname = "MD kick+drift (fused force + kick + exp drift" + (", links read as two rows)" if leg["r12"] else ")")
    return contract_line(
        args, "link updates/s, 48^3x96 FP32-mixed HMC", value, "link updates/s", args.steps, leg["seconds"],
        "f32", cfg, roofline_from(leg["kick_drift"], kname), cpu,
        traj_per_s=round(args.steps / leg["seconds"], 3), acceptance=leg["acceptance"],
        mean_abs_dH=float(sum(abs(x) for x in leg["dh"]) / len(leg["dh"])),
        plaquette_per_trajectory=[round(p, 10) for p in leg["plaquette"]])


def bench_mh3d(args, lqcd):
    F = lqcd._ffi
    L, beta, spread = 128, 6.0, 0.1
    st = lqcd.LatticeStateDefault.new_determinist(1.0, beta, L, args.seed, 3, lqcd.F64)
    mh = lqcd.MetropolisHastingsSweep(1, spread, args.seed)
    ctx = st._dev.ctx
    for _ in range(args.warmup):
        mh.next_element(st)
    ctx.profile_enable(1 << F.KERNEL_METROPOLIS)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mh.next_element(st)
    ctx.synchronize()
    seconds = time.perf_counter() - t0
    nlink = st.lattice().number_of_canonical_links_space()
    # one class launch: every link read once (staples; 12 of 18 reals when the links are read
    # as two rows), its nlink/6 links written, p and flag stored
    r12 = st.md_reconstruct_last()
    per_launch = nlink * (12 if r12 else 18) * 8 + nlink // 6 * (18 * 8 + 12)
    rate = kernel_rate(ctx, F.KERNEL_METROPOLIS, per_launch)
    ctx.profile_enable(0)
    value = nlink * args.steps / seconds
    cfg = {"workload": ("D=3 128^3 FP64 checkerboard Metropolis sweeps, beta=6.0, spread 0.1, 1 update per "
                        "link, hot start, proposals and accept draws generated on the device"),
           "lattice": [L] * 3, "links": nlink, "beta": beta, "spread": spread}
    cpu = None
    if args.cpu_md_steps > 0:
        import oracle as O
        ls = 32
        ol = O.OracleLattice(3, ls)
        u = ol.hot_start(args.seed)
        rng = O.OracleRng(args.seed + 1)
        n, t = 0, 0.0
        while t < args.cpu_seconds:
            t0 = time.perf_counter()
            props = ol.mh_proposals(args.seed + 7 + n, spread, 1)
            unis = np.array([rng.next_u64() for _ in range(ol.nlink)], dtype=np.uint64)
            u, _ = ol.metropolis_sweep(u, beta, props, unis, O.ORDER_LEXICOGRAPHIC)
            t += time.perf_counter() - t0
            n += 1
        cpu = {"value": ol.nlink * n / t, "unit": "link updates/s", "cores": 1, "kind": "port",
               "sample": (f"{n} lexicographic single-threaded sweeps (MetropolisHastingsSweep restatement, "
                          f"proposal and uniform draws included) of the oracle on 32^3, {t:.1f} s")}
    return contract_line(
        args, "link updates/s, D=3 128^3 FP64 Metropolis", value, "link updates/s", args.steps, seconds, "f64",
        cfg, roofline_from(rate, "k_metropolis<3,double,RNG> (one checkerboard class"
                           + (", staple links read as two rows)" if r12 else ")")), cpu,
        sweeps_per_s=round(args.steps / seconds, 3), prob_replace_mean=mh.prob_replace_mean(),
        number_replace_last=mh.number_replace_last())


def bench_fields(args, lqcd):
    """F_mu_nu (LinkMatrix::f_mu_nu, field.rs:822-834) of every site of the 32^4
    f32 lattice into a device buffer: the clover stencil of the observables."""
    F = lqcd._ffi
    L = args.extent
    st = lqcd.LatticeStateDefault.new_determinist(1.0, args.beta, L, args.seed, 4, lqcd.F32)
    nsite = st.lattice().number_of_points()
    n = nsite * 16 * 18
    buf = F.DeviceBuffer(8 * n)
    ctx = st._dev.ctx

    def step():
        lqcd._raise(F.lib().lqcd_field_strength_device(st.handle, F.FIELD_F_MU_NU, buf.ptr, n))

    for _ in range(args.warmup):
        step()
    ctx.profile_enable(1 << F.KERNEL_FIELDS)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    seconds = time.perf_counter() - t0
    # compulsory: the 4 links of a site (72 B each in f32) read, 16 matrices of 144 B written
    per_launch = nsite * (4 * 72 + 16 * 144)
    rate = kernel_rate(ctx, F.KERNEL_FIELDS, per_launch)
    ctx.profile_enable(0)
    cfg = {"workload": ("F_mu_nu at every site of a 32^4 D=4 lattice (f32 links, double arithmetic, "
                        "16 matrices per site in the reference's host layout, into device memory)"),
           "lattice": [L] * 4, "sites": nsite}
    cpu = None
    if args.cpu_md_steps > 0:
        import oracle as O
        ol = O.OracleLattice(4, 16)
        u = ol.hot_start(args.seed)
        O.set_threads(16)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            ol.f_mu_nu_field(u)
            k += 1
        t = time.perf_counter() - t0
        cpu = {"value": ol.nsite * k / t, "unit": "sites/s", "cores": 16, "kind": "port",
               "sample": f"{k} F_mu_nu fields of the oracle on 16^4, 16 OpenMP threads, {t:.1f} s"}
    return contract_line(
        args, "F_mu_nu sites/s, 32^4", nsite * args.steps / seconds, "sites/s", args.steps, seconds, "f64",
        cfg, roofline_from(rate, "k_clover_fields<4,float,1> (clover, F_mu_nu)"), cpu)


def bench_sweeps(args, lqcd):
    """A quenched update of the 32^4 D=4 f64 lattice: HeatBathSweep
    (heat_bath.rs:73-123), normalize_link_matrices, then four
    OverrelaxationSweepRotation (overrelaxation.rs:66-91), as the reference's
    HybridMethodVec would chain them; checkerboard class launches."""
    F = lqcd._ffi
    L, beta = args.extent, args.beta
    st = lqcd.LatticeStateDefault.new_determinist(1.0, beta, L, args.seed, 4, lqcd.F64)
    hb = lqcd.HeatBathSweep(args.seed)
    orx = lqcd.OverrelaxationSweepRotation()
    ctx = st._dev.ctx

    def step():
        st.monte_carlo_step(hb)
        st.normalize_link_matrices()
        for _ in range(4):
            st.monte_carlo_step(orx)

    for _ in range(args.warmup):
        step()
    kinds = (F.KERNEL_HEAT_BATH, F.KERNEL_OVERRELAX, F.KERNEL_REUNIT)
    ctx.profile_enable(sum(1 << k for k in kinds))
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    seconds = time.perf_counter() - t0
    nlink = st.lattice().number_of_canonical_links_space()
    # one class launch: every link read once for the staples (144 B each), the
    # class's nlink / 8 links written (and the heat bath's per-link flag)
    hb_rate = kernel_rate(ctx, F.KERNEL_HEAT_BATH, nlink * 144 + nlink // 8 * (144 + 4))
    or_rate = kernel_rate(ctx, F.KERNEL_OVERRELAX, nlink * 144 + nlink // 8 * 144)
    re_rate = kernel_rate(ctx, F.KERNEL_REUNIT, nlink * REALS_REUNIT * 8)
    ctx.profile_enable(0)
    value = nlink * 5 * args.steps / seconds
    cfg = {"workload": ("32^4 D=4 FP64 quenched update per step: HeatBathSweep, normalize_link_matrices, "
                        "4 OverrelaxationSweepRotation (checkerboard), beta 6.0, hot start"),
           "lattice": [L] * 4, "links": nlink, "beta": beta}
    cpu = None
    if args.cpu_md_steps > 0:
        import oracle as O
        ol = O.OracleLattice(4, 16)
        u = ol.hot_start(args.seed)
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            u, _ = ol.heat_bath_sweep(u, beta, args.seed, k)
            u = ol.normalize_link_matrices(u)
            for _ in range(4):
                u = ol.overrelaxation_sweep(u)
            k += 1
        t = time.perf_counter() - t0
        cpu = {"value": ol.nlink * 5 * k / t, "unit": "link updates/s", "cores": 1, "kind": "port",
               "sample": (f"{k} steps (heat bath, normalisation, 4 overrelaxation sweeps) of the oracle "
                          f"on 16^4, single-threaded sweeps as in the reference, {t:.1f} s")}
    dominant, name = ((hb_rate, "k_heat_bath<4,double> (one checkerboard class)")
                      if hb_rate["avg_us"] * 8 >= or_rate["avg_us"] * 32
                      else (or_rate, "k_overrelax<4,double,false> (one checkerboard class)"))
    return contract_line(
        args, "link updates/s, 32^4 FP64 heat bath + overrelaxation", value, "link updates/s", args.steps,
        seconds, "f64", cfg, roofline_from(dominant, name), cpu,
        kernels={"heat_bath": hb_rate, "overrelaxation": or_rate, "reunitarize": re_rate},
        sweeps_per_s=round(5 * args.steps / seconds, 3))


def bench_stats(args, lqcd):
    """statistics::mean_and_variance (statistics/mod.rs:375-399, two passes:
    the sum, then sum (x - mean)^2) of 2^27 f64 values resident in device
    memory; each pass is one reduction kernel reading the values once."""
    from lqcd import statistics as S
    F = lqcd._ffi
    n = 1 << 27
    x = np.random.default_rng(args.seed & 0xFFFFFFFF).uniform(-1.0, 1.0, n)
    vals = S.DeviceValues.from_numpy(x)
    ctx = lqcd.Context.default()

    def step():
        return S.mean_and_variance(vals, ctx)

    for _ in range(args.warmup):
        step()
    ctx.profile_enable(1 << F.KERNEL_STATS)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    ctx.synchronize()
    seconds = time.perf_counter() - t0
    rate = kernel_rate(ctx, F.KERNEL_STATS, 8 * n)
    ctx.profile_enable(0)
    cfg = {"workload": "mean_and_variance (two-pass) of 2^27 f64 values in device memory", "values": n}
    cpu = None
    if args.cpu_md_steps > 0:
        import oracle as O
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            O.stats_mean_and_variance(x)
            k += 1
        t = time.perf_counter() - t0
        cpu = {"value": n * k / t, "unit": "values/s", "cores": 1, "kind": "port",
               "sample": f"{k} two-pass mean_and_variance of the same 2^27 values on the oracle, 1 thread, {t:.1f} s"}
    return contract_line(
        args, "statistics values/s, mean_and_variance of 2^27 f64", n * args.steps / seconds, "values/s",
        args.steps, seconds, "f64", cfg, roofline_from(rate, "k_stats_partials<double> (one pass of the values)"),
        cpu, result=[float(res[0]), float(res[1])])


def bench_state(args, lqcd):
    """LatticeStateEFSyncDefault written and read back in the serde bincode
    layout (state.rs:1046-1062): device download + encode + file, and file +
    decode + upload, on the 32^4 f32 lattice (files under the temp dir)."""
    import tempfile
    F = lqcd._ffi
    L = args.extent
    st = lqcd.LatticeStateDefault.new_determinist(1.0, args.beta, L, args.seed, 4, lqcd.F32)
    ef = lqcd.LatticeStateEFSyncDefault.new_random_e_state(st, args.seed, 0, project=False)
    nlink = st.lattice().number_of_canonical_links_space()
    nbytes = 8 + nlink * 64 + 8 + 32 + nlink * 144
    with tempfile.TemporaryDirectory() as d:
        path = f"{d}/state.bin"
        ef.write(path)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ef.write(path)
        tw = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            back = lqcd.LatticeStateEFSyncDefault.read(path, 4, F.FORMAT_BINCODE, lqcd.F32)
            del back
        tr = time.perf_counter() - t0
    cfg = {"workload": ("LatticeStateEFSyncDefault of a 32^4 D=4 f32 lattice written and read back in the "
                        "serde bincode layout (links and E as f64, 872 MB per file), temp-dir file"),
           "lattice": [L] * 4, "file_bytes": nbytes}
    return contract_line(
        args, "state checkpoint write+read GB/s, 32^4", 2 * nbytes * args.steps / (tw + tr) / 1e9, "GB/s",
        args.steps, tw + tr, "f64", cfg, None, None,
        write_GB_s=round(nbytes * args.steps / tw / 1e9, 3), read_GB_s=round(nbytes * args.steps / tr / 1e9, 3))


def bench_small(args, lqcd):
    F = lqcd._ffi
    L, beta, spread = 8, 6.0, 0.1
    n_traj, md, dt, n_sweeps = 10, 20, 0.01, 100
    st = lqcd.LatticeStateDefault.new_determinist(1.0, beta, L, args.seed, 4, lqcd.F64)
    nlink = st.lattice().number_of_canonical_links_space()
    hmc = lqcd.HybridMonteCarloDiagnostic(dt, md, lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE,
                                                                        lqcd.SCHEME_REFERENCE), args.seed,
                                          project_gauss=False)
    mh = lqcd.MetropolisHastingsSweep(1, spread, args.seed)
    ctx = st._dev.ctx
    hmc.next_element(st)
    mh.next_element(st)
    ctx.synchronize()
    t0 = time.perf_counter()
    plaq0 = st.average_trace_plaquette().real / 3.0
    acc = 0
    for _ in range(n_traj):
        hmc.next_element(st)
        acc += hmc.last_stats.accepted
    t1 = time.perf_counter()
    for k in range(n_sweeps):
        mh.next_element(st)
        if (k + 1) % 10 == 0:
            st.normalize_link_matrices()
    ctx.synchronize()
    t2 = time.perf_counter()
    updates = nlink * (n_traj * md + n_sweeps)
    value = updates / (t2 - t0)
    cfg = {"workload": ("8^4 FP64 beta=6.0 hot start: average_trace_plaquette, 10 HMC trajectories of 20 "
                        "kick-drift-kick steps dt=0.01 (additive drift), 100 checkerboard Metropolis sweeps "
                        "(spread 0.1) with normalize_link_matrices every 10 sweeps"),
           "lattice": [L] * 4, "links": nlink}
    cpu = None
    if args.cpu_md_steps > 0:
        import oracle as O
        ol = O.OracleLattice(4, L)
        u = ol.hot_start(args.seed)
        rng = O.OracleRng(args.seed + 1)
        c0 = time.perf_counter()
        for k in range(n_traj):
            e = ol.random_efield(args.seed + 100 + k, 0.5 / beta)
            u, _ = ol.hmc_trajectory(u, e, beta, dt, md, rng.next_u64(), O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
        for k in range(n_sweeps):
            props = ol.mh_proposals(args.seed + 7 + k, spread, 1)
            unis = np.array([rng.next_u64() for _ in range(ol.nlink)], dtype=np.uint64)
            u, _ = ol.metropolis_sweep(u, beta, props, unis, O.ORDER_LEXICOGRAPHIC)
            if (k + 1) % 10 == 0:
                u = ol.normalize_link_matrices(u)
        c = time.perf_counter() - c0
        cpu = {"value": updates / c, "unit": "link updates/s", "cores": O.get_threads(), "kind": "port",
               "sample": (f"the whole workload on the oracle (MD over {O.get_threads()} OpenMP threads, "
                          f"sweeps single-threaded, lexicographic), {c:.1f} s")}
    return contract_line(
        args, "link updates/s, 8^4 FP64 HMC + Metropolis", value, "link updates/s", 1, t2 - t0, "f64", cfg,
        None, cpu, hmc_seconds=round(t1 - t0, 4), mh_seconds=round(t2 - t1, 4), hmc_acceptance=acc / n_traj,
        plaquette_start=plaq0, plaquette_end=st.average_trace_plaquette().real / 3.0)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    import lqcd
    F = lqcd._ffi
    if args.copy_only:
        print(json.dumps({"stream_copy_GB_s": round(stream_copy_gbs(lqcd), 1),
                          "bytes": 1 << 30, "iters": 3}))
        return
    if args.workload != "hmc32":
        run = {"batched": bench_batched, "large": bench_large, "mh3d": bench_mh3d,
               "small": bench_small, "fields": bench_fields,
               "state": bench_state, "stats": bench_stats, "sweeps": bench_sweeps}[args.workload]
        print(json.dumps(run(args, lqcd)))
        return

    mixed = run_leg(lqcd, lqcd.F32, args, args.steps, args.warmup, profile=True)
    f64 = run_leg(lqcd, lqcd.F64, args, args.f64_steps, args.warmup, profile=True)

    nlink = mixed["nlink"]
    link_updates = nlink * args.md_steps * args.steps
    value = link_updates / mixed["seconds"]
    ms_per_step = 1e3 * mixed["seconds"] / args.steps

    def kernel_line(leg, kind, reals, elem):
        n, ms = leg["kernels"].get(kind, (0, 0.0))
        if n == 0:
            return None
        avg_s = ms / n / 1e3
        gbs = nlink * reals * elem / avg_s / 1e9
        return {"launches": n, "avg_us": round(avg_s * 1e6, 2), "GB_s": round(gbs, 1),
                "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}

    r12 = mixed["r12"]
    reals_kd = REALS_KICK_DRIFT_R12 if r12 else REALS_KICK_DRIFT
    kd = kernel_line(mixed, F.KERNEL_KICK_DRIFT, reals_kd, 4)
    variant = mixed["md_variant"]
    traffic = measured_traffic(variant, r12) if args.extent == 32 else None
    roofline = {
        "bound": "hbm",
        "achieved": kd["GB_s"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kd["hbm_frac"],
        "traffic": traffic,
        "kernel": (f"k_md_{'site' if variant >= 2 else 'link'}_xcd<4,float,float,EXP> "
                   f"(fused force + momentum kick + exp link drift, MD variant {variant}"
                   f"{', links read as two rows' if r12 else ''})"),
        "bytes_per_launch": nlink * reals_kd * 4,
    }
    if variant >= 2:
        # the second roof of this kernel: vector ALU issue (no MFMA, no packed FP32)
        floor_s = (nlink // 4) / 64 * VALU_PER_SITE[r12] / VALU_WAVE_INSTR_PER_S
        roofline["valu"] = {"instr_per_site": VALU_PER_SITE[r12], "floor_us": round(floor_s * 1e6, 1),
                            "frac": round(floor_s * 1e6 / kd["avg_us"], 4)}
    kernels = {}
    for name, kind, reals in [("kick_drift", F.KERNEL_KICK_DRIFT, REALS_KICK_DRIFT),
                              ("kick", F.KERNEL_KICK, REALS_KICK),
                              # H_U and H_E sums in one kernel (plaquette + E energy)
                              ("hamiltonian_sums", F.KERNEL_PLAQUETTE, REALS_PLAQ + REALS_EENERGY),
                              ("reunitarize", F.KERNEL_REUNIT, REALS_REUNIT),
                              ("momenta", F.KERNEL_MOMENTA, REALS_MOMENTA)]:
        reals32 = {"kick_drift": reals_kd, "kick": REALS_KICK_R12 if r12 else REALS_KICK}.get(name, reals)
        reals64 = {"kick_drift": REALS_KICK_DRIFT_R12 if f64["r12"] else REALS_KICK_DRIFT,
                   "kick": REALS_KICK_R12 if f64["r12"] else REALS_KICK}.get(name, reals)
        kernels["f32_" + name] = kernel_line(mixed, kind, reals32, 4)
        kernels["f64_" + name] = kernel_line(f64, kind, reals64, 8)

    cpu = cpu_baseline(args) if args.cpu_md_steps > 0 else None
    copy_gbs = stream_copy_gbs(lqcd)
    gauss = gauss_projection_cost(lqcd, args) if args.gauss_checks > 0 else None
    precision = (precision_table(lqcd, args, args.precision_trajectories)
                 if args.precision_trajectories > 0 else None)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"HMC {args.extent}^4 D=4 beta={args.beta}, cold start, "
                         f"{args.md_steps} leapfrog steps dt={args.dt}, exp link update, "
                         "FP32 links/E with FP64 dH/plaquette/accept/reunitarisation per trajectory, "
                         "momenta N(0, 0.5/beta) without the Gauss projection (timed separately)"),
            "lattice": [args.extent] * 4,
            "links": nlink,
            "md_steps": args.md_steps,
            "dt": args.dt,
            "beta": args.beta,
        },
        "traj_per_s": {"f32_mixed": round(mixed["traj_per_s"], 3),
                       "f64": round(f64["traj_per_s"], 3)},
        "mixed_vs_f64_speedup": round(mixed["traj_per_s"] / f64["traj_per_s"], 3),
        "mean_abs_dH": {"f32_mixed": mixed["mean_abs_dh"], "f64": f64["mean_abs_dh"]},
        "acceptance": {"f32_mixed": mixed["acceptance"], "f64": f64["acceptance"]},
        "plaquette": {"f32_mixed": mixed["plaquette"], "f64": f64["plaquette"]},
        "dH_f32_minus_f64_first_trajectories": [
            a - b for a, b in zip(mixed["dh"][: len(f64["dh"])], f64["dh"])][:3],
        "roofline": roofline,
        "gauss_projection": gauss,
        "mixed_precision_errors": precision,
        "stream_copy_GB_s": round(copy_gbs, 1),
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
