#!/usr/bin/env python3
"""Headline benchmark: HMC on a 32^4 lattice (D = 4) on one MI355X.

Workload (one "step" = one HMC trajectory, everything resident in HBM):
  beta = 6.0, hot start (random SU(3) per link from the device stream), 40
  leapfrog steps of dt = 0.025 with the su(3)-exponential link update, links
  and E field stored in FP32, FP64 for the Hamiltonian / plaquette sums, the
  accept decision and the reunitarisation that follows every trajectory.
  Momenta N(0, 0.5/beta) are drawn on the device each trajectory.  H_old and
  H_new are summed inside the first MD launch and the final half kick of each
  trajectory (per-link traces in the storage precision, sums in FP64).

value = link updates/s = links * MD steps * trajectories / wall time.
The same trajectories are run with FP64 storage for the traj/s comparison at
the same seeds.  roofline = algorithmic HBM bytes of the fused kick+drift
kernel over its average duration from HIP events recorded around every launch
in the timed region.  The algorithmic bytes of one leapfrog step are those of
one force evaluation (read U and E, write E: 34 reals per link) plus one link
drift (read U and E, write U: 44 reals), 78 reals per link; the fused kernel
does both in one pass, which needs 52 reals per link (40 when U is read and
written as its first two rows), reported beside it.
mixed_precision_errors_vs_oracle = one trajectory of the workload from the
oracle's f64 hot start and momenta, in the FP32-mixed and FP64 device legs and
in the oracle (f64): per-observable differences.  The oracle's trajectory is
timed too and is the cpu_baseline (same workload, f64, OpenMP over all host
threads).

Other workloads (--workload), one JSON line each, same keys:
  adjoint  16^4 FP64, beta 5.7, hot start: Wilson action forward + gauge-force
           backward pairs, then 50 HMC trajectories of 20 kick-drift-kick
           steps dt = 0.01 with the reference's additive drift.
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
  variants every MD work decomposition (lqcd_md_variant_set 0..4) on the
           headline lattice: per-launch time of the fused kick+drift kernel.
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
# Algorithmic reals per link: each input and output counted once, staple
# neighbour re-reads count zero.  One force evaluation reads U (18) and E (8)
# and writes E (8); one link drift reads U and E and writes U (18).
REALS_FORCE = 34
REALS_DRIFT = 44
REALS_STEP = REALS_FORCE + REALS_DRIFT   # one leapfrog step: a force and a drift
# what the fused kick+drift kernel itself moves in one pass: read U and E,
# write E and U; 40 when the links are in SU(3) and U is read and written as
# its first two rows (12 reals each way)
REALS_KICK_DRIFT = 52
REALS_KICK_DRIFT_R12 = 40
REALS_KICK = REALS_FORCE
REALS_KICK_R12 = 28
# VALU instructions per site of the fused f32 exp-drift kernels by (MD variant,
# two-row links), counted in their gfx950 assembly (hipcc -O3 --cuda-device-only -S
# of k_md.hip: k_md_site_xcd<4, float, float, 1, R12, 0> for variants 2 and 3,
# k_md_site_tile<4, float, float, 1, 0> for variant 4; every branch counted, so an
# upper bound of the executed count), and the issue rate of
# the vector ALU: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles per SIMD
VALU_PER_SITE = {(2, True): 6971, (3, True): 6971, (4, True): 7703, (2, False): 6842, (3, False): 6842}
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
    p.add_argument("--cpu-md-steps", type=int, default=5,
                   help="kick-drift-kick MD steps of the oracle in the reference's integrator form "
                        "(0 = skip; also skips the CPU samples of the other workloads)")
    p.add_argument("--seed", type=int, default=0x457893F44AB067F0)
    p.add_argument("--copy-only", action="store_true",
                   help="only run the 1 GiB streaming copy (byte-counter calibration)")
    p.add_argument("--workload", default="hmc32",
                   choices=["hmc32", "adjoint", "batched", "large", "mh3d", "small", "fields", "state", "stats",
                            "sweeps", "variants"])
    p.add_argument("--oracle-trajectory", type=int, default=1,
                   help="1: the per-observable error table against the oracle's f64 trajectory, "
                        "which is also the CPU baseline (0 = skip)")
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


def stream_read_gbs(lqcd, nbytes=1 << 30, iters=3):
    import ctypes
    out = ctypes.c_double()
    ctx = lqcd.Context.default()
    lqcd._raise(lqcd._ffi.lib().lqcd_stream_read(ctx.handle, nbytes, iters, ctypes.byref(out)))
    return out.value


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_leg(lqcd, precision, args, steps, warmup, profile):
    F = lqcd._ffi
    L = args.extent
    st = lqcd.LatticeStateDefault.new_determinist(1.0, args.beta, L, args.seed, 4, precision)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    hmc = lqcd.HybridMonteCarloDiagnostic(args.dt, args.md_steps, integ, args.seed, project_gauss=False)
    ctx = st._dev.ctx

    def trajectory():
        hmc.next_element(st)
        st.normalize_link_matrices()
        return hmc.last_stats

    plaq0 = st.average_trace_plaquette().real / 3.0
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
        "plaquette_start": plaq0,
        "plaquette": plaq,
        "nlink": nlink,
        "kernels": kern,
        "dh": dh,
    }


def measured_traffic(variant, r12):
    """HBM bytes per kick+drift launch (fetched + written) from the committed
    rocprofv3 counter summary of the same kernel build, when it covers the
    selected MD variant and link-read mode (counters cannot be read inside this
    run: rocprofv3 --pmc runs the program on its own)."""
    path = os.path.join(ROOT, "profiles", "kick_drift_32x4_f32_pmc.md")
    key = f"traffic_bytes_per_launch variant={variant}{' r12' if r12 else ''}:"
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_reference_integrator(args, n_md=5):
    """The reference's own integrator form on the oracle: kick-drift-kick
    steps with the additive drift (SymplecticEulerRayon::integrate_symplectic),
    f64, on the headline lattice."""
    import oracle as O
    L = args.extent
    ol = O.OracleLattice(4, L)
    u = ol.hot_start(args.seed)
    e = ol.random_efield(args.seed + 1, 0.5 / args.beta)
    threads = O.get_threads()
    t0 = time.perf_counter()
    ol.simulate_symplectic_n(u, e, args.dt, n_md, O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
    t = time.perf_counter() - t0
    return {"value": ol.nlink * n_md / t, "unit": "link updates/s", "cores": threads,
            "sample": (f"{n_md} kick-drift-kick steps (additive drift, f64) of the oracle on the {L}^4 "
                       f"lattice, {threads} OpenMP threads, {t:.1f} s")}


def oracle_error_table(lqcd, args):
    """One trajectory of the headline workload from the oracle's f64 hot start
    and momenta: the oracle (f64), the FP32-mixed device leg and the FP64 device
    leg from identical inputs.  Per observable: the device value, the oracle's
    and their difference.  The oracle's trajectory is timed: it is the CPU
    baseline of the same workload."""
    import oracle as O
    L, beta, dt, md = args.extent, args.beta, args.dt, args.md_steps
    uniform = 1 << 63
    ol = O.OracleLattice(4, L)
    u = ol.hot_start(args.seed)
    e = ol.random_efield(args.seed + 1, 0.5 / beta)
    threads = O.get_threads()
    t0 = time.perf_counter()
    h_old = ol.hamiltonian_links(u, beta) + ol.hamiltonian_efield(e, beta)
    u_new, e_new = ol.simulate_symplectic_n(u, e, dt, md, O.DRIFT_EXP, O.SCHEME_LEAPFROG)
    h_e = ol.hamiltonian_efield(e_new, beta)
    h_new = ol.hamiltonian_links(u_new, beta) + h_e
    t_cpu = time.perf_counter() - t0
    ref = {"h_old": h_old, "h_new": h_new, "dh": h_new - h_old, "h_e": h_e,
           "plaquette": ol.average_trace_plaquette(u_new).real / 3.0}
    ref_acc = O.bernoulli_accept(O.replacement_probability(h_old, h_new), uniform)
    legs = {}
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    for name, prec in (("f32_mixed", lqcd.F32), ("f64", lqcd.F64)):
        st = lqcd.LatticeStateDefault(lqcd.LatticeCyclic(1.0, L, 4), beta, u, prec)
        ef = lqcd.LatticeStateEFSyncDefault(st, e)
        g_old = ef.hamiltonian_total()
        ef.simulate_symplectic_n(integ, dt, md)
        got = {"h_old": g_old, "h_new": ef.hamiltonian_total(), "h_e": ef.hamiltonian_efield(),
               "plaquette": st.average_trace_plaquette().real / 3.0}
        got["dh"] = got["h_new"] - got["h_old"]
        leg = {k: {"device": got[k], "abs_err": abs(got[k] - ref[k]),
                   "rel_err": abs(got[k] - ref[k]) / abs(ref[k])} for k in ref}
        leg["max_abs_dU"] = float(np.max(np.abs(st.link_matrix() - u_new)))
        st.set_link_matrix(u)
        ef.set_e_field(e)
        hmc = lqcd.HybridMonteCarloDiagnostic(dt, md, integ, 1, project_gauss=False)
        s = hmc.trajectory(st, uniform)
        leg["accepted"] = bool(s.accepted)
        leg["same_decision"] = bool(s.accepted) == bool(ref_acc)
        legs[name] = leg
        del st, ef
    table = {"oracle": dict(ref, accepted=bool(ref_acc)), **legs,
             "inputs": (f"oracle hot start (Xoshiro256++ random SU(3), seed {args.seed:#x}) and momenta "
                        f"N(0, 0.5/beta) (seed + 1), {L}^4, beta {beta}, {md} leapfrog steps dt {dt}, "
                        f"exp link update; accept word 2^63")}
    cpu = {"value": ol.nlink * md / t_cpu, "unit": "link updates/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": (f"the oracle's f64 trajectory of the same workload ({md} leapfrog steps with the "
                      f"su(3)-exponential link update on the {L}^4 hot start, H before and after), "
                      f"{threads} OpenMP threads, {t_cpu:.1f} s")}
    return table, cpu


def gauss_projection_cost(lqcd, args):
    """One project_to_gauss (field.rs:1265-1294) of fresh momenta on the
    headline lattice after a trajectory: the refresh step of the reference's
    HMC, which the timed trajectories leave out."""
    st = lqcd.LatticeStateDefault.new_determinist(1.0, args.beta, args.extent, args.seed, 4, lqcd.F32)
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
    if _COPY_GBS is not None:
        line["stream_copy_GB_s"] = round(_COPY_GBS, 1)
    if _READ_GBS is not None:
        line["stream_read_GB_s"] = round(_READ_GBS, 1)
    line.update(extra)
    return line


# the run's own streaming rates of 1 GiB, copy (read + write) and read only; a
# kernel row above the one that matches its access mix is served from the
# caches, not HBM
_COPY_GBS = None
_READ_GBS = None


def kernel_rate(ctx, kind, bytes_per_launch, read_only=False):
    n, ms = ctx.profile_read(kind)
    if n == 0:
        return None
    avg_s = ms / n / 1e3
    gbs = bytes_per_launch / avg_s / 1e9
    out = {"launches": n, "avg_us": round(avg_s * 1e6, 2), "GB_s": round(gbs, 1),
           "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": int(bytes_per_launch)}
    ref = _READ_GBS if read_only else _COPY_GBS
    if ref is not None and gbs > ref:
        out["cache_served"] = True
    return out


def roofline_from(rate, kernel):
    out = {"bound": "hbm", "achieved": rate["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": rate["hbm_frac"], "traffic": None, "kernel": kernel,
           "bytes_per_launch": rate["bytes_per_launch"]}
    if rate.get("cache_served"):
        out["cache_served"] = True
    return out


def md_roofline(rate_kd, nlink, elem, r12, kernel, rate_kick=None):
    """Roofline of the MD force/drift kernels at the algorithmic bytes of the
    work they do: a fused kick+drift launch is one force evaluation and one
    link drift (REALS_STEP per link), a kick launch one force evaluation;
    over the sum of their launch times.  Beside it the bytes the fused kernel
    itself must move in one pass (REALS_KICK_DRIFT) and those it reads with
    two-row link reads."""
    alg = rate_kd["launches"] * nlink * REALS_STEP * elem
    one_pass = rate_kd["launches"] * nlink * REALS_KICK_DRIFT * elem
    actual = rate_kd["launches"] * nlink * (REALS_KICK_DRIFT_R12 if r12 else REALS_KICK_DRIFT) * elem
    t = rate_kd["launches"] * rate_kd["avg_us"] * 1e-6
    if rate_kick is not None:
        alg += rate_kick["launches"] * nlink * REALS_FORCE * elem
        one_pass += rate_kick["launches"] * nlink * REALS_FORCE * elem
        actual += rate_kick["launches"] * nlink * (REALS_KICK_R12 if r12 else REALS_KICK) * elem
        t += rate_kick["launches"] * rate_kick["avg_us"] * 1e-6
    gbs = alg / t / 1e9
    out = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
           "bytes_per_launch": nlink * REALS_STEP * elem,
           "accounting": (f"algorithmic bytes: a force evaluation ({REALS_FORCE} reals per link: read U, E; "
                          f"write E) plus a link drift ({REALS_DRIFT} reals: read U, E; write U) per fused "
                          f"launch, {elem}-byte reals"),
           "one_pass": {"reals_per_link": REALS_KICK_DRIFT, "GB_s": round(one_pass / t / 1e9, 1),
                        "frac": round(one_pass / t / 1e9 / HBM_PEAK_GBS, 4)},
           "bytes_read_actual": {"reals_per_link": REALS_KICK_DRIFT_R12 if r12 else REALS_KICK_DRIFT,
                                 "GB_s": round(actual / t / 1e9, 1)}}
    if _COPY_GBS is not None and actual / t / 1e9 > _COPY_GBS:
        out["cache_served"] = True
    return out


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


def bench_adjoint(args, lqcd):
    """16^4 FP64, beta 5.7, hot start: Wilson action forward + hand-written
    backward (the gauge force as dS/domega, recomputing the staples), then 50
    HMC trajectories of 20 kick-drift-kick steps dt = 0.01 with the reference's
    additive drift (hybrid_monte_carlo.rs:465-471,584-613)."""
    F = lqcd._ffi
    L, beta, dt, md, n_traj = 16, 5.7, 0.01, 20, 50
    st = lqcd.LatticeStateDefault.new_determinist(1.0, beta, L, args.seed, 4, lqcd.F64)
    nlink = st.lattice().number_of_canonical_links_space()
    ctx = st._dev.ctx
    for _ in range(max(args.warmup, 1)):
        st.wilson_forward()
        st.wilson_backward()
    ctx.profile_enable((1 << F.KERNEL_PLAQUETTE) | (1 << F.KERNEL_FORCE))
    n_fb = 50
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_fb):
        st.wilson_forward()
        st.wilson_backward()
    t_fb = time.perf_counter() - t0
    fwd = kernel_rate(ctx, F.KERNEL_PLAQUETTE, nlink * REALS_PLAQ * 8, read_only=True)
    # backward: read U (18 reals), write the 8 gradient components in f64
    bwd = kernel_rate(ctx, F.KERNEL_FORCE, nlink * (REALS_PLAQ + 8) * 8)
    ctx.profile_enable(0)
    # the same pairs with the gradient left in a device buffer
    gbuf = F.DeviceBuffer(nlink * 8 * 8)
    st.wilson_backward_device(gbuf)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_fb):
        st.wilson_forward()
        st.wilson_backward_device(gbuf)
    t_fb_dev = time.perf_counter() - t0
    del gbuf
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE, lqcd.SCHEME_REFERENCE)
    # kernel times from profiled trajectories on a copy of the state (with
    # profiling on, the reference scheme launches eagerly, not from a graph)
    prof = lqcd.LatticeStateDefault(st.lattice(), beta, st.link_matrix(), lqcd.F64)
    hmc_p = lqcd.HybridMonteCarloDiagnostic(dt, md, integ, args.seed + 1, project_gauss=False)
    pctx = prof._dev.ctx
    pctx.profile_enable((1 << F.KERNEL_KICK_DRIFT) | (1 << F.KERNEL_KICK))
    for _ in range(3):
        hmc_p.next_element(prof)
    kd = kernel_rate(pctx, F.KERNEL_KICK_DRIFT, nlink * REALS_STEP * 8)
    kick = kernel_rate(pctx, F.KERNEL_KICK, nlink * REALS_FORCE * 8)
    pctx.profile_enable(0)
    del prof
    hmc = lqcd.HybridMonteCarloDiagnostic(dt, md, integ, args.seed, project_gauss=False)
    hmc.next_element(st)
    ctx.synchronize()
    dh, acc = [], 0
    t0 = time.perf_counter()
    for _ in range(n_traj):
        hmc.next_element(st)
        dh.append(hmc.last_stats.delta_h)
        acc += hmc.last_stats.accepted
    ctx.synchronize()
    seconds = time.perf_counter() - t0
    value = nlink * md * n_traj / seconds
    cfg = {"workload": ("16^4 D=4 FP64 beta=5.7, hot start (device stream): 50 Wilson forward + backward "
                        "pairs (action to the host, gradient [link][8] to the host; then again with the gradient "
                        "left in a device buffer), then 50 HMC trajectories "
                        "of 20 kick-drift-kick steps dt=0.01 with the reference's additive drift, momenta "
                        "N(0, 0.5/beta) without the Gauss projection"),
           "lattice": [L] * 4, "links": nlink, "md_steps": md, "dt": dt, "beta": beta}
    cpu = None
    if args.cpu_md_steps > 0:
        import oracle as O
        ol = O.OracleLattice(4, L)
        u = ol.hot_start(args.seed)
        k, c0 = 0, time.perf_counter()
        while time.perf_counter() - c0 < args.cpu_seconds:
            e = ol.random_efield(args.seed + k, 0.5 / beta)
            u, _ = ol.hmc_trajectory(u, e, beta, dt, md, 1 << 63, O.DRIFT_ADDITIVE, O.SCHEME_REFERENCE)
            k += 1
        c = time.perf_counter() - c0
        cpu = {"value": nlink * md * k / c, "unit": "link updates/s", "cores": O.get_threads(), "kind": "port",
               "cpu_model": cpu_model(),
               "sample": (f"{k} HMC trajectories of the same form (20 kick-drift-kick steps, additive drift, "
                          f"f64) of the oracle on 16^4, {O.get_threads()} OpenMP threads, {c:.1f} s")}
    return contract_line(
        args, "link updates/s, 16^4 FP64 HMC (reference integrator)", value, "link updates/s", n_traj, seconds,
        "f64", cfg, md_roofline(kd, nlink, 8, False, "k_md_*<4,double,ADDITIVE> kick(dt/2)+drift and kick(dt/2) of "
                                "the reference kick-drift-kick step", rate_kick=kick), cpu,
        traj_per_s=round(n_traj / seconds, 3), acceptance=acc / n_traj,
        mean_abs_dH=float(np.mean(np.abs(dh))), plaquette=st.average_trace_plaquette().real / 3.0,
        forward_backward_per_s=round(n_fb / t_fb, 2),
        forward_backward_ms=round(1e3 * t_fb / n_fb, 3),
        forward_backward_device_per_s=round(n_fb / t_fb_dev, 2),
        forward_backward_device_ms=round(1e3 * t_fb_dev / n_fb, 3),
        kernels={"wilson_forward": fwd, "wilson_backward": bwd, "kick_drift": kd, "kick": kick})


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
    kd = kernel_rate(ctx, F.KERNEL_KICK_DRIFT, nlink * REALS_STEP * 4)
    kick = kernel_rate(ctx, F.KERNEL_KICK, nlink * REALS_FORCE * 4)
    sums = kernel_rate(ctx, F.KERNEL_PLAQUETTE, nlink * (REALS_PLAQ + REALS_EENERGY) * 4, read_only=True)
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
        cfg, md_roofline(kd, nlink, 4, r12, "k_batch_md<4,float,EXP> (fused force + kick + exp drift, all chains"
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
    kd = kernel_rate(ctx, F.KERNEL_KICK_DRIFT, nlink * REALS_STEP * elem)
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
        "f32", cfg, md_roofline(leg["kick_drift"], leg["nlink"], 4, leg["r12"], kname), cpu,
        traj_per_s=round(args.steps / leg["seconds"], 3), acceptance=leg["acceptance"],
        acceptance_note=(None if leg["acceptance"] > 0 else
                         "every trajectory was rejected (|dH| grows with the volume at fixed dt with momenta "
                         "of standard deviation 0.5/beta): the cold start does not thermalise, the plaquette "
                         "stays at 1"),
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
        cfg, roofline_from(rate, "k_f_mu_nu_blocks<4,float> (clover, F_mu_nu, one thread per site and pair)"), cpu)


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


def bench_variants(args, lqcd):
    """Every MD work decomposition (lqcd_md_variant_set 0..4) timed on the
    headline lattice: the 40 fused kick+drift launches of one leapfrog
    integration (f32, exp drift, links read and written as two rows),
    replayed from a captured graph that has run before (the first replay of
    a freshly instantiated graph runs slower), per-launch time from HIP
    events; two alternated rounds, the better kept per variant.  The links are those the
    headline's timed trajectories start from: the hot start after --warmup
    HMC trajectories of the headline workload.  The line's value is the link
    updates/s of the fastest decomposition."""
    F = lqcd._ffi
    L = args.extent
    st = lqcd.LatticeStateDefault.new_determinist(1.0, args.beta, L, args.seed, 4, lqcd.F32)
    integ = lqcd.SymplecticEulerGpu(lqcd.DRIFT_EXP, lqcd.SCHEME_LEAPFROG)
    hmc = lqcd.HybridMonteCarloDiagnostic(args.dt, args.md_steps, integ, args.seed, project_gauss=False)
    for _ in range(args.warmup):
        hmc.next_element(st)
        st.normalize_link_matrices()
    ef = lqcd.LatticeStateEFSyncDefault(st)
    ctx = st._dev.ctx
    nlink = st.lattice().number_of_canonical_links_space()
    u0 = st.link_matrix()
    names = {0: "k_md_link (2-D grid)", 1: "k_md_link_xcd", 2: "k_md_site_xcd", 3: "k_md_site_xcd, slab order",
             4: "k_md_site_tile (LDS tile of x_0/x_1 rows)"}
    best = {}
    for _ in range(2):
        for v in range(5):
            st.set_link_matrix(u0)
            ef.reset_e_field(args.seed, 0, 0.5 / args.beta)
            st.set_md_variant(v)
            # capture and warm up the graphs of the timed integration: one per
            # starting buffer (an even number of steps moves the links two
            # buffers on, so three integrations start from each of the three),
            # so that the timed one replays a graph that has run before
            for _ in range(3):
                ef.simulate_symplectic_n(integ, args.dt, args.md_steps)
            st.set_link_matrix(u0)
            ef.reset_e_field(args.seed, 0, 0.5 / args.beta)
            ctx.profile_enable(1 << F.KERNEL_KICK_DRIFT)
            ctx.synchronize()
            ef.simulate_symplectic_n(integ, args.dt, args.md_steps)
            n, ms = ctx.profile_read(F.KERNEL_KICK_DRIFT)
            ctx.profile_enable(0)
            us = ms / n * 1e3
            if v not in best or us < best[v]["avg_us"]:
                best[v] = {"avg_us": round(us, 2), "launches": n, "kernel": names[v],
                           "GB_s_alg": round(nlink * REALS_STEP * 4 / (us * 1e-6) / 1e9, 1),
                           "frac_alg": round(nlink * REALS_STEP * 4 / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                           "two_row_reads": st.md_reconstruct_last()}
    fastest = min(best, key=lambda v: best[v]["avg_us"])
    us = best[fastest]["avg_us"]
    rate = {"launches": best[fastest]["launches"], "avg_us": us}
    cfg = {"workload": (f"{L}^4 D=4 FP32 links, exp-drift leapfrog, beta {args.beta}, hot start after "
                        f"{args.warmup} HMC trajectories: {args.md_steps} fused kick+drift launches per MD "
                        f"work decomposition"),
           "lattice": [L] * 4, "links": nlink}
    roof = md_roofline(rate, nlink, 4, best[fastest]["two_row_reads"], names[fastest])
    return contract_line(args, "MD kick+drift link updates/s by work decomposition, 32^4 FP32",
                         nlink / (us * 1e-6), "link updates/s", args.md_steps, us * 1e-3 * args.md_steps / 1e3,
                         "f32", cfg, roof, None, variants={str(v): best[v] for v in sorted(best)},
                         fastest_variant=fastest)


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
    rate = kernel_rate(ctx, F.KERNEL_STATS, 8 * n, read_only=True)
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
    # kernel times from two profiled trajectories on a copy of the state (with
    # profiling on the reference scheme launches eagerly, not from a graph)
    prof = lqcd.LatticeStateDefault(st.lattice(), beta, st.link_matrix(), lqcd.F64)
    hmc_p = lqcd.HybridMonteCarloDiagnostic(dt, md, lqcd.SymplecticEulerGpu(lqcd.DRIFT_ADDITIVE,
                                                                          lqcd.SCHEME_REFERENCE), args.seed + 1,
                                            project_gauss=False)
    pctx = prof._dev.ctx
    pctx.profile_enable((1 << F.KERNEL_KICK_DRIFT) | (1 << F.KERNEL_KICK))
    for _ in range(2):
        hmc_p.next_element(prof)
    kd = kernel_rate(pctx, F.KERNEL_KICK_DRIFT, nlink * REALS_STEP * 8)
    kick = kernel_rate(pctx, F.KERNEL_KICK, nlink * REALS_FORCE * 8)
    pctx.profile_enable(0)
    del prof
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
        md_roofline(kd, nlink, 8, False, "k_md<4,double,ADDITIVE> kick(dt/2)+drift and kick(dt/2) of the "
                    "reference kick-drift-kick step (a 2.4 MB lattice: launch-bound, cache-resident)",
                    rate_kick=kick), cpu, hmc_seconds=round(t1 - t0, 4), mh_seconds=round(t2 - t1, 4), hmc_acceptance=acc / n_traj,
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
    global _COPY_GBS, _READ_GBS
    _COPY_GBS = stream_copy_gbs(lqcd)
    _READ_GBS = stream_read_gbs(lqcd)
    if args.workload != "hmc32":
        run = {"adjoint": bench_adjoint, "batched": bench_batched, "large": bench_large, "mh3d": bench_mh3d,
               "small": bench_small, "fields": bench_fields,
               "state": bench_state, "stats": bench_stats, "sweeps": bench_sweeps,
               "variants": bench_variants}[args.workload]
        print(json.dumps(run(args, lqcd)))
        return

    mixed = run_leg(lqcd, lqcd.F32, args, args.steps, args.warmup, profile=True)
    f64 = run_leg(lqcd, lqcd.F64, args, args.f64_steps, args.warmup, profile=True)

    nlink = mixed["nlink"]
    link_updates = nlink * args.md_steps * args.steps
    value = link_updates / mixed["seconds"]
    ms_per_step = 1e3 * mixed["seconds"] / args.steps

    def kernel_line(leg, kind, reals_alg, reals_actual, elem):
        n, ms = leg["kernels"].get(kind, (0, 0.0))
        if n == 0:
            return None
        avg_s = ms / n / 1e3
        gbs = nlink * reals_alg * elem / avg_s / 1e9
        out = {"launches": n, "avg_us": round(avg_s * 1e6, 2), "GB_s": round(gbs, 1),
               "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        actual = nlink * reals_actual * elem / avg_s / 1e9
        if reals_actual != reals_alg:
            out["actual_GB_s"] = round(actual, 1)
        if actual > _COPY_GBS:
            out["cache_served"] = True
        return out

    r12 = mixed["r12"]
    variant = mixed["md_variant"]
    n_kd, ms_kd = mixed["kernels"][F.KERNEL_KICK_DRIFT]
    rate_kd = {"launches": n_kd, "avg_us": ms_kd / n_kd * 1e3}
    kbase = {0: "k_md_link", 1: "k_md_link_xcd", 2: "k_md_site_xcd", 3: "k_md_site_xcd", 4: "k_md_site_tile"}
    kname = (f"{kbase.get(variant, 'k_md')}<4,float,float,EXP> "
             f"(fused force + momentum kick + exp link drift, MD variant {variant}"
             f"{', links read as two rows' if r12 else ''})")
    roofline = md_roofline(rate_kd, nlink, 4, r12, kname)
    traffic = measured_traffic(variant, r12) if args.extent == 32 else None
    roofline["traffic"] = traffic
    if traffic is not None:
        roofline["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same kernel build, "
                                      "scaled by the copy-kernel calibration (profiles/kick_drift_32x4_f32_pmc.md)")
    roofline["avg_us"] = round(rate_kd["avg_us"], 2)
    if (variant, r12) in VALU_PER_SITE:
        # the second roof of this kernel: vector ALU issue (no MFMA, no packed FP32)
        valu = VALU_PER_SITE[(variant, r12)]
        floor_s = (nlink // 4) / 64 * valu / VALU_WAVE_INSTR_PER_S
        roofline["valu"] = {"instr_per_site": valu, "floor_us": round(floor_s * 1e6, 1),
                            "frac": round(floor_s * 1e6 / rate_kd["avg_us"], 4)}
    kernels = {}
    for leg, tag, elem in ((mixed, "f32_", 4), (f64, "f64_", 8)):
        lr12 = leg["r12"]
        rows = [("kick_drift", F.KERNEL_KICK_DRIFT, REALS_STEP, REALS_KICK_DRIFT_R12 if lr12 else REALS_KICK_DRIFT),
                ("kick", F.KERNEL_KICK, REALS_FORCE, REALS_KICK_R12 if lr12 else REALS_KICK),
                # H_U and H_E sums in one kernel (plaquette + E energy)
                ("hamiltonian_sums", F.KERNEL_PLAQUETTE, REALS_PLAQ + REALS_EENERGY, REALS_PLAQ + REALS_EENERGY),
                ("reunitarize", F.KERNEL_REUNIT, REALS_REUNIT, REALS_REUNIT),
                ("momenta", F.KERNEL_MOMENTA, REALS_MOMENTA, REALS_MOMENTA)]
        for name, kind, ra, rr in rows:
            kernels[tag + name] = kernel_line(leg, kind, ra, rr, elem)

    gauss = gauss_projection_cost(lqcd, args) if args.gauss_checks > 0 else None
    errors, cpu = (oracle_error_table(lqcd, args) if args.oracle_trajectory else (None, None))
    cpu_ref = cpu_reference_integrator(args, args.cpu_md_steps) if args.cpu_md_steps > 0 else None
    if cpu is not None and cpu_ref is not None:
        cpu["reference_integrator"] = cpu_ref

    def acceptance_note(leg):
        if leg["acceptance"] > 0:
            return None
        return (f"every timed trajectory was rejected: mean |dH| {leg['mean_abs_dh']:.1f} with momenta of "
                f"standard deviation 0.5/beta; |dH| grows with the volume at fixed dt, so the kept links "
                f"only change through the reunitarisation after each trajectory")

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
            "workload": (f"HMC {args.extent}^4 D=4 beta={args.beta}, hot start (device stream, seed "
                         f"{args.seed:#x}), {args.md_steps} leapfrog steps dt={args.dt}, exp link update, "
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
        "acceptance_note": {"f32_mixed": acceptance_note(mixed), "f64": acceptance_note(f64)},
        "plaquette": {"start": mixed["plaquette_start"], "f32_mixed": mixed["plaquette"],
                      "f64": f64["plaquette"]},
        "dH_f32_minus_f64_first_trajectories": [
            a - b for a, b in zip(mixed["dh"][: len(f64["dh"])], f64["dh"])][:3],
        "roofline": roofline,
        "gauss_projection": gauss,
        "mixed_precision_errors_vs_oracle": errors,
        "stream_copy_GB_s": round(_COPY_GBS, 1),
        "stream_read_GB_s": round(_READ_GBS, 1),
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
