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
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "lattice-qcd-rs_amd"))

METRIC = ("link updates/s + % HBM peak; FP32-mixed vs FP64 HMC traj/s at matched ΔH, 32^4")
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, AMD specification
REALS_KICK_DRIFT = 52   # read U (18) + E (8), write E (8) + U (18) per link
REALS_KICK = 34         # read U (18) + E (8), write E (8)
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
    p.add_argument("--cpu-md-steps", type=int, default=3,
                   help="MD steps of the CPU baseline sample (0 = skip)")
    p.add_argument("--seed", type=int, default=0x457893F44AB067F0)
    p.add_argument("--copy-only", action="store_true",
                   help="only run the 1 GiB streaming copy (byte-counter calibration)")
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
    hmc = lqcd.HybridMonteCarloDiagnostic(args.dt, args.md_steps, integ, args.seed)
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
        "seconds": t1 - t0,
        "traj_per_s": steps / (t1 - t0),
        "mean_abs_dh": float(sum(abs(x) for x in dh) / len(dh)),
        "acceptance": acc / steps,
        "plaquette": plaq,
        "nlink": nlink,
        "kernels": kern,
        "dh": dh,
    }


def measured_traffic(variant):
    """HBM bytes per kick+drift launch (fetched + written) from the committed
    rocprofv3 counter summary, when it covers the selected MD variant."""
    path = os.path.join(ROOT, "profiles", "kick_drift_32x4_f32_pmc.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch variant={variant}:"):
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

    kd = kernel_line(mixed, F.KERNEL_KICK_DRIFT, REALS_KICK_DRIFT, 4)
    variant = mixed["md_variant"]
    traffic = measured_traffic(variant) if args.extent == 32 else None
    roofline = {
        "bound": "hbm",
        "achieved": kd["GB_s"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kd["hbm_frac"],
        "traffic": traffic,
        "kernel": (f"k_md_{'site' if variant >= 2 else 'link'}_xcd<4,float,float,EXP> "
                   f"(fused force + momentum kick + exp link drift, MD variant {variant})"),
        "bytes_per_launch": nlink * REALS_KICK_DRIFT * 4,
    }
    kernels = {}
    for name, kind, reals in [("kick_drift", F.KERNEL_KICK_DRIFT, REALS_KICK_DRIFT),
                              ("kick", F.KERNEL_KICK, REALS_KICK),
                              ("plaquette", F.KERNEL_PLAQUETTE, REALS_PLAQ),
                              ("efield_energy", F.KERNEL_EFIELD_ENERGY, REALS_EENERGY),
                              ("reunitarize", F.KERNEL_REUNIT, REALS_REUNIT),
                              ("momenta", F.KERNEL_MOMENTA, REALS_MOMENTA)]:
        kernels["f32_" + name] = kernel_line(mixed, kind, reals, 4)
        kernels["f64_" + name] = kernel_line(f64, kind, reals, 8)

    cpu = cpu_baseline(args) if args.cpu_md_steps > 0 else None
    copy_gbs = stream_copy_gbs(lqcd)

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
                         "FP32 links/E with FP64 dH/plaquette/accept/reunitarisation per trajectory"),
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
        "stream_copy_GB_s": round(copy_gbs, 1),
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
