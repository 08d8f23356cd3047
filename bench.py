"""Benchmark of the nonbonded hot path on one GPU.

Headline: the 64-job FEP lambda-window serving mix (1,000-30,000-atom SPC
water jobs with 3-30 soft-core atoms, seeded Poisson arrivals, 2,000 md steps
each) through the continuous-batching server in mixed precision: aggregate pair
interactions/s, ns/day and p50/p99 step latency, against the same jobs run one
at a time on the plain FP64 GPU path.  One md step = Forcefield::apply (GPU
charge-group pair list every 5 steps, LJ + reaction-field pair loops, RF
excluded terms, soft-core corrections) + leap-frog + SETTLE + E_kin, the
sequence of gromos-rs md (md.rs:1143-1286).  Pairs are counted as the
reference evaluates them (n_i * n_j per charge-group pair, 9 per water pair).

Beside it: the 32,768-water box in both precisions, the 128-replica exchange
run, force-error and drift comparisons of the mixed path, the roofline of the
dominant kernel (its launches timed with HIP events on the batch's stream) and
the CPU baseline (the oracle's OpenMP restatement of the rayon loops on a
bounded sample, all threads of the host share and one thread).
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 60          # algorithmic count per pair interaction (r, r2, 1/r2, r^-6,
                            # E_LJ, F_LJ, 1/r, q, E_CRF, F_CRF, F, accumulate, energies, virial)
VIRIAL_FLOP = 18            # the 3x3 virial's share of it (9 mul + 9 add)
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix are the same figure)
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 (packed; this build keeps FP32 un-packed: 78.6 effective)
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--latency-steps", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=15, help="timed oracle MD steps of the CPU baseline (about 10-15 s)")
    ap.add_argument("--cpu-steps-1core", type=int, default=5,
                    help="timed oracle MD steps of the single-core CPU baseline (one pair-list cycle, about 10 s)")
    ap.add_argument("--only", choices=["fep-mixed", "water", "remd"], default=None,
                    help="run one part only (for profiling): the FEP mix's batched mixed replay, the water box, "
                         "or the replica-exchange batch (--remd-steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (HBM counter calibration) and exit")
    ap.add_argument("--fep-steps", type=int, default=2000, help="md steps per job of the 64-job FEP mix")
    ap.add_argument("--fep-alone-steps", type=int, default=200,
                    help="steps per job timed for the one-job-at-a-time plain FP64 comparison")
    ap.add_argument("--remd-steps", type=int, default=10000, help="steps of the 128-replica exchange run")
    ap.add_argument("--fep-admit-interval", type=int, default=200,
                    help="steps between batch layout changes of the FEP server (a multiple of 5); the mixed "
                         "server also runs at 50 for the latency side of the trade-off")
    ap.add_argument("--no-remd", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=2000, help="NVE steps of the equilibrated drift comparison")
    ap.add_argument("--drift-replicas", type=int, default=8,
                    help="independent copies in the ensemble drift comparison (0: skip)")
    return ap.parse_args()


def copy_calibration(nbytes):
    """One torch device copy of known size: FETCH_SIZE / WRITE_SIZE of this
    dispatch calibrate the counters against algorithmic bytes."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "read": n * 4, "written": n * 4}))


def cpu_baseline(s, nsteps, nsteps_1core):
    """The oracle (CPU restatement of the reference algorithm) on the same
    32,768-water system: initial pair list + one warm step untimed, then
    `nsteps` timed steps (a pair-list rebuild every 5 steps, as on the GPU) on
    the host's thread share, and `nsteps_1core` on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    out = {}
    for key, nthr, n in (("all", threads, nsteps), ("one", 1, nsteps_1core)):
        if n <= 0:
            continue
        oracle.set_threads(nthr)
        o = oracle.OracleSystem(s)
        md = oracle.MD(o, use_settle=True, parallel=True)
        md.run(1)
        p0 = md.pair_count
        t0 = time.perf_counter()
        md.run(n)
        dt = time.perf_counter() - t0
        pairs = md.pair_count - p0
        out[key] = dict(value=pairs / dt, unit="pair interactions/s", cores=nthr, kind="port",
                        sample=f"{n} MD steps of the same 32,768-water system (steps 2-{n + 1}, the O(N^2) standard pair "
                               f"list rebuilt every 5 steps), OpenMP fold/reduce restatement of the rayon loops on "
                               f"{nthr} thread(s); {dt:.2f} s wall, {dt / n * 1e3:.1f} ms/step")
    oracle.set_threads(threads)
    res = dict(out.get("all", out.get("one")))
    res["cpu_model"] = cpu_model()
    if "one" in out and "all" in out:
        res["single_core"] = out["one"]
    return res


def traffic_from_profile(kernel, tag):
    """HBM bytes per step of the solvent pair kernels measured by rocprofv3
    (FETCH_SIZE, WRITE_SIZE scaled by a copy of known size), read from the
    committed summary: the far-pair kernel's dispatch plus, where the summary
    has it, the near-pair kernel's."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{tag}.md")
    try:
        for line in open(path):
            if line.startswith("machine-readable:"):
                kv = dict(t.split("=", 1) for t in line.split(":", 1)[1].split() if "=" in t)
                total = float(kv["traffic_bytes"]) + float(kv.get("traffic_bytes_near", 0.0))
                return total, os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        pass
    return None, None


def run_path(rt, s, prec, steps, warmup, latency_steps):
    import numpy as np
    import torch
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    if warmup:
        ds.md(warmup)
    ds.kernel_timing(enable=1)
    p0 = ds.pair_count
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = ds.md(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms, launches = ds.kernel_timing(enable=0)
    kern_pairs = ds.kernel_pairs
    pairs = ds.pair_count - p0
    lat = []
    for _ in range(latency_steps):
        t = time.perf_counter()
        ds.md(1)
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat) if lat else np.array([float("nan")])
    ds.close()
    tot = e[:, 2] + e[:, 3]
    t_ps = np.arange(len(tot)) * s.dt
    drift = float(np.polyfit(t_ps, tot, 1)[0]) if len(tot) > 1 else float("nan")
    return dict(drift=drift, pairs_per_s=pairs / wall, ms_per_step=wall / steps * 1e3, wall=wall,
                ns_per_day=steps * s.dt * 1e-3 * 86400.0 / wall,
                p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                kern_ms=kern_ms / max(launches, 1), launches=launches, kern_pairs=kern_pairs,
                pairs_per_step=pairs / steps, e_first=float(e[0, 2] + e[0, 3]), e_last=float(e[-1, 2] + e[-1, 3]))


def force_error(rt, s):
    """mixed vs plain FP64 GPU forces on the same configuration (the plain path
    matches the CPU oracle to 1e-10 in the parity tests)"""
    import numpy as np
    out = {}
    for prec in (rt.FP64, rt.MIXED):
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        out[prec] = ds.forcefield_apply()
        ds.close()
    f64, fmx = out[rt.FP64]["forces"], out[rt.MIXED]["forces"]
    dn = np.linalg.norm(fmx - f64, axis=1)
    d = dn / np.maximum(np.linalg.norm(f64, axis=1), 1.0)
    rms = float(np.sqrt(np.mean(np.sum(f64 * f64, axis=1))))
    e64 = out[rt.FP64]["e_lj"] + out[rt.FP64]["e_crf"]
    emx = out[rt.MIXED]["e_lj"] + out[rt.MIXED]["e_crf"]
    # max / median of |dF_i| / max(|F_i|, 1 kJ/mol/nm) per atom (the stated measure), and
    # for reference the largest |dF_i| against the RMS force of the system
    return dict(max=float(d.max()), median=float(np.median(d)), f_floor=1.0,
                max_abs_over_rms_force=float(dn.max() / rms), rms_force=rms,
                energy_rel=float(abs(emx - e64) / abs(e64)))


def progress(msg):
    """a line on stderr as each part of the run ends (stdout carries only the result)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); independent jobs shard across GPUs "
              "without per-step communication, see the design notes", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    gen = importlib.import_module("gromos-rs_amd.generators")
    rt = importlib.import_module("gromos-rs_amd.runtime")
    if a.only == "fep-mixed":
        # one batched mixed-precision replay of the FEP mix, nothing else (for profiling)
        fep = fep_mix_bench(rt, gen, a.fep_steps, a.fep_alone_steps, a.fep_admit_interval, only_mixed=True)
        print(json.dumps({"fep_mix": fep}), flush=True)
        return
    if a.only == "remd":
        # the 128-replica exchange batch only (for profiling)
        print(json.dumps({"remd": remd_bench(rt, gen, a.remd_steps)}), flush=True)
        return
    if a.only == "water":
        # the mixed path on the water box only (for profiling)
        print(json.dumps({"water_32768_mixed": run_path(rt, gen.water_32768(), rt.MIXED, a.steps, a.warmup, 0)}),
              flush=True)
        return
    s = gen.water_32768()
    water = water_bench(rt, gen, s, a)
    progress("water_32768 done")
    fep = fep_mix_bench(rt, gen, a.fep_steps, a.fep_alone_steps, a.fep_admit_interval, drift_steps=a.drift_steps)
    bm = fep["batched_mixed"]
    kms, klaunch, kpairs = bm["kernel_timing"]
    flop_per_pair = FLOP_PER_PAIR - VIRIAL_FLOP      # the md steps' solvent launches take no virial
    achieved = kpairs * flop_per_pair / (kms * 1e-3) / 1e12 if kms > 0 else None
    traffic, traffic_src = traffic_from_profile("k_vv_mx", "fep_mix_64jobs")
    out = {
        "metric": "pair interactions/s + aggregate ns/day; p50/p99 step latency vs plain FP64 GPU",
        "value": bm["pairs_per_s"],
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": bm["session_steps"],
        "warmup": a.warmup,
        "ms_per_step": bm["wall_s"] / bm["session_steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+int64",
        "data": "synthetic",
        "config": {"workload": "fep_mix: 64 FEP lambda-window jobs of SPC water, 1,000-30,000 atoms (log-uniform), "
                               "3-30 perturbed atoms each (soft-core alpha_LJ 0.5, alpha_CRF 0.5 nm^2), lambda from 21 "
                               "windows, seeded Poisson arrivals, rc 1.4 nm, NSNB 5, 2,000 md steps per job at 2 fs, "
                               "continuous batching on one GPU, mixed precision",
                   "jobs": fep["jobs"], "atoms_total": fep["atoms"], "steps_per_job": fep["steps_per_job"],
                   "precision": "mixed", "steps_meaning": "session steps (one md step of every resident job)"},
        "ns_per_day": bm["ns_per_day"],
        "step_latency_ms": {"p50": bm["step_ms_p50"], "p99": bm["step_ms_p99"]},
        "speedup_vs_plain_fp64_gpu_one_at_a_time": fep["speedup_batched_mixed_vs_one_at_a_time_fp64"],
        "plain_fp64_gpu_one_at_a_time": fep["one_at_a_time_fp64"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS if achieved else None, "traffic": traffic,
                     "kernel": "k_vv_mx + k_vv_nr (solvent-solvent pair kernels of one md step: far pairs in "
                               "FP32, near pairs in FP64; mixed precision, no virial)",
                     "avg_launch_ms": kms / max(klaunch, 1), "launches": klaunch,
                     "flop_per_launch": kpairs * flop_per_pair / max(klaunch, 1),
                     "flop_per_pair": flop_per_pair, "traffic_source": traffic_src,
                     "note": "vector-ALU bound pair kernels (no dense contraction, no MFMA); each step's far-pair "
                             "launch (batch's stream) and near-pair launch (side stream, beside it) timed with HIP "
                             "events as one span, from the first start to the last end, over the timed replay; "
                             "work = the advancing jobs' solvent-solvent pair interactions x 42 FLOP (60 minus the "
                             "18 virial FLOP); traffic = both kernels' HBM bytes per step; peak = MI355X FP32 157.3 "
                             "TFLOP/s (packed; this build keeps FP32 un-packed: 78.6 reachable)"},
        "fep_mix": fep,
        "water_32768": water,
    }
    if not a.no_remd:
        out["remd"] = remd_bench(rt, gen, a.remd_steps)
        progress("remd done")
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s, a.cpu_steps, a.cpu_steps_1core)
        progress("cpu baseline done")
    print(json.dumps(out), flush=True)


def water_bench(rt, gen, s, a):
    """32,768 SPC waters: device-resident MD in both precisions, the mixed path's
    solvent kernel timed per launch, force error and drift comparisons."""
    res = {name: run_path(rt, s, rt.FP64 if name == "fp64" else rt.MIXED, a.steps, a.warmup, a.latency_steps)
           for name in ("fp64", "mixed")}
    h, p = res["mixed"], res["fp64"]
    flop_per_pair = FLOP_PER_PAIR - VIRIAL_FLOP
    flop = h["kern_pairs"] / max(h["launches"], 1) * flop_per_pair
    achieved = flop / (h["kern_ms"] * 1e-3) / 1e12
    traffic, traffic_src = traffic_from_profile("k_vv_mx", "water32768_mixed")
    out = {"workload": f"water_32768: 32,768 SPC waters (98,304 atoms), cubic box {s.box[0]:.4f} nm (0.997 g/cm3), "
                       "single-range cutoff 1.4 nm, RF eps_rf 61, pair list every 5 steps, SETTLE, dt 2 fs, "
                       "device-resident MD",
           "steps": a.steps, "warmup": a.warmup,
           "mixed": {"pairs_per_s": h["pairs_per_s"], "ms_per_step": h["ms_per_step"], "ns_per_day": h["ns_per_day"],
                     "step_latency_ms": {"p50": h["p50"], "p99": h["p99"]}, "pairs_per_step": h["pairs_per_step"]},
           "plain_fp64_gpu": {"pairs_per_s": p["pairs_per_s"], "ms_per_step": p["ms_per_step"],
                              "ns_per_day": p["ns_per_day"], "step_latency_ms": {"p50": p["p50"], "p99": p["p99"]},
                              "vv_kernel_ms": p["kern_ms"],
                              "vv_kernel_tflops": p["kern_pairs"] / max(p["launches"], 1) * flop_per_pair /
                              (p["kern_ms"] * 1e-3) / 1e12},
           "speedup_mixed_vs_plain_fp64_gpu": h["pairs_per_s"] / p["pairs_per_s"],
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                        "kernel": "k_vv_mx + k_vv_nr (one system: run one after the other on one stream)",
                        "avg_launch_ms": h["kern_ms"], "launches": h["launches"], "flop_per_launch": flop,
                        "flop_per_pair": flop_per_pair, "traffic_source": traffic_src},
           "energy_drift_kj_mol_ps": {"fp64": p["drift"], "mixed": h["drift"],
                                      "rel_diff": abs(h["drift"] - p["drift"]) / abs(p["drift"]),
                                      "steps": a.steps, "start": "lattice"},
           "mixed_force_error_vs_fp64": force_error(rt, s)}
    progress("water_32768 timed runs done")
    if a.drift_replicas > 0:
        out["energy_drift_ensemble"] = drift_ensemble_bench(rt, gen, a.drift_steps, a.drift_replicas)
    return out


def fep_mix_bench(rt, gen, steps, alone_steps, admit_interval=200, only_mixed=False, drift_steps=2000):
    """The 64-job FEP lambda-window mix: seeded Poisson arrivals, continuous
    batching (mixed precision, then plain FP64), against the same jobs run one at
    a time on the plain FP64 single-job path; a wall-clock serving run (arrivals
    while the GPU works); the mixed path's force error on every job and its NVE
    drift against FP64 on a subset."""
    import numpy as np
    srv = importlib.import_module("gromos-rs_amd.server")
    mix = gen.fep_mix()
    systems = [gen.fep_job(m["n_atoms"], m["n_pert"], m["lam"], m["seed"]) for m in mix]
    # synthetic lattice waters relax first (untimed), then run as NVT windows:
    # Berendsen 300 K, tau 0.1 ps, dof 3N - 3 per rigid water
    progress("fep_mix jobs generated")
    systems = srv.equilibrate(systems, [300.0] * len(systems))
    progress("fep_mix jobs equilibrated")
    nve = [q.copy() for q in systems]
    for q in systems:
        q.thermostat = (300.0, 0.1, 3 * q.n_atoms - 3 * q.nsm)
    sched = [dict(key=k, system=systems[k], nsteps=steps, arrival=m["arrival"]) for k, m in enumerate(mix)]
    out = dict(jobs=len(mix), atoms=int(sum(q.n_atoms for q in systems)), steps_per_job=steps,
               arrival_rate_per_step=0.05, seed=3, rc=1.4, nsnb=5, thermostat="Berendsen 300 K, tau 0.1 ps",
               admit_interval=admit_interval,
               arrivals_last_step=int(max(m["arrival"] for m in mix)))
    runs = [("batched_mixed", rt.MIXED, admit_interval)]
    if not only_mixed:
        runs.append(("batched_fp64", rt.FP64, admit_interval))
        if admit_interval != 50:
            runs.append(("batched_mixed_admit50", rt.MIXED, 50))
    for name, prec, ai in runs:
        server = srv.Server(precision=prec, nsnb=5, admit_interval=ai)
        timed = name == "batched_mixed"
        if timed:
            server.sess.kernel_timing(enable=1)
        r = srv.replay(server, sched)
        sim_ns = len(mix) * steps * 0.002 * 1e-3
        out[name] = dict(pairs_per_s=r["pairs"] / r["wall"], wall_s=r["wall"], ns_per_day=sim_ns * 86400.0 / r["wall"],
                         step_ms_p50=float(np.percentile(r["step_ms"], 50)),
                         step_ms_p99=float(np.percentile(r["step_ms"], 99)),
                         job_latency_s_p50=float(np.percentile(r["job_latency_s"], 50)),
                         job_latency_s_p99=float(np.percentile(r["job_latency_s"], 99)),
                         admit_interval=ai, layouts=r["layouts"], layout_ms_total=r["layout_ms"],
                         layout_concat_ms_total=r["layout_concat_ms"], session_steps=r["clock"],
                         failed_jobs=sum(1 for v in server.results.values() if v["failed"]))
        if timed:
            out[name]["kernel_timing"] = server.sess.kernel_timing(enable=0)
        server.sess.close()
        progress(f"fep_mix {name} done ({r['wall']:.2f} s)")
    if only_mixed:
        return out
    # wall-clock serving: a client thread submits each job at its arrival time
    # (arrival step x the batched mixed path's mean step time) while the server
    # thread steps the resident batch
    sps = out["batched_mixed"]["wall_s"] / out["batched_mixed"]["session_steps"]
    ts = srv.ThreadedServer(precision=rt.MIXED, nsnb=5, admit_interval=admit_interval)
    rr = srv.replay_realtime(ts, sched, sps)
    progress("fep_mix realtime done")
    out["realtime_mixed"] = dict(pairs_per_s=rr["pairs"] / rr["wall"], wall_s=rr["wall"],
                                 seconds_per_arrival_step=sps, layouts=rr["layouts"],
                                 job_latency_s_p50=float(np.percentile(rr["job_latency_s"], 50)),
                                 job_latency_s_p99=float(np.percentile(rr["job_latency_s"], 99)),
                                 admit_wait_s_p50=float(np.percentile(rr["admit_wait_s"], 50)),
                                 admit_wait_s_p99=float(np.percentile(rr["admit_wait_s"], 99)),
                                 failed_jobs=sum(1 for v in ts.results.values() if v["failed"]),
                                 mode="arrivals in wall-clock time from a client thread; the server thread owns the "
                                      "session and admits at its 200-step boundaries")
    # one at a time, plain FP64 (rate over alone_steps per job, pair-list cadence included)
    pairs, wall = 0, 0.0
    soft, per_step = 0, 0.0
    import torch
    for q in systems:
        ds = rt.DeviceSystem(q, precision=rt.FP64)
        ds.pairlist_update()
        # soft-core pairs (first atom perturbed) of the starting pair list, untimed
        pf = np.asarray(q.pert.flag) != 0
        pl = ds.pairlist()
        soft += sum(int(pf[pl[k][:, 0]].sum()) for k in ("solute_short", "solute_long") if len(pl[k]))
        p0 = ds.pair_count
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ds.md(alone_steps)
        wall += time.perf_counter() - t0
        pairs += ds.pair_count - p0
        per_step += (ds.pair_count - p0) / alone_steps
        ds.close()
    out["one_at_a_time_fp64"] = dict(pairs_per_s=pairs / wall, steps_per_job_timed=alone_steps, wall_s=wall)
    progress("fep_mix one at a time done")
    # soft-core pairs reported separately (they are part of the pair counts above)
    out["softcore_pairs"] = dict(per_step_all_jobs=soft, fraction_of_pairs=soft / per_step,
                                 source="starting pair lists of the 64 jobs (pairs whose first atom is perturbed)")
    out["speedup_batched_mixed_vs_one_at_a_time_fp64"] = out["batched_mixed"]["pairs_per_s"] / (pairs / wall)
    out["speedup_batched_fp64_vs_one_at_a_time_fp64"] = out["batched_fp64"]["pairs_per_s"] / (pairs / wall)
    if "batched_mixed_admit50" in out:
        out["speedup_batched_mixed_admit50_vs_one_at_a_time_fp64"] = (out["batched_mixed_admit50"]["pairs_per_s"] /
                                                                      (pairs / wall))
    out["mixed_force_error_vs_fp64"] = mix_force_error(rt, systems)
    progress("fep_mix force error done")
    if drift_steps > 0:
        out["nve_drift"] = nve_drift(rt, nve, drift_steps, "all 64 jobs of the mix, equilibrated start, "
                                     "thermostat off, batched")
    return out


def mix_force_error(rt, systems, start="each job's state after 300 untimed Berendsen steps"):
    """Mixed vs plain FP64 GPU forces (the plain path matches the CPU oracle to
    1e-10 in the parity tests) on each system: per atom |dF_i| / max(|F_i|,
    1 kJ/mol/nm), over all atoms of all systems."""
    import numpy as np
    errs, pert_errs, job_max = [], [np.zeros(0)], []
    for q in systems:
        out = {}
        for prec in (rt.FP64, rt.MIXED):
            ds = rt.DeviceSystem(q, precision=prec)
            ds.pairlist_update()
            out[prec] = ds.forcefield_apply()["forces"]
            ds.close()
        d = np.linalg.norm(out[rt.MIXED] - out[rt.FP64], axis=1) / np.maximum(np.linalg.norm(out[rt.FP64], axis=1), 1.0)
        errs.append(d)
        if q.pert is not None:
            pert_errs.append(d[np.asarray(q.pert.flag) != 0])
        job_max.append(float(d.max()))
    e = np.concatenate(errs)
    pe = np.concatenate(pert_errs)
    return dict(max=float(e.max()), median=float(np.median(e)), p99_9=float(np.percentile(e, 99.9)),
                atoms=int(e.size), systems=len(systems), systems_over_1e_5=int(sum(m > 1e-5 for m in job_max)),
                perturbed_atoms_max=float(pe.max()) if pe.size else None, f_floor=1.0, start=start)


def nve_drift(rt, systems, steps, what):
    """NVE total-energy slope (least squares over `steps` steps) of each system in
    FP64 and in mixed precision from the same start, all batched in one session
    per precision; ensemble means with standard errors and the paired
    difference per system."""
    import numpy as np
    slopes = {}
    for name, prec in (("fp64", rt.FP64), ("mixed", rt.MIXED)):
        sess = rt.Session(precision=prec, nsnb=5)
        jobs = []
        for q in systems:
            q = q.copy()
            q.thermostat = None
            jobs.append(sess.admit(q, steps))
        sess.step(steps)
        sl = []
        for j, q in zip(jobs, systems):
            st, done = sess.query(j)
            if st != rt.DONE or done != steps:
                raise rt.GromosAmdError(f"drift job {j} ended in state {st} after {done} steps")
            e = sess.energies(j, 0, steps)
            sl.append(float(np.polyfit(np.arange(steps) * q.dt, e[:, 2] + e[:, 3], 1)[0]))
        sess.close()
        slopes[name] = np.array(sl)
    f, m = slopes["fp64"], slopes["mixed"]
    sem = lambda v: float(v.std(ddof=1) / np.sqrt(len(v))) if len(v) > 1 else float("nan")
    paired = (m - f) / np.abs(f)
    return {"systems": what, "n": len(systems), "steps": steps,
            "fp64_mean": float(f.mean()), "fp64_sem": sem(f), "mixed_mean": float(m.mean()), "mixed_sem": sem(m),
            "rel_diff_of_means": float(abs(m.mean() - f.mean()) / abs(f.mean())),
            "paired_rel_diff_mean": float(paired.mean()), "paired_rel_diff_sem": sem(paired),
            "fp64": f.tolist(), "mixed": m.tolist()}


def remd_bench(rt, gen, steps):
    """128 replicas of 4,000 SPC waters (12,000 atoms, rc 1.2 nm), temperatures
    280 * (450/280)^(k/127) K, NVE segments with temperature exchange every 500
    steps (odd-even pairs, velocity rescaling), all replicas batched."""
    import numpy as np
    remd = importlib.import_module("gromos-rs_amd.remd")
    n = 128
    temps = [280.0 * (450.0 / 280.0) ** (k / (n - 1)) for k in range(n)]
    systems = [gen.water_replica(seed=100 + k, temperature=temps[k], n_waters=4000, rc=1.2) for k in range(n)]
    srv = importlib.import_module("gromos-rs_amd.server")
    systems = srv.equilibrate(systems, temps)          # untimed relaxation of the lattice starts
    progress("remd equilibrated")
    sess = rt.Session(precision=rt.MIXED, nsnb=5)
    rx = remd.ReplicaExchange(sess, systems, temps, exchange_interval=500, scheme="odd_even", seed=7,
                              total_steps=steps)
    sess.step_times(enable=1)
    i0 = sess.info()
    t0 = time.perf_counter()
    rx.run(steps)
    wall = time.perf_counter() - t0
    i1 = sess.info()
    ms = sess.step_times(enable=0)
    seg = min(500, steps)
    drift, failed, final = [], 0, []
    for j, q in zip(rx.jobs, systems):
        st, done = sess.query(j)
        failed += st == rt.FAILED
        e = sess.energies(j, 0, done)
        tot = e[:, 2] + e[:, 3]
        # slope of E_tot within each NVE segment (exchanges rescale velocities
        # between segments), averaged per replica
        sl = [np.polyfit(np.arange(seg) * 0.002, tot[a:a + seg], 1)[0] for a in range(0, done - seg + 1, seg)]
        drift.append(float(np.mean(sl)) if sl else float("nan"))
        x, v = sess.fetch(j)
        f = q.copy()
        f.pos, f.vel = x, v
        final.append(f)
    sess.close()
    drift = np.array(drift)
    ferr = mix_force_error(rt, final, start="every replica's final state")
    return dict(replicas=n, atoms_per_replica=systems[0].n_atoms, box_nm=float(systems[0].box[0]), rc=1.2,
                steps=steps, exchange_interval=500, pairs_per_s=(i1["pair_count"] - i0["pair_count"]) / wall,
                ns_per_day=n * steps * 0.002 * 1e-3 * 86400.0 / wall, wall_s=wall,
                step_ms_p50=float(np.percentile(ms, 50)), step_ms_p99=float(np.percentile(ms, 99)),
                exchange_attempts=rx.attempts, exchange_accepted=rx.accepted,
                failed_replicas=int(failed),
                nve_drift_kj_mol_ps_per_replica=dict(median=float(np.median(drift)), min=float(drift.min()),
                                                     max=float(drift.max()),
                                                     median_per_atom=float(np.median(drift)) / systems[0].n_atoms,
                                                     note="mean over the replica's 500-step NVE segments of the "
                                                          "least-squares E_tot slope"),
                mixed_force_error_vs_fp64=ferr,
                precision="mixed", start="lattice waters relaxed 300 steps under Berendsen tau 0.01 ps (untimed)")


def drift_ensemble_bench(rt, gen, steps, replicas):
    """Drift over an ensemble: `replicas` copies of the 32,768-water system
    (lattice seeds 2, 3, ...) relaxed together (300 Berendsen steps, tau 0.01 ps,
    300 K), then NVE for `steps` steps in FP64 and in mixed precision from the
    same states.  A single trajectory's slope carries sampling noise
    (trajectories from nearby states separate within a few hundred steps), so
    the comparison is the ensemble mean with its standard error, and per
    replica the paired difference mixed - FP64."""
    srv = importlib.import_module("gromos-rs_amd.server")
    systems = [gen.water_32768(seed=2 + k) for k in range(replicas)]
    eq = srv.equilibrate(systems, [300.0] * replicas)
    return nve_drift(rt, eq, steps, "lattice seeds 2..%d relaxed 300 steps under Berendsen tau 0.01 ps" % (1 + replicas))


if __name__ == "__main__":
    main()
