#!/usr/bin/env python3
"""DEM-MPH explicit step on one MI355X: particle-steps/s on the large single scene.

Workload (``dam_break_c2``): 3D dam break in a 0.8 x 0.4 x 0.6 m tank at
d = 0.004 m -- 562,500 fluid particles, 340,908 wall particles (three layers,
open top) and 64 rigid debris bodies of 10x10x10 particles with seeded
densities in [500, 900] kg/m^3 -- 967,408 particles, dt = 2e-5 s.  FP32 pair
arithmetic on cell-relative positions, FP64 rigid-body reductions/integration
and FP64 time.  Inputs are resident in HBM before the timed region.

One JSON line on stdout:
  value      particle-steps per second over the timed steps (whole job = one GPU)
  roofline   the dominant kernel group's (the pair sweeps') algorithmic HBM
             bytes per step over their average device time per step (HIP events
             on the library's stream), vs 8 TB/s
  cpu_baseline  the FP64 oracle (this package's C restatement of the reference,
             OpenMP on the host cores) on a bounded sample of the same scene
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s and sim-s/wall-s on 1 MI355X, plain vs mixed/batched/learned"

# Algorithmic HBM bytes per unit and launch (FP32 state), the per-particle
# contract figures of the step: each unit's own inputs read once and outputs
# written once; neighbour gathers re-read arrays already counted and are
# served from cache in the ideal.
#   sweep1 (every particle): P-density, divergence, A-density, GC, alpha
#           32 read + 28 written                                           = 60
#   sweep2 (fluid + solid slots): all forces and the fused integration
#           56 read + 28 written (force 12, velocity 12, PressureP copy 4) = 84
#   dem (solid + wall slots): pos 12 + vel 12 + type 4 + key 4 + body 4 + force 12
#           + four coefficients 16 read; force 12 + coefficients 16 written = 64 + 28 = 92
# The whole step: 244 B per particle (binning 24, rank/scatter 8, reorder 72,
# sweep 1 60, sweep 2 84 -- the last two counted for every particle).
ALGO_BYTES = {"sweep1": 60.0, "sweep2": 84.0, "dem": 92.0}
STEP_BYTES = 244.0
# The CPU baseline is this package's FP64 restatement (oracle/); the reference
# solver built from its own sources ran slower on the same input files in the
# build container (8 threads, oracle/calibrate_cpu.py): C2 8.48 s/step against
# 4.39 s/step, C1 171 ms/step against 97 ms/step.
PORT_OVER_REFERENCE = {"dam_break_c2": 1.93, "dam_break_c1": 1.77}


def build_workload(seed: int = 1):
    from dem_mph_amd import scenes as S
    sc = S.dam_break_c2(seed=seed, rigid=True)
    return sc, S.params_c2()


def cpu_baseline(sc, params, budget_s: float = 20.0):
    """The FP64 oracle (OpenMP, all host threads) on the same scene: set-up is
    not timed; steps are run until about `budget_s` seconds of wall time.
    Returns the baseline record and the oracle, left at that step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    if not os.path.exists(P.LIB_PATH):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except Exception:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", threads))
    o = P.Oracle(P.params_from_dict(params), sc, sort_mode=P.SORT_CANONICAL, rigid=True, threads=threads)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 50:
            break
    rec = {"value": sc.n * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": "FP64 oracle, rigid debris, same %d-particle scene, %d steps after set-up (%.1f s)"
                     % (sc.n, steps, el),
           "port_over_reference_speed": PORT_OVER_REFERENCE["dam_break_c2"],
           "calibration": "the reference built from its own sources ran this scene (loose grains) at 8.48 s/step, "
                          "the port at 4.39 s/step, 8 threads each, same files (oracle/calibrate_cpu.py)"}
    return rec, o, steps


def _nrel(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def balance(mass, vel, typ):
    """Total mass and momentum of the moving particles (fluid and debris)."""
    mv = (typ < 4)
    return float(mass[mv].sum()), (mass[mv, None] * vel[mv]).sum(axis=0)


# the solver variants timed on C2: options of dem_mph_amd.Solver
VARIANTS = {
    "mixed": dict(precision=1, compensated=True, rigid_precision=0),    # FP32 pairs, FP64 sums, FP64 bodies
    "plain": dict(precision=1, compensated=False, rigid_precision=1),   # FP32 everywhere but the clock
    "fp64": dict(precision=0),                                           # FP64 everywhere (the parity mode)
}


GPU_FIELDS = ("pos", "vel", "vol_strain_p", "pressure_p", "mass", "force")


def gpu_state(s, lo=0, hi=None, b0=0, nb=None):
    """The fields the errors compare, for particles [lo, hi) of a handle (one
    scene of a batch) and its bodies [b0, b0 + nb), in original order."""
    st = {f: s.field(f)[lo:hi] for f in GPU_FIELDS}
    b = s.bodies()
    st["body_x"] = b["x"][b0:None if nb is None else b0 + nb]
    st["body_v"] = b["v"][b0:None if nb is None else b0 + nb]
    return st


def field_errors(st, o, sc, steps):
    """Per-field errors of a GPU state (gpu_state) against the FP64 oracle at
    the same step, plus mass and momentum balance of both."""
    typ = o.field("type")
    fl, so = typ < 2, (typ >= 2) & (typ < 4)
    dx = st["pos"] - o.field("pos")
    L = sc.domain_max - sc.domain_min
    dx -= L * np.round(dx / L)
    n0p = o.scalars()["n0p"]
    nd, ndr = st["vol_strain_p"] + n0p, o.field("vol_strain_p") + n0p
    m, mr = st["mass"], o.field("mass")
    F, Fr = st["force"], o.field("force")
    v, vr = st["vel"], o.field("vel")
    bx, bv = o.get("body_x").reshape(-1, 3), o.get("body_v").reshape(-1, 3)
    M, P = balance(m, v, typ)
    Mr, Pr = balance(mr, vr, typ)
    return {
        "steps": steps,
        "position_max_abs_m": float(np.abs(dx).max()),
        "velocity_norm_rel": _nrel(v, vr),
        "number_density_max_rel": float(np.max(np.abs(nd - ndr) / np.abs(ndr))),
        "pressure_p_norm_rel": _nrel(st["pressure_p"][fl], o.field("pressure_p")[fl]),
        "fluid_accel_norm_rel": _nrel(F[fl] / m[fl, None], Fr[fl] / mr[fl, None]),
        "solid_force_norm_rel": _nrel(F[so], Fr[so]),
        "body_position_max_abs_m": float(np.abs(st["body_x"] - bx).max()) if len(bx) else 0.0,
        "body_velocity_norm_rel": _nrel(st["body_v"], bv) if len(bv) else 0.0,
        "mass_total": M, "mass_total_oracle": Mr,
        "momentum": [float(x) for x in P], "momentum_oracle": [float(x) for x in Pr],
        "momentum_rel_diff": float(np.linalg.norm(P - Pr) / max(np.linalg.norm(Pr), 1e-300)),
    }


def errors_vs_oracle(D, sc, params, variant, o, steps, cell_scale=1.0):
    """Per-field errors of a fresh GPU run of `steps` steps (the variant's cell
    size) against the FP64 oracle at the same step, plus mass and momentum
    balance of both."""
    s = D.Solver(D.Params.from_dict(params), sc, rigid=True, cell_scale=cell_scale, **VARIANTS[variant])
    s.step(steps)
    s.sync()
    out = field_errors(gpu_state(s), o, sc, steps)
    s.close()
    return out


def ensemble_errors_vs_oracle(scenes, snaps, params, steps):
    """The checker for the batched ensemble: each sampled scene of the batched
    mixed handle (its state after `steps` steps, taken from the batch) against
    the FP64 oracle run on that scene alone for the same steps (its own skin
    schedule; the batch rebins when any scene needs it, which moves only the
    rebinning steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as P
    out = {}
    for k, st in snaps.items():
        o = P.Oracle(P.params_from_dict(params), scenes[k], sort_mode=P.SORT_CANONICAL, rigid=True)
        o.step(steps)
        out["scene_%d" % k] = dict(field_errors(st, o, scenes[k], steps), particles=int(scenes[k].n))
        o.close()
    return out


def copy_probe(mib: int) -> None:
    from dem_mph_amd.profiling import copy_probe as probe
    probe(mib)
    print("bench: copy probe of %d MiB done" % mib, file=sys.stderr, flush=True)


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "dam_break_c2_fp32_traffic.json")


def _params_c3():
    from dem_mph_amd import scenes as S
    return S.params_c3()


def kernel_utilisation(kernel_name: str):
    """VALU issue and vector-L1 lookup utilisation of a kernel from the
    committed rocprofv3 counters (the same summary as the traffic), or None."""
    try:
        js = json.load(open(TRAFFIC_JSON))
    except Exception:
        return None
    out = {}
    for sec, keys in (("counters", ("valu_issue_util",)), ("l1", ("l1_lookup_util", "lookups_per_vmem"))):
        for name, v in js.get(sec, {}).items():
            if kernel_name in name:
                out.update({k: v[k] for k in keys if k in v})
    return out or None


def kernel_traffic(kernel_name: str):
    """HBM bytes per launch of a kernel group (every kernel whose name contains
    `kernel_name`, summed) from the committed rocprofv3 summary (FETCH_SIZE +
    WRITE_SIZE, calibrated by a plain copy), or None."""
    try:
        js = json.load(open(TRAFFIC_JSON))
    except Exception:
        return None
    hits = [v["bytes"] for name, v in js.get("traffic", {}).items() if kernel_name in name]
    return sum(hits) if hits else None


def time_variant(D, sc, params, variant, steps, warmup, cell_scale=-1.0):
    """Timed region: `steps` replays of the captured step (hipGraph), bracketed
    by device synchronisation.  Then the same number of steps launched one by
    one with HIP events around each kernel group, on the library's stream, for
    the per-kernel durations.  FP32 variants tune the cell size at creation by
    default (cell_scale < 0: each candidate width on a handle of its own)."""
    s = D.Solver(D.Params.from_dict(params), sc, rigid=True, graph=True, cell_scale=cell_scale, **VARIANTS[variant])
    info = s.info()
    s.step(warmup)
    s.sync()
    lanes = s.autotune(5)   # lanes per particle and workgroup size of the two sweeps, timed on this scene
    s.step(warmup)
    s.sync()
    reb0 = s.info().rebuilds
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    el = time.perf_counter() - t0
    reb = s.info().rebuilds
    s.set_timing(True)
    t1 = time.perf_counter()
    s.step(steps)
    s.sync()
    el_ev = time.perf_counter() - t1
    tm = s.timing()
    s.set_timing(False)
    blocks = list(s.info().block)
    wall_lanes = int(s.info().wall_lanes)
    s.close()
    return {"seconds": el, "timing": tm, "info": info, "rebuilds": int(reb - reb0), "seconds_timed_launches": el_ev,
            "lanes": list(lanes), "blocks": blocks, "wall_lanes": wall_lanes, "cell_scale": float(info.cell_scale)}


def full_config(D, sc, params, variant, steps, cell_scale=1.0):
    """The whole C2 configuration as stated: `steps` steps from t = 0 (every
    rebuild of the run included), after autotuning on the initial state, with
    the cell size the variant's tuning chose."""
    s = D.Solver(D.Params.from_dict(params), sc, rigid=True, graph=True, cell_scale=cell_scale, **VARIANTS[variant])
    lanes = s.autotune(3)
    s.step(1)   # the graph capture, outside the clock; the run below starts at step 1
    s.sync()
    reb0 = s.info().rebuilds
    t0 = time.perf_counter()
    s.step(steps - 1)
    s.sync()
    el = time.perf_counter() - t0
    reb = s.info().rebuilds - reb0
    t_end = s.time()
    s.close()
    dt = params["Dt"]
    return {"steps": steps, "timed_steps": steps - 1, "ms_per_step": el / (steps - 1) * 1e3,
            "particle_steps_per_s": sc.n * (steps - 1) / el, "sim_s_per_wall_s": dt * (steps - 1) / el,
            "rebuilds": int(reb), "simulated_s": t_end, "autotuned_lanes": list(lanes), "cell_scale": cell_scale}


def ensemble_bench(D, nscenes: int, steps: int, warmup: int, seed: int, sample: int):
    """The batched ensemble: `nscenes` scenes in one handle (one graph per step
    for all of them) against the plain FP32 solver advancing the same scenes one
    handle at a time (timed on `sample` evenly spaced scenes, scaled to all)."""
    from dem_mph_amd import scenes as S
    t0 = time.perf_counter()
    scenes = S.ensemble_c3(nscenes, seed=seed)
    t_gen = time.perf_counter() - t0
    params = S.params_c3()
    dt = params["Dt"]
    n = sum(sc.n for sc in scenes)
    out = {"workload": "ensemble_c3: %d dam-break scenes in the C1 tank at d = 0.01 m, fluid 2k-60k (log-uniform), "
                       "1-32 rigid cuboids of side 3-8, dt = 5e-5 s" % nscenes,
           "scenes": nscenes, "particles": n, "steps": steps, "generation_s": t_gen}
    res = {}
    snaps = {}   # sampled scenes of the batched mixed handle after its last timed step, for the checker
    sample_k = sorted({0, nscenes // 3, (2 * nscenes) // 3, nscenes - 1})
    cs = -1.0   # the mixed handle tunes the cell size at creation; the plain one reuses its choice
    for name, var in (("batched_mixed", "mixed"), ("batched_plain", "plain")):
        t0 = time.perf_counter()
        s = D.Solver(D.Params.from_dict(params), scenes, rigid=True, cell_scale=cs, **VARIANTS[var])
        t_create = time.perf_counter() - t0
        cs = float(s.info().cell_scale)
        s.step(warmup)
        s.sync()
        lanes = s.autotune(3)
        s.step(2)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        el = time.perf_counter() - t0
        if name == "batched_mixed":
            off = s.scene_offsets
            nbod = [int(sc.body.max()) + 1 if sc.body is not None and (sc.body >= 0).any() else 0 for sc in scenes]
            for k in sample_k:
                snaps[k] = gpu_state(s, off[k], off[k + 1], sum(nbod[:k]), nbod[k])
        wall_lanes = int(s.info().wall_lanes)
        s.close()
        res[name] = {"particle_steps_per_s": n * steps / el, "ms_per_step": el / steps * 1e3,
                     "sim_s_per_wall_s": dt * steps / el, "create_s": t_create, "autotuned_lanes": list(lanes),
                     "autotuned_wall_lanes": wall_lanes, "cell_scale": cs}
        print("bench: ensemble %s %.2f ms/step" % (name, el / steps * 1e3), file=sys.stderr, flush=True)
    # plain FP32 solver, one scene per handle, scenes one after another (each autotuned)
    idx = np.linspace(0, nscenes - 1, min(sample, nscenes)).round().astype(int)
    el = 0.0
    n_s = 0
    for k in idx:
        s = D.Solver(D.Params.from_dict(params), scenes[k], rigid=True, **VARIANTS["plain"])
        s.step(warmup)
        s.sync()
        s.autotune(2)
        s.step(2)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        el += time.perf_counter() - t0
        n_s += scenes[k].n
        s.close()
    scale = n / n_s   # sample time scaled by particle count to the whole ensemble
    el_all = el * scale
    res["sequential_plain"] = {"particle_steps_per_s": n * steps / el_all, "ms_per_step": el_all / steps * 1e3,
                               "sim_s_per_wall_s": dt * steps / el_all,
                               "sample": "%d of %d scenes (%d particles), each autotuned, time scaled by particle "
                                         "count" % (len(idx), nscenes, n_s)}
    out.update(res)
    out["batched_mixed_over_sequential"] = (res["batched_mixed"]["sim_s_per_wall_s"] /
                                            res["sequential_plain"]["sim_s_per_wall_s"])
    out["batched_mixed_over_batched_plain"] = (res["batched_mixed"]["sim_s_per_wall_s"] /
                                               res["batched_plain"]["sim_s_per_wall_s"])
    out["_snapshots"] = (snaps, warmup + 2 + steps)
    return out, scenes


SURROGATE_WEIGHTS = os.path.join(ROOT, "dem_mph_amd", "surrogate_c3.npz")


def judge_summary(j, tol):
    """The judge's checks of a hybrid run against the true error at each check."""
    return {"checks": len(j), "fell_back_checks": sum(1 for x in j if x["fell_back"]),
            "true_error_m_mean": float(np.mean([x["true_error_m"] for x in j])) if j else None,
            "estimate_m_mean": float(np.mean([x["estimate_m"] for x in j])) if j else None,
            "kept_with_true_error_over_tol": sum(1 for x in j if not x["fell_back"] and x["true_error_m"] > tol),
            "fell_with_true_error_within_tol": sum(1 for x in j if x["fell_back"] and x["true_error_m"] <= tol)}


def channel_bench(D, steps: int, warmup: int, seed: int, surrogate: bool, config_steps: int = 1000):
    """The large single scene (C5 channel, 12.9M particles): the mixed solver
    timed over `steps` replays of the captured step after autotuning; the
    learned surrogate's rollout throughput on the same particles; and the
    surrogate with its plain-solver fallback run as deployed over the
    configuration's `config_steps` steps from t = 0 (judged once, halfway),
    with its debris-centroid error against the mixed-precision solver's
    trajectory.  The network is trained on C3 scenes (d = 0.01 m, the C1 tank);
    here it runs at this scene's spacing and in this channel."""
    from dem_mph_amd import scenes as S
    t0 = time.perf_counter()
    sc = S.channel_c5(seed=seed)
    t_gen = time.perf_counter() - t0
    params = S.params_c5()
    dt = params["Dt"]
    n = sc.n
    cnt = sc.counts()
    out = {"workload": "channel_c5: 1.6x0.8x0.8 m channel, d = 0.002 m, %d fluid (released column), %d debris "
                       "in %d rigid cuboids of side 4-12, %d wall, dt = 1e-5 s" % (cnt["fluid"], cnt["solid"],
                                                                                  cnt["bodies"], cnt["wall"]),
           "particles": n, "steps": steps, "generation_s": t_gen}
    t0 = time.perf_counter()
    s = D.Solver(D.Params.from_dict(params), sc, precision=D.PREC_FP32, rigid=True, compensated=True, graph=True)
    out["create_s"] = time.perf_counter() - t0
    s.step(warmup)
    s.sync()
    out["autotuned_lanes"] = list(s.autotune(3))
    s.step(2)
    s.sync()
    reb0 = s.info().rebuilds
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    el = time.perf_counter() - t0
    out["rebuilds_in_timed_region"] = int(s.info().rebuilds - reb0)
    s.close()
    out["mixed"] = {"particle_steps_per_s": n * steps / el, "ms_per_step": el / steps * 1e3,
                    "sim_s_per_wall_s": dt * steps / el}
    print("bench: channel mixed %.2f ms/step" % (el / steps * 1e3), file=sys.stderr, flush=True)
    if surrogate and os.path.exists(SURROGATE_WEIGHTS):
        from dem_mph_amd import surrogate as SG
        # the graph radius follows this scene's spacing (3 d), as in training
        r = SG.advance_batch(SG.load(SURROGATE_WEIGHTS).at_spacing(sc.d), [sc], 2, dt)
        t = r["seconds_per_step"]
        out["surrogate"] = {"ms_per_step": t * 1e3, "stride_dt": SG.STRIDE, "finite": r["finite"],
                            "sim_s_per_wall_s": SG.STRIDE * dt / t, "particle_steps_per_s": n * SG.STRIDE / t,
                            "vs_mixed": (SG.STRIDE * dt / t) / (dt * steps / el),
                            }
        print("bench: channel surrogate %.1f ms/step" % (t * 1e3), file=sys.stderr, flush=True)
        if config_steps > 0:
            # as deployed: solver warm-up, surrogate strides, one judgement by a
            # solver probe halfway, fallback from the last anchored solver state
            strides = config_steps // SG.STRIDE - (SG.HIST - 1)
            tol = 0.01 * 1.6   # 1 % of the channel length
            sur = SG.load(SURROGATE_WEIGHTS).at_spacing(sc.d).in_tank((0.0, 0.0, 0.0), (1.6, 0.8, 0.8))
            print("bench: channel surrogate + fallback, %d strides" % strides, file=sys.stderr, flush=True)
            hyb = SG.hybrid_report(sur, [sc], params, strides, reanchor=strides // 2, tol=tol)
            j = hyb.pop("judgements")
            hyb["judge"] = dict(judge_summary(j, tol), invalid_states=hyb.pop("invalid_states", 0))
            out["surrogate"]["hybrid"] = hyb
            print("bench: channel surrogate + fallback %.1f s (plain solver %.1f s)"
                  % (hyb["seconds"], hyb["plain_batched_seconds"]), file=sys.stderr, flush=True)
    return out


def surrogate_bench(D, scenes, seq_ms_per_dt: float, batched_plain_ms_per_dt: float, chunk: int, steps: int,
                    eval_scenes: int, eval_steps: int, hybrid_strides: int, seed: int):
    """The learned surrogate on the ensemble.
    - hybrid: the surrogate with its plain-solver fallback as a deployment runs
      it (dem_mph_amd.surrogate.hybrid: solver warm-up, re-anchoring and a
      probe-based judge every 200 strides, fallen-back scenes on the solver),
      timed end to end from t = 0 on `eval_scenes` held-out scenes for
      `hybrid_strides` strides, against the plain FP32 solver batch advancing
      the same scenes over the same simulated time; accuracy against the
      mixed-precision solver's trajectory.
    - rollout_512: the surrogate alone on all scenes (in chunks), per stride.
    - accuracy_vs_baselines: 100-stride rollouts of the held-out scenes re-
      anchored on the FP64 solver's trajectory (an evaluation of the network,
      not a deployment), next to rollouts without a network."""
    from dem_mph_amd import scenes as S
    from dem_mph_amd import surrogate as SG
    if not os.path.exists(SURROGATE_WEIGHTS):
        return {"skipped": "no trained weights at dem_mph_amd/surrogate_c3.npz"}
    sur = SG.load(SURROGATE_WEIGHTS)
    params = S.params_c3()
    dt = params["Dt"]
    ev = S.ensemble_c3(eval_scenes, seed=seed + 1000)
    print("bench: surrogate + fallback on %d held-out scenes, %d strides" % (len(ev), hybrid_strides),
          file=sys.stderr, flush=True)
    hyb = SG.hybrid_report(sur, ev, params, hybrid_strides)
    hyb.pop("judgements_detail", None)
    j = hyb.pop("judgements")
    hyb["judge"] = dict(judge_summary(j, 0.008), invalid_states=hyb.pop("invalid_states", 0))
    # the scene-by-scene plain solver on these scenes, estimated from its
    # measured cost per particle-step on the ensemble (sequential_plain)
    n_ev = sum(sc.n for sc in ev)
    n_all = sum(sc.n for sc in scenes)
    seq_s = seq_ms_per_dt * 1e-3 * (n_ev / n_all) * hyb["simulated_s"] / dt
    hyb["sequential_plain_seconds_estimate"] = seq_s
    hyb["vs_sequential_plain_estimate"] = seq_s / hyb["seconds"]
    t_step = 0.0
    finite = True
    nch = (len(scenes) + chunk - 1) // chunk   # equal chunks of at most `chunk` scenes
    cuts = np.linspace(0, len(scenes), nch + 1).round().astype(int)
    for k in range(nch):
        r = SG.advance_batch(sur, scenes[cuts[k]:cuts[k + 1]], steps, dt)
        t_step += r["seconds_per_step"]
        finite = finite and r["finite"]
        print("bench: surrogate chunk %d: %.1f ms/step for %d nodes" % (k, r["seconds_per_step"] * 1e3, r["nodes"]),
              file=sys.stderr, flush=True)
    n = sum(sc.n for sc in scenes)
    roll = {"scenes": len(scenes), "surrogate_ms_per_stride": t_step * 1e3, "stride_dt": SG.STRIDE,
            "sim_s_per_wall_s": SG.STRIDE * dt / t_step, "particle_steps_per_s": n * SG.STRIDE / t_step,
            "finite": finite,
            "vs_sequential_plain": (SG.STRIDE * dt / t_step) / (dt / (seq_ms_per_dt * 1e-3)),
            "vs_batched_plain": (SG.STRIDE * dt / t_step) / (dt / (batched_plain_ms_per_dt * 1e-3))}
    acc = SG.evaluate(sur, ev, params, 2000, eval_steps, 200, 0.008, fp64=True)
    return {"hybrid": hyb, "rollout_512": roll, "accuracy_vs_baselines": acc}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain variant (FP32 throughout)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 variant")
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--config-steps", type=int, default=2000,
                    help="steps of the whole-configuration run from t = 0 (0: skip)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-scenes", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=40)
    ap.add_argument("--ensemble-sample", type=int, default=32)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate measurement")
    ap.add_argument("--surrogate-chunk", type=int, default=96,
                    help="scenes per surrogate batch (96: ~7.8M nodes, within the 4 GB row offsets of the edge gathers)")
    ap.add_argument("--surrogate-eval-scenes", type=int, default=64)
    ap.add_argument("--hybrid-strides", type=int, default=400,
                    help="surrogate strides of the timed surrogate + fallback run (two re-anchor periods)")
    ap.add_argument("--no-channel", action="store_true", help="skip the large-channel (C5) measurement")
    ap.add_argument("--channel-steps", type=int, default=100)
    ap.add_argument("--channel-config-steps", type=int, default=1000,
                    help="steps of the channel's surrogate-with-fallback run from t = 0 (0: skip)")
    ap.add_argument("--cell-scale", type=float, default=-1.0,
                    help="cell width of the C2 handles over the reference's (default: tuned at creation over 1, "
                         "1.25 and 1.5; a fixed value keeps the tuner's extra handles out of a profile)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (counter calibration)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print("bench.py measures one GPU: the DEM-MPH step here runs on a single MI355X "
              "(multi-GPU slab decomposition is designed, not built); use --gpus 1", file=sys.stderr)
        return 2

    import dem_mph_amd as D
    if a.copy_probe:
        copy_probe(a.copy_probe)
    sc, params = build_workload(a.seed)
    prec = D.PREC_FP32 if a.precision == "fp32" else D.PREC_FP64
    head = "mixed" if prec == D.PREC_FP32 else "fp64"
    n = sc.n
    dt = params["Dt"]
    print("bench: %d particles, headline variant %s" % (n, head), file=sys.stderr, flush=True)
    mixed = time_variant(D, sc, params, head, a.steps, a.warmup, a.cell_scale)
    variants = {head: mixed}
    if not a.no_plain and prec == D.PREC_FP32:
        print("bench: plain FP32 (sums and rigid bodies in FP32)", file=sys.stderr, flush=True)
        variants["plain"] = time_variant(D, sc, params, "plain", a.steps, a.warmup, a.cell_scale)
    if not a.no_fp64 and prec == D.PREC_FP32:
        print("bench: FP64 (the parity mode)", file=sys.stderr, flush=True)
        variants["fp64"] = time_variant(D, sc, params, "fp64", a.fp64_steps, a.warmup)
    info = mixed["info"]
    el = mixed["seconds"]
    pss = n * a.steps / el
    groups = {k: v for k, v in mixed["timing"].items() if v[1] > 0}
    # The dominant kernel group: the pair sweeps.  Without surface tension they
    # run as pass A (sweep 1's densities and divergence with the viscous,
    # drag and buoyancy forces) and pass B (the fluid pressure force and the
    # velocity update), so their algorithmic bytes are counted together: 60 B
    # per particle (sweep 1) + 84 B per fluid and solid slot (sweep 2), over
    # the two groups' device time per step.
    ms1, l1 = groups["sweep1"]
    ms2, l2 = groups["sweep2"]
    avg_s = (ms1 / l1 + ms2 / l2) / 1e3
    fs = info.solid_end - info.fluid_begin
    algo_bytes = ALGO_BYTES["sweep1"] * n + ALGO_BYTES["sweep2"] * fs
    achieved = algo_bytes / avg_s / 1e9
    # the kernels the autotuner picked (the profile holds every candidate's launches)
    lanes = mixed["lanes"]
    knames = ["k_pass_a_fw<%d, %d, true>" % (lanes[0], mixed["wall_lanes"]), "k_pass_a_s<%d, true>" % lanes[0],
              "k_pass_b<%d, true>" % lanes[1]]
    traffic = [kernel_traffic(k) for k in knames]
    step_achieved = pss * STEP_BYTES / 1e9
    out = {
        "metric": METRIC,
        "value": pss,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == D.PREC_FP32 else "f64",
        "data": "synthetic (seeded dam-break scene, jitter +-0.01 d)",
        "config": {"workload": "dam_break_c2: 3D tank 0.8x0.4x0.6 m, d=0.004 m, %d particles "
                               "(562500 fluid, 64 rigid debris bodies of 10^3, 340908 wall), dt=2e-5 s" % n,
                   "particles": n, "bodies": info.nbodies, "precision": "mixed: FP32 pairs, FP64 sums, "
                   "FP64 rigid bodies and time" if prec == D.PREC_FP32 else "fp64",
                   "sort": "canonical", "rebuild": "skin 0.05 d"},
        "sim_s_per_wall_s": dt * a.steps / el,
        "rebuilds_in_timed_region": mixed["rebuilds"],
        "kernel_ms_per_step": {k: v[0] / a.steps for k, v in groups.items()},
        "ms_per_step_timed_launches": mixed["seconds_timed_launches"] / a.steps * 1e3,
        "variants": {k: {"particle_steps_per_s": n * (a.fp64_steps if k == "fp64" else a.steps) / v["seconds"],
                         "ms_per_step": v["seconds"] / (a.fp64_steps if k == "fp64" else a.steps) * 1e3,
                         "sim_s_per_wall_s": dt * (a.fp64_steps if k == "fp64" else a.steps) / v["seconds"],
                         "steps": a.fp64_steps if k == "fp64" else a.steps,
                         "options": {kk: vv for kk, vv in VARIANTS[k].items()},
                         "autotuned_lanes": v["lanes"], "autotuned_blocks": v["blocks"],
                         "autotuned_wall_lanes": v["wall_lanes"], "autotuned_cell_scale": v["cell_scale"]}
                     for k, v in variants.items()},
        "roofline": {"bound": "hbm", "kernel": "sweeps: pass A (%s, %s) + pass B (%s)" % tuple(knames),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": sum(traffic) if all(t is not None for t in traffic) else None,
                     "algorithmic_bytes": algo_bytes,
                     "algorithmic_bytes_per_unit": {"per particle (sweep 1)": ALGO_BYTES["sweep1"],
                                                    "per fluid or solid slot (sweep 2)": ALGO_BYTES["sweep2"]},
                     "units": {"particles": n, "fluid_and_solid_slots": fs},
                     "avg_ms_per_step": avg_s * 1e3,
                     # what bounds the sweeps below the HBM roof: VALU issue and
                     # vector-L1 tag lookups, from the same committed profile
                     "utilisation": {k: kernel_utilisation(k) for k in knames},
                     "whole_step": {"algorithmic_bytes_per_particle": STEP_BYTES, "achieved": step_achieved,
                                    "frac": step_achieved / HBM_PEAK_GBS}},
    }
    if a.config_steps > 0 and prec == D.PREC_FP32:
        print("bench: the whole C2 configuration, %d steps from t = 0" % a.config_steps, file=sys.stderr, flush=True)
        out["full_config"] = {v: full_config(D, sc, params, v, a.config_steps,
                                             variants[v]["cell_scale"] if v in variants else 1.0)
                              for v in ("mixed", "plain")}
    if not a.no_cpu_baseline:
        print("bench: timing the CPU baseline (bounded sample)", file=sys.stderr, flush=True)
        rec, o, k = cpu_baseline(sc, params, a.cpu_budget)
        out["cpu_baseline"] = rec
        print("bench: errors against the FP64 oracle after %d steps" % k, file=sys.stderr, flush=True)
        out["errors_vs_fp64_oracle"] = {v: errors_vs_oracle(D, sc, params, v, o, k, variants[v]["cell_scale"])
                                        for v in variants}
        o.close()
    else:
        out["cpu_baseline"] = None
    # the secondary workloads: a failure is recorded in its own entry (with the
    # error) and the headline line is still printed; the process exits non-zero
    failed = []

    def section(name, fn):
        try:
            return fn()
        except Exception as e:   # noqa: BLE001 -- recorded, and the exit status reports it
            import traceback
            traceback.print_exc()
            failed.append(name)
            return {"error": "%s: %s" % (type(e).__name__, e)}

    if not a.no_ensemble:
        print("bench: batched ensemble", file=sys.stderr, flush=True)
        res = section("ensemble", lambda: ensemble_bench(D, a.ensemble_scenes, a.ensemble_steps, 5, a.seed,
                                                         a.ensemble_sample))
        ens, scenes = res if isinstance(res, tuple) else (res, None)
        snaps = ens.pop("_snapshots", None)
        out["ensemble"] = ens
        if snaps is not None and not a.no_cpu_baseline:
            print("bench: ensemble scenes against the FP64 oracle", file=sys.stderr, flush=True)
            ens["batched_mixed"]["errors_vs_fp64_oracle"] = section(
                "ensemble_errors", lambda: ensemble_errors_vs_oracle(scenes, snaps[0], _params_c3(), snaps[1]))
        if not a.no_surrogate and scenes is not None:
            print("bench: learned surrogate", file=sys.stderr, flush=True)
            ens["surrogate"] = section("surrogate", lambda: surrogate_bench(
                D, scenes, ens["sequential_plain"]["ms_per_step"], ens["batched_plain"]["ms_per_step"],
                a.surrogate_chunk, 3, a.surrogate_eval_scenes, 100, a.hybrid_strides, a.seed))
        del scenes
    if not a.no_channel:
        print("bench: large channel", file=sys.stderr, flush=True)
        out["channel"] = section("channel", lambda: channel_bench(D, a.channel_steps, 5, a.seed, not a.no_surrogate,
                                                                  a.channel_config_steps))
    if failed:
        out["failed_sections"] = failed
    print(json.dumps(out))
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
