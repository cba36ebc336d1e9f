"""Benchmark: grid-point updates per second of the FP64 HEVI short step on one
MI355X, with the HBM roofline fraction of the HEVI stage and the CPU oracle
timed beside it.

A step is one full dynamical short step on the device-resident state:
numerical filter (num_diff) + four HEVI stages (RK4) + the RK4 combination,
ghost levels and periodic halos.  Workload: 512x512x128 (DX = 1 km, DZ 50..500 m
stretched, linear shear 0..20 m/s over 10 km, theta noise 0.1 K, UD3, dt = 2 s),
synthetic and seeded; inputs are resident in HBM before timing starts.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size IxJxK]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
# algorithmic bytes per interior point (FP64, flat terrain, divdmp_coef = 0):
# one HEVI stage reads 5 state0 + 5 state + 5 tendencies + DPRES0, RT2P,
# REF_dens, REF_rhot + 15 num_diff (34 arrays, MOMZ0 counted in state0) and
# writes 5 stage outputs + 3 mflx + 3 tflx: 45 arrays x 8 B.
STAGE_BYTES_PER_POINT = 45 * 8
# short step = numfilter (7 read + 15 written) + 4 stages + RK4 combination (15 + 18 read)
STEP_BYTES_PER_POINT = 22 * 8 + 4 * STAGE_BYTES_PER_POINT + 33 * 8  # = 1880
# the RK4 combination (15 stage outputs + 18 flux arrays read) runs inside the
# last stage launch, so one short step's four stage launches move
# 4 * 360 + 264 B per point: 426 B per launch on average
COMBINE_BYTES_PER_POINT = 33 * 8
STAGE_LAUNCH_BYTES_PER_POINT = STAGE_BYTES_PER_POINT + COMBINE_BYTES_PER_POINT / 4
# mixed precision: DPRES0, RT2P, REF_dens, REF_rhot stay FP64, the other 41 arrays are FP32
MIXED_STAGE_BYTES_PER_POINT = 41 * 4 + 4 * 8


def progress(msg):
    """One line per finished leg on stderr, so a long run shows it is alive."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def theta_check(case, nsteps):
    """Mixed precision against FP64 from the same state after nsteps large
    steps (short step and the FCT tracers): max and rms of the potential
    temperature difference, and per prognostic and tracer the largest
    difference, absolute and relative to the field's FP64 range."""
    from scale_amd import hip
    import numpy as np
    g = case.grid
    sl = (slice(g.KS, g.KE + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
    names = list(hip.STATE) + [f"QTRC{iq}" for iq in range(case.qtrc.shape[-1])]
    res = {}
    for mixed in (False, True):
        ctx = hip.model_from_case(case, mixed_precision=mixed)
        ctx.model_large_step(nsteps)
        ctx.synchronize()
        res[mixed] = {n: ctx.model_get(n if n in hip.STATE else hip.QTRC_BASE + int(n[4:]))[sl] for n in names}
        ctx.close()
        progress(f"precision check: {nsteps} large steps, {'mixed' if mixed else 'FP64'} done")
    th = {m: res[m]["RHOT"] / res[m]["DENS"] for m in res}
    diff = np.abs(th[True] - th[False])
    fields = {}
    for n in names:
        a, b = res[False][n], res[True][n]
        rng = float(max(a.max() - a.min(), np.abs(a).max() * 1e-3, 1e-300))
        d = float(np.abs(a - b).max())
        fields[n] = {"max_abs": d, "max_over_range": d / rng}
    return {"steps": nsteps, "large_steps_with_tracers": True, "max_abs_K": float(diff.max()),
            "rms_K": float(np.sqrt((diff * diff).mean())), "fields": fields}


def ensemble_leg(steps=5, tune_steps=2):
    """The 64-member ensemble of uneven domains (24 x 96x96x48, 24 x
    128x128x64, 16 x 192x192x64; mixed precision, one FCT tracer each): large
    steps of all members one model after the other, against the same steps
    as one ensemble (one filter and four stage launches over every member's
    tiles, autotuned shapes, the step replayed from a hipGraph).  Value:
    member grid-point large-step updates per second."""
    from scale_amd import hip, inputs
    cases = inputs.ensemble_members()
    pts = sum(c.grid.imax * c.grid.jmax * c.grid.kmax for c in cases)
    t0 = time.perf_counter()
    ctxs = [hip.model_from_case(c, mixed_precision=True) for c in cases]
    setup_s = time.perf_counter() - t0
    progress(f"ensemble: {len(ctxs)} members set up in {setup_s:.1f} s")
    try:
        for ctx in ctxs:  # warm-up
            ctx.model_large_step(1)
        for ctx in ctxs:
            ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for ctx in ctxs:
                ctx.model_large_step(1)
        for ctx in ctxs:
            ctx.synchronize()
        seq = (time.perf_counter() - t0) / steps
        progress(f"ensemble: members one after the other {seq * 1e3:.2f} ms per large step")
        with hip.Ensemble(ctxs) as ens:
            best_ms, report = ens.tune(tune_steps)
            progress(f"ensemble: tuned, best {best_ms:.2f} ms per large step")
            shape = ens.get_shape()
            ens.set_graph(True)
            ens.large_step(1)  # capture + first replay
            ens.synchronize()
            t0 = time.perf_counter()
            ens.large_step(steps)
            ens.synchronize()
            batched = (time.perf_counter() - t0) / steps
            ens.set_graph(False)
            t0 = time.perf_counter()
            ens.large_step(steps)
            ens.synchronize()
            direct = (time.perf_counter() - t0) / steps
    finally:
        for ctx in ctxs:
            ctx.close()
    return {"workload": "64 members: 24 x 96x96x48, 24 x 128x128x64, 16 x 192x192x64; DX=500 m, DZ=250 m, dt=1 s; "
                        "theta noise N(0, 0.5 K) and u ~ U(5, 15) m/s per member; 1 FCT tracer; mixed precision",
            "members": len(cases), "points": pts, "steps": steps, "setup_s": setup_s,
            "sequential_ms_per_large_step": seq * 1e3, "ensemble_graph_ms_per_large_step": batched * 1e3,
            "ensemble_direct_ms_per_large_step": direct * 1e3, "tuned_best_ms": best_ms,
            "tuned_shape": {"stage_bpc": shape[0], "nf_bpc": shape[1], "nf_tj": shape[2]}, "tune_report": report,
            "member_updates_per_s": pts / batched, "speedup_vs_sequential": seq / batched}


def measured_traffic(size):
    """HBM bytes per HEVI-stage launch from the committed counter profile
    (profiles/hevi_stage_hbm_traffic.json: FETCH_SIZE and WRITE_SIZE averaged
    over the stage launches of one short step, calibrated on the copy probe),
    or None when it was measured on another workload."""
    path = os.path.join(ROOT, "profiles", "hevi_stage_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("workload") != size:
        return None
    return t.get("bytes_per_launch")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", default="512x512x128", help="IMAXxJMAXxKMAX")
    p.add_argument("--cpu-sample", default="128x128x128", help="sub-domain timed on the host with the oracle")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-large-step", action="store_true", help="skip the informational large step with tracers")
    p.add_argument("--no-copy-probe", action="store_true", help="skip the HBM copy probe")
    p.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision (FP32 + FP64 solve) timing")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    p.add_argument("--no-ensemble", action="store_true", help="skip the 64-member ensemble leg")
    p.add_argument("--batched-adjoint", default=None, metavar="MEMBERSxSIZExKMAXxSTEPSxEVERYxTRACERS",
                   help="run the batched terrain adjoint leg, e.g. 32x128x64x100x10x5 (off by default)")
    p.add_argument("--tracer-scheme", type=int, default=None, choices=[1, 2, 3, 4, 5, 6, 31],
                   help="also time large steps whose 3 FCT tracers use this scheme "
                        "(UD1=1, CD2=2, UD3=3, CD4=4, UD5=5, CD6=6, UD3KOREN1993=31; 5 and 6 on a lateral halo of 3)")
    p.add_argument("--adjoint-dirs", type=int, default=64, help="finite-difference directions of the adjoint check")
    p.add_argument("--theta-check", type=int, default=600, metavar="N",
                   help="run N large steps (with the 3 FCT tracers) in FP64 and in mixed precision and report the "
                        "potential-temperature and per-field differences (0 skips it)")
    return p.parse_args()


def adjoint_leg(nsteps, every, ndirs, eps_list, seed=20261202 + 2000):
    """Adjoint against forward cost and finite differences on the warm bubble
    at 128x128x64 (DX = 500 m, DZ = 250 m, u = 10 m/s, theta noise 0.01 K,
    dt = 1 s, 3 tracers with FCT), started from the state after two large
    steps (the initial MOMY = MOMZ = 0 sit on the kink of the upwind switch).
    J = sum over the 16x16x16 box around the bubble of (theta - REF_pott)^2 after
    nsteps large steps; its gradient with respect to the interior state comes
    from the checkpointed forward (every `every` steps) and the adjoint sweep,
    and is compared with central differences of the forward along ndirs
    seeded directions d (standard normal per point times each field's largest
    magnitude, perturbation eps*d): second-order differences at eps = 1e-6 and
    fourth-order ones at eps = 1e-5, whose truncation error is negligible and
    whose rounding error is ten times smaller."""
    import numpy as np
    from scale_amd import hip, inputs
    case = inputs.make_case(kmax=64, imax=128, jmax=128, dx=500.0, dy=500.0, dz=250.0, seed=seed, u0=10.0,
                            noise=0.01, ntracer=3, tracer_kind=["blob", "tophat", "uniform"], dt=1.0, fct=True)
    g = case.grid
    # the 16x16x16 box centred on the bubble (horizontal centre of the domain, z = 0..4 km)
    I, J = g.IS + 64, g.JS + 64
    box = (slice(g.KS, g.KS + 16), slice(I - 8, I + 8), slice(J - 8, J + 8))
    ref_pott = case.ref["REF_pott"]
    names = hip.STATE

    def objective(st):
        th = st["RHOT"] / st["DENS"] - ref_pott
        xb = {k: np.zeros(g.shape, order="F") for k in names}
        xb["RHOT"][box] = 2.0 * th[box] / st["DENS"][box]
        xb["DENS"][box] = -2.0 * th[box] * st["RHOT"][box] / st["DENS"][box] ** 2
        return float(np.sum(th[box] ** 2)), xb

    ctx = hip.model_from_case(case)
    try:
        ctx.model_large_step(2)
        x0 = {k: ctx.model_get(k) for k in names}
        q0 = [ctx.model_get(hip.QTRC_BASE + iq) for iq in range(3)]

        def reset():
            for k in names:
                ctx.model_set(k, x0[k])
            for iq in range(3):
                ctx.model_set(hip.QTRC_BASE + iq, q0[iq])

        # forward cost, and the adjoint: checkpointed forward + reverse sweep.
        # All three are timed alike, by device events around the whole call
        # and none between its kernels (profiling mode 2); the host transfers
        # of the five adjoint fields fall outside the timed regions.
        reset()
        ctx.model_large_step(nsteps)  # warm-up
        reset()
        ctx.synchronize()
        ctx.profile(2)
        t0 = time.perf_counter()
        ctx.model_large_step(nsteps)
        ctx.synchronize()
        fwd_host_ms = (time.perf_counter() - t0) * 1e3
        _, fwd_ms = ctx.profile_query("model_large_step")
        ctx.profile(False)
        def adjoint_run(c, tape):
            """checkpointed forward + reverse sweep with `tape` taped segments
            (0: as many as half the free HBM holds); (fwd_ckpt_ms, rev_ms, xb0, J)"""
            c.tune_set(hip.TUNE_ADJ_TAPE, tape)
            c.profile(2)
            c.model_forward_checkpointed(nsteps, every)
            J, xbn = objective({k: c.model_get(k) for k in names})
            xb = c.model_adjoint(xbn)
            _, f_ms = c.profile_query("model_forward_checkpointed")
            _, r_ms = c.profile_query("model_adjoint")
            c.profile(False)
            return f_ms, r_ms, xb, J

        # the default: the forward keeps every segment's tape in HBM (they fit:
        # 400 B per point and step), so the reverse sweep recomputes nothing
        reset()
        fck_ms, rev_ms, xb0, J0 = adjoint_run(ctx, 0)
        # the classic checkpoint schedule: only the last segment taped, every
        # other segment recomputed once from its checkpoint
        reset()
        fck1_ms, rev1_ms, _, _ = adjoint_run(ctx, 1)
        ctx.tune_set(hip.TUNE_ADJ_TAPE, 0)
        progress(f"adjoint: forward {fwd_ms:.1f} ms, all segments taped {fck_ms + rev_ms:.1f} ms, "
                 f"last segment taped {fck1_ms + rev1_ms:.1f} ms")
        # like for like without tracers: the same dynamics state in a model
        # with no tracers (the reverse sweep has no tracer work either way)
        dyn = {}
        case0 = inputs.make_case(kmax=64, imax=128, jmax=128, dx=500.0, dy=500.0, dz=250.0, seed=seed, u0=10.0,
                                 noise=0.01, ntracer=0, dt=1.0, fct=True)
        ctx0 = hip.model_from_case(case0)
        try:
            def reset0():
                for k in names:
                    ctx0.model_set(k, x0[k])
            reset0()
            ctx0.model_large_step(nsteps)  # warm-up
            reset0()
            ctx0.synchronize()
            ctx0.profile(2)
            ctx0.model_large_step(nsteps)
            _, fwd0_ms = ctx0.profile_query("model_large_step")
            ctx0.profile(False)
            for tape, key in ((0, "all_segments_taped"), (1, "last_segment_taped")):
                reset0()
                f0, r0, _, _ = adjoint_run(ctx0, tape)
                dyn[key] = {"forward_checkpointed_ms": f0, "reverse_ms": r0, "adj_over_fwd": (f0 + r0) / fwd0_ms}
            dyn["fwd_ms"] = fwd0_ms
        finally:
            ctx0.close()
        progress(f"adjoint, dynamics only: forward {fwd0_ms:.1f} ms, adj/fwd {dyn['all_segments_taped']['adj_over_fwd']:.2f} "
                 f"(all taped), {dyn['last_segment_taped']['adj_over_fwd']:.2f} (last segment taped)")
        # gradient w.r.t. the interior: fold the halo and ghost-level adjoints
        grad = inputs.interior_gradient(g, xb0)
        rng = np.random.default_rng(seed)
        scale = {k: float(np.abs(x0[k]).max()) for k in names}
        ads, fd2, fd4, dirs = [], [], [], []
        sweep_dirs = 8
        e2, e4 = eps_list

        def J_at(d, t):
            for k in names:
                ctx.model_set(k, np.asfortranarray(x0[k] + t * d[k]))
            ctx.model_large_step(nsteps)
            d1, r1 = ctx.model_get("DENS"), ctx.model_get("RHOT")
            return float(np.sum((r1[box] / d1[box] - ref_pott[box]) ** 2))

        for idir in range(ndirs):
            if idir % 16 == 0:
                progress(f"adjoint: finite differences, direction {idir} of {ndirs}")
            d = {}
            for k in names:
                a = np.zeros(g.shape, order="F")
                ke = g.KE - 1 if k == "MOMZ" else g.KE
                sl = (slice(g.KS, ke + 1), slice(g.IS, g.IE + 1), slice(g.JS, g.JE + 1))
                a[sl] = rng.standard_normal(a[sl].shape) * scale[k]
                d[k] = inputs.fill_ghost_and_halo(g, a, k)
            ads.append(sum(float(np.sum(grad[k] * d[k])) for k in names))
            # second-order central difference at eps = e2
            fd2.append((J_at(d, e2) - J_at(d, -e2)) / (2.0 * e2))
            # fourth-order central difference at eps = e4
            fd4.append((8.0 * (J_at(d, e4) - J_at(d, -e4)) - (J_at(d, 2 * e4) - J_at(d, -2 * e4))) / (12.0 * e4))
            dirs.append(d)
        # eps sweep of the second-order difference on the first directions:
        # the error falls with eps until no face velocity crosses zero (the
        # upwind switch is a kink of J), then rounding takes over
        sweep = {}
        for e in ((1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9) if dirs else ()):
            errs = []
            for q in range(min(sweep_dirs, len(dirs))):
                fd = (J_at(dirs[q], e) - J_at(dirs[q], -e)) / (2.0 * e)
                errs.append(abs(fd - ads[q]) / abs(ads[q]))
            sweep[f"{e:.0e}"] = {"median_rel_err": float(np.median(errs)), "max_rel_err": float(np.max(errs))}
        progress("adjoint: eps sweep done")
    finally:
        ctx.close()
    return dict(workload="warm bubble 128x128x64, DX=500 m, DZ=250 m, u=10 m/s, dt=1 s, 3 FCT tracers, FP64, UD3",
                steps=nsteps, checkpoint_every=every, J=J0, fwd_ms=fwd_ms, fwd_host_ms=fwd_host_ms,
                tapes="every segment's tape kept in HBM by the checkpointed forward (the default where they fit)",
                forward_checkpointed_ms=fck_ms, reverse_ms=rev_ms, adj_ms=fck_ms + rev_ms,
                adj_over_fwd=(fck_ms + rev_ms) / fwd_ms, reverse_over_fwd=rev_ms / fwd_ms,
                last_segment_taped={"forward_checkpointed_ms": fck1_ms, "reverse_ms": rev1_ms,
                                    "adj_over_fwd": (fck1_ms + rev1_ms) / fwd_ms},
                dynamics_only=dyn,
                fd=fd_summary(ads, fd2, fd4, eps_list) if ndirs > 0 else None,
                fd_eps_sweep=sweep if ndirs > 0 else None)


def batched_adjoint_leg(nmem, size, kmax, nsteps, every, ntracer):
    """The batched terrain adjoint: nmem members of size x size x kmax over a
    bell mountain (h = 500 m, half-width 5 km; DX = DY = 500 m, DZ = 250 m,
    N = 0.012 1/s, u = 10 m/s, theta noise N(0, 0.1 K) per member seed,
    ntracer FCT tracers, FP64, dt = 1 s), nsteps large steps with checkpoints
    every `every`; per member J = sum over the 16x16x16 box above the
    mountain top of (theta - REF_pott)^2 and its gradient with respect to the
    member's state at the start.  Times, by the host clock around calls that
    end with a device synchronisation: the forward of every member
    (interleaved step by step over the members' streams), the batched
    checkpointed forward and the batched reverse sweep (the latter with the
    host transfers of the 5 adjoint fields of each member each way)."""
    import numpy as np
    from scale_amd import hip, inputs
    members = inputs.mountain_members(nmem, imax=size, jmax=size, kmax=kmax, ntracer=ntracer)
    names = hip.STATE
    ctxs = []
    try:
        for case, zs in members:
            ctxs.append(hip.model_from_case(case, zsfc=zs))
        g = members[0][0].grid
        ref_pott = members[0][0].ref["REF_pott"]
        I, J = g.IS + size // 2, g.JS + size // 2
        box = (slice(g.KS + 2, g.KS + 18), slice(I - 8, I + 8), slice(J - 8, J + 8))
        x0 = [{k: c.model_get(k) for k in names} for c in ctxs]
        q0 = [[c.model_get(hip.QTRC_BASE + iq) for iq in range(ntracer)] for c in ctxs]

        def reset():
            for c, x, q in zip(ctxs, x0, q0):
                for k in names:
                    c.model_set(k, x[k])
                for iq in range(ntracer):
                    c.model_set(hip.QTRC_BASE + iq, q[iq])

        def sync():
            for c in ctxs:
                c.synchronize()

        for c in ctxs:  # warm-up
            c.model_large_step(1)
        sync()
        reset()
        sync()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            for c in ctxs:
                c.model_large_step(1)
        sync()
        fwd_s = time.perf_counter() - t0
        progress(f"batched adjoint: forward of {nmem} members x {nsteps} steps {fwd_s:.2f} s")
        reset()
        sync()
        t0 = time.perf_counter()
        hip.batch_forward_checkpointed(ctxs, nsteps, every)
        fck_s = time.perf_counter() - t0
        Js, xbs = [], []
        for c in ctxs:
            st = {k: c.model_get(k) for k in ("DENS", "RHOT")}
            th = st["RHOT"] / st["DENS"] - ref_pott
            xb = {k: np.zeros(g.shape, order="F") for k in names}
            xb["RHOT"][box] = 2.0 * th[box] / st["DENS"][box]
            xb["DENS"][box] = -2.0 * th[box] * st["RHOT"][box] / st["DENS"][box] ** 2
            Js.append(float(np.sum(th[box] ** 2)))
            xbs.append(xb)
        progress(f"batched adjoint: checkpointed forward {fck_s:.2f} s")
        t0 = time.perf_counter()
        got = hip.batch_adjoint(ctxs, xbs)
        rev_s = time.perf_counter() - t0
        norms = [float(np.sqrt(sum(np.sum(inputs.interior_gradient(g, x)[k] ** 2) for k in names))) for x in got]
        progress(f"batched adjoint: reverse sweep {rev_s:.2f} s, adj/fwd {(fck_s + rev_s) / fwd_s:.2f}")
    finally:
        for c in ctxs:
            c.close()
    return dict(workload=f"{nmem} members of {size}x{size}x{kmax} over a bell mountain h=500 m, a=5 km, "
                         f"DX=500 m, DZ=250 m, N=0.012 1/s, u=10 m/s, {ntracer} FCT tracers, FP64, UD3, dt=1 s",
                members=nmem, steps=nsteps, checkpoint_every=every, fwd_s=fwd_s, forward_checkpointed_s=fck_s,
                reverse_s=rev_s, adj_over_fwd=(fck_s + rev_s) / fwd_s,
                member_J=Js, member_gradient_norms=norms,
                precision="FP64 throughout (the mixed-precision adjoint is not built)")


def fd_summary(ads, fd2, fd4, eps_list):
    import numpy as np
    ad = np.array(ads)
    rms = float(np.sqrt(np.mean(ad * ad)))
    out = {"directions": len(ads), "rms_directional_derivative": rms}
    for name, fd, e in (("central_2nd_order", fd2, eps_list[0]), ("central_4th_order", fd4, eps_list[1])):
        err = np.abs(np.array(fd) - ad)
        out[name] = {"eps": e, "median_rel_err": float(np.median(err / np.abs(ad))),
                     "max_rel_err": float(np.max(err / np.abs(ad))), "max_err_over_rms": float(np.max(err) / rms)}
    return out


def host_cpu():
    """(logical CPUs the OS reports, CPU model name) of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


# host threads of the CPU baseline: the GPU box gives one GPU's job a share of
# 16 host threads, whatever the OS reports, so the baseline uses at most 16
CPU_THREADS_MAX = 16


def cpu_baseline(ni, nj, nk, reps=5):
    """The oracle (C, OpenMP) on a bounded sub-domain of the same workload:
    one warm-up short step, then the median of `reps` timed short steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from scale_amd import inputs
    nproc, model = host_cpu()
    threads = min(nproc, CPU_THREADS_MAX)
    po.set_threads(threads)
    case = inputs.shear_512x512x128(imax=ni, jmax=nj, kmax=nk, ntracer=0)
    m = po.Model(case)
    m.step(1)  # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        m.step(1)
        times.append(time.perf_counter() - t0)
    times.sort()
    dt = times[len(times) // 2]
    # tracers off: one large step = one short step (numfilter + RK4) plus the cheap preparation
    pts = ni * nj * nk
    return dict(value=pts / dt, unit="grid-pt short-step updates/s", cores=threads, kind="port",
                host_nproc=nproc, host_cpu=model, step_s_median=dt, step_s_all=times,
                sample=f"oracle (C, OpenMP {threads} threads of {nproc} reported; one GPU's job gets a 16-thread share) "
                       f"on a {ni}x{nj}x{nk} sub-domain of the workload, median of {reps} short steps after 1 warm-up, "
                       f"each incl. the large-step preparation, no tracers")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only (see the design notes)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    ni, nj, nk = (int(x) for x in a.size.split("x"))
    from scale_amd import hip, inputs

    case = inputs.shear_512x512x128(imax=ni, jmax=nj, kmax=nk)
    g = case.grid
    ctx = hip.model_from_case(case)
    ctx.model_prepare()
    ctx.model_short_step(a.warmup)
    ctx.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    ctx.model_short_step(a.steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kern = {}
    for name in ("hevi_stage", "numfilter", "halo_fill", "rk4_combine"):
        cnt, ms = ctx.profile_query(name)
        kern[name] = dict(count=cnt, avg_ms=(ms / cnt if cnt else None))
    ctx.profile(False)
    # large step incl. 3 tracers with FCT (informational), then the same
    # steps with the kernel timers on for the tracer kernels' averages
    large_ms = None
    tracer_kern = None
    if not a.no_large_step:
        ctx.synchronize()
        t2 = time.perf_counter()
        nls = 2
        ctx.model_large_step(nls)
        ctx.synchronize()
        large_ms = (time.perf_counter() - t2) * 1e3 / nls
        ctx.profile(True)
        ctx.model_large_step(nls)
        ctx.synchronize()
        tracer_kern = {}
        for name in ("tracer_nofct", "fct_philo", "fct_ratio", "fct_update"):
            cnt, ms = ctx.profile_query(name)
            if cnt:
                tracer_kern[name] = ms / cnt
        ctx.profile(False)
    # achievable HBM bandwidth: streaming copy of 2 x 2 GiB, and the other
    # probes (non-temporal copy, read-only, write-only, hipMemcpy D2D)
    copy_gbs = None
    probes = None
    if not a.no_copy_probe:
        n_copy = 1 << 28
        copy_ms = ctx.hbm_copy_probe(n_copy, 10)
        copy_gbs = 16.0 * n_copy / (copy_ms * 1e-3) / 1e9
        probes = {}
        for kind in hip.Context.HBM_PROBES:
            ms = ctx.hbm_probe(kind, n_copy, 10)
            nbytes = (8 if kind in ("read", "write") else 16) * n_copy
            probes[kind] = {"ms": ms, "GBs": nbytes / (ms * 1e-3) / 1e9}
    ctx.close()
    progress(f"FP64 short steps: {elapsed * 1e3 / a.steps:.3f} ms/step; large step {large_ms} ms")
    # informational: large steps with the tracers on another flux scheme
    tracer_leg = None
    if a.tracer_scheme is not None:
        ts = a.tracer_scheme
        tcase = inputs.shear_512x512x128(imax=ni, jmax=nj, kmax=nk, scheme_tracer=ts, halo=3 if ts in (5, 6) else 2)
        tctx = hip.model_from_case(tcase)
        tctx.model_prepare()
        tctx.model_large_step(1)
        tctx.synchronize()
        t3 = time.perf_counter()
        tctx.model_large_step(2)
        tctx.synchronize()
        tracer_leg = {"scheme": ts, "halo": tcase.grid.IS, "large_step_ms": (time.perf_counter() - t3) * 1e3 / 2}
        tctx.close()
        print(f"tracer scheme {ts}: large step {tracer_leg['large_step_ms']:.2f} ms", file=sys.stderr, flush=True)

    # mixed precision: the same short steps with FP32 prognostics, fluxes and
    # tendencies and the FP64 vertical solve / reference state
    mixed = None
    if not a.no_mixed:
        mctx = hip.model_from_case(case, mixed_precision=True)
        mctx.model_prepare()
        mctx.model_short_step(a.warmup)
        mctx.synchronize()
        mctx.profile(True)
        t0 = time.perf_counter()
        mctx.model_short_step(a.steps)
        mctx.synchronize()
        m_el = time.perf_counter() - t0
        m_kern = {}
        for name in ("hevi_stage", "numfilter", "halo_fill"):
            cnt, ms = mctx.profile_query(name)
            m_kern[name] = ms / cnt if cnt else None
        mctx.profile(False)
        mctx.close()
        progress(f"mixed short steps: {m_el * 1e3 / a.steps:.3f} ms/step")
        mixed = dict(value=ni * nj * nk * a.steps / m_el, ms_per_step=m_el * 1e3 / a.steps,
                     stage_ms=m_kern["hevi_stage"], kernels_avg_ms=m_kern, speedup_vs_fp64=elapsed / m_el,
                     stage_bytes_per_point=MIXED_STAGE_BYTES_PER_POINT)
    adj = None
    if not a.no_adjoint:
        adj = adjoint_leg(40, 8, a.adjoint_dirs, (1e-6, 1e-5))
        progress(f"adjoint: adj/fwd {adj['adj_over_fwd']:.3f}")
    ens = None
    if not a.no_ensemble:
        ens = ensemble_leg()
        progress(f"ensemble: {ens['ensemble_graph_ms_per_large_step']:.2f} ms/large step "
                 f"({ens['speedup_vs_sequential']:.2f}x the sequential members)")
    badj = None
    if a.batched_adjoint:
        nm, sz, kz, ns, ev, nt = (int(x) for x in a.batched_adjoint.split("x"))
        badj = batched_adjoint_leg(nm, sz, kz, ns, ev, nt)
    theta = None
    if a.theta_check > 0:
        theta = theta_check(case, a.theta_check)
        progress(f"precision check: max |dtheta| {theta['max_abs_K']:.3e} K after {a.theta_check} large steps")

    pts = ni * nj * nk
    value = pts * a.steps / elapsed
    ms_step = elapsed * 1e3 / a.steps
    st_ms = kern["hevi_stage"]["avg_ms"]
    achieved = STAGE_LAUNCH_BYTES_PER_POINT * pts / (st_ms * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                traffic=measured_traffic(a.size),
                kernel="hevi_tile (HEVI stage; the 4th launch of a step includes the RK4 combination)",
                bytes_per_point=STAGE_LAUNCH_BYTES_PER_POINT,
                stage_ms=st_ms, kernels_avg_ms={k: v["avg_ms"] for k, v in kern.items()},
                copy_probe_GBs=copy_gbs,
                bandwidth_probes=probes,
                frac_of_copy=(achieved / copy_gbs if copy_gbs else None))
    out = {
        "metric": "grid-pt updates/s short step (%HBM roofline); FP32 vs FP64 speedup; adjoint/fwd",
        "value": value,
        "unit": "grid-pt short-step updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded shear flow + theta noise, hydrostatic base state)",
        "config": {"workload": f"HEVI short step, FP64, {ni}x{nj}x{nk}, DX=1km, DZ 50-500 m stretched, "
                               f"linear shear 0-20 m/s, UD3, RK4, numfilter 1e-4, dt=2 s",
                   "grid": [ni, nj, nk], "short_step_bytes_per_point": STEP_BYTES_PER_POINT,
                   "short_step_hbm_frac": value * STEP_BYTES_PER_POINT / 1e9 / HBM_PEAK_GBS,
                   "large_step_with_3_fct_tracers_ms": large_ms,
                   "tracer_kernels_avg_ms": tracer_kern,
                   **({"large_step_other_tracer_scheme": tracer_leg} if tracer_leg else {})},
        "roofline": roof,
        "fp32_vs_fp64": mixed,
        "adjoint": adj,
        "ensemble": ens,
    }
    if theta is not None:
        out["theta_check"] = theta
    if badj is not None:
        out["batched_terrain_adjoint"] = badj
    if not a.no_cpu_baseline:
        cn, cj, ck = (int(x) for x in a.cpu_sample.split("x"))
        out["cpu_baseline"] = cpu_baseline(cn, cj, ck)
        progress("CPU baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
