#!/usr/bin/env python3
"""Benchmark of the 3D viscoelastic velocity-stress loop on one MI355X.

Workload (default, ``--workload production``): 1024 x 1024 x 512 cells,
dx = dy = dz = 0.25 km, 3 Zener mechanisms, 20-cell ADE-CFS PML, free surface
on top, layered background with a depth gradient Vs 2.0 -> 4.5 km/s
(Vp = sqrt(3) Vs, Qs = 50 Vs, Qp = 2 Qs) plus seeded von Karman heterogeneity
(eps 5 %, a = 5 km, kappa = 0.5, seed 3), one seeded moment-tensor source at
10 km depth with a Kupper wavelet of 1 s, 256 seeded surface receivers.
``--workload plain`` is the 128^3 example grid.

A step is one pass of the time loop (receivers, stress interior + PML,
stress glut, velocity interior + PML); steps run as replays of a HIP graph
of whole steps.  Inputs are resident in HBM before the timed region; the
timed region is bracketed by HIP events on the library's stream.

Prints one JSON line (metric, value = Gcell-updates/s, roofline of the
dominant kernel, step roofline, CPU baseline).  After the production grid,
the default run measures the other forms on the same GPU and adds them under
"forms", each with its rate over the plain form's: the 64-member batched
Green's-function workload (autotuned, 100 graph-replayed steps, and the same
members stepped one after another), the elastic 256x256x192 adjoint gradient
(2000 steps, checkpoints every 45) and one source of the 384x384x256
multi-source adjoint (2500 steps, NM = 3).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from openswpc_amd.workloads import (MULTI_LAYERS, adjoint_multi_geometry, adjoint_multi_spec,  # noqa: E402
                                    adjoint_specs, batch_specs, plain_spec, production_spec)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
METRIC = "Gcell-updates/s & %HBM roofline: plain 3D, batched Green's fns, adjoint grad"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# placement of the per-cell arrays over the HBM regions (swpc_stripe_reserve)
# ---------------------------------------------------------------------------
STRIPE = {"on": True}


def stripe_on(bytes_per_region):
    """Reserve the striped arena for the contexts created next; returns a
    description of the placement (hipMalloc when striping is off or fails)."""
    from openswpc_amd.gpu import SwpcError, stripe_reserve
    if not STRIPE["on"]:
        return "hipMalloc"
    try:
        t = time.perf_counter()
        stripe_reserve(int(bytes_per_region))
        log(f"[bench] striped arena: {bytes_per_region / 2**30:.1f} GiB in each of 3 HBM regions "
            f"({time.perf_counter() - t:.2f}s)")
        return "striped over 3 HBM regions"
    except SwpcError as e:
        log(f"[bench] striped arena not reserved ({e}); hipMalloc placement")
        return "hipMalloc"


def stripe_off(placement):
    from openswpc_amd.gpu import stripe_release
    if placement != "hipMalloc":
        stripe_release()


def device_free_bytes():
    h = C.CDLL("libamdhip64.so")
    fr, tot = C.c_size_t(), C.c_size_t()
    assert h.hipMemGetInfo(C.byref(fr), C.byref(tot)) == 0
    return int(fr.value)


# ---------------------------------------------------------------------------
# runs of the workloads (openswpc_amd/workloads.py)
# ---------------------------------------------------------------------------
def measure_batch(steps, warmup, block_req):
    """The batched launch against the same members stepped one after
    another, each as graph replays.  Returns the measured figures."""
    from openswpc_amd.gpu import SwpcBatch
    from openswpc_amd.host import build_inputs, make_simulation
    from openswpc_amd.gpu import stripe_bytes_per_region
    t0 = time.perf_counter()
    specs = batch_specs()
    placement = stripe_on(sum(stripe_bytes_per_region(sp.nx, sp.ny, sp.nz, sp.nm, sp.na) for sp in specs)
                          + (1 << 30))
    sims = []
    for sp in specs:
        inp = build_inputs(sp)
        sims.append(make_simulation(inp))
        del inp
    cells = sum(sp.nx * sp.ny * sp.nz for sp in specs)
    log(f"[bench] batch of {len(specs)} members, {cells / 1e6:.1f} M cells, set-up {time.perf_counter() - t0:.1f}s")
    b = SwpcBatch(sims)
    t0 = time.perf_counter()
    kc, by, width, tune_ms = b.tune(trial_steps=3)
    packing = b.packing()
    log(f"[bench] autotuned kchunk={kc} by={by} width={width} packing={packing} ({tune_ms:.2f} ms/step in trials, "
        f"{time.perf_counter() - t0:.1f}s)")
    ev = HipEvents()
    block = min(block_req, max(steps, 1))
    it = 1
    b.step(it, warmup, use_graph=True, block=min(block, max(warmup, 1)))
    it += warmup
    b.synchronize()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, b.stream())
    b.step(it, steps, use_graph=True, block=block)
    ev.record(e1, b.stream())
    ms_batch = ev.elapsed_ms(e0, e1) / steps
    it += steps
    # the same members one after another (each its own graph of `block`
    # steps), each with the plain defaults rather than the batch's tuned shape
    b.close()
    for s, sp in zip(sims, specs):
        s.set_tuning(-1, 4, 1 | 8)
    for s in sims:
        s.step(it, block, use_graph=True, block=block)
    for s in sims:
        s.synchronize()
    it += block
    nseq = max(block, (steps // block) * block)
    t_seq = 0.0
    for s in sims:
        a0, a1 = ev.create(), ev.create()
        ev.record(a0, s.stream())
        s.step(it, nseq, use_graph=True, block=block)
        ev.record(a1, s.stream())
        t_seq += ev.elapsed_ms(a0, a1)
    ms_seq = t_seq / nseq
    nm = 2
    n_int = sum(cell_counts(sp.nx, sp.ny, sp.nz, sp.na)[0] for sp in specs)
    n_pml = cells - n_int
    b_si, b_vi, b_sp, b_vp = bytes_model(nm)
    bytes_step = n_int * (b_si + b_vi) + n_pml * (b_sp + b_vp)
    for s in sims:
        s.close()
    stripe_off(placement)
    return dict(members=len(specs), cells=cells, block=block, kc=kc, by=by, width=width, packing=packing,
                ms_batch=ms_batch, ms_seq=ms_seq, bytes_step=bytes_step, placement=placement)


def batch_summary(r):
    """the batched form as one object of the default benchmark line"""
    cells, ms_b, ms_s = r["cells"], r["ms_batch"], r["ms_seq"]
    return {"workload": f"{r['members']} uneven grids 96..256 x 96..256 x 96..192 (seed 6), NM=2, 10-cell PML, "
                        f"unit single forces, 8 receivers each",
            "cells": cells, "ms_per_step": round(ms_b, 4), "value": round(cells / (ms_b * 1e-3) / 1e9, 4),
            "unit": "Gcell-updates/s",
            "step_roofline_frac": round(r["bytes_step"] / (ms_b * 1e-3) / HBM_PEAK, 4),
            "sequential_ms_per_step": round(ms_s, 4), "batched_over_sequential": round(ms_s / ms_b, 3),
            "autotuned": {"kchunk": r["kc"], "by": r["by"], "tile_width": r["width"], "packing": r["packing"]},
            "placement": r["placement"]}


def run_batch(args):
    """--workload batch: the batched launch against the same members stepped
    one after another, each as graph replays."""
    r = measure_batch(args.steps, args.warmup, args.block)
    cells, ms_batch, ms_seq, bytes_step = r["cells"], r["ms_batch"], r["ms_seq"], r["bytes_step"]
    block, kc, by, width = r["block"], r["kc"], r["by"], r["width"]
    value = cells / (ms_batch * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batch, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched Green's functions: {r['members']} uneven grids 96..256 x 96..256 x "
                               f"96..192 (seed 6), NM=2, 10-cell PML, unit single forces, 8 receivers each, 4 "
                               f"layered profiles", "members": r["members"], "cells": cells, "graph_block": block,
                   "autotuned": {"kchunk": kc, "by": by, "tile_width": width, "packing": r["packing"]},
                   "placement": r["placement"]},
        "step_roofline": {"bytes_per_step": bytes_step, "achieved": round(bytes_step / (ms_batch * 1e-3) / 1e9, 2),
                          "unit": "GB/s", "frac": round(bytes_step / (ms_batch * 1e-3) / HBM_PEAK, 4)},
        "sequential": {"ms_per_step": round(ms_seq, 4), "value": round(cells / (ms_seq * 1e-3) / 1e9, 4),
                       "speedup_batched": round(ms_seq / ms_batch, 3)},
    }
    print(json.dumps(out), flush=True)



def measure_adjoint(nt, block):
    """Misfit gradient of the elastic 256 x 256 x 192 workload with
    checkpoints every 45 steps, and the plain forward loop on the same grid."""
    from openswpc_amd.host import build_inputs, make_simulation
    t0 = time.perf_counter()
    true_spec, init_spec, info = adjoint_specs(nt=nt)
    # contexts, adjoint workspace and checkpoints (45 states and 45 step
    # records, about 75 GB) from the striped arena
    placement = stripe_on(30 << 30)
    sim_true = make_simulation(build_inputs(true_spec))
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, sim_true.stream())
    sim_true.step(1, nt, use_graph=True, block=min(block, nt))
    ev.record(e1, sim_true.stream())
    ms_fwd = ev.elapsed_ms(e0, e1) / nt
    obs = sim_true.seismograms("vel")
    sim_true.close()
    sim = make_simulation(build_inputs(init_spec))
    log(f"[bench] adjoint workload set-up and true data {time.perf_counter() - t0:.1f}s")
    t = time.perf_counter()
    J, gl, gm, gr = sim.adjoint_gradient(nt, obs, ckp=45)
    wall = time.perf_counter() - t
    sim.close()
    stripe_off(placement)
    inner = (slice(3, -3),) * 3
    return dict(cells=init_spec.nx * init_spec.ny * init_spec.nz, nt=nt, wall=wall, ms_fwd=ms_fwd, misfit=J,
                placement=placement,
                sign=info["sign"], grad_norms={"lam": float(np.linalg.norm(gl[inner])),
                                               "mu": float(np.linalg.norm(gm[inner])),
                                               "rho": float(np.linalg.norm(gr[inner]))})


def adjoint_summary(r):
    """the adjoint-gradient form as one object of the default benchmark line"""
    cells, nt, wall, ms_fwd = r["cells"], r["nt"], r["wall"], r["ms_fwd"]
    return {"workload": "elastic adjoint gradient 256x256x192, dx=0.2 km, 20-cell PML, 1 moment tensor, "
                        f"64 receivers, {nt} steps, checkpoints every 45",
            "ms_per_step": round(wall * 1e3 / nt, 4), "value": round(cells * nt / wall / 1e9, 4),
            "unit": "Gcell-updates/s of gradient (forward, recomputation and adjoint of every step)",
            "forward_ms_per_step": round(ms_fwd, 4), "forward_value": round(cells / (ms_fwd * 1e-3) / 1e9, 4),
            "gradient_over_forward_time": round(wall * 1e3 / nt / ms_fwd, 3), "placement": r["placement"]}


def run_adjoint(args):
    """--workload adjoint: misfit gradient with checkpoints every 45 steps,
    against the plain forward loop on the same grid"""
    nt = args.adj_steps
    r = measure_adjoint(nt, args.block)
    cells, wall, ms_fwd = r["cells"], r["wall"], r["ms_fwd"]
    value = cells * nt / wall / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s (gradient: forward, recomputation and adjoint of every step)",
        "n_gpus": 1, "steps": nt, "warmup": 0, "ms_per_step": round(wall * 1e3 / nt, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "adjoint gradient, elastic 256x256x192, dx=0.2 km, 20-cell PML, layered model; "
                               "data from layered + Gaussian anomaly (8 %, radius 3 km, seed 7); 1 moment tensor, "
                               "64 receivers; checkpoints every 45 steps", "anomaly_sign": r["sign"]},
        "forward_plain": {"ms_per_step": round(ms_fwd, 4), "value": round(cells / (ms_fwd * 1e-3) / 1e9, 4)},
        "gradient_over_forward_time": round(wall * 1e3 / nt / ms_fwd, 3),
        "misfit": r["misfit"],
        "grad_norms": r["grad_norms"],
    }
    print(json.dumps(out), flush=True)


def measure_adjoint_multi(nt, ns, block):
    """Gradients of the first ns seeded moment-tensor sources on the shared
    384 x 384 x 256 grid (dx 0.2 km, NM = 3, 20-cell PML), each with its own
    128 receivers, summed on the device under the two-level checkpoint
    scheme the planner fits into the free device memory."""
    import dataclasses
    from openswpc_amd.host import build_inputs, make_simulation, set_acquisition
    geo = adjoint_multi_geometry()
    t0 = time.perf_counter()
    # the checkpoint planner fills the device: the arena keeps all but 8 GiB
    # of the free memory, and the planner counts what it keeps free
    placement = stripe_on((device_free_bytes() - (8 << 30)) // 3)
    spec_true = adjoint_multi_spec(geo, nt=nt, true=True)
    inp_true = build_inputs(spec_true)
    sim = make_simulation(inp_true)
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    obs, ms_fwd = [], None
    for s in range(ns):
        sp = dataclasses.replace(spec_true, sources=[geo["sources"][s]], stations=geo["stations"][s], xi=None)
        set_acquisition(sim, inp_true, sp)
        ev.record(e0, sim.stream())
        sim.step(1, nt, use_graph=True, block=min(block, nt))
        ev.record(e1, sim.stream())
        ms = ev.elapsed_ms(e0, e1) / nt
        ms_fwd = ms if ms_fwd is None else min(ms_fwd, ms)
        obs.append(sim.seismograms("vel"))
    sim.close()
    del inp_true
    spec = adjoint_multi_spec(geo, nt=nt)
    inp = build_inputs(spec)
    sim = make_simulation(inp)
    c1, c2, nbytes, nfwd = sim.adjoint_plan(nt)
    log(f"[bench] adjoint_multi set-up and data of {ns} sources {time.perf_counter() - t0:.1f}s; "
        f"checkpoints every {c1} / {c2} steps, {nbytes / 2**30:.1f} GiB, {nfwd} forward steps per source")
    misfits = []
    sim.synchronize()
    t = time.perf_counter()
    sim.adjoint_reset()
    for s in range(ns):
        if s:
            set_acquisition(sim, inp, dataclasses.replace(spec, sources=[geo["sources"][s]],
                                                          stations=geo["stations"][s]))
        misfits.append(sim.adjoint_accumulate(nt, obs[s], c1, c2))
        log(f"[bench] source {s}: misfit {misfits[-1]:.4e}, {time.perf_counter() - t:.1f}s")
    gl, gm, gr = sim.adjoint_sums()
    wall = time.perf_counter() - t
    held = sim.device_bytes()
    sim.close()
    stripe_off(placement)
    inner = (slice(3, -3),) * 3
    return dict(geo=geo, cells=spec.nx * spec.ny * spec.nz, nt=nt, ns=ns, wall=wall, ms_fwd=ms_fwd, c1=c1, c2=c2,
                nbytes=nbytes, nfwd=nfwd, held=held, misfits=misfits, placement=placement,
                grad_norms={"lam": float(np.linalg.norm(gl[inner])), "mu": float(np.linalg.norm(gm[inner])),
                            "rho": float(np.linalg.norm(gr[inner]))})


def adjoint_multi_summary(r):
    """the multi-source adjoint form as one object of the default benchmark line"""
    cells, nt, ns, wall, ms_fwd = r["cells"], r["nt"], r["ns"], r["wall"], r["ms_fwd"]
    return {"workload": f"multi-source adjoint 384x384x256, dx=0.2 km, NM=3, 20-cell PML, {nt} steps, {ns} of 16 "
                        "seeded moment tensors timed, 128 receivers each, gradients summed on the device",
            "ms_per_step": round(wall * 1e3 / (nt * ns), 4), "value": round(cells * nt * ns / wall / 1e9, 4),
            "unit": "Gcell-updates/s of gradient per source-step",
            "checkpoints": {"outer": r["c1"], "inner": r["c2"], "device_bytes": r["nbytes"],
                            "forward_steps": r["nfwd"]},
            "forward_ms_per_step": round(ms_fwd, 4), "forward_value": round(cells / (ms_fwd * 1e-3) / 1e9, 4),
            "gradient_over_forward_time": round(wall * 1e3 / (nt * ns) / ms_fwd, 3), "placement": r["placement"]}


def run_adjoint_multi(args):
    """--workload adjoint_multi: gradients of seeded moment-tensor sources on a
    shared 384 x 384 x 256 grid (dx 0.2 km, NM = 3, 20-cell PML, 2500 steps),
    each with its own 128 receivers, summed on the device under the two-level
    checkpoint scheme the planner fits into the free device memory.  The
    first --adj-sources of the 16 seeded sources are timed; the gradient
    rate is per source-step, the same for every source."""
    nt, ns = args.adj_steps, args.adj_sources
    r = measure_adjoint_multi(nt, ns, args.block)
    geo, wall, ms_fwd, held = r["geo"], r["wall"], r["ms_fwd"], r["held"]
    c1, c2, nbytes, nfwd, misfits = r["c1"], r["c2"], r["nbytes"], r["nfwd"], r["misfits"]
    cells = r["cells"]
    value = cells * nt * ns / wall / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s (gradient: forward, recomputation and adjoint of every step of every source)",
        "n_gpus": 1, "steps": nt, "warmup": 0, "ms_per_step": round(wall * 1e3 / (nt * ns), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"multi-source adjoint, 384x384x256, dx=0.2 km, NM=3, 20-cell PML, {nt} steps; "
                               f"{ns} of 16 seeded moment tensors timed, 128 receivers each; data from layered + "
                               "Gaussian anomaly (8 %, radius 5 km, seed 21); gradients summed on the device",
                   "checkpoints": {"outer": c1, "inner": c2, "device_bytes": nbytes, "forward_steps": nfwd},
                   "device_bytes_held": held},
        "forward_plain": {"ms_per_step": round(ms_fwd, 4), "value": round(cells / (ms_fwd * 1e-3) / 1e9, 4)},
        "gradient_over_forward_time": round(wall * 1e3 / (nt * ns) / ms_fwd, 3),
        "misfits": misfits,
        "grad_norms": r["grad_norms"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_adjoint(geo)
    print(json.dumps(out), flush=True)


def cpu_baseline_adjoint(geo, nt=300, nsrc=1):
    """The FP64 oracle adjoint on the 64^3 subsampling of the multi-source
    grid (dx = dy = 1.2 km, dz = 0.8 km, 10-cell PML, dt 0.04 s, nt steps,
    nsrc sources): its gradient rate, and the relative L2 error of the GPU's
    device-summed gradient on the same case (interior and PML cells).  By
    300 steps the waves have crossed the anomaly and the residuals are a few
    per cent of the traces; much earlier they are ~1e-4 of them and the
    FP32 forward's rounding, not the adjoint, dominates the comparison."""
    import dataclasses
    from oracle import oracle as O
    from openswpc_amd.host import build_inputs, make_simulation, set_acquisition
    n, d, dt, na = (64, 64, 64), (1.2, 1.2, 0.8), 0.04, 10
    sp_true = adjoint_multi_spec(geo, n=n, d=d, dt=dt, nt=nt, na=na, true=True)
    sp = adjoint_multi_spec(geo, n=n, d=d, dt=dt, nt=nt, na=na)
    inp_t, inp = build_inputs(sp_true), build_inputs(sp)
    st, sg = make_simulation(inp_t), make_simulation(inp)
    obs = []
    sg.adjoint_reset()
    for s in range(nsrc):
        acq = dict(sources=[geo["sources"][s]], stations=geo["stations"][s])
        set_acquisition(st, inp_t, dataclasses.replace(sp_true, xi=None, **acq))
        st.step(1, nt)
        obs.append(st.seismograms("vel"))
        if s:
            set_acquisition(sg, inp, dataclasses.replace(sp, **acq))
        sg.adjoint_accumulate(nt, obs[s], 10, 3)
    gpu = sg.adjoint_sums()
    st.close()
    sg.close()
    want = [0.0, 0.0, 0.0]
    t = time.perf_counter()
    for s in range(nsrc):
        cfg = O.Config(nx=n[0], ny=n[1], nz=n[2], dx=d[0], dy=d[1], dz=d[2], dt=dt, nt=nt, na=na, nm=3, zbeg=-1.0,
                       vmodel="lhm", layers=MULTI_LAYERS, sources=[geo["sources"][s]], stations=geo["stations"][s],
                       wav_flags=1)
        # the GPU's float32 medium: the two set-ups differ by a few ulp in the
        # relaxed moduli, which is not what this compares
        m = dataclasses.replace(O.build_model(cfg), lam=inp.arrays["lam"], mu=inp.arrays["mu"],
                                rho=inp.arrays["rho"])
        d_obs = np.asarray(obs[s], np.float64).transpose(2, 0, 1)   # (nt, nst, 3)
        _, g, _ = O.misfit_and_gradient(m, nt, d_obs)
        want = [w + gg for w, gg in zip(want, g)]
    wall = time.perf_counter() - t
    inner = (slice(3, -3),) * 3
    err = {k: float(np.linalg.norm(a[inner].astype(np.float64) - b[inner]) / np.linalg.norm(b[inner]))
           for k, a, b in zip(("lam", "mu", "rho"), gpu, want)}
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"value": round(n[0] * n[1] * n[2] * nt * nsrc / wall / 1e9, 6),
            "unit": "Gcell-updates/s (gradient)", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle adjoint (oracle/adjoint.c, OpenMP {cores} threads), 64^3 subsampling of the "
                      f"grid (1.2 x 1.2 x 0.8 km cells, 10-cell PML, NM=3), {nsrc} source(s), {nt} steps, "
                      f"{wall:.1f} s",
            "gpu_rel_l2_vs_fp64": err}


def run_rmedia(args):
    """--workload rmedia: the production grid's von Karman field (the (k,i,j)
    halo array of 1024x1024x512) generated on the device, against the
    oracle's port of the same algorithm on a bounded sample"""
    from openswpc_amd.gpu import rmedia_3d
    shape = (518, 1030, 1030)
    cells = shape[0] * shape[1] * shape[2]
    kw = dict(d=(0.25, 0.25, 0.25), a=(5.0, 5.0, 5.0), epsil=0.05, kappa=0.5)
    rmedia_3d("vonkarman", (64, 64, 64), seed=1, **kw)   # first use: library and FFT plan set-up
    times = []
    for rep in range(3):
        t = time.perf_counter()
        f = rmedia_3d("vonkarman", shape, seed=3 + rep, **kw)
        times.append(time.perf_counter() - t)
        log(f"[bench] rmedia {shape} in {times[-1]:.3f}s rms={float(np.sqrt(np.mean(f.astype(np.float64)**2))):.5f}")
        del f
    wall = min(times)
    out = {"metric": "random-media cells/s (m_rmedia.f90 method: noise, FFT, PSDF filter, inverse FFT, scaling)",
           "value": round(cells / wall / 1e9, 4), "unit": "Gcells/s, host array out", "n_gpus": 1,
           "steps": len(times), "warmup": 1, "ms_per_step": round(wall * 1e3, 2), "higher_is_better": True,
           "scaling": "strong", "vs_baseline": None, "dtype": "f64 FFT, f32 out", "data": "synthetic",
           "config": {"workload": "von Karman 3D, a 5 km, kappa 0.5, eps 5 %, 518x1030x1030 (the (k,i,j) halo "
                                  "array of the 1024x1024x512 production grid), dx 0.25 km"}}
    if not args.no_cpu_baseline:
        sys.path.insert(0, ROOT)
        from oracle import rmedia as R
        sh = (256, 256, 256)
        t = time.perf_counter()
        R.generate(3, sh, kw["d"], kw["a"], kw["epsil"], kw["kappa"], seed=3)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(sh[0] * sh[1] * sh[2] / dt / 1e9, 5), "unit": "Gcells/s", "cores": 1,
                               "kind": "port", "sample": f"oracle/rmedia.py generate(), numpy FFT, "
                                                         f"{sh[0]}x{sh[1]}x{sh[2]}, {dt:.2f} s"}
    print(json.dumps(out), flush=True)


def cell_counts(nx, ny, nz, na):
    n_int = (nx - 2 * na) * (ny - 2 * na) * (nz - na)
    return n_int, nx * ny * nz - n_int


def bytes_model(nm):
    """Algorithmic bytes per cell and phase (FP32; every array read or written
    once per phase, stencil neighbours free)."""
    stress_int = 4 * (3 + 4 + 2 * 6 + 2 * 6 * nm)   # V3 read; lam,mu,taup,taus; S6 RMW; R(6 nm) RMW
    vel_int = 4 * (6 + 1 + 2 * 3)                   # S6 read; rho; V3 RMW
    stress_pml = 4 * (3 + 2 + 2 * 6 + 2 * 9)        # V3; lam, mu; S6 RMW; 9 ADE RMW
    vel_pml = 4 * (6 + 1 + 2 * 3 + 2 * 9)           # S6; rho; V3 RMW; 9 ADE RMW
    return stress_int, vel_int, stress_pml, vel_pml


# ---------------------------------------------------------------------------
# HIP events on the library's stream
# ---------------------------------------------------------------------------
class HipEvents:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


# ---------------------------------------------------------------------------
# CPU baseline: the oracle (restated reference scheme, FP32, OpenMP)
# ---------------------------------------------------------------------------
def cpu_baseline(target_s=12.0):
    """The FP32 oracle, vectorised for AVX2 and threaded with OpenMP, on a
    384 x 384 x 256 sub-grid of the production workload: its cell size, time
    step, layered gradient, NM = 3 and 20-cell PML (the random perturbation
    does not change the work per cell); timed for about target_s seconds."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    nx, ny, nz = 384, 384, 256
    cfg = O.Config(nx=nx, ny=ny, nz=nz, dx=0.25, dy=0.25, dz=0.25, dt=0.0117, nt=2000, na=20, nm=3, zbeg=-1.0,
                   vmodel="lgm", layers=[[0.0, 2.3, 3.464, 2.0, 200, 100], [60.0, 3.2, 7.794, 4.5, 450, 225]],
                   sources=[dict(x=0.0, y=0.0, z=10.0, t0=0.0, tr=1.0, mo=1e15, strike=30.0, dip=60.0, rake=90.0)])
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32, vectorised=True)
    run.run(2)                                   # warm-up (thread pool, page faults)
    t = time.perf_counter()
    run.run(2)
    per = (time.perf_counter() - t) / 2
    steps = max(4, min(int(target_s / max(per, 1e-6)), cfg.nt - 4))
    t = time.perf_counter()
    run.run(steps)
    dt = time.perf_counter() - t
    val = nx * ny * nz * steps / dt / 1e9
    return {"value": round(val, 5), "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP32 restatement of the reference scheme (gcc -O3 -march=x86-64-v3, no FMA "
                      f"contraction, OpenMP {cores} threads), the production scheme on a {nx}x{ny}x{nz} sub-grid "
                      f"(NM=3, 20-cell PML, layered gradient), {steps} steps, {dt:.1f} s"}


def read_traffic(workload):
    """HBM bytes per launch of the dominant kernel from committed rocprofv3
    --pmc summaries (FETCH_SIZE + WRITE_SIZE), if present."""
    p = os.path.join(ROOT, "profiles", f"k_stress_int_{workload}_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["production", "plain", "batch", "adjoint", "adjoint_multi", "rmedia"],
                    default="production")
    ap.add_argument("--adj-steps", type=int, default=None,
                    help="time steps of the adjoint workloads (default 2000 single-source, 2500 multi-source)")
    ap.add_argument("--adj-sources", type=int, default=2, help="sources timed by the multi-source adjoint")
    ap.add_argument("--block", type=int, default=50, help="steps per captured HIP graph")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="production workload: skip the batched and adjoint forms measured after it")
    ap.add_argument("--form-steps", type=int, default=100, help="timed steps of the batched form")
    ap.add_argument("--host-medium", action="store_true",
                    help="build the medium with the host library and upload it, instead of on the device")
    ap.add_argument("--kernel-steps", type=int, default=10, help="steps timed per kernel with events")
    ap.add_argument("--no-stripe", action="store_true",
                    help="plain hipMalloc placement instead of the arrays striped over the HBM regions "
                         "(no arena reserved, SWPC_STRIPE=0 for the library)")
    ap.add_argument("--kchunk", type=int, default=0)
    ap.add_argument("--by", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1)
    args = ap.parse_args()
    STRIPE["on"] = not args.no_stripe
    if args.no_stripe:
        os.environ["SWPC_STRIPE"] = "0"     # no private arena for large contexts either

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this library runs on one GPU only: the whole domain lives on one "
                                   "device; sharding across GPUs (x/y slabs with halo exchange) is not built",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "adjoint":
        args.adj_steps = args.adj_steps or 2000
        run_adjoint(args)
        return
    if args.workload == "adjoint_multi":
        args.adj_steps = args.adj_steps or 2500
        run_adjoint_multi(args)
        return
    if args.workload == "rmedia":
        run_rmedia(args)
        return

    from openswpc_amd.gpu import stripe_bytes_per_region
    from openswpc_amd.host import build_inputs, make_simulation, make_simulation_device

    t0 = time.perf_counter()
    spec = production_spec() if args.workload == "production" else plain_spec()
    fused = args.variant >= 0 and (args.variant & 128) != 0     # a second field set (fused steps)
    placement = stripe_on(stripe_bytes_per_region(spec.nx, spec.ny, spec.nz, spec.nm, spec.na,
                                                  field_sets=2 if fused else 1) + (512 << 20))
    if args.host_medium:
        inp = build_inputs(spec)
        log(f"[bench] host set-up {time.perf_counter() - t0:.1f}s")
        sim = make_simulation(inp)
    else:
        # the medium is built on the device (swpc_medium_layered), from a
        # random medium generated into the context's staging buffer
        sim, inp = make_simulation_device(spec)
    setup_s = time.perf_counter() - t0
    setup_again = None
    if not args.host_medium:
        # the same set-up once more in this process: the first one includes
        # the FFT library's compilation of its kernels (once per machine and
        # user, then read from its cache)
        sim.close()
        t1 = time.perf_counter()
        sim, inp = make_simulation_device(spec)
        setup_again = time.perf_counter() - t1
    log(f"[bench] set-up {setup_s:.2f}s, again {setup_again if setup_again is None else round(setup_again, 2)}s "
        f"({'host' if args.host_medium else 'device'} medium) dt={spec.dt} "
        f"vmin={inp.scalars['vmin']:.3f} vmax={inp.scalars['vmax']:.3f}, {sim.device_bytes() / 2**30:.1f} GiB on device")
    if args.kchunk or args.by or args.variant >= 0:
        sim.set_tuning(args.kchunk, args.by, args.variant)
    del inp

    ev = HipEvents()
    stream = sim.stream()
    use_graph = not args.no_graph
    block = min(args.block, max(args.steps, 1))
    it = 1
    sim.step(it, args.warmup, use_graph=use_graph, block=min(block, max(args.warmup, 1)))
    it += args.warmup
    sim.synchronize()

    e0, e1 = ev.create(), ev.create()
    w0 = time.perf_counter()
    ev.record(e0, stream)
    sim.step(it, args.steps, use_graph=use_graph, block=block)
    ev.record(e1, stream)
    total_ms = ev.elapsed_ms(e0, e1)
    wall = time.perf_counter() - w0
    it += args.steps
    ms_per_step = total_ms / args.steps

    # per-launch durations (HIP events between the launches of whole steps)
    tm = sim.step_timed(it, args.kernel_steps)
    it += args.kernel_steps
    kms = {k: v / args.kernel_steps for k, v in tm.items()}
    vmax = sim.kernel_vmax()

    nx, ny, nz, na, nm = spec.nx, spec.ny, spec.nz, spec.na, spec.nm
    n_int, n_pml = cell_counts(nx, ny, nz, na)
    n_bot = (nx - 2 * na) * (ny - 2 * na) * na      # bottom PML band of the interior columns
    b_si, b_vi, b_sp, b_vp = bytes_model(nm)
    bytes_step = n_int * (b_si + b_vi) + n_pml * (b_sp + b_vp)
    # dominant launch: interior columns of the stress phase = kernel__update_stress
    # over the interior plus absorb__update_stress over their bottom band
    bytes_launch = n_int * b_si + n_bot * b_sp
    achieved = bytes_launch / (kms["stress"] * 1e-3)
    value = nx * ny * nz / (ms_per_step * 1e-3) / 1e9
    step_bw = bytes_step / (ms_per_step * 1e-3)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the baseline must not hide the GPU result
            cpu = {"error": str(e)}

    workload = "production" if args.workload == "production" else "plain"
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
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
            "workload": ("plain 3D viscoelastic FDM, 1024x1024x512 cells, dx=0.25 km, NM=3, 20-cell ADE-CFS PML, "
                         "layered gradient Vs 2.0->4.5 km/s + von Karman (eps 5%, a 5 km, kappa 0.5, seed 3), "
                         "1 moment-tensor source, 256 surface receivers" if workload == "production" else
                         "plain 3D viscoelastic FDM, 128^3 cells, dx=0.5 km, NM=3, 20-cell ADE-CFS PML, two-layer "
                         "medium, 1 moment-tensor source, 16 surface receivers"),
            "grid": [nx, ny, nz], "dt": spec.dt, "nm": nm, "na": na, "graph_block": block if use_graph else 0,
            "parallelism": "single GPU, whole domain resident in HBM",
            "placement": placement,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_stress_int<3,true> (kernel__update_stress + bottom band of absorb__update_stress)",
            "achieved": round(achieved / 1e9, 2),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": read_traffic(workload),
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": round(kms["stress"], 4),
        },
        "step_roofline": {
            "bytes_per_step": bytes_step,
            "achieved": round(step_bw / 1e9, 2),
            "unit": "GB/s",
            "frac": round(step_bw / HBM_PEAK, 4),
            "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        },
        "cpu_baseline": cpu,
        "setup_s": round(setup_s, 2),
        "setup_s_repeat": None if setup_again is None else round(setup_again, 2),
        "wall_s_timed": round(wall, 3),
        "vmax_end": [float(v) for v in vmax],
    }
    sim.close()
    stripe_off(placement)
    if workload == "production" and not args.no_forms:
        out["forms"] = measure_forms(args, value)
    print(json.dumps(out), flush=True)


def measure_forms(args, plain_value):
    """The batched Green's-function, adjoint-gradient and multi-source adjoint
    forms, each against the plain form's rate measured above (the production
    grid's memory is released first).  A form that fails reports its error
    and leaves the plain figures standing."""
    forms = {}
    legs = [("batched_greens", lambda: batch_summary(measure_batch(args.form_steps, 5, args.form_steps))),
            ("adjoint_gradient", lambda: adjoint_summary(measure_adjoint(2000, args.block))),
            ("adjoint_multi", lambda: adjoint_multi_summary(measure_adjoint_multi(2500, 1, args.block)))]
    for name, leg in legs:
        t = time.perf_counter()
        try:
            r = leg()
            r["over_plain_form"] = round(r["value"] / plain_value, 4)
        except Exception as e:  # one form must not hide the others
            r = {"error": f"{type(e).__name__}: {e}"}
        log(f"[bench] form {name}: {time.perf_counter() - t:.1f}s {r}")
        forms[name] = r
    return forms


if __name__ == "__main__":
    main()
