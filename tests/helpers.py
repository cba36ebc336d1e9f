"""Shared builders for the parity tests: stage inputs, perturbations, metrics."""
from __future__ import annotations

import numpy as np

import pyoracle as po
from scale_amd import inputs

VARS = ("DENS", "MOMZ", "MOMX", "MOMY", "RHOT")


def small_case(kmax=20, imax=12, jmax=10, seed=1234, **kw):
    args = dict(kmax=kmax, imax=imax, jmax=jmax, dx=500.0, dy=500.0, dz=250.0, seed=seed, noise=0.05,
                bubble_r=1500.0, bubble_zc=1500.0, ntracer=2, tracer_kind=["blob", "tophat"], momz_noise=0.05,
                u0=8.0, v0=-3.0)
    args.update(kw)
    return inputs.make_case(**args)


def perturbed(case, seed, scale=1.0e-3):
    """A state near case.state (as an RK stage input would be) with valid halos."""
    g = case.grid
    rng = np.random.Generator(np.random.PCG64(seed))
    out = {}
    for k, v in case.state.items():
        a = np.array(v, order="F", copy=True)
        s = scale * (np.abs(v).max() + 1.0)
        a[:, g.IS:g.IE + 1, g.JS:g.JE + 1] += s * rng.standard_normal(a[:, g.IS:g.IE + 1, g.JS:g.JE + 1].shape)
        inputs.halo_fill(g, a)
        out[k] = a
    out["MOMZ"][:g.KS] = 0.0
    out["MOMZ"][g.KE:] = 0.0
    return out


def terrain_metrics(case, seed=7, amp=0.05):
    """Non-trivial GSQRT/J13G/J23G/MAPF (smooth random), periodic in x and y."""
    g = case.grid
    rng = np.random.Generator(np.random.PCG64(seed))
    m = dict(case.metric)
    gs = 1.0 + amp * rng.uniform(-1, 1, size=g.shape + (7,))
    j13 = amp * rng.uniform(-1, 1, size=g.shape + (7,))
    j23 = amp * rng.uniform(-1, 1, size=g.shape + (7,))
    mapf = 1.0 + amp * rng.uniform(-1, 1, size=(g.IA, g.JA, 2, 4))
    for c in range(7):
        for a in (gs, j13, j23):
            inputs.halo_fill(g, a[..., c])
    for n in range(2):
        for c in range(4):
            t = mapf[:, :, n, c][None]
            inputs.halo_fill(g, t)
    m["GSQRT"] = np.asfortranarray(gs)
    m["J13G"] = np.asfortranarray(j13)
    m["J23G"] = np.asfortranarray(j23)
    m["MAPF"] = np.asfortranarray(mapf)
    return m


def stage_inputs(case, seed=11, tend=True, coriolis=False):
    """All inputs of one HEVI stage: state0 = case.state, state = perturbed."""
    g = case.grid
    prep = inputs.large_step_prep(case)
    s1 = perturbed(case, seed)
    nd = po.numfilter_coef(case, case.state, case.dt)
    inp = {}
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    for v in VARS:
        inp[v + "0"] = case.state[v]
        inp[v] = s1[v]
        t = g.zeros()
        if tend:
            t[:, g.IS:g.IE + 1, g.JS:g.JE + 1] = 1.0e-4 * rng.standard_normal((g.KA, g.IE - g.IS + 1, g.JE - g.JS + 1))
            inputs.halo_fill(g, t)
        inp[v + "_t"] = t
    inp.update(prep)
    inp["num_diff"] = nd
    inp["DDIV"] = g.zeros()
    inp["REF_dens"] = case.ref["REF_dens"]
    if coriolis:
        inp["CORIOLI"] = np.asfortranarray(np.full((1, g.IA, g.JA), 1.0e-4))
    return inp


def defined(a):
    return ~np.isnan(a)


def compare(ref, got, what, rtol=0.0, atol=0.0):
    """Elementwise comparison on the elements the oracle defines; returns max |diff|."""
    mask = defined(ref)
    assert not np.isnan(got[mask]).any(), f"{what}: NaN where the oracle has values"
    d = np.abs(ref[mask] - got[mask])
    scale = np.abs(ref[mask])
    bad = d > atol + rtol * scale
    assert not bad.any(), f"{what}: {bad.sum()} mismatches, max |diff| {d.max():.3e} (max |ref| {scale.max():.3e})"
    return float(d.max()) if d.size else 0.0
