"""Benchmark of the cubed-sphere cell-update path on one MI355X.

Workload (default, C5): one spherical shell of 6 x 512^2 x 128 zones (201 M
owned zones, r in [1, 1.5], NGhost 4) advanced by the reference's P1 scheme
(WENO-ADP: multi-D WENO with the flattener and the pressure-positivity
limiter, the ADP constrained least squares for the zone B), HLLD face
fluxes, the 4-state HLL edge fields with arc-length constrained transport,
and IMEX-SSP2(2,2,2) with the radial gravity g = 1/r^2 implicit; hydrostatic
state with 1e-2 density noise.  A step is one full IMEX time step (the
source-only stage, two flux stages, CLEANUP, EVALUATE_TIMESTEP) of all six
sectors, captured once in a hipGraph and replayed; inputs are resident in HBM
before the timed region.

One JSON line on stdout; details on stderr.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, GB/s
FP64_VEC_TFLOPS = 78.6    # MI355X vector FP64
FP32_VEC_TFLOPS = 157.3

METRIC = "cell-updates/s on 1 MI355X, mixed vs FP64 at ≤1e-6 rel L1; % HBM roofline"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def word_bytes(prec):
    return 8 if prec == 0 else 4


def algorithmic_bytes_per_unit(cs, kernel, prec, surface):
    """Compulsory HBM bytes per work item of a kernel: every input it must
    read once and every output it must write once, with data shared between
    neighbouring items (a zone read by two faces, a face flux read by two
    zones) counted once.  Words of the primitive array V are FP32 in mixed
    precision, everything else is FP64.  With one radial layer (surface) the
    2-D metric tables are per zone and count; in a 3-D shell they are reused
    by every layer and do not.
    """
    w = word_bytes(prec)
    t2 = 1 if surface else 0
    if kernel in (cs.K_FLUX_X, cs.K_FLUX_Y, cs.K_FLUX_Z, cs.K_FLUX_XY):
        # V of one zone, face B, face frame (normal, radial dir, angle), flux (9) + speeds (2) out
        return 9 * w + 8 + t2 * 7 * 8 + 11 * 8
    if kernel == cs.K_PRIMS:
        return 9 * 8 + 9 * w                      # U + entropy in, V out
    if kernel in (cs.K_EMF_X, cs.K_EMF_Y, cs.K_EMF_Z, cs.K_EMF):
        # v and B of one zone, two face B, four speeds, edge frame (5), EMF out
        return 6 * w + 2 * 8 + 4 * 8 + t2 * 5 * 8 + 8
    if kernel == cs.K_UPDATE_CELLS:
        # zone update + cleanup + primitives of the output.
        # Surface: lateral fluxes of the 6 components the face-average cleanup
        # keeps (2 directions), stage input U + entropy (9), U0 of those 6,
        # output U + entropy (9), V in and out, 2 lateral face B, the two wall
        # z-faces, zone metric (solid angle, area vector, frame, zone-B
        # operator: 17).
        # Shell: fluxes of all 9 components (the induction fluxes advance the
        # zone B the least-squares cleanup replaces; 3 directions), stage input,
        # U^n (read by every implicit-explicit stage) and output U + entropy
        # (9 each), V out, 3 face B; the 2-D metric is reused by every layer.
        if surface:
            return 12 * 8 + 9 * 8 + 6 * 8 + 9 * 8 + 18 * w + 2 * 8 + 2 * 8 + 17 * 8
        return 27 * 8 + 9 * 8 + 9 * 8 + 9 * 8 + 9 * w + 3 * 8
    if kernel == cs.K_UPDATE_FACES:
        return 4 * 3 * 8 + t2 * 3 * 8             # in, U0, out B and the EMF of 3 families; areas
    if kernel == cs.K_HALO:
        return 2 * 9 * 8 + 2 * 9 * w              # one ghost zone: U + entropy and V read and written
    return 0


def describe(cs, cfg, problem):
    """Scheme and initial state of a workload, in words."""
    rs = "HLLD" if cfg.riemann == cs.RS_HLLD else "HLL"
    rec = "WENO5" if cfg.recon == cs.RECON_WENO5 else "PCM"
    if cfg.imex:
        tint = "IMEX-SSP2(2,2,2) with implicit gravity 1/r^2"
    else:
        tint = "SSP-RK3" if cfg.rk_order == 3 else "Heun RK2"
        if cfg.grav:
            tint += " with explicit gravity 1/r^2"
    ic = {0: "coordinate fields", 1: "Gaussian bump + solid-body rotation", 2: "magnetised blast",
          3: "hydrostatic + 1e-2 density noise"}[problem]
    ic = ("surface, " if cfg.nr == 1 else "shell r in [%g, %g], " % (cfg.rmin, cfg.rmax)) + ic
    return f"{rec} + {rs} + 2-D HLL/LLF EMF + CT, {tint}, {ic}"


TRAFFIC_PROFILES = ("c2_surface256_hbm_traffic.txt", "c5_shell512x128_p1adp_counters.md")


def measured_traffic(kernel, prec, workload):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter read
    for a copy kernel of known size in the same summary), or None when none
    holds such a line."""
    for name in TRAFFIC_PROFILES:
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        cal, rec = {}, None
        for line in open(path):
            f = line.split()
            if not f or f[0].startswith("#"):
                continue
            if f[0].startswith("calibration_"):
                cal[f[0]] = float(f[1])
            elif f[0] == "kernel" and f[1:4] == [kernel, prec, workload]:
                rec = {f[4]: float(f[5]), f[6]: float(f[7])}
        if rec is None or len(cal) != 3:
            continue
        copy = cal["calibration_copy_bytes"]
        return (rec["fetch_kib"] * 1024 * copy / (cal["calibration_fetch_kib"] * 1024) +
                rec["write_kib"] * 1024 * copy / (cal["calibration_write_kib"] * 1024))
    return None


_STATES = {}


def initial_state(cs, name):
    """The workload's host initial state (the same for both precisions), built
    once per process: the large shell holds about 20 GB of host arrays."""
    if name not in _STATES:
        t0 = time.perf_counter()
        cfg, prob, _ = cs.workload(name, cs.FP64)
        _STATES[name] = cs.init_problem(cfg, prob, seed=20261202)
        log(f"initial state of {name}: {time.perf_counter() - t0:.1f} s")
    return _STATES[name]


def run_gpu(cs, name, prec, steps, warmup, profile_steps):
    """Time `steps` graph-replayed steps (the headline), then run
    `profile_steps` more steps launch by launch with a hipEvent pair around
    every launch on the library's stream, for the per-kernel durations as the
    kernels run inside a step."""
    cfg, prob, _ = cs.workload(name, prec)
    st = initial_state(cs, name)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    dev.dt = 0.0
    dev.evaluate_timestep()
    dev.advance(warmup, use_graph=True)
    tuned = dev.autotune(3)
    log("autotuned workgroups:", {k: tuned[k] for k in ("halo", "flux_xy", "flux_z", "emf", "update_faces",
                                                         "update_cells")})
    ms = dev.time_steps(steps, use_graph=True)
    kern = {}
    for k, (ms_avg, calls) in dev.profile_steps(profile_steps).items():
        kern[cs.KERNEL_NAMES[k]] = (ms_avg, dev.kernel_units(k), k)
    divb = dev.max_divb()
    sums = dev.conservation_sums()
    dt = dev.dt
    dev.close()
    return cfg, ms, kern, divb, sums, dt


def precision_gate(cs, name, steps):
    """FP64 and mixed from the same initial state for `steps` steps: relative
    L1 of each conserved variable, volume weighted, and L-infinity."""
    out = {}
    for prec in (cs.FP64, cs.MIXED):
        cfg, prob, _ = cs.workload(name, prec)
        st = initial_state(cs, name)
        dev = cs.CubeSphere(cfg)
        dev.upload(st)
        dev.dt = 0.0
        dev.evaluate_timestep()
        dev.advance(steps, use_graph=True)
        out[prec] = dev.download()
        dev.close()
    vol = cs.metric(cfg)[0]
    g, N, NR = cfg.nghost, cfg.ntp, cfg.nr
    nc, nrc = N + 2 * g, NR + 2 * g
    V = vol.reshape(6, nrc, nc, nc)[:, g:g + NR, g:g + N, g:g + N]
    U64, _ = out[cs.FP64].interior()
    U32, _ = out[cs.MIXED].interior()
    rel_l1 = []
    for q in range(8):
        den = float((V * np.abs(U64[:, q])).sum())
        num = float((V * np.abs(U32[:, q] - U64[:, q])).sum())
        rel_l1.append(num / den if den > 0 else num)
    linf = float(np.abs(U32 - U64).max() / np.abs(U64).max())
    return {"steps": steps, "rel_l1_per_var": rel_l1, "max_rel_l1": max(rel_l1), "rel_linf": linf}


def cpu_baseline(cs, name, max_seconds):
    """The C oracle (single thread) on a bounded sample of the same workload."""
    from oracle import oracle as orc
    cfg, prob, _ = cs.workload(name, cs.FP64)
    st = initial_state(cs, name).copy()  # the oracle advances its state in place
    G = orc.Grid(cfg)
    dt = G.cfl(st, 0.0)
    n_int = 6 * cfg.ntp * cfg.ntp * cfg.nr
    t0 = time.perf_counter()
    steps = 0
    while True:
        dt = G.advance(st, 1, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > max_seconds or steps >= 50:
            break
    return {"value": n_int * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} SSP-RK3 steps of the full {name} grid ({n_int} zones) by the C oracle, "
                      f"single thread, {el:.1f} s"}


def c4_workload(cs, prec, seed=20261202):
    """48 independent shells: panel edge N from {32, 48, 64, 96, 128, 192},
    radial zones from {16, 32, 64}, drawn with a seeded generator; blast and
    hydrostatic (with 1/r^2 gravity) initial states alternate."""
    rng = np.random.default_rng(seed)
    cfgs, probs = [], []
    for b in range(48):
        N = int(rng.choice([32, 48, 64, 96, 128, 192]))
        nr = int(rng.choice([16, 32, 64]))
        prob = 2 if b % 2 == 0 else 3
        cfgs.append(cs.make_config(N, nr, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, prec, 3, cs.RBC_PAD,
                                   cs.RMESH_UNIFORM, 1.0, 1.5, grav=1.0 if prob == 3 else 0.0))
        probs.append(prob)
    return cfgs, probs


def run_batch(cs, prec, steps, warmup, states=None):
    """The 48-shell batch: autotuned workgroup sizes, graph-replayed steps;
    then every shell on its own (its own per-shell graph), for comparison."""
    from concurrent.futures import ThreadPoolExecutor
    cfgs, probs = c4_workload(cs, prec)
    if states is None:
        with ThreadPoolExecutor(16) as ex:
            states = list(ex.map(lambda a: cs.init_problem(a[0], a[1], seed=20261202 + a[2]),
                                 [(c, p, i) for i, (c, p) in enumerate(zip(cfgs, probs))]))
    t0 = time.perf_counter()
    B = cs.Batch(cfgs)
    for sh, st in zip(B.shells, states):
        sh.upload(st)
        sh.dt = 0.0
    B.evaluate_timestep()
    tuned = B.autotune(2)
    B.advance(warmup, use_graph=True)
    ms = B.time_steps(steps)
    ms_single = sum(sh.time_steps(steps, use_graph=True) for sh in B.shells)
    divb = max(sh.max_divb() for sh in B.shells)
    B.close()
    cells = sum(6 * c.ntp * c.ntp * c.nr for c in cfgs)
    log(f"batch {'fp64' if prec == 0 else 'mixed'}: {len(cfgs)} shells, {cells} zones, {ms / steps:.3f} ms/step "
        f"batched, {ms_single / steps:.3f} ms/step shell by shell; workgroups {tuned}; max rel divB {divb:.2e} "
        f"({time.perf_counter() - t0:.1f} s)")
    return dict(cells=cells, ms=ms, ms_single=ms_single, tuned=tuned, divb=divb, states=states)


def bench_batch(cs, args):
    res = {}
    states = None
    for prec in (cs.FP64, cs.MIXED):
        r = run_batch(cs, prec, args.steps, args.warmup, states)
        states = r.pop("states")
        res[prec] = r
    r = res[cs.FP64]
    cups = r["cells"] * args.steps / (r["ms"] * 1e-3)
    line = {
        "metric": METRIC, "value": cups, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "C4: 48 shells, N in {32..192}, nr in {16, 32, 64}, WENO5 + HLLD + CT, SSP-RK3, "
                               "blast / hydrostatic alternating; one launch per phase over all shells, one graph "
                               "per step, autotuned workgroups", "zones": r["cells"]},
        "roofline": None, "cpu_baseline": None,
        "batched_vs_shell_by_shell": r["ms_single"] / r["ms"],
        "mixed": {"cell_updates_per_s": res[cs.MIXED]["cells"] * args.steps / (res[cs.MIXED]["ms"] * 1e-3),
                  "speedup_vs_fp64": r["ms"] / res[cs.MIXED]["ms"]},
        "autotuned_workgroups": {"fp64": r["tuned"], "mixed": res[cs.MIXED]["tuned"]},
        "max_rel_divb": max(r["divb"], res[cs.MIXED]["divb"]),
    }
    print(json.dumps(line))


def bench_geometry(cs, args):
    """The P1 quadrature geometry (vertices and mask, face quadrature with the
    B-field weights and the scale length, volume quadrature, basis moments) of
    the 6 x 192^2 x 64 shell built on the GPU; HBM bytes per item are the
    outputs written plus the compulsory inputs read (x/y faces 32 words, z-faces
    44, zones 108 for the volume quadrature)."""
    cfg = cs.make_config(192, 64, 3, cs.RECON_WENO5, cs.RS_HLLD, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                         cs.RMESH_UNIFORM, 1.0, 1.5)
    z = cs.geometry_sizes(cfg)
    runs = [cs.build_geometry(cfg, download=False)[1] for _ in range(max(1, args.steps))]
    ms = [min(r[q] for r in runs) for q in range(4)]
    items = [z["vert"], z["xf"] + z["yf"] + z["zf"], z["zone"], z["zone"]]
    nbytes = [24 * z["vert"] + 4 * z["zone"],
              256 * (z["xf"] + z["yf"]) + 352 * z["zf"] + 8 * z["zone"] + 24 * z["vert"],
              864 * z["zone"] + 24 * z["vert"],
              (292 + 32) * z["zone"]]
    names = ["vertices+mask", "faces+scale", "volume", "moments"]
    kern = {n: {"ms": t, "items": it, "GBps": b / (t * 1e-3) / 1e9, "frac": b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
            for n, t, it, b in zip(names, ms, items, nbytes)}
    for n, k in kern.items():
        log(f"   {n:14s} {k['ms']:8.3f} ms  {k['items']:11d} items  {k['GBps']:7.1f} GB/s ({100 * k['frac']:.0f} % of HBM)")
    total = sum(ms)
    dom = max(kern, key=lambda n: kern[n]["ms"])
    line = {
        "metric": "P1 geometry zones/s", "value": z["zone"] / (total * 1e-3), "unit": "zones/s", "n_gpus": 1,
        "steps": len(runs), "warmup": 0, "ms_per_step": total, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "P1 quadrature geometry of the 6x192^2x64 shell (ghosts included): vertices, "
                               "face and volume quadrature, scale length, basis moments", "zones": z["zone"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kern[dom]["frac"], "traffic": None},
        "cpu_baseline": None, "kernels": kern,
    }
    print(json.dumps(line))


def p1_patch_on_gpu(cs, N, NR, g=4, noise=0.3, seed=7, device="cuda:0", geo=None, sector=0, front_axis=None):
    """One sector of an N^2 x NR shell as a LIMIT_ALL_SLOPES_P1 patch on the
    GPU: quadrature geometry from cs_build_geometry, the stencil table, and a
    state with smooth flow, a front across the patch and log-normal noise on
    density, pressure and entropy (so the flattener and the positivity limiter
    act), generated with torch on the device."""
    import torch
    cfg = cs.make_config(N, NR, g, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                         cs.RMESH_UNIFORM, 1.0, 1.6)
    if geo is None:
        geo, _ = cs.build_geometry(cfg, skip=("v_pts", "v_wts", "v_tan", "moments"))
    z = {k: v // 6 for k, v in cs.geometry_sizes(cfg).items()}  # one sector
    dev = torch.device(device)
    sec = lambda a, n, m: torch.from_numpy(a[sector * n * m:(sector + 1) * n * m]).to(dev)
    T = dict(vert=sec(geo["vert"], z["vert"], 3), mask=sec(geo["mask"], z["zone"], 1),
             stencil=torch.from_numpy(cs.stencil_indexing(cfg)).to(dev), scale=sec(geo["scale"], z["zone"], 1),
             v_centroid=sec(geo["v_cen"], z["zone"], 3))
    T["v_volume"] = sec(geo["v_vol"], z["zone"], 1)
    for f in ("xf", "yf", "zf"):
        n = z[f]
        T[f + "_normal"] = sec(geo[f]["normal"], n, 15 if f == "zf" else 3)
        T[f + "_pts"] = sec(geo[f]["pts"], n, 15)
        T[f + "_area"] = sec(geo[f]["area"], n, 1)
        T[f + "_centroid"] = sec(geo[f]["centroid"], n, 3)
        b = f[0] + "b"
        T[b + "_pts"] = sec(geo[b]["pts"], n, 15)
        T[b + "_wts"] = sec(geo[b]["wts"], n, 5)
        T[b + "_normal"] = sec(geo[b]["normal"], n, 15 if f == "zf" else 3)
    del geo
    cnt = cs.p1_counts(N, NR, g)
    nz = cnt["mask"]
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    cen = T["v_centroid"].view(3, nz)
    r = cen.norm(dim=0).clamp_min(1e-30)
    x, y, w = cen / r
    gam = 5.0 / 3.0
    rho = 1.0 + 0.3 * torch.sin(3.0 * x + 1.0) * torch.cos(2.0 * y) + 0.2 * (r - 1.0)
    prs = 1.0 + 0.2 * torch.cos(2.5 * w - 0.5 * x)
    vel = torch.stack([0.3 * y - 0.2 * w, -0.3 * x + 0.1 * w, 0.2 * x * y])
    act = T["mask"] == 1
    if front_axis is None:  # the front through this sector's centre
        c0 = torch.stack([x[act].mean(), y[act].mean(), w[act].mean()])
    else:  # one front for every sector of a shell
        c0 = torch.tensor(front_axis, dtype=torch.float64, device=dev)
    c0 = c0 / c0.norm()
    t = torch.linalg.cross(c0, torch.tensor([0.3, 0.5, 0.8], dtype=torch.float64, device=dev), dim=0)
    t = t / t.norm()
    sfr = x * t[0] + y * t[1] + w * t[2] - 0.1
    front = sfr > 0
    rho = torch.where(front, 3.5 * rho, rho) * torch.exp(noise * torch.randn(nz, generator=gen, device=dev,
                                                                             dtype=torch.float64))
    prs = torch.where(front, 12.0 * prs, prs) * torch.exp(noise * torch.randn(nz, generator=gen, device=dev,
                                                                              dtype=torch.float64))
    sgn = torch.where(front, 1.0, -1.0).to(torch.float64)
    vel = vel - 2.5 * sgn[None] * t[:, None] * (sfr.abs() < 0.12)[None]
    Bc = torch.stack([0.5 + 0.2 * y, -0.3 + 0.1 * w * x, 0.2 * torch.cos(2 * x)])
    U = torch.zeros(4, 8, nz, dtype=torch.float64, device=dev)
    U[0, 0] = rho
    U[0, 1:4] = rho * vel
    U[0, 4] = prs / (gam - 1.0) + 0.5 * rho * (vel ** 2).sum(0) + (Bc ** 2).sum(0) / (8.0 * math.pi)
    U[0, 5:8] = Bc
    ent = torch.zeros(4, nz, dtype=torch.float64, device=dev)
    ent[0] = prs / rho ** (gam - 1.0) * torch.exp(noise * torch.randn(nz, generator=gen, device=dev,
                                                                      dtype=torch.float64))
    T["U"] = U.view(-1)
    T["ent"] = ent.view(-1)
    nfc = cnt["Bxf"]
    ar = torch.arange(nfc, device=dev, dtype=torch.float64)
    for q, nm in enumerate(("Bxf", "Byf", "Bzf")):
        T[nm] = 0.4 * torch.sin(ar * (0.37 + 0.11 * q)) + 0.1 * q
    cntf = cs.p1f_counts(N, NR, g)
    for k in cs.P1_ARRAYS + cs.P1F_OUT:
        if k not in T:
            T[k] = torch.zeros(cntf[k], dtype=torch.float64, device=dev)
    scal = dict(gamma=gam, molewt=1.0, rhofloor=1e-12, prsfloor=1e-9, smallnum=1e-12, flatten_coef=0.4,
                prs_flatten_coef=1.5)
    T["gamma_sound"].fill_(gam)  # as CLEANUP_AFTER_UPDATE leaves it
    if dev.type == "cuda":
        torch.cuda.synchronize()
    return cfg, scal, T


# algorithmic HBM bytes per item of the P1 stages (words read + written once):
# WENO + positivity per zone in its range: V means 8 + entropy 1 of the zone
# (its 6 neighbours' are re-reads, counted once), flattener 1, centroid 3,
# scale 1, the 27-node geometry of the zone (vertex 3 and face centroids
# 3 x 3, counted once per zone), V slopes 24 + entropy slopes 3 + B_zc 12
# written; CLSQ + E per zone: V B modes 12 read and written, U 4 read and
# written, scale 1, centroid 3, the zone's 3 own faces (B-field points 15,
# normals 3 or 15, centroids 3, face B 1, face-point B 5), E 3 written.
# CLEANUP_AFTER_UPDATE per zone with ghosts: U mode 1 read 8, U written 32 (the
# higher modes cleared), entropy read 1 and written 4, V written 32, gamma_eos,
# gamma_sound, mean_mol_wt written 3, the zone's own x, y face normals 6 and
# z-face centroid 3, its 3 own face B, B_zc written 3, the mask 4 bytes.
# EVALUATE_TIMESTEP per owned zone: U read and written 16, entropy 1, V written
# 8, gamma_sound 1, scale 1, msonic written 1, the mask 4 bytes.
# MAKE_EDGE_EFIELDS_P1 per owned zone (one edge of each family): the zone's
# mode-1 E 3, the four edge points of its three own faces 12, its six face
# speeds 6, its vertex 3; plus the three edge arrays written whole (zeros off
# the owned edges), counted per entry.
P1_EDGE_BYTES = 8 * (3 + 12 + 6 + 3)
# BUILD_NON_STIFF_SOURCES per zone with ghosts: Source written 8, and on the
# zones of its range centroid 3, V mode 1 density and velocity 4, dU_dt read
# and written 16, the mask 4 bytes (counted over all zones: the range holds
# all but the outer ghost layers).
P1_SOURCE_BYTES = 8 * (8 + 3 + 4 + 16) + 4
P1_CLEANUP_BYTES = 8 * (8 + 32 + 1 + 4 + 32 + 3 + 6 + 3 + 3 + 3) + 4
P1_TIMESTEP_BYTES = 8 * (16 + 1 + 8 + 1 + 1 + 1) + 4
P1_BYTES = {"weno+positivity": 8 * (8 + 1 + 1 + 3 + 1 + 12 + 24 + 3 + 12),
            "clsq+E": 8 * (24 + 8 + 1 + 3 + 2 * (15 + 3 + 3 + 1 + 5) + (15 + 15 + 3 + 1 + 5) + 3)}


def bench_p1(cs, args):
    """LIMIT_ALL_SLOPES_P1 on one sector of the 6 x 192^2 x 64 shell as a
    patch (NGhost 4): owned-zone reconstructions per second, per-stage times;
    then the rest of the P1 stage on the same patch, each call timed:
    MAKE_FACIAL_FLUXES, MAKE_EDGE_EFIELDS_P1, BUILD_NON_STIFF_SOURCES (point
    mass and rotation on), UPDATE_IMEX_RK_STEP_P1, CLEANUP_AFTER_UPDATE,
    EVALUATE_TIMESTEP."""
    import torch
    N, NR, g = 192, 64, 4
    cfg, scal, T = p1_patch_on_gpu(cs, N, NR, g)
    U0 = T["U"].clone()
    e0 = T["ent"].clone()
    sc0 = T["scale"].clone()
    Bf0 = [T[k].clone() for k in ("Bxf", "Byf", "Bzf")]

    def reset():
        T["U"].copy_(U0)
        T["ent"].copy_(e0)
        T["scale"].copy_(sc0)
        for k, b in zip(("Bxf", "Byf", "Bzf"), Bf0):
            T[k].copy_(b)

    T["Source"] = torch.zeros(8 * cs.p1_counts(N, NR, g)["mask"], dtype=torch.float64, device=T["U"].device)
    T["flag_entprs"] = torch.zeros(cs.p1_counts(N, NR, g)["mask"], dtype=torch.float64, device=T["U"].device)
    uscal = dict(scal, dtmove=1e-4)
    dt_out = torch.empty(1, dtype=torch.float64, device=T["U"].device)

    def chain():
        """the rest of one P1 stage after the reconstruction, each call timed"""
        out = []
        for fn in (lambda: cs.make_facial_fluxes_p1(N, NR, g, scal, T),
                   lambda: cs.make_edge_efields_p1(N, NR, g, T),
                   lambda: cs.build_non_stiff_sources_p1(N, NR, g, 1.0, 0.8, 1e-12, (0.0, 0.0, 0.1), T),
                   lambda: cs.update_imex_rk_step_p1(N, NR, g, uscal, T),
                   lambda: cs.cleanup_after_update_p1(N, NR, g, scal, T),
                   lambda: cs.evaluate_timestep_p1(N, NR, g, scal, T, 0.4, 1.0, out=dt_out)):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            out.append(time.perf_counter() - t0)
        return out

    for _ in range(max(1, args.warmup)):
        reset()
        cs.limit_all_slopes_p1(N, NR, g, scal, T)
        chain()
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 50))
    tot = 0.0
    tchain = [0.0] * 6
    stage = None
    for _ in range(steps):
        reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vmin, vmax, ms = cs.limit_all_slopes_p1(N, NR, g, scal, T, timed=True)
        torch.cuda.synchronize()
        tot += time.perf_counter() - t0
        tchain = [a + b for a, b in zip(tchain, chain())]
        stage = ms if stage is None else [a + b for a, b in zip(stage, ms)]
    ms_call = tot / steps * 1e3
    ms_flux, ms_edge, ms_src, ms_upd, ms_clean, ms_dt = (t / steps * 1e3 for t in tchain)
    stage = [s / steps for s in stage]
    own = N * N * NR
    cnt = cs.p1_counts(N, NR, g)
    wr = (N + 6) ** 2 * (NR + 6)
    kern = {}
    for name, t in zip(cs.P1_STAGES, stage):
        k = {"ms": t}
        if name in P1_BYTES:
            items = wr if name == "weno+positivity" else cnt["mask"]
            k["items"] = items
            k["GBps"] = P1_BYTES[name] * items / (t * 1e-3) / 1e9
            k["frac"] = k["GBps"] / HBM_PEAK_GBS
        kern[name] = k
        log(f"   {name:20s} {t:8.3f} ms" + (f"  {k['GBps']:7.1f} GB/s algorithmic" if "GBps" in k else ""))
    rest = {}
    nfc = cnt["Bxf"]
    for name, t, nb, items in (("MAKE_FACIAL_FLUXES", ms_flux, None, own),
                               ("MAKE_EDGE_EFIELDS_P1", ms_edge, P1_EDGE_BYTES + 3 * 8 * nfc / own, own),
                               ("BUILD_NON_STIFF_SOURCES", ms_src, P1_SOURCE_BYTES, cnt["mask"]),
                               ("UPDATE_IMEX_RK_STEP_P1", ms_upd, None, own),
                               ("CLEANUP_AFTER_UPDATE", ms_clean, P1_CLEANUP_BYTES, cnt["mask"]),
                               ("EVALUATE_TIMESTEP", ms_dt, P1_TIMESTEP_BYTES, own)):
        r = {"ms": t, "zones_per_s": own / (t * 1e-3)}
        if nb:
            r["GBps"] = nb * items / (t * 1e-3) / 1e9
            r["frac"] = r["GBps"] / HBM_PEAK_GBS
        rest[name] = r
        log(f"   {name:22s} {t:8.3f} ms  ({own / (t * 1e-3):.3e} zones/s)"
            + (f"  {r['GBps']:7.1f} GB/s algorithmic" if nb else ""))
    dom = max(kern, key=lambda n: kern[n]["ms"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_p1(args.cpu_seconds)
    line = {
        "metric": "P1 zone reconstructions/s", "value": own / (ms_call * 1e-3), "unit": "zones/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": ms_call, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"P1: LIMIT_ALL_SLOPES_P1 on one sector of the 6x{N}^2x{NR} shell as a patch, "
                               f"NGhost {g} ({own} owned zones, {cnt['mask']} with ghosts); front + log-normal "
                               "noise 0.3", "zones": own},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom].get("GBps"), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kern[dom].get("frac"), "traffic": None},
        "cpu_baseline": cpu, "kernels": kern,
        "facial_fluxes": rest["MAKE_FACIAL_FLUXES"], "stage_chain": rest,
    }
    print(json.dumps(line))


def bench_p12(cs, args):
    """The third-order LIMIT_ALL_SLOPES_P12 (P1 WENO, WENO-AO hybridisation
    with the 18-zone quadratic stencil, positivity with the quadratic modes,
    face-point B, the 30-unknown constrained least squares of the zone B, the
    zone E with its WENO-AO, the diagnostics) on one sector of the
    6 x 192^2 x 64 shell as a patch (NGhost 4): owned-zone reconstructions
    per second; the oracle and the reference's own routine (built at
    NUMORDER 3) timed on a small patch."""
    import torch
    N, NR, g = 192, 64, 4
    cfg, scal, T = p1_patch_on_gpu(cs, N, NR, g)
    dev = T["U"].device
    cnt = cs.p1_counts(N, NR, g)
    nz = cnt["mask"]
    vol = cs.build_volume_p12(cfg, skip=("v_pts", "v_wts"))
    sec = lambda a, m: torch.from_numpy(a[:nz * m].copy()).to(dev)  # sector 0
    T["v_centroid"] = sec(vol["v_cen"], 3)
    mom = sec(vol["moments"], 10)
    del vol
    U4 = T["U"].view(4, 8 * nz)
    U = torch.zeros(10, 8 * nz, dtype=torch.float64, device=dev)
    U[0] = U4[0]
    T["U"] = U.view(-1)
    T["V"] = torch.zeros(80 * nz, dtype=torch.float64, device=dev)
    for k in ("B1_zc", "B2_zc", "B3_zc", "Ex_zc", "Ey_zc", "Ez_zc"):
        T[k] = torch.zeros(10 * nz, dtype=torch.float64, device=dev)
    U0, e0, sc0 = T["U"].clone(), T["ent"].clone(), T["scale"].clone()

    def reset():
        T["U"].copy_(U0)
        T["ent"].copy_(e0)
        T["scale"].copy_(sc0)

    for _ in range(max(1, args.warmup)):
        reset()
        cs.limit_all_slopes_p12(N, NR, g, scal, T, mom)
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 50))
    tot = 0.0
    for _ in range(steps):
        reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cs.limit_all_slopes_p12(N, NR, g, scal, T, mom)
        torch.cuda.synchronize()
        tot += time.perf_counter() - t0
    ms_call = tot / steps * 1e3
    own = N * N * NR
    log(f"LIMIT_ALL_SLOPES_P12 on a {N}^2 x {NR} patch: {ms_call:.3f} ms per call, {own / (ms_call * 1e-3):.3e} zones/s")
    # the stages of the same call, HIP events on its stream, a few calls more
    stage = None
    nst = max(1, min(steps, 5))
    for _ in range(nst):
        reset()
        _, _, ms = cs.limit_all_slopes_p12(N, NR, g, scal, T, mom, timed=True)
        stage = ms if stage is None else [a + b for a, b in zip(stage, ms)]
    stage = {name: t / nst for name, t in zip(cs.P1_STAGES, stage)}
    for name, t in stage.items():
        log(f"   {name:20s} {t:8.3f} ms")
    # the dominant stage: 8.d's constrained least squares (k_clsq12) on every
    # active zone of the box -2..N+3, P12_CLSQ_FLOP algorithmic FP64 flops each
    box = _box_zones(N, NR, g, -2, N + 3, -2, NR + 3).to(T["mask"].device)
    nclsq = int(torch.count_nonzero(T["mask"].view(-1)[box]))
    t_clsq = stage["clsq+E"]
    ach = P12_CLSQ_FLOP * nclsq / (t_clsq * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "clsq+E (k_clsq12, k_e12)", "achieved": ach, "peak": FP64_PEAK_TFS,
            "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFS, "traffic": None, "ms_per_launch": t_clsq,
            "flop_per_zone": P12_CLSQ_FLOP, "zones_per_launch": nclsq,
            "note": "FP64 VALU (one wave per zone); 78.6 TFLOP/s is the MI355X FP64 peak, vector and matrix alike"}
    log(f"   k_clsq12: {nclsq} zones, {ach:.2f} TFLOP/s algorithmic ({ach / FP64_PEAK_TFS:.3f} of FP64 peak)")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_p12(args.cpu_seconds)
    line = {
        "metric": "P12 zone reconstructions/s", "value": own / (ms_call * 1e-3), "unit": "zones/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": ms_call, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"P12: LIMIT_ALL_SLOPES_P12 (third-order WENO-AO, positivity, constrained "
                               f"least-squares B, zone E) on "
                               f"one sector of the 6x{N}^2x{NR} shell as a patch, NGhost {g} ({own} owned zones, "
                               f"{nz} with ghosts); front + log-normal noise 0.3", "zones": own},
        "roofline": roof, "cpu_baseline": cpu, "stages_ms": stage,
    }
    print(json.dumps(line))


# LIMIT_ALL_SLOPES_P12 step 8.d per zone, algorithmic FP64 flops: the
# symmetric Gram of the 30 point rows of length 30 (30 x 465 multiply-adds),
# their right-hand side and the 6 face-mean constraint rows (2 x 900), the
# 36 x 36 KKT system by LU without pivoting (2/3 36^3) and its two triangular
# solves (2 x 36^2)
P12_CLSQ_FLOP = 2 * 30 * 465 + 2 * 2 * 900 + (2 * 36 ** 3) // 3 + 2 * 36 ** 2
FP64_PEAK_TFS = 78.6


def _box_zones(N, NR, g, i0, i1, k0, k1):
    """flat zone indices of the box i, j in [i0, i1], k in [k0, k1] of a patch"""
    import torch
    nz = N + 2 * g
    ii = torch.arange(i0, i1 + 1)
    kk = torch.arange(k0, k1 + 1)
    zi = (ii[None, None, :] - 1 + g) + nz * ((ii[None, :, None] - 1 + g) + nz * (kk[:, None, None] - 1 + g))
    return zi.reshape(-1)


def cpu_baseline_p12(max_seconds):
    """The oracle's LIMIT_ALL_SLOPES_P12 (one host core) and the reference's
    own routine built at NUMORDER 3 (one host core) on a 24^2 x 8 patch."""
    from oracle import oracle as orc
    N, NR = 24, 8
    prob, A0 = orc.p12_problem(N, NR, 4, sector=1, seed=3, shock=True, noise=0.3)
    own = N * N * NR
    out = {"unit": "zones/s", "cores": 1, "kind": "port", "sample": f"{N}^2 x {NR} patch, NGhost 4"}
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < max_seconds / 2 or n == 0:
        orc.p12_oracle(prob, orc.p1_copy(A0))
        n += 1
    out["value"] = own * n / (time.perf_counter() - t0)
    if orc.rlib3() is not None:
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < max_seconds / 2 or n == 0:
            orc.p12_reference(prob, orc.p1_copy(A0))
            n += 1
        out["reference"] = own * n / (time.perf_counter() - t0)
    return out


def bench_p1_shell(cs, args):
    """The reference's own time step on the 6 x 192^2 x 64 shell (the C3
    size): six P1 patches, the driver's loop (cubesphere_amd.P1Shell) of
    exchange, LIMIT_ALL_SLOPES_P1, MAKE_FACIAL_FLUXES, MAKE_EDGE_EFIELDS_P1,
    BUILD_NON_STIFF_SOURCES per patch, the defect correction over the shared
    faces and edges (cs_sync_shared_p1), UPDATE_IMEX_RK_STEP_P1,
    CLEANUP_AFTER_UPDATE per patch, Heun RK2, EVALUATE_TIMESTEP per step.
    Cell-updates per second of whole steps."""
    import torch
    N, NR, g = 192, 64, 4
    cfg0 = cs.make_config(N, NR, g, cs.RECON_WENO5, cs.RS_HLL, cs.EMF_LLF, cs.FP64, 2, cs.RBC_PAD,
                          cs.RMESH_UNIFORM, 1.0, 1.6)
    geo, _ = cs.build_geometry(cfg0, skip=("v_pts", "v_wts", "v_tan", "moments"))
    patches = []
    for s in range(6):
        cfg, scal, T = p1_patch_on_gpu(cs, N, NR, g, geo=geo, sector=s, front_axis=(0.3, 0.5, 0.8), seed=7 + s)
        T["Source"] = torch.zeros(8 * cs.p1_counts(N, NR, g)["mask"], dtype=torch.float64, device=T["U"].device)
        T["flag_entprs"] = torch.zeros(cs.p1_counts(N, NR, g)["mask"], dtype=torch.float64, device=T["U"].device)
        patches.append(T)
    del geo
    shell = cs.P1Shell(cfg, patches, scal, sources=(1.0, 0.0, (0.0, 0.0, 0.0)))
    try:
        dt = shell.timestep(0.4, 1.0)
        for _ in range(max(1, args.warmup)):
            shell.step(dt)
            dt = shell.timestep(0.4, dt)
        steps = max(1, min(args.steps, 10))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            shell.step(dt)
            dt = shell.timestep(0.4, dt)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
    finally:
        shell.close()
    cells = 6 * N * N * NR
    log(f"   P1 shell step {ms:8.2f} ms  ({cells / (ms * 1e-3):.3e} cell-updates/s), dt {dt:.3e}")
    cpu = None if args.no_cpu else cpu_baseline_p1_shell(args.cpu_seconds)
    line = {
        "metric": "cell-updates/s on 1 MI355X, the reference's P1 scheme", "value": cells / (ms * 1e-3),
        "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"P1 shell: 6x{N}^2x{NR} shell (C3 size), six P1 patches, NGhost {g}, the driver's "
                               "Heun RK2 step of the reference's routines with its defect correction and outflow radial pads; front + log-normal "
                               "noise 0.3", "zones": cells},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def cpu_baseline_p1_shell(max_seconds):
    """The oracle's restatement of the same step (the reference's routines 450,
    470 and 484 do not build from its sources, so no reference build of the
    whole step exists) on a 6 x 16^2 x 8 shell, one host thread."""
    import cubesphere_amd as cs
    from oracle import oracle as orc
    N, NR = 16, 8
    cfg, probs, As = orc.p1_shell_problem(N, NR, 4, seed=7, noise=0.3)
    grid = orc.Grid(cfg)
    pairings = cs.p1_shell_pairings(As, N, NR, 4)
    dt = min(orc.p1_timestep_oracle(p, orc.p1_copy(A), 0.4, 1.0) for p, A in zip(probs, As))
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.p1_shell_step_oracle(grid, probs, As, dt, pairings=pairings)
        steps += 1
        el = time.perf_counter() - t0
        if el > max_seconds or steps >= 50:
            break
    cells = 6 * N * N * NR
    return {"value": cells * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of the oracle's shell step on a 6x16^2x8 shell (six P1 patches, defect "
                      f"correction), single thread, {el:.1f} s"}


def cpu_baseline_p1(max_seconds):
    """The reference's own LIMIT_ALL_SLOPES_P1 (built from its sources) on the
    host, single thread, on a 32^2 x 16 patch of the same kind of data; the
    oracle's restatement when that build is absent."""
    from oracle import oracle as orc
    prob, A0 = orc.p1_problem(32, 16, 4, sector=0, seed=7, shock=True, noise=0.3)
    kind = "reference" if orc.rlib() is not None else "port"
    fn = orc.p1_reference if kind == "reference" else orc.p1_oracle
    calls, t0 = 0, time.perf_counter()
    while True:
        fn(prob, orc.p1_copy(A0))
        calls += 1
        el = time.perf_counter() - t0
        if el > max_seconds or calls >= 200:
            break
    own = 32 * 32 * 16
    return {"value": own * calls / el, "unit": "zones/s", "cores": 1, "kind": kind,
            "sample": f"{calls} calls on a 32^2x16 patch (NGhost 4), single thread, {el:.1f} s"}


# Algorithmic HBM bytes per work item of the P1 stage's launch groups (the
# words an item must read or write once, data shared by neighbouring items
# counted once): the work items are zones of the stated range, faces or edges.
#   prim          zone (ghosts incl.): U + entropy read and written (18), V means
#                 (9) and the magnetosonic speed (1) written
#   flatten       zone: speed, velocity, centroid, scale, pressure, three flag
#                 passes (14 words)
#   weno          zone of the WENO range: V means 9, centroid 3, scale 1,
#                 flattener 1 read, slopes 27 written
#   positivity    zone of the WENO range: V means 9 and slopes 27 read,
#                 centroid 3, scale 1, flattener 1 (rewritten slopes not
#                 counted: only reduced zones write them)
#   face_points   face: the B modes of its two zones (each zone's 12 shared by
#                 its two faces of the family: 12), face B 1, five points written
#   clsq          zone of the WENO range: B modes 12, zone B and energy 4, its
#                 three own face B and six own face-point values, centroid 3,
#                 scale 1 read; B modes 12 and U 4 written
#   flux          face: its two zones' 36 V words (each zone's shared by its
#                 two faces of the family: 36), centroid and scale 4; flux x
#                 area 9 and speeds 2 written
#   edge          edge: v and B of its zones (6, each zone's shared by its
#                 four edges of the family: 6), 4 face-point values and 4
#                 speeds; the field written
#   update_faces  owned face: B of the stage input and of U^n, the edge
#                 fields (each shared by 4 faces: 1), B written
#   update_zones  owned zone: the fluxes of its six faces (each shared by
#                 two zones: 27), stage input and U^n (9 + 9), its three face
#                 B, centroid 3, volume 1; U + entropy written (9)
#   imex          zone (ghosts incl.): U + entropy read and written, centroid
#                 3, face B copied (3 + 3)
#   cfl           owned zone: U + entropy 9 and scale 1 read, U 8 written
P1K_BYTES = {"prim": 8 * 28, "flatten": 8 * 14, "weno": 8 * 41, "positivity": 8 * 41, "face_points": 8 * 18,
             "clsq": 8 * 45, "flux": 8 * 51, "edge": 8 * 15, "update_faces": 8 * 4, "update_zones": 8 * 61,
             "imex": 8 * 27, "cfl": 8 * 18}


def p1k_units(cfg):
    """Work items of one launch (group) of each P1 stage kernel over the six
    sectors: zones with ghosts, the WENO range (-2 .. N+3), faces of the flux
    and face-point ranges, edges and owned faces and zones."""
    N, NR, g = cfg.ntp, cfg.nr, cfg.nghost
    zones_all = 6 * (N + 2 * g) ** 2 * (NR + 2 * g)
    zones_wr = 6 * (N + 6) ** 2 * (NR + 6)
    own = 6 * N * N * NR
    fpts = 6 * 3 * (N + 7) * (N + 6) * (NR + 6)
    fflux = 6 * 3 * (N + 3) * (N + 2) * (NR + 2)
    edges = 6 * 3 * N * (N + 1) * (NR + 1)
    ofaces = 6 * 3 * (N + 1) * N * NR
    return {"prim": zones_all, "flatten": zones_all, "weno": zones_wr, "positivity": zones_wr,
            "face_points": fpts, "clsq": zones_wr, "flux": fflux, "edge": edges, "update_faces": ofaces,
            "update_zones": own, "imex": zones_all, "cfl": own}


# compulsory HBM bytes per cell-update of a whole step (the state of 12 FP64
# words read 5 times and written 3 times: IMEX-SSP2(2,2,2) reads U^n in its
# source stage and both flux stages and the two stage inputs, and writes U1,
# U2 and U^n+1; SSP-RK3 moves the same 8 copies); the metric is separable
# (O(N^2 + nr) tables), so it adds nothing for nr >= 16
B_ALG_3D = 8 * 12 * 8
# the surface (nr = 1): the same plus 3 stages x 26 metric words per zone
B_ALG_SURFACE = B_ALG_3D + 3 * 26 * 8


def host_cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def _p1adp_cpu_worker(args):
    """One host process: the oracle's step of the same scheme on its own small
    shell, steps counted for max_seconds (returns (cell-updates, seconds))."""
    name, N, NR, seed, max_seconds = args
    import cubesphere_amd as cs
    from oracle import oracle as orc
    cfg0, prob, _ = cs.workload(name, cs.FP64)
    cfg = cs.make_config(N, 1 if cfg0.nr == 1 else NR, cfg0.nghost, cfg0.recon, cfg0.riemann, cfg0.emf, cs.FP64,
                         cfg0.rk_order, cfg0.radial_bc, cfg0.radial_mesh, cfg0.rmin, cfg0.rmax, grav=cfg0.grav,
                         imex=cfg0.imex)
    st = cs.HostState(cfg)
    st.U[:], st.ent[:], st.Bx[:], st.By[:], st.Bz[:] = orc.init_problem_p1(cfg, P1ADP_PROBLEM[name], seed=seed)
    G = orc.Grid(cfg)
    As = orc.p1s_patches(cfg)
    dt = orc.p1s_cfl_oracle(cfg, As, st, 1.0)
    steps, t0 = 0, time.perf_counter()
    while True:
        dt = orc.p1s_step_oracle(G, cfg, As, st, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > max_seconds or steps >= 50:
            break
    return 6 * cfg.ntp * cfg.ntp * cfg.nr * steps, el, steps


def cpu_baseline_p1adp(cs, name, max_seconds):
    """The oracle's step of the same scheme (its restatement of the
    reference's routines; 450, 470 and 484 do not build from the reference's
    sources, so no reference build of a whole step exists) on the host: one
    small shell per core, as independent processes on every core the box
    gives the job (at most 16), their throughputs summed."""
    import multiprocessing as mp
    info = host_cpu_info()
    cores = max(1, min(info["nproc"] or 1, 16))
    N, NR = (32, 16) if name != "C2" else (48, 1)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_p1adp_cpu_worker, [(name, N, NR, 100 + q, max_seconds) for q in range(cores)])
    wall = time.perf_counter() - t0
    value = sum(c / el for c, el, _ in res)
    steps = sum(n for _, _, n in res)
    log(f"cpu baseline: {cores} processes, {steps} oracle steps in {wall:.1f} s, {value:.3e} cell-updates/s")
    return {"value": value, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "cpu_model": info["model"], "nproc": info["nproc"],
            "sample": f"the oracle's step of the same scheme (its C restatement of the reference's routines, "
                      f"driven per sector) on a 6x{N}^2x{NR if name != 'C2' else 1} shell per process, "
                      f"{cores} independent processes for about {max_seconds:.0f} s each ({steps} steps in all)"}


def reference_cpu_leg(seconds=10.0, N=48, NR=24):
    """The reference's own LIMIT_ALL_SLOPES_P1 built from its sources with its
    OpenMP directives (flang -fopenmp) on a P1 patch, zones per second, on the
    cores the box gives the job; None when that build is absent."""
    from oracle import oracle as orc
    info = host_cpu_info()
    cores = max(1, min(info["nproc"] or 1, 16))
    os.environ["OMP_NUM_THREADS"] = str(cores)  # read when the OpenMP runtime loads
    R = orc.rlib(openmp=True)
    if R is None:
        return None
    prob, A0 = orc.p1_problem(N, NR, 4, sector=0, seed=7, shock=True, noise=0.3)
    calls, t0 = 0, time.perf_counter()
    while True:
        orc.p1_reference(prob, orc.p1_copy(A0), lib=R)
        calls += 1
        el = time.perf_counter() - t0
        if el > seconds or calls >= 500:
            break
    return {"routine": "LIMIT_ALL_SLOPES_P1 (410...f:9) with 440 and 120", "value": N * N * NR * calls / el,
            "unit": "zones/s", "cores": cores, "kind": "reference",
            "sample": f"{calls} calls on a {N}^2x{NR} patch (NGhost 4), flang -O2 -fopenmp, OMP_NUM_THREADS={cores}"}


# the device problem of each P1 workload (cs_init_problem_p1): the surface
# blast with B = 0.5 east (C2), the isothermal hydrostatic shell with gravity
# 1/r^2 and 1e-2 density noise (C5)
P1ADP_PROBLEM = {"C2": 102, "C5": 103}


def bench_p1adp(cs, args, cpu):
    """C5 (or C2) with the reference's P1 scheme: FP64 steps replayed from one
    hipGraph per step, per-launch times from an eagerly run step with a
    hipEvent pair around every launch group, the dominant group's roofline,
    the whole step's share of the HBM roofline, div B and the conservation
    integrals."""
    name = args.workload
    cfg, prob, _ = cs.workload(name, cs.FP64)
    t0 = time.perf_counter()
    dev = cs.CubeSphere(cfg)
    # the initial state on the device (INIT_PATCH_CUBE_SPHERE with the
    # workload's problem): face B from the vector potential's circulation,
    # so div B starts at rounding on the P1 geometry
    dev.init_problem_p1(P1ADP_PROBLEM[name], seed=20261202)
    log(f"{name}: handle created and initial state built on the device in {time.perf_counter() - t0:.1f} s")
    try:
        dev.dt = 0.0
        dev.evaluate_timestep()
        divb0 = dev.max_divb()
        sums0 = dev.conservation_sums()
        dev.advance(args.warmup, use_graph=True)
        ms = dev.time_steps(args.steps, use_graph=True)
        prof = dev.profile_steps_p1(args.profile_steps)
        divb = dev.max_divb()
        sums = dev.conservation_sums()
        dt = dev.dt
    finally:
        dev.close()
    n_int = 6 * cfg.ntp * cfg.ntp * cfg.nr
    cups = n_int * args.steps / (ms * 1e-3)
    log(f"{name} fp64: {ms / args.steps:.2f} ms/step, {cups:.3e} cell-updates/s, dt {dt:.3e}, "
        f"max rel div B {divb0:.2e} -> {divb:.2e}")
    units = p1k_units(cfg)
    kern = {}
    for k, (t, calls) in sorted(prof.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        e = {"ms": t, "launches_per_step": calls / args.profile_steps}
        if k in P1K_BYTES:
            e["units"] = units[k]
            e["alg_GBps"] = P1K_BYTES[k] * units[k] / (t * 1e-3) / 1e9
        kern[k] = e
        log(f"   {k:14s} {t:9.3f} ms x {e['launches_per_step']:4.1f}"
            + (f"  units {units[k]:11d}  alg {e['alg_GBps']:8.1f} GB/s" if k in P1K_BYTES else ""))
    step_ms_prof = sum(e["ms"] * e["launches_per_step"] for e in kern.values())
    dom = max((k for k in kern if k in P1K_BYTES), key=lambda k: kern[k]["ms"] * kern[k]["launches_per_step"])
    kd = kern[dom]
    alg = P1K_BYTES[dom] * units[dom]
    roof = {"bound": "hbm", "kernel": dom, "achieved": kd["alg_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": kd["alg_GBps"] / HBM_PEAK_GBS, "traffic": measured_traffic(dom, "fp64", name),
            "alg_bytes_per_launch": alg,
            "alg_bytes_per_unit": P1K_BYTES[dom], "units_per_launch": units[dom],
            "avg_launch_us": kd["ms"] * 1e3,
            "launch_time": f"hipEvent pair around each launch group on the library's stream, "
                           f"{args.profile_steps} step(s) run launch by launch"}
    b_alg = B_ALG_SURFACE if cfg.nr == 1 else B_ALG_3D
    step_gbs = cups * b_alg / 1e9
    # the radial boundaries of a shell are open (outflow pads): its mass
    # changes by what crosses them; the surface between walls is closed
    rel_mass = abs(sums[0] / sums0[0] - 1.0) if sums0[0] else None
    mass_key = "rel_mass_change_open_radial_boundaries" if cfg.nr > 1 else "rel_mass_change_closed"
    ref_leg = None
    if not args.no_cpu:
        try:
            ref_leg = reference_cpu_leg()
        except Exception as e:
            log("reference CPU leg failed:", e)
    line = {
        "metric": METRIC,
        "value": cups,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{name}: 6x{cfg.ntp}x{cfg.ntp}x{cfg.nr} zones, the reference's P1 scheme "
                               f"(WENO-ADP) + {'HLLD' if cfg.riemann == cs.RS_HLLD else 'HLL'} + 4-state HLL "
                               "edge fields + arc-length CT, "
                               + ("IMEX-SSP2(2,2,2) with implicit gravity 1/r^2, shell r in [1, 1.5], isothermal "
                                  "hydrostatic state + seeded 1e-2 density noise, B from the vector potential "
                                  "0.05/r^2 (yz, zx + xy/2, xy)" if cfg.imex else
                                  "SSP-RK3, surface between conducting walls, magnetised blast (p 10 in a 0.1 rad "
                                  "cap, B = 0.5 east)")
                               + "; initial state built on the device by INIT_PATCH_CUBE_SPHERE "
                                 "(cs_init_problem_p1)",
                   "zones": n_int, "precision": "fp64", "graph": True},
        "roofline": roof,
        "cpu_baseline": cpu,
        "reference_cpu_leg": ref_leg,
        "fp64": {"cell_updates_per_s": cups, "ms_per_step": ms / args.steps,
                 "max_rel_divb_start": divb0, "max_rel_divb_end": divb, mass_key: rel_mass},
        "mixed": None,
        "precision_gate": None,
        "step_hbm": {"b_alg_per_cell_update": b_alg, "achieved_GBps": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
        "kernels": kern,
        "profiled_step_ms": step_ms_prof,
    }
    print(json.dumps(line))



def probe(cs, name, prec, kernel, iters):
    """Hardware-counter run: `iters` launches of one kernel on the workload's
    state, then one plain device copy of 1 GiB (read) + 1 GiB (write), whose
    counter readings calibrate the byte counters."""
    import torch
    cfg, prob, _ = cs.workload(name, prec)
    st = initial_state(cs, name)
    dev = cs.CubeSphere(cfg)
    dev.upload(st)
    dev.dt = 0.0
    dev.evaluate_timestep()
    dev.advance(2, use_graph=False)
    kid = cs.KERNEL_NAMES.index(kernel)
    ms = dev.time_kernel(kid, iters)
    units = dev.kernel_units(kid)
    dev.close()
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"probe {kernel} ({'fp64' if prec == 0 else 'mixed'}): {iters} launches, {ms * 1e3:.2f} us avg, "
        f"{units} units; calibration copy {a.numel() * 8} bytes read and written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="C5")
    ap.add_argument("--profile-steps", type=int, default=1,
                    help="steps run launch by launch for the per-kernel durations")
    ap.add_argument("--probe-iters", type=int, default=5, help="launches of the --probe kernel")
    ap.add_argument("--gate-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", default=None, metavar="KERNEL",
                    help="only launch this kernel --profile-iters times plus a calibration copy "
                         "(for rocprofv3 --pmc runs); prints no JSON line")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64", help="precision of --probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note only (see the sharding paragraph of the design document)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import cubesphere_amd as cs
    p1adp = args.workload in ("C2", "C5") and not args.probe
    cpu = None
    if p1adp and not args.no_cpu:
        # before any HIP call: the host processes are forked from a process
        # that has not touched the GPU
        try:
            cpu = cpu_baseline_p1adp(cs, args.workload, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, not required
            log("cpu baseline failed:", e)
    if args.workload in ("P1", "P1SHELL", "P12"):
        # the P1 legs keep their patches in torch tensors: torch's HIP runtime has
        # to be initialised before the library's first HIP call
        import torch
        torch.zeros(1, device="cuda")
    log("device:", cs.device_info())
    if p1adp:
        bench_p1adp(cs, args, cpu)
        return
    if args.probe:
        probe(cs, args.workload, cs.FP64 if args.precision == "fp64" else cs.MIXED, args.probe,
              args.probe_iters)
        return
    if args.workload == "C4":
        bench_batch(cs, args)
        return
    if args.workload == "G3":
        bench_geometry(cs, args)
        return
    if args.workload == "P1":
        bench_p1(cs, args)
        return
    if args.workload == "P1SHELL":
        bench_p1_shell(cs, args)
        return
    if args.workload == "P12":
        bench_p12(cs, args)
        return
    res = {}
    for prec in (cs.FP64, cs.MIXED):
        try:
            cfg, ms, kern, divb, sums, dt = run_gpu(cs, args.workload, prec, args.steps, args.warmup,
                                                    args.profile_steps)
        except cs.CsError as e:
            if prec == cs.FP64:
                raise
            log("mixed-precision run failed:", e)
            continue
        n_int = 6 * cfg.ntp * cfg.ntp * cfg.nr
        cups = n_int * args.steps / (ms * 1e-3)
        res[prec] = dict(cfg=cfg, ms=ms, kern=kern, divb=divb, sums=sums, dt=dt, cups=cups, n_int=n_int)
        log(f"prec={'fp64' if prec == 0 else 'mixed'}: {ms / args.steps:.4f} ms/step, "
            f"{cups:.3e} cell-updates/s, max rel divB {divb:.2e}, dt {dt:.3e}")
        surface = cfg.nr == 1 and cfg.radial_bc == cs.RBC_WALL
        for k, (t, u, kid) in sorted(kern.items(), key=lambda kv: -kv[1][0]):
            b = algorithmic_bytes_per_unit(cs, kid, prec, surface) * u
            log(f"   {k:14s} {t * 1e3:9.2f} us  units {u:10d}  alg {b / (t * 1e-3) / 1e9 if t > 0 else 0:8.1f} GB/s")

    try:
        gate = precision_gate(cs, args.workload, args.gate_steps)
    except cs.CsError as e:
        gate = {"steps": args.gate_steps, "error": str(e), "max_rel_l1": None}
    log("precision gate:", json.dumps(gate))
    mixed_ok = cs.MIXED in res and gate.get("max_rel_l1") is not None and gate["max_rel_l1"] <= 1e-6
    head = cs.MIXED if mixed_ok else cs.FP64
    r = res[head]

    # dominant kernel of the headline precision
    cfg = r["cfg"]
    surface = cfg.nr == 1 and cfg.radial_bc == cs.RBC_WALL
    kname, (kt, ku, kid) = max(r["kern"].items(), key=lambda kv: kv[1][0])
    alg = algorithmic_bytes_per_unit(cs, kid, head, surface) * ku
    achieved = alg / (kt * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(kname, "fp64" if head == cs.FP64 else "mixed", args.workload),
            "alg_bytes_per_launch": alg, "avg_launch_us": kt * 1e3,
            "launch_time": f"hipEvent pair around each launch, {args.profile_steps} steps run launch by launch"}
    # whole-step view: the compulsory bytes of every kernel of a stage, times
    # the stages of a step, per zone update
    stage_bytes = sum(algorithmic_bytes_per_unit(cs, kid_, head, surface) * u_
                      for (_, u_, kid_) in r["kern"].values())
    b_alg = stage_bytes * cfg.rk_order / r["n_int"]
    step_gbs = r["cups"] * b_alg / 1e9

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(cs, args.workload, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, not required
            log("cpu baseline failed:", e)
    line = {
        "metric": METRIC,
        "value": r["cups"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == cs.MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: 6x{cfg.ntp}x{cfg.ntp}x{cfg.nr} zones, {describe(cs, cfg, cs.workload(args.workload)[1])}",
                   "zones": r["n_int"], "precision": "mixed" if head == cs.MIXED else "fp64",
                   "graph": True},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64": {"cell_updates_per_s": res[cs.FP64]["cups"], "ms_per_step": res[cs.FP64]["ms"] / args.steps,
                 "max_rel_divb": res[cs.FP64]["divb"]},
        "mixed": ({"cell_updates_per_s": res[cs.MIXED]["cups"], "ms_per_step": res[cs.MIXED]["ms"] / args.steps,
                   "max_rel_divb": res[cs.MIXED]["divb"],
                   "speedup_vs_fp64": res[cs.MIXED]["cups"] / res[cs.FP64]["cups"]}
                  if cs.MIXED in res else None),
        "precision_gate": gate,
        "step_hbm": {"b_alg_per_cell_update": b_alg, "achieved_GBps": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
