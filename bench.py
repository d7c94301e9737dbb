#!/usr/bin/env python3
"""RK3 substep throughput of libmeltgpu.so on one MI355X.

Workload (default): 256^3 periodic cube of side 2*pi with stochastic HIT
forcing at |k| <= 2.5 (hit.f90, regenerated on the device every step), a
melting sphere of radius 0.5 meshed as a level-5 icosphere (10242 vertices,
20480 faces), Stefan number 2 and Newton-Euler coupling on, FP64.  A "step"
is one time step of the reference's loop: the HIT forcing and one RK3 step
(three substeps of tsch.f90) of the whole field and body; inputs are resident
in HBM before the timed region.  Inputs are synthetic and seeded (no datasets
exist for this path).

Prints one JSON line: value = Mcell-steps/s (whole job), plus
 - roofline: algorithmic HBM bytes of the dominant kernel over its average
   launch time (HIP events on the library's stream, over the timed region),
   against the 8 TB/s HBM3E peak, and the same for the whole substep;
 - cpu_baseline: the FP64 C restatement of the reference path on this host's
   cores, median substep time at the 64^3 reference case and at this workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s

# Algorithmic HBM bytes per cell of each fused-path kernel (FP64, 8 B per value;
# a half-spectrum complex field counts 8 B per cell).  Minimal traffic of the
# pass: each field it must read or write, once.
KERNEL_BYTES = {
    "explicit_xsolve": (13 + 8) * 8,   # u,v,w,T,p, ru x4, VOFp, forc x3 -> x-solved rhs x4, ru x4
    "solve_y": (4 + 4 + 4) * 8,         # y and z solves in one pass: rhs x4, q x4 -> q x4
    "solve_z_update": (4 + 4 + 4) * 8,  # (two-pass form only) tmp x4, q x4 -> q x4
    "divg_fft_x": (3 + 1) * 8,
    "fft_y": 2 * 8,
    "fft_z_solve": 2 * 8,
    "ifft_y": 2 * 8,
    "ifft_x": 2 * 8,
    "project": (5 + 4) * 8,            # dph, u,v,w,p -> u,v,w,p
}
DOMINANT = "explicit_xsolve"   # pass A: the largest share of every kernel trace of this workload
B_ALG_CELL = 480  # algorithmic bytes per cell per substep of the whole path (FP64)
# HIT forcing (HITForcing.in): TL, EPSSTAR, KF/KMIN = 2.5 (|k| <= 2.5, 5 wave numbers per direction)
HIT_PARAMS = (0.5, 0.05, 2.5)
# precision form: the same passes at 4 B per value (240 B) plus its two FP64
# residual reductions over u, v, w (provisional and projected field, 12 B each)
B_ALG_CELL_MIXED = 240 + 24


def workload(n: int):
    from melt_ibm_amd.api import MeltConfig
    st = 2.0
    cfg = MeltConfig(n1m=n, n2m=n, n3m=n, ren=1000.0, prandtl=1.0, betagz=0.0,
                     tmelt=0.0, tliq=1.0, tsol=0.0, cpliquid=1.0, latheat=1.0 / st,
                     dens_ratio=1.0, imlsfor=1, imlsstr=1, imelt=1, iremesh=0, forcing=1)
    return cfg, 5, 0.5


def c3_workload(n: int = 512, perc_ethresh: float = 0.38):
    """The precision-form configuration: the melting workload at n^3 with a
    prolate spheroid (semi-axes 0.3, 0.3, 0.6, icosphere frequency 55: 60500
    faces) and remeshing on.  Its edges span 0.40 h to 1.17 h at 512^3, so
    E_thresh = 0.38 h starts them above the threshold, as the reference's deck
    does (level-5 sphere of radius 0.2 on a 128^3 unit cube: edges near 0.9 h
    against 0.7 h); coarsening then follows the melting.  Returns (cfg, mesh,
    volume-equivalent radius)."""
    from melt_ibm_amd import geometry
    cfg, _, _ = workload(n)
    cfg.iremesh = 1
    cfg.perc_ethresh = perc_ethresh
    r_eq = (0.3 * 0.3 * 0.6) ** (1.0 / 3.0)   # volume-equivalent radius: the mesh is scaled by it, inqpr uses it
    return cfg, geometry.spheroid(0.3 / r_eq, 0.6 / r_eq, 55), r_eq


def run_gpu(args, precision=0, keep_fields=False):
    from melt_ibm_amd.api import MeltContext
    cfg, level, rad = workload(args.n)
    cfg.precision = precision
    ctx = MeltContext(cfg)
    from melt_ibm_amd import geometry, inputs
    # temperature from inqpr, then the seeded solenoidal velocity and forcing
    m = geometry.sphere_level(level)
    ctx.set_geometry(m.xyz, m.vert_of_edge, m.edge_of_face, rad)
    ctx.inqpr()
    u, v, w = inputs.solenoidal_velocity(cfg, seed=0, kmax=4.0, rms=0.1)
    for name, a in (("vx", u), ("vy", v), ("vz", w)):
        ctx.upload(name, a)
    dt = inputs.cfl_dt(cfg, u, v, w, cfl=0.5)
    del u, v, w
    ctx.set_dt(dt)
    ctx.hit_init(*HIT_PARAMS, seed=1)
    ctx.tagCells()
    ctx.sync()
    tag = "fp64" if precision == 0 else "mixed"
    print(f"# {tag} setup done: N={args.n} dt={dt:.4e} faces={m.counts[2]}", file=sys.stderr, flush=True)

    def step():   # the time loop's order (gcurv.f90:163-172): HIT forcing, then the RK3 step
        ctx.hit_force()
        ctx.tschem()
    for _ in range(args.warmup):
        step()
    ctx.sync()
    # the timed loop records HIP events around the dominant kernel only; events
    # around every slot cost about 5% of a step, so the per-slot breakdown
    # comes from further steps afterwards
    ctx.timing(True, slots=(DOMINANT,))
    ctx.timing_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    tm = ctx.timing_read()
    ctx.timing(False)
    print(f"# {tag} timed {args.steps} steps: {(t1 - t0) / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    qmax = ctx.divgck()
    body = ctx.body()
    fields = {q: ctx.download(q) for q in ("vx", "vy", "vz", "temp", "pr", "VOFp")} if keep_fields else None
    stats = ctx.refinement_stats() if precision else None
    tm_all = {}
    if args.slot_steps > 0:
        ctx.timing(True)
        ctx.timing_reset()
        for _ in range(args.slot_steps):
            step()
        ctx.sync()
        tm_all = ctx.timing_read()
        ctx.timing(False)
    ctx.close()
    return cfg, (t1 - t0), tm, tm_all, qmax, body, fields, stats


def workload_c1():
    """The CPU-runnable reference case: 64^3, sphere of radius 0.6 (level-3
    icosahedron, 1280 faces), Re 500, Pr 1, St 1, no forcing."""
    from melt_ibm_amd.api import MeltConfig
    cfg = MeltConfig(n1m=64, n2m=64, n3m=64, ren=500.0, prandtl=1.0, betagz=0.0, tmelt=0.0, tliq=1.0, tsol=0.0, cpliquid=1.0,
                     latheat=1.0, dens_ratio=1.0, imlsfor=1, imlsstr=1, imelt=1, iremesh=0, forcing=0)
    return cfg, 3, 0.6


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_substeps(cfg, level, rad, forcing: bool, warm: int, timed: int) -> list[float]:
    """Wall time of each of `timed` oracle substeps after `warm` untimed ones
    (substeps cycle 1, 2, 3 as in tschem), excluding set-up and I/O."""
    import oracle
    from melt_ibm_amd import geometry, inputs
    o = oracle.context(cfg)
    m = geometry.sphere_level(level)
    o.set_geometry(m.xyz, m.vert_of_edge, m.edge_of_face, rad)
    o.inqpr()
    u, v, w = inputs.solenoidal_velocity(cfg, seed=0, kmax=4.0, rms=0.1)
    for name, a in (("vx", u), ("vy", v), ("vz", w)):
        o.upload(name, a)
    o.set_dt(inputs.cfl_dt(cfg, u, v, w, cfl=0.5))
    del u, v, w
    if forcing:
        o.hit_init(*HIT_PARAMS, seed=1)
        o.hit_force()      # the forcing field of the first step (outside the substep timings)
    o.tagCells()
    times = []
    for i in range(warm + timed):
        t0 = time.perf_counter()
        o.substep(i % 3 + 1)
        if i >= warm:
            times.append(time.perf_counter() - t0)
    o.close()
    return times


def cpu_baseline(n: int, warm: int = 3, timed: int = 20):
    """The oracle (the FP64 C restatement of the reference path, OpenMP over
    this host's cores) on the reference case (64^3) and on the benchmark
    workload (n^3): median substep time of `timed` substeps after `warm`."""
    import oracle
    lib = oracle.load()
    threads = int(lib.orc_threads())
    out = {}
    for tag, (cfg, level, rad), forcing in (("c1_64", workload_c1(), False), (f"c2_{n}", workload(n), True)):
        print(f"# cpu baseline {tag}: {warm} + {timed} oracle substeps", file=sys.stderr, flush=True)
        t = np.array(_oracle_substeps(cfg, level, rad, forcing, warm, timed))
        cells = cfg.n1m * cfg.n2m * cfg.n3m
        med = float(np.median(t))
        out[tag] = {"median_substep_s": round(med, 5), "min_substep_s": round(float(t.min()), 5),
                    "max_substep_s": round(float(t.max()), 5), "mcell_steps_per_s": round(cells / (3.0 * med) / 1e6, 3)}
    main_tag = f"c2_{n}"
    return {"value": out[main_tag]["mcell_steps_per_s"], "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "implementation": "oracle/ (the FP64 C restatement of the reference path; the reference's MPI Fortran does not "
                              "build without MPI/HDF5/FFTW modules), gcc -O2 -ffp-contract=off -fopenmp",
            "host": {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
                     "omp_threads": threads},
            "sample": f"median of {timed} substeps after {warm} warm-up substeps (immersed boundary included, set-up and I/O "
                      f"excluded) on {threads} OpenMP threads, at the 64^3 reference case and at the {n}^3 workload; "
                      f"value = {n}^3 cells / (3 x median substep)",
            "runs": out}


PROFILES = ROOT / "profiles"
TRAFFIC_JSON = PROFILES / "hbm_traffic_256cube_melting_sphere.json"


def _kernel_short(name: str) -> str:
    return name.replace("(anonymous namespace)::", "").split("(")[0]


def _pmc_avg(db: str, counter: str) -> dict:
    """Average value per dispatch of `counter` for every kernel in a rocprofv3 --pmc database."""
    import sqlite3
    con = sqlite3.connect(db)
    out = {}
    for name, val, n in con.execute("select kernel_name, sum(value), count(*) from counters_collection where counter_name = ? "
                                    "group by kernel_name", (counter,)):
        out[_kernel_short(name)] = (val / n, n)
    return out


def summarize_pmc(fetch_db, write_db, cfetch_db, cwrite_db, command: str, out_md: Path, out_json: Path) -> None:
    """HBM bytes per launch from FETCH_SIZE / WRITE_SIZE (KiB per dispatch), each
    scaled by what the same counter reads for a plain 1 GiB device copy."""
    f, w = _pmc_avg(fetch_db, "FETCH_SIZE"), _pmc_avg(write_db, "WRITE_SIZE")
    cf, cw = _pmc_avg(cfetch_db, "FETCH_SIZE"), _pmc_avg(cwrite_db, "WRITE_SIZE")
    copy_bytes = float(1 << 30)
    kf = copy_bytes / (cf["__amd_rocclr_copyBuffer"][0] * 1024.0)
    kw = copy_bytes / (cw["__amd_rocclr_copyBuffer"][0] * 1024.0)
    rows, data = [], {}
    for k in sorted(f, key=lambda k: -f[k][0]):
        if k not in w:
            continue
        rd, wr = f[k][0] * 1024.0 * kf, w[k][0] * 1024.0 * kw
        data[k] = {"read_bytes": rd, "write_bytes": wr, "bytes_per_launch": rd + wr, "dispatches": f[k][1]}
        if rd + wr > 1e6:
            rows.append(f"| `{k}` | {f[k][1]} | {rd / 1e9:.3f} | {wr / 1e9:.3f} | {(rd + wr) / 1e9:.3f} |")
    out_json.write_text(json.dumps({"read_scale": kf, "write_scale": kw, "kernels": data}, indent=1))
    out_md.write_text("\n".join([
        "# HBM traffic per launch, 256^3 melting-sphere workload (FP64)", "",
        "Command (GPU box):", "", "```", command, "```", "",
        f"Calibration: a plain 1 GiB device copy (`bench.py --copy-probe`, source evicted from the last-level cache) "
        f"reads {1 / kf:.3f} FETCH_SIZE KiB per KiB and writes {1 / kw:.3f} WRITE_SIZE KiB per KiB, so reads = "
        f"FETCH_SIZE x {kf:.3f}, writes = WRITE_SIZE x {kw:.3f}.", "",
        "| kernel | dispatches | read GB | write GB | total GB per launch |", "|---|---|---|---|---|", *rows, ""]))


def summarize_trace(db: str, command: str, out_md: Path, top: int = 30, leg: str = "fp64") -> None:
    """Per-kernel average duration from a rocprofv3 --kernel-trace database.
    The FP64 leg runs before the precision-form leg in one process, so the
    first dispatch of an FP32 kernel splits the trace: leg "fp64" keeps what
    comes before it, leg "mixed" what comes after (its FP64 immersed-boundary
    and Lagrangian kernels included)."""
    import sqlite3
    con = sqlite3.connect(db)
    split = con.execute("select min(start) from kernels where name like '%<float%'").fetchone()[0]
    where = ""
    if split is not None:
        where = f"where start < {split}" if leg == "fp64" else f"where start >= {split}"
    elif leg != "fp64":
        raise SystemExit("no precision-form kernels in this trace")
    rows = list(con.execute(f"select name, count(*), avg(duration), sum(duration) from kernels {where} "
                            "group by name order by sum(duration) desc"))
    tot = sum(r[3] for r in rows)
    title = "# Kernel times, 256^3 melting-sphere workload" + (", precision form" if leg != "fp64" else "")
    lines = [title, "", "Command (GPU box):", "", "```", command, "```", "",
             "| kernel | calls | avg us | total us | share |", "|---|---|---|---|---|"]
    for name, n, avg, sm in rows[:top]:
        lines.append(f"| `{_kernel_short(name)}` | {n} | {avg / 1e3:.1f} | {sm / 1e3:.1f} | {100 * sm / tot:.1f}% |")
    out_md.write_text("\n".join(lines + [""]))


KERNEL_SYMBOL = {"explicit_xsolve": "k_explicit_zmarch", "solve_y": "k_green_yz", "solve_z_update": "k_green<2",
                 "divg_fft_x": "k_fft_x_fwd_reg", "fft_y": "k_fft_y_reg", "fft_z_solve": "k_fft_z_solve_reg", "ifft_x": "k_fft_x_bwd_reg",
                 "project": "k_project"}


def measured_traffic(slot: str, n: int):
    """HBM bytes per launch of the kernel behind `slot` from the committed
    counter summary (256^3 FP64 workload only), else None."""
    if n != 256 or not TRAFFIC_JSON.exists():
        return None
    data = json.loads(TRAFFIC_JSON.read_text())["kernels"]
    for k, v in data.items():
        if k.removeprefix("void ").startswith(KERNEL_SYMBOL.get(slot, "?")) and "double" in k:
            return round(v["bytes_per_launch"])
    return None


def precision_study(args):
    """The precision-form case of the benchmark configurations (c3_workload):
    a prolate spheroid (semi-axes 0.3, 0.3, 0.6, 60500 faces) in a 512^3
    periodic cube, melting (St 2) with Newton-Euler coupling and remeshing,
    seeded solenoidal velocity synthesised on the GPU and HIT forcing
    regenerated every step.  FP64 and the precision form advance side by side
    from the same state; at every
    checkpoint the max-norm relative error of each field is reported over the
    whole field and away from interface cells whose tag differs (see the
    mixed leg), with the melted-volume error and the time per substep."""
    from scipy.ndimage import binary_dilation

    from melt_ibm_amd import inputs
    from melt_ibm_amd.api import MeltContext
    n = args.n if args.n != 256 else 512
    cfg, mesh, r_eq = c3_workload(n, args.study_ethresh)
    if args.study_no_remesh:
        cfg.iremesh = 0
    print(f"# building the {n}^3 initial state", file=sys.stderr, flush=True)
    u, v, w, dt = inputs.solenoidal_velocity_device(cfg, seed=0, kmax=4.0, rms=0.1)
    dt = dt * args.study_cfl / 0.5
    ctxs = []
    for prec in (0, 1):
        cfg.precision = prec
        c = MeltContext(cfg)
        c.set_geometry(mesh.xyz, mesh.vert_of_edge, mesh.edge_of_face, r_eq)
        c.inqpr()
        for name, a in (("vx", u), ("vy", v), ("vz", w)):
            c.upload_device(name, a.data_ptr())
        c.set_dt(dt)
        c.hit_init(*HIT_PARAMS, seed=3)
        c.tagCells()
        c.sync()
        ctxs.append(c)
        print(f"# precision {prec} set up: N={n} faces={mesh.edge_of_face.shape[0]} dt={dt:.4e}", file=sys.stderr, flush=True)
    del u, v, w
    qs = ("vx", "vy", "vz", "temp", "pr")
    t_run = [0.0, 0.0]
    dts = []
    checkpoints = []
    done = 0
    from melt_ibm_amd.api import MeltError
    stopped = None
    while done < args.study_steps:
        k = min(args.study_every, args.study_steps - done)
        try:
            for p_, c in enumerate(ctxs):
                c.sync()
                for i in range(k):
                    if args.study_cfl_control:
                        # the reference's CFL mode (IDTV = 1, gcurv.f90:151-157): dt
                        # from the FP64 run's CFL number before every step, the
                        # same sequence replayed in the precision form
                        if p_ == 0:
                            dts.append(c.get_dt(args.study_cfl))
                        c.set_dt(dts[done + i])
                    c.hit_force()
                    c.sync()
                    t0 = time.perf_counter()
                    c.tschem()
                    c.sync()
                    t_run[p_] += time.perf_counter() - t0
        except MeltError as e:   # a terminal condition of the reference (e.g. the body left its window)
            stopped = {"after_step": done, "precision": ("fp64", "mixed")[p_], "error": str(e)}
            break
        done += k
        f = [{q: c.download(q) for q in qs + ("VOFp",)} for c in ctxs]
        flip = np.abs(f[1]["VOFp"] - f[0]["VOFp"]) > 1e-6
        near = binary_dilation(flip, iterations=2, border_value=0) if flip.any() else flip
        cp = {"step": done, "tag_flip_cells": int(flip.sum()), "field_rel_err": {}, "field_rel_err_off_tag_flips": {}}
        for q in qs:
            d = np.abs(f[1][q] - f[0][q])
            ref = max(float(np.max(np.abs(f[0][q]))), 1e-300)
            cp["field_rel_err"][q] = float(f"{float(d.max()) / ref:.3e}")
            cp["field_rel_err_off_tag_flips"][q] = float(f"{float(np.where(near, 0.0, d).max()) / ref:.3e}")
        v0, v1 = ctxs[0].body().volume, ctxs[1].body().volume
        cp["volume_rel_err"] = float(f"{abs(v1 - v0) / v0:.3e}")
        cp["fp64_umax"] = float(f"{max(float(np.max(np.abs(f[0][q]))) for q in ('vx', 'vy', 'vz')):.4f}")
        cp["fp64_cfl"] = float(f"{ctxs[0].cfl() * ctxs[0].dt:.3f}")
        cp["dt"] = float(f"{ctxs[0].dt:.4e}")
        cp["remesh"] = {("fp64", "mixed")[p_]: {k_: c.remesh_stats()[k_] for k_ in ("events", "total_collapses")}
                        for p_, c in enumerate(ctxs)}
        checkpoints.append(cp)
        del f, flip, near
        print(f"# step {done}: {cp}", file=sys.stderr, flush=True)
    stats = ctxs[1].refinement_stats()
    for c in ctxs:
        c.close()
    sub = [t / (3 * max(done, 1)) * 1e3 for t in t_run]
    print(json.dumps({"study": "precision form vs FP64", "grid": [n, n, n], "faces": int(mesh.edge_of_face.shape[0]),
                      "forcing": f"HIT (TL, eps*, kf/kmin) = {HIT_PARAMS}, regenerated every step", "cfl": args.study_cfl,
                      "dt": ("from get_dt before every step (CFL mode), the FP64 run's sequence replayed in the precision form"
                             if args.study_cfl_control else dt),
                      "solid": f"prolate spheroid, semi-axes 0.3/0.3/0.6, icosphere frequency 55, melting St 2, Newton-Euler on, "
                               + (f"remeshing with E_thresh = {args.study_ethresh} h" if not args.study_no_remesh else "no remeshing"),
                      "steps": done, "substep_ms": {"fp64": round(sub[0], 4), "mixed": round(sub[1], 4)},
                      "speedup": round(sub[0] / sub[1], 3), "refinement": {"checks": stats[0], "refined": stats[1], "last_ratio": stats[2]},
                      "checkpoints": checkpoints, "stopped": stopped}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slot-steps", type=int, default=5, help="steps with per-slot HIP-event timing after the timed loop")
    ap.add_argument("--n", type=int, default=256, help="cells per side")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the precision-form leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ens-steps", type=int, default=100, help="timed ensemble steps")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-window", type=float, default=4.0, help="arrival window of the serving leg (s)")
    ap.add_argument("--variant-remesh", choices=("on", "off"), default="on", help="remeshing in the ensemble and serving legs")
    ap.add_argument("--serve-steps-scale", type=float, default=1.0, help="scale of the 50-500 step counts in the serving leg")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH_DB", "WRITE_DB", "COPY_FETCH_DB", "COPY_WRITE_DB"),
                    help="write the HBM-traffic summary under profiles/ from rocprofv3 --pmc databases")
    ap.add_argument("--summarize-trace", metavar="TRACE_DB", help="write the kernel-time summary under profiles/")
    ap.add_argument("--trace-leg", choices=("fp64", "mixed"), default="fp64", help="which leg of the trace to summarize")
    ap.add_argument("--command", default="", help="command line recorded in a summary")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run plain device copies of 1 GiB (calibrates the HBM byte counters of a profiler)")
    ap.add_argument("--precision-study", action="store_true",
                    help="precision form vs FP64 side by side on the 512^3 spheroid case (errors per field over --study-steps)")
    ap.add_argument("--study-steps", type=int, default=100)
    ap.add_argument("--study-every", type=int, default=20, help="steps between error checkpoints")
    ap.add_argument("--study-ethresh", type=float, default=0.38, help="PERCT_ETHRESH of the remeshing (E_thresh / h)")
    ap.add_argument("--study-cfl", type=float, default=0.5, help="CFL number (of every step, or of the fixed dt)")
    ap.add_argument("--study-no-remesh", action="store_true", help="the precision study without remeshing (IREMESH = 0)")
    ap.add_argument("--study-fixed-dt", dest="study_cfl_control", action="store_false",
                    help="keep the first step's dt instead of the reference's CFL control")
    args = ap.parse_args()
    if args.precision_study:
        precision_study(args)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, args.command, PROFILES / "hbm_traffic_256cube_melting_sphere.md", TRAFFIC_JSON)
        return
    if args.summarize_trace:
        name = "kernel_times_256cube_melting_sphere" + ("_precision_form" if args.trace_leg == "mixed" else "")
        summarize_trace(args.summarize_trace, args.command, PROFILES / f"{name}.md", leg=args.trace_leg)
        return
    if args.copy_probe:
        import torch
        # 1 GiB buffers (4x the 256 MiB last-level cache); the source is written
        # first and then pushed out of the cache by filling a third buffer
        a = torch.rand(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        c = torch.empty_like(a)
        for _ in range(5):
            c.fill_(1.0)
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": a.numel() * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs one domain per GPU on a single-GPU box; "
                                   "--gpus > 1 is not supported (multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    cfg, elapsed, tm, tm_all, qmax, body, f64, _ = run_gpu(args, 0, keep_fields=not args.no_mixed)
    cells = cfg.n1m * cfg.n2m * cfg.n3m
    ms_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed / 1e6
    substeps = 3 * args.steps
    # the dominant kernel's average launch over the timed region
    dom = DOMINANT
    avg_s = tm[dom][0] / tm[dom][1] * 1e-3
    bytes_launch = KERNEL_BYTES[dom] * cells
    achieved = bytes_launch / avg_s
    sub_time = elapsed / substeps
    out = {
        "metric": "RK3 Mcell-steps/s @256^3 FP64 vs mixed (HBM roofline %); ensemble p99 step ms",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded solenoidal velocity |k|<=4 rms 0.1, stochastic HIT forcing |k|<=2.5 regenerated every step, "
                "inqpr temperature, level-5 icosphere)",
        "config": {"workload": f"{cfg.n1m}^3 periodic 2pi cube, HIT forcing, melting sphere R=0.5 (level-5 icosphere, 20480 faces), "
                               f"St=2, Newton-Euler on, Re=1000, FP64",
                   "grid": [cfg.n1m, cfg.n2m, cfg.n3m], "faces": 20480, "precision": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": measured_traffic(dom, args.n),
                     "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_s * 1e6, 2),
                     "substep_frac": round(B_ALG_CELL * cells / sub_time / HBM_PEAK, 4),
                     "substep_ms": round(sub_time * 1e3, 4)},
        "kernels_ms_per_substep": {k: round(v[0] / max(v[1], 1), 4) for k, v in tm_all.items() if v[1] > 0},
        "kernels_sample": f"HIP events around every slot over {args.slot_steps} further steps (outside the timed loop)",
        "checks": {"divgck": qmax, "volume": body.volume},
    }
    if not args.no_mixed:
        _, el32, _, tm32, q32, body32, f32, stats = run_gpu(args, 1, keep_fields=True)
        sub32 = el32 / substeps
        qs = ("vx", "vy", "vz", "temp", "pr")
        errs = {q: float(np.max(np.abs(f32[q] - f64[q])) / max(np.max(np.abs(f64[q])), 1e-300)) for q in qs}
        # the same away from interface cells whose narrow-band tag differs between
        # the two runs (and their 5x5x5 neighbourhood): the nearest-face sign rule
        # of the tagging is discontinuous at triangle edges, so geometries that
        # differ at round-off can resolve a near-tie differently
        from scipy.ndimage import binary_dilation
        flip = np.abs(f32["VOFp"] - f64["VOFp"]) > 1e-6
        near = binary_dilation(flip, iterations=2, border_value=0) if flip.any() else flip
        errs_off = {q: float(np.max(np.where(near, 0.0, np.abs(f32[q] - f64[q]))) / max(np.max(np.abs(f64[q])), 1e-300)) for q in qs}
        n_flip = int(flip.sum())
        del f32, f64, flip, near
        out["mixed"] = {
            "value": round(cells * args.steps / el32 / 1e6, 3), "unit": "Mcell-steps/s", "dtype": "f32 storage, f64 line reductions / sums / Lagrangian",
            "ms_per_step": round(el32 / args.steps * 1e3, 4), "speedup_vs_fp64": round(elapsed / el32, 3),
            "substep_ms": round(sub32 * 1e3, 4), "substep_frac": round(B_ALG_CELL_MIXED * cells / sub32 / HBM_PEAK, 4),
            "bytes_per_cell_substep": B_ALG_CELL_MIXED,
            "field_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
            "field_rel_err_off_tag_flips": {k: float(f"{v:.3e}") for k, v in errs_off.items()},
            "tag_flip_cells": n_flip,
            "volume_rel_err_vs_fp64": float(f"{abs(body32.volume - body.volume) / body.volume:.3e}"),
            "divgck": q32,
            "refinement": {"checks": stats[0], "refined": stats[1], "last_ratio": stats[2]},
            "kernels_ms_per_substep": {k: round(v[0] / max(v[1], 1), 4) for k, v in tm32.items() if v[1] > 0},
            "error_sample": f"both forms from the same state after {args.warmup + args.steps} RK3 steps",
        }
    if not args.no_ensemble:
        # batched form: 48 melting simulations of uneven grids, graph-replayed
        from melt_ibm_amd.ensemble import Ensemble, ensemble_members
        print("# ensemble: building 48 simulations", file=sys.stderr, flush=True)
        ens = Ensemble(ensemble_members(48, seed=0, remesh=args.variant_remesh == "on"), precision=1, tune=True)
        st = ens.run(steps=args.ens_steps, warmup=2)
        ens.close()
        del ens
        st["workload"] = ("48 independent melting spheres (R 0.3-0.7, level-3/4 icosahedra), grids uniform from {48,64,96,128}^3, "
                          "St in [0.5,4], Re in [200,1000], seeded; FP32 bulk, FP64 Lagrangian; one graph replay per simulation "
                          "and step on its own stream")
        out["ensemble"] = st
        print(f"# ensemble: p50 {st['step_ms_p50']:.2f} ms, p99 {st['step_ms_p99']:.2f} ms", file=sys.stderr, flush=True)
    if not args.no_serve:
        from melt_ibm_amd.serve import Server, request_stream
        reqs = request_stream(args.serve_window, rate=4.0, seed=7, steps_scale=args.serve_steps_scale,
                              remesh=args.variant_remesh == "on")
        print(f"# serving: {len(reqs)} requests over {args.serve_window} s", file=sys.stderr, flush=True)
        rep = Server(200e9, precision=1, tune=False).run(reqs, log=sys.stderr)
        scale = "" if args.serve_steps_scale == 1.0 else f" x {args.serve_steps_scale}"
        rep["workload"] = (f"Poisson arrivals (4/s, seed 7) over {args.serve_window} s; grids from {{64,128,256,384}}^3; spheres or "
                           f"spheroids (aspect 1-3, 2k-60k faces); St in [0.5,4]; 50-500 RK3 steps{scale}; 200 GB budget; "
                           f"FP32 bulk, FP64 Lagrangian; initial states synthesised on the GPU")
        out["serving"] = rep
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.n)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
