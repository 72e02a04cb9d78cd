"""Benchmark of the real-space Kohn–Sham SCF hot path on one MI355X.

One step = one cold SCF of water (H2O, O at the origin, O-H 0.96 A, H-O-H 104.5 deg; the O pseudised by the generated
pseudopotential filtered for the grid, the two H all-electron: 8 valence electrons, 4 + 4 bands plus Fermi-Dirac
witnesses) on a 95^3 grid at h = 0.15 Bohr, LSDA, Fermi–Dirac occupations at kT = 5 mHa, converged to
delta_rho < 1e-6 in MIXED precision (f32 grid fields, f64 Rayleigh–Ritz / DIIS / energies). The same SCF in FP64 is
timed once for the mixed-vs-FP64 ratio and the per-result errors; the all-electron water of the same grid is timed
beside it. The roofline object is the block stencil H.X on 192^3 x 64 bands (f32, V ~ U[-5, 0] Ha,
X ~ N(0, 1)), timed with HIP events on the stream it runs on. The CPU baseline times the f64 oracle (a restatement of
the reference's single-threaded CPU solver) on a bounded sample of the same water problem.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "atomic-sandbox_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

BOHR_PER_ANGSTROM = 1.8897261  # scenes/aspirin.scene:3
METRIC = "SCF s to 1e-6 Ha (water 95³) mixed vs FP64 on 1 GPU; H·X GB/s % of HBM peak"
HBM_PEAK_GBS = 8000.0


def water():
    """All-electron water: positions, Z, electrons per spin."""
    r = 0.96 * BOHR_PER_ANGSTROM
    half = math.radians(104.5 / 2)
    pos = [[0.0, 0.0, 0.0], [r * math.sin(half), r * math.cos(half), 0.0], [-r * math.sin(half), r * math.cos(half), 0.0]]
    return pos, [8, 1, 1], 5, 5


def hx_roofline(torch, lira, n=192, nb=64, reps=20, dtype="float32", variant=None):
    """Times lira_hamiltonian on n^3 x nb bands with HIP events on torch's current stream (the tile variant the
    autotuner picks, or the given one)."""
    dt = getattr(torch, dtype)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((nb, n ** 3), generator=g, device="cuda", dtype=torch.float32).to(dt)
    v = (torch.rand(n ** 3, generator=g, device="cuda", dtype=torch.float32) * -5.0).to(dt)
    out = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    esize = x.element_size()
    if variant is None:
        tuned, _ = lira.autotune_stencil(esize, n, nb)
    else:
        tuned = variant
        lira.lib().lira_set_stencil_variant(esize, variant)
    for _ in range(3):
        lira.hamiltonian(n, 0.2, v, x, out, nb, stream=sp)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        lira.hamiltonian(n, 0.2, v, x, out, nb, stream=sp)
    stop.record(stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    # the practical ceiling of a read + write kernel on this device: a plain device copy of the same block
    for _ in range(2):
        out.copy_(x)
    start.record(stream)
    for _ in range(reps):
        out.copy_(x)
    stop.record(stream)
    stop.synchronize()
    torch_copy_gbs = 2 * x.numel() * esize / (start.elapsed_time(stop) / reps * 1e-3) / 1e9
    # the hand-written streaming kernels on the same buffers (16-byte vector loads and stores, plain and nontemporal
    # stores): the device's read, write and copy ceilings on this box
    nbytes = x.numel() * esize
    stream_gbs = {"copy": lira.stream_bandwidth(lira.STREAM_COPY, x, out, nbytes, reps, stream=sp),
                  "copy_nontemporal": lira.stream_bandwidth(lira.STREAM_COPY, x, out, nbytes, reps, True, stream=sp),
                  "read": lira.stream_bandwidth(lira.STREAM_READ, x, None, nbytes, reps, stream=sp),
                  "write": lira.stream_bandwidth(lira.STREAM_WRITE, None, out, nbytes, reps, stream=sp),
                  "write_nontemporal": lira.stream_bandwidth(lira.STREAM_WRITE, None, out, nbytes, reps, True,
                                                             stream=sp),
                  "torch_copy": torch_copy_gbs}
    copy_gbs = max(stream_gbs["copy"], stream_gbs["copy_nontemporal"], torch_copy_gbs)
    N = n ** 3
    algo_bytes = N * (2 * esize * nb + esize)  # x in, HX out per band, V once per block
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    del x, v, out
    torch.cuda.empty_cache()
    return {"ms": ms, "bytes": algo_bytes, "gbs": achieved, "variant": lira.stencil_variants()[tuned],
            "variant_index": tuned, "frac": achieved / HBM_PEAK_GBS, "copy_gbs": copy_gbs, "stream_gbs": stream_gbs}


def poisson_256(lira, n=256, h=0.15, seed=3):
    """Poisson MG-PCG on 256^3: a neutral density of 32 seeded normalised Gaussians (centres in the inner 60 % of the
    box, sigma U[0.3, 1.0] A, charges +1 x 16 and -1 x 16), solved to a relative residual of 1e-8 in MIXED precision.
    Time of the device solve (the density upload excluded), after one untimed solve that builds the levels."""
    import workloads
    rho = workloads.neutral_blobs(n, h, seed)
    s = lira.GpuSolver(n, h, lira.MIXED)
    s.poisson(rho, tol=1e-8)
    _, st = s.poisson(rho, tol=1e-8)
    it = max(st.iterations, 1)
    # 162 algorithmic bytes per point per MG-PCG iteration in f32 (stencil, V-cycle, dots, axpys)
    gbs = 162.0 * n ** 3 * it / st.seconds / 1e9 if st.seconds > 0 else None
    return {"seconds": st.seconds, "pcg_iterations": st.iterations, "corrections": st.outer, "residual": st.residual,
            "grid": n, "h_bohr": h, "gbs_algorithmic": gbs}


def pseudo_species(lira, h):
    """The ionic setup of the headline water: the O generated and filtered for h (lira_pseudo_generate: Z_v = 6, s
    projector, p local, LFC partial core); H stays all-electron."""
    return [lira.generate_pseudo(8, h)]


def water_pair(lira, n=95, h=0.15, pseudised=False):
    """A cold MIXED and a cold FP64 solve of water (each after an untimed one), Fermi-Dirac kT 5 mHa,
    delta_rho < 1e-6: all-electron (5 + 5 electrons), or with the O pseudised (4 + 4)."""
    pos, z, nu, nd = water()
    settings = lira.default_settings(h, lira.FERMI_DIRAC)
    species = pseudo_species(lira, h) if pseudised else []
    if pseudised:
        nu, nd = 4, 4
    out = {"o_table": {k: species[0].record[k] for k in ("f", "local", "z_v", "r_nlc", "r_core")}} if pseudised else {}
    for name, prec in (("mixed", lira.MIXED), ("fp64", lira.FP64)):
        s = lira.GpuSolver(n, h, prec)
        if species:
            s.set_ionic(species)
        s.solve(pos, z, nu, nd, settings)
        t0 = time.perf_counter()
        st = s.solve(pos, z, nu, nd, settings)
        out[name] = {"seconds": time.perf_counter() - t0, "steps": st.iterations, "energy": st.energy.total,
                     "lobpcg_iterations": st.lobpcg_iterations}
        out[name + "_forces"] = s.forces(pos, z, h)
        del s
    out["dE"] = out["mixed"]["energy"] - out["fp64"]["energy"]
    out["max_dF"] = float(np.abs(out.pop("mixed_forces") - out.pop("fp64_forces")).max())
    out["speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    return out


def batch_scf(lira, count=256, workers=4, seed=0, seq_count=32):
    """Batched SCF: `count` all-electron molecules drawn from H2, LiH, H2O, NH3, CH4 (equilibrium geometry plus a seeded
    uniform +-0.1 A per nucleus and coordinate), each on an n^3 grid with n drawn from 48, 56, ..., 96 at h = 0.25 Bohr,
    MIXED precision, cold SCF (delta_rho < 1e-6, Fermi-Dirac kT 5 mHa) + forces, over `workers` host threads that
    each own a solver. An untimed pass over one molecule of every (species, grid) pair first runs the one-time stencil
    tuning. Reports converged SCFs/s of the whole batch, p50/p99 of the per-molecule time (its solve and forces on its
    worker, plus a solver rebuild when the worker's grid changed), and the same first `seq_count` molecules run one at
    a time (one worker) for comparison."""
    import molecules as M
    mols = M.sample(count, seed=seed)
    seen, warm = set(), []
    for m in mols:
        if (m.name, m.n) not in seen:
            seen.add((m.name, m.n))
            warm.append(m)
    lira.solve_batch(warm, precision=lira.MIXED, workers=workers)
    t0 = time.perf_counter()
    out = lira.solve_batch(mols, precision=lira.MIXED, workers=workers)
    wall = time.perf_counter() - t0
    ok = [o for o in out if o.status == lira.OK]
    per = np.array([o.end - o.start for o in ok])
    names = sorted({m.name for m in mols})
    res = {"molecules": count, "workers": workers, "converged": len(ok), "seconds": wall,
            "scf_per_s": len(ok) / wall, "p50_s": float(np.percentile(per, 50)) if len(per) else None,
            "p99_s": float(np.percentile(per, 99)) if len(per) else None,
            "steps_mean": float(np.mean([o.stats.iterations for o in ok])) if ok else None,
            "mix": {k: sum(m.name == k for m in mols) for k in names},
            "grids": sorted({m.n for m in mols}), "h_bohr": mols[0].h}
    if seq_count > 0:
        t1 = time.perf_counter()
        seq = lira.solve_batch(mols[:seq_count], precision=lira.MIXED, workers=1)
        seq_wall = time.perf_counter() - t1
        t2 = time.perf_counter()
        lira.solve_batch(mols[:seq_count], precision=lira.MIXED, workers=workers)
        sub_wall = time.perf_counter() - t2
        res["one_at_a_time"] = {"molecules": seq_count, "seconds": seq_wall,
                                "scf_per_s": sum(o.status == lira.OK for o in seq) / seq_wall,
                                "batched_seconds_same_set": sub_wall, "speedup": seq_wall / sub_wall}
    return res


def aspirin_leg(lira, n=160, h=0.2, seed=0, fd_components=3, max_iter=80):
    """Aspirin C9H8O4 (the reference's scene geometry with seeded +-0.05 A per nucleus and coordinate): C and O
    pseudised by the generated tables filtered for h, H all-electron, 68 valence electrons (34 + 34), 160^3 at
    h = 0.2 Bohr, MIXED, Fermi-Dirac kT 5 mHa, the SCF tolerance raised to the f32 floor of 68 electrons as the
    reference's provider does (max(1e-6, 2 * 5e-8 * 68) = 6.8e-6). Cold SCF time, steps and block size, then the
    Hellmann-Feynman forces (the hand-written dE/dR kernels) of the three largest force components against central
    differences of the converged FP64 energy at +-1e-3 A (an FP64 cold solve, then two warm solves per component),
    for the FP64 and the mixed HF forces."""
    import molecules as M
    pos, z = M.aspirin(seed)
    s = lira.GpuSolver(n, h, lira.MIXED)
    species = [lira.generate_pseudo(6, h), lira.generate_pseudo(8, h)]
    s.set_ionic(species)
    tol = lira.effective_gpu_tol(1e-6, 68)
    st = lira.default_settings(h, lira.FERMI_DIRAC, tol=tol, max_iter=max_iter)
    steps = [0]

    def progress():  # polled before each SCF step: a progress line per 10 steps, never an abort
        steps[0] += 1
        if steps[0] % 10 == 0:
            print(f"# aspirin: step {steps[0]}, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        return False

    t0 = time.perf_counter()
    r = s.solve(pos, z, 34, 34, st, abort=progress)
    wall = time.perf_counter() - t0
    print(f"# aspirin 160^3: {wall:.2f} s, {r.iterations} steps, E {r.energy.total:.8f}, block {list(r.n_stored)}, "
          f"lobpcg {r.lobpcg_iterations}", file=sys.stderr, flush=True)
    f = s.forces(pos, z, h)
    out = {"seconds": wall, "steps": r.iterations, "energy": r.energy.total, "tol": tol,
           "lobpcg_iterations": r.lobpcg_iterations, "stored": list(r.n_stored), "grid": n, "h_bohr": h,
           "valence_electrons": 68, "max_force": float(np.abs(f).max())}
    # the force check on the FP64 path (a difference of mixed energies carries the f32 solve's energy noise, about
    # 1e-7 Ha over the 3.8e-3 Bohr step): an FP64 cold solve, its HF forces, and per component two warm FP64 solves at
    # +-1e-3 A; the mixed HF force is compared with the same difference
    del s
    s64 = lira.GpuSolver(n, h, lira.FP64)
    s64.set_ionic(species)
    t1 = time.perf_counter()
    r64 = s64.solve(pos, z, 34, 34, st, abort=progress)
    out["fp64"] = {"seconds": time.perf_counter() - t1, "steps": r64.iterations, "energy": r64.energy.total}
    f64 = s64.forces(pos, z, h)
    d = 1e-3 * BOHR_PER_ANGSTROM
    checks = []
    for flat in np.argsort(-np.abs(f64).ravel())[:fd_components]:
        atom, ax = divmod(int(flat), 3)
        e = []
        for sgn in (1.0, -1.0):
            p = pos.copy()
            p[atom, ax] += sgn * d
            rr = s64.solve(p, z, 34, 34, st, warm=True, abort=progress)
            e.append(rr.energy.total)
        fd = -(e[0] - e[1]) / (2 * d)
        checks.append({"nucleus": atom, "z": z[atom], "axis": int(ax), "hf_fp64": float(f64[atom, ax]),
                       "hf_mixed": float(f[atom, ax]), "fd_fp64": fd, "error": abs(float(f64[atom, ax]) - fd),
                       "error_mixed": abs(float(f[atom, ax]) - fd)})
        print(f"# aspirin force check: nucleus {atom} (Z {z[atom]}) axis {ax}: HF FP64 {f64[atom, ax]:+.7f}, mixed "
              f"{f[atom, ax]:+.7f}, FD FP64 {fd:+.7f} Ha/Bohr", file=sys.stderr, flush=True)
    out["force_checks"] = checks
    out["max_component_error"] = max(c["error"] for c in checks)
    out["max_component_error_mixed"] = max(c["error_mixed"] for c in checks)
    return out


def traffic_probe(esize, variant, n=192, nb=64):
    """Child of hbm_traffic, run under rocprofv3 --pmc: two copies of a known size (the counter calibration), then
    two H.X launches of the given stencil variant on the roofline problem."""
    import torch
    torch.cuda.set_device(0)
    import lira
    dt = torch.float32 if esize == 4 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((nb, n ** 3), generator=g, device="cuda", dtype=torch.float32).to(dt)
    v = (torch.rand(n ** 3, generator=g, device="cuda", dtype=torch.float32) * -5.0).to(dt)
    out = torch.empty_like(x)
    sp = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        out.copy_(x)
    torch.cuda.synchronize()
    lira.lib().lira_set_stencil_variant(esize, variant)
    for _ in range(2):
        lira.hamiltonian(n, 0.2, v, x, out, nb, stream=sp)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": x.numel() * esize}), flush=True)


def hbm_traffic(esize, variant, n=192, nb=64):
    """HBM bytes per H.X launch from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each
    scaled by what the same counter reads for a device copy of known size in the same process. None if rocprofv3 is
    absent or a pass fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    copy_bytes = x_bytes = n ** 3 * nb * esize
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="lira_pmc_")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--traffic-probe", str(esize), str(variant)]
        try:
            subprocess.run(cmd, timeout=90, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=False)
        except subprocess.TimeoutExpired:
            shutil.rmtree(d, ignore_errors=True)
            return None
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f, newline="") as fh:
                rows += [r for r in csv.DictReader(fh) if r.get("Counter_Name") == counter]
        shutil.rmtree(d, ignore_errors=True)
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        sten = [float(r["Counter_Value"]) for r in rows if "lira::k_stencil" in r["Kernel_Name"]]
        if not copy or not sten:
            return None
        kib_copy = sum(copy) / len(copy)
        kib_sten = sum(sten) / len(sten)
        # a copy reads and writes copy_bytes: the counter's scale is copy_bytes / (KiB counted * 1024)
        factor = copy_bytes / (kib_copy * 1024.0)
        per[counter] = {"kib_per_dispatch": kib_sten, "copy_kib": kib_copy, "factor": factor,
                        "bytes": kib_sten * 1024.0 * factor}
    del x_bytes
    return {"bytes": per["FETCH_SIZE"]["bytes"] + per["WRITE_SIZE"]["bytes"], "fetch": per["FETCH_SIZE"],
            "write": per["WRITE_SIZE"]}


def cpu_share():
    """The host threads this process may use: the OpenMP setting the GPU box exports (its CPU share per GPU), else the
    CPU affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(state, n, h, pos, z, species, fp64_counts):
    """The oracle (the C f64 restatement of the reference's single-threaded CPU solver, crates/qm) on this host: a
    bounded sample of the same water problem, one f64 Poisson solve of the converged density to the SCF's Poisson
    tolerance and LOBPCG (with the pseudised O's nonlocal operator) at the oracle's own block size on the converged
    potential (setup alone, then two iterations), timed on 1 thread (the reference's CPU solver) and on the host's CPU
    share, and scaled to one whole SCF with the oracle's own counts on this workload
    (tests/golden/oracle_pseudo_water95_counts.json: its steps, LOBPCG iterations and Poisson CG iterations, from one
    full oracle solve; when that file is absent, the FP64 GPU run's counts, and the sample says so)."""
    import oracle as O
    path = os.path.join(ROOT, "tests", "golden", "oracle_pseudo_water95_counts.json")
    if os.path.exists(path):
        with open(path) as fh:
            counts = json.load(fh)
        source = "the oracle's own full solve"
    else:
        counts = dict(fp64_counts)
        source = "the FP64 GPU run's counts (the oracle's full-solve counts are not recorded)"
    counts["source"] = source
    rho, vh, vnuc = state
    v = vnuc + vh
    k = int(counts["k_up"])
    guess = O.initial_orbitals(n, h, pos, k, 1)
    nl = O.Nonlocal(n, h, pos, z, species)
    runs = {}
    share = cpu_share()
    for label, threads in (("threads_1", 1), ("threads_all", share)):
        O.set_threads(threads)
        t0 = time.perf_counter()
        _, it_p, _ = O.poisson_solve(n, h, rho, tol=1e-10, max_iter=5000)
        t_poisson = time.perf_counter() - t0
        t0 = time.perf_counter()
        nl.lobpcg(v, guess, 1e-30, 0, precond_lambda=-1.0)
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        nl.lobpcg(v, guess, 1e-30, 2, precond_lambda=-1.0)
        t_iter = max(time.perf_counter() - t0 - t_setup, 0.0) / 2
        per_cg = t_poisson / max(it_p, 1)
        calls = 2 * counts["steps"]  # one eigensolve per spin channel and step
        est = per_cg * counts["cg_iterations"] + t_iter * counts["lobpcg_iterations"] + t_setup * calls
        runs[label] = {"seconds_per_scf": est, "threads": threads, "poisson_seconds_per_cg_iteration": per_cg,
                       "lobpcg_seconds_per_iteration": t_iter, "lobpcg_setup_seconds": t_setup,
                       "sample_seconds": t_poisson + t_setup * 2 + t_iter * 2}
        print(f"# cpu baseline {label} ({threads} threads): {est:.1f} s per SCF (Poisson {per_cg * 1e3:.1f} ms/CG it, "
              f"LOBPCG {t_iter:.2f} s/it)", file=sys.stderr, flush=True)
    O.set_threads(share)
    top = runs["threads_all"]
    return {"value": top["seconds_per_scf"], "unit": "s", "cores": share, "kind": "port",
            "sample": f"oracle (C f64 restatement of crates/qm) on the same pseudised water 95^3 problem: one Poisson solve and "
                      f"LOBPCG setup + 2 iterations at k = {k}, timed on 1 and {share} threads, scaled with {source} "
                      f"({counts['steps']} SCF steps, {counts['lobpcg_iterations']} LOBPCG iterations, "
                      f"{counts['cg_iterations']} CG iterations)",
            "threads_1": runs["threads_1"], "threads_all": runs["threads_all"], "cpu_share": share,
            "nproc": os.cpu_count(), "oracle_counts": counts}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the all-electron water and Poisson 256^3 measurements")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc passes")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched small-molecule SCF measurement")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-workers", type=int, default=4)
    ap.add_argument("--batch-only", action="store_true", help="only the batched SCF, without the one-at-a-time run")
    ap.add_argument("--aspirin", action="store_true", help="add the 160^3 aspirin SCF and its force check")
    ap.add_argument("--aspirin-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--roofline-only", action="store_true", help="only the H.X roofline measurement (for profiling)")
    ap.add_argument("--poisson-only", action="store_true", help="only the Poisson 256^3 measurement (for profiling)")
    ap.add_argument("--pseudo-only", action="store_true", help="only the pseudised-water measurement")
    ap.add_argument("--variant", type=int, default=None, help="stencil tile variant of the roofline (default: tuned)")
    ap.add_argument("--traffic-probe", nargs=2, type=int, metavar=("ESIZE", "VARIANT"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(*args.traffic_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the SCF on a single MI355X (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    import lira

    if args.aspirin_only:
        print(json.dumps({"aspirin_only": True, **aspirin_leg(lira)}))
        return
    if args.batch_only:
        b = batch_scf(lira, count=args.batch_count, workers=args.batch_workers, seq_count=0)
        print(json.dumps({"batch_only": True, **b}))
        return
    if args.poisson_only:
        print(json.dumps({"poisson_only": True, **poisson_256(lira)}))
        return
    if args.pseudo_only:
        print(json.dumps({"pseudo_only": True, **water_pair(lira, pseudised=True)}, default=float))
        return
    if args.roofline_only:
        r = hx_roofline(torch, lira, variant=args.variant)
        print(json.dumps({"roofline_only": True, "ms_per_launch": r["ms"], "achieved_gbs": r["gbs"],
                          "variant": r["variant"], "variant_index": r["variant_index"], "copy_gbs": r["copy_gbs"],
                          "stream_gbs": r["stream_gbs"]}))
        return
    n, h = 95, 0.15
    pos, z, _, _ = water()
    nu, nd = 4, 4  # valence electrons per spin: 6 of the pseudised O, 1 + 1 of the H
    species = pseudo_species(lira, h)
    settings = lira.default_settings(h, lira.FERMI_DIRAC)
    solver = lira.GpuSolver(n, h, lira.MIXED)
    solver.set_ionic(species)
    for _ in range(args.warmup):
        solver.solve(pos, z, nu, nd, settings)
    torch.cuda.synchronize()
    times, stats = [], None
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = solver.solve(pos, z, nu, nd, settings)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        print(f"# mixed SCF: {times[-1]:.3f} s, {stats.iterations} steps, E {stats.energy.total:.8f}, "
              f"lobpcg {stats.lobpcg_iterations}, cg {stats.cg_iterations}, readbacks {stats.readbacks}",
              file=sys.stderr, flush=True)
    mixed_s = float(np.mean(times))
    f_mixed = solver.forces(pos, z, h)
    ev_mixed, occ_mixed = solver.levels(0)
    result = {"metric": METRIC, "value": mixed_s, "unit": "s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": mixed_s * 1e3, "higher_is_better": False, "scaling": "strong",
              "vs_baseline": None, "dtype": "f32 grid / f64 control", "data": "synthetic",
              "config": {"workload": "water H2O, O pseudised by the generated pseudopotential (f = 1.4, p local, LFC "
                                     "partial core, filtered for h) and H all-electron, 95^3 grid at h = 0.15 Bohr, "
                                     "LSDA, 4 + 4 bands plus witnesses, FD kT 5 mHa, delta_rho < 1e-6",
                         "grid": n, "h_bohr": h, "nuclei": 3, "valence_electrons": nu + nd,
                         "o_table": {k: species[0].record[k] for k in ("f", "local", "z_v", "r_nlc", "r_core")},
                         "precision": "mixed", "parallelism": "single GPU"},
              "mixed": {"seconds": mixed_s, "steps": stats.iterations, "energy": stats.energy.total,
                        "lobpcg_iterations": stats.lobpcg_iterations, "cg_iterations": stats.cg_iterations,
                        "readbacks": stats.readbacks, "seconds_poisson": stats.seconds_poisson,
                        "seconds_eigen": stats.seconds_eigen, "stored": list(stats.n_stored)}}
    gpu_fp64 = None
    if not args.no_fp64:
        s64 = lira.GpuSolver(n, h, lira.FP64)
        s64.set_ionic(species)
        for _ in range(args.warmup):  # as the mixed solver: the one-time stencil tuning stays out of the timing
            s64.solve(pos, z, nu, nd, settings)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st64 = s64.solve(pos, z, nu, nd, settings)
        torch.cuda.synchronize()
        fp64_s = time.perf_counter() - t0
        f64 = s64.forces(pos, z, h)
        ev64, occ64 = s64.levels(0)
        kk = min(len(ev64), len(ev_mixed))
        # the levels that hold charge in either run (occupation >= 1e-6, the partial-level floor); the stored block's
        # top columns beyond them are witnesses and the untested buffer column, converged to a looser bound or not at all
        held = [i for i in range(kk) if max(occ_mixed[i], occ64[i]) >= 1e-6]
        gpu_fp64 = {"seconds": fp64_s, "steps": st64.iterations, "energy": st64.energy.total,
                    "lobpcg_iterations": st64.lobpcg_iterations, "cg_iterations": st64.cg_iterations,
                    "k": int(st64.n_stored[0])}
        result["fp64"] = gpu_fp64
        result["mixed_vs_fp64"] = {"speedup": fp64_s / mixed_s, "dE": stats.energy.total - st64.energy.total,
                                   "max_deps": float(np.abs(ev_mixed[held] - ev64[held]).max()),
                                   "levels_compared": len(held),
                                   "levels_mixed": [float(e) for e in ev_mixed[:kk]],
                                   "levels_fp64": [float(e) for e in ev64[:kk]],
                                   "max_dF": float(np.abs(f_mixed - f64).max())}
        print(f"# fp64 SCF: {fp64_s:.3f} s, {st64.iterations} steps, E {st64.energy.total:.8f}", file=sys.stderr,
              flush=True)
        state = (s64.field(2), s64.field(3), s64.field(4))
        del s64
    del solver
    torch.cuda.empty_cache()
    if not args.no_roofline:
        r = hx_roofline(torch, lira, variant=args.variant)
        traffic = None if args.no_traffic else hbm_traffic(4, r["variant_index"])
        result["roofline"] = {"bound": "hbm", "achieved": r["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": r["frac"], "traffic": None if traffic is None else traffic["bytes"],
                              "kernel": "lira::k_stencil* (ST_HAM, f32) H.X on 192^3 x 64 bands",
                              "ms_per_launch": r["ms"], "algorithmic_bytes": r["bytes"], "variant": r["variant"],
                              "copy_gbs": r["copy_gbs"], "frac_of_copy": r["gbs"] / r["copy_gbs"],
                              "stream_gbs": r["stream_gbs"],
                              "traffic_detail": traffic}
        r64 = hx_roofline(torch, lira, dtype="float64")
        result["roofline_fp64"] = {"achieved": r64["gbs"], "frac": r64["frac"], "ms_per_launch": r64["ms"],
                                   "copy_gbs": r64["copy_gbs"]}
        print(f"# H.X 192^3 x 64 f32: {r['ms']:.3f} ms, {r['gbs']:.0f} GB/s ({100 * r['frac']:.1f} % of peak); "
              f"f64 {r64['gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    if not args.no_extra:
        result["water_all_electron"] = water_pair(lira)
        print(f"# all-electron water: mixed {result['water_all_electron']['mixed']['seconds']:.3f} s,"
              f" fp64 {result['water_all_electron']['fp64']['seconds']:.3f} s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
        result["poisson_256"] = poisson_256(lira)
        print(f"# Poisson 256^3 to 1e-8: {result['poisson_256']['seconds'] * 1e3:.1f} ms, "
              f"{result['poisson_256']['pcg_iterations']} PCG iterations", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_batch:
        b = batch_scf(lira, count=args.batch_count, workers=args.batch_workers)
        result["batch"] = b
        print(f"# batch of {b['molecules']} molecules, {b['workers']} workers: {b['seconds']:.2f} s, "
              f"{b['scf_per_s']:.1f} SCF/s, p50 {b['p50_s'] * 1e3:.0f} ms, p99 {b['p99_s'] * 1e3:.0f} ms; "
              f"one at a time {b['one_at_a_time']['scf_per_s']:.1f} SCF/s "
              f"(x{b['one_at_a_time']['speedup']:.2f} batched on the same {b['one_at_a_time']['molecules']})",
              file=sys.stderr, flush=True)
    if args.aspirin:
        torch.cuda.empty_cache()
        result["aspirin"] = aspirin_leg(lira)
    if not args.no_cpu and gpu_fp64 is not None:
        result["cpu_baseline"] = cpu_baseline(state, n, h, pos, z, species, {
            "steps": gpu_fp64["steps"], "lobpcg_iterations": gpu_fp64["lobpcg_iterations"],
            "cg_iterations": gpu_fp64["cg_iterations"], "k_up": gpu_fp64["k"]})
    print(json.dumps(result))


if __name__ == "__main__":
    main()
