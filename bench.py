"""Benchmark of the suzerain spectral-DNS substep hot path on one MI355X.

Default workload (--workload rk3): one full FP64 SMR91 low-storage RK3 step
(3 substeps) of the compressible channel at Nx x Ny x Nz = 96 x 96 x 96 with
a 144 x 144 dealiased transform grid, k = 8 B-splines on 90 tanh-stretched
breakpoints (htdelta 3), Lx = 4 pi, Ly = 2, Lz = 4 pi/3, gamma 1.4, Re 3000,
Ma 1.5, Pr 0.7, isothermal walls; state = isothermal laminar profile plus
seeded Fourier perturbations of amplitude 1e-2 |k|^-2 for |k| <= 8.  Each
substep runs the explicit operator (41 derivative fields, 46 backward and 4
forward dealiased transforms, references, pointwise flux, step candidates on
the first substep) and the hybrid implicit operator (apply/accumulate
M + phi L, assembly with isothermal walls, batched banded LU and solve for
the 4,560 retained (kx, kz) systems).  The state is resident in HBM.

--workload solve: the batched FP64 implicit solve alone (S=5, n=96, k=8,
N = 480, KL = KU = 34, 1152 (kx, kz) systems, one right hand side each).

Prints one JSON line.  `roofline` reports the dominant kernel, the batched
LU factorization, timed with HIP events on the stream it runs on;
`cpu_baseline` times the oracle's restatement of the reference CPU path on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=("rk3", "solve"), default="rk3")
    p.add_argument("--steps", type=int, default=40)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    p.add_argument("--pmc-child", choices=("rk3",), default=None,
                   help=argparse.SUPPRESS)
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device-to-device copy of MIB MiB "
                        "three times (reference reading for HBM traffic "
                        "counters) and exit")
    return p.parse_args()


def copy_calibration(mib):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings give
    the factor between counter units and bytes moved."""
    import torch
    nbytes = mib << 20
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 3}),
          flush=True)


def host_threads():
    """(threads used, cores the OS reports): worker pools stay within the
    16-core host share of one GPU whatever the machine reports."""
    n = os.cpu_count() or 1
    return max(1, min(16, n)), n


def cpu_baseline(inp, op, nsample, threads):
    """The reference CPU path of one system -- pack (oracle's restatement of
    bsmbsm_pack.def), permute, LAPACK ZGBTRF + ZGBTRS (scipy's OpenBLAS,
    the reference's MKL calls), un-permute -- over `nsample` systems on
    `threads` host threads (each system single-threaded, as the reference's
    ranks work through their own wavenumbers); returns (solves/s, seconds)."""
    import concurrent.futures
    import oracle
    import oracle.model as om
    from oracle import lapack
    from oracle.bspline import band_of
    S, n = inp["S"], inp["n"]
    kl = max(op.ku[:3])
    ku = max(op.kl[:3])
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    N = S * n
    ldab = 2 * KL + KU + 1
    M, D1, D2 = op.dense(0), op.dense(1), op.dense(2)
    idx = np.linspace(0, len(inp["kx"]) - 1, nsample).astype(int)
    # submatrix bands prepared outside the timed region
    bands = {}
    for s in idx:
        bands[s] = [[band_of(om.submatrix(ih, jh, M, D1, D2, inp["kx"][s],
                                          inp["kz"][s], inp["cij"], inp["uij"],
                                          inp["phidt"], inp["nu"]),
                             kl, ku).reshape(-1) for jh in range(S)]
                    for ih in range(S)]

    def one(s):
        ab = np.zeros(N * ldab, dtype=np.complex128)
        view = ab[KL:]
        for ih in range(S):
            for jh in range(S):
                oracle.zpack(S, n, ih, jh, kl, ku, 1.0, bands[s][ih][jh],
                             kl + 1 + ku, KL, KU, view, ldab)
        pb = np.zeros(N, dtype=np.complex128)
        oracle.zaPxpby("N", S, n, 1, inp["rhs"][s], 0, pb)
        info, ipiv = lapack.zgbtrf(N, KL, KU, ab, ldab)
        lapack.zgbtrs("N", N, KL, KU, ab, ldab, ipiv, pb)
        x = np.zeros(N, dtype=np.complex128)
        oracle.zaPxpby("T", S, n, 1, pb, 0, x)
        return info

    with lapack.single_threaded():
        one(idx[0])
        t0 = time.perf_counter()
        if threads == 1:
            infos = [one(s) for s in idx]
        else:
            with concurrent.futures.ThreadPoolExecutor(threads) as ex:
                infos = list(ex.map(one, idx))
        dt = time.perf_counter() - t0
    assert all(i == 0 for i in infos)
    return nsample / dt, dt


def lu_roofline(ms, nsys, N, KL, KU, kernel="k_gbtrf<double>", rhs=False):
    """Algorithmic HBM bytes of one factorization launch (read the band of A,
    write the LU storage and the pivots; rhs: also read b and write y) over
    its average duration."""
    ldab = 2 * KL + KU + 1
    bytes_sys = N * (KL + KU + 1) * 16 + N * ldab * 16 + N * 4
    if rhs:
        bytes_sys += 2 * N * 16
    achieved = bytes_sys * nsys / (ms * 1e-3) / 1e9
    flops_sys = 8.0 * N * KL * (KL + KU)
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": round(ms, 4),
            "bytes_per_launch": bytes_sys * nsys,
            "fp64_tflops": round(flops_sys * nsys / (ms * 1e-3) / 1e12, 3)}


def lu_kernel_name(KL, KU, fused_ut=False):
    """The factorization kernel the library dispatches for this band shape
    (the one-wave register kernel for the channel shapes unless
    SZ_GBTRF_KERNEL selects the LDS window kernel); fused_ut: the variant
    that also solves U^T y = b (a fresh factorization followed by one
    transposed solve, as in every implicit substep)."""
    sel = os.environ.get("SZ_GBTRF_KERNEL", "")
    if sel.startswith("l") or (KL, KU) not in ((34, 34), (24, 24), (34, 20)):
        return "k_gbtrf<double, 256>"
    if fused_ut:
        return f"k_gbtrf_wave<double, {KL + 1}, {KL + KU}, 4, true>"
    bs = {"v2": 2, "v4": 4}.get(sel, 8)
    return f"k_gbtrf_wave<double, {KL + 1}, {KL + KU}, {bs}>"


# HBM bytes per LU launch measured with FETCH_SIZE/WRITE_SIZE PMC passes and
# scaled by a plain-copy calibration (profiles/gbtrf_4560_hbm_traffic_pmc.md),
# keyed by (kernel, systems per launch).
RK3_CONFIG = dict(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0)


def cpu_baseline_rk3(ch, W0, threads):
    """The oracle's restatement of one full SMR91 step (3 substeps: explicit
    operator with references and step candidates, M + phi L apply and
    accumulate, assembly with isothermal walls and LAPACK ZGBTRF/ZGBTRS per
    retained wavenumber) at the full benchmark size, the per-wavenumber work
    spread over `threads` host threads.  Returns (substeps/s, seconds)."""
    import math
    import oracle
    from oracle import channel as och
    from oracle import lapack
    ops = oracle.Ops(bands=ch.operators())
    g = och.Grid(ch.Nx, ch.Nz, ch.dNx, ch.dNz, ch.Lx, ch.Lz, ops)
    l1y, l2y = ch.lambdas()
    l1x, l1z = math.pi * ch.Nx / ch.Lx, math.pi * ch.Nz / ch.Lz
    lam = (l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y)
    with lapack.single_threaded():
        t0 = time.perf_counter()
        och.smr91_step(g, W0.copy(), ch.scenario, (ch.lower, ch.upper), lam,
                       threads=threads, lapack=True)
        el = time.perf_counter() - t0
    return 3 / el, el


def stage_breakdown(ch):
    """Device time of each stage of one more (untimed) step, replaying
    lowstorage::step call by call with HIP events on the channel's stream."""
    import torch
    from suzerain_amd.channel import SMR91
    st = torch.cuda.current_stream()
    acc = {"explicit_operator": 0.0, "operator_apply": 0.0,
           "implicit_solve": 0.0, "state_update": 0.0}

    def timed(key, fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        r = fn()
        e1.record(st)
        e1.synchronize()
        acc[key] += e0.elapsed_time(e1)
        return r

    chi = 1.0 / (ch.dNx * ch.dNz)
    timed("state_update", lambda: ch.copy(0, 1))
    cand = timed("explicit_operator", lambda: ch.nonlinear(1, True))
    dt = float(np.nanmin(cand))
    a, b, g, z = SMR91
    timed("operator_apply", lambda: ch.apply_operator(dt * a[0], 0))
    timed("state_update", lambda: ch.add_scaled(chi * dt * g[0], 1, 0))
    timed("implicit_solve", lambda: ch.invert_operator(-dt * b[0], 0))
    for i in (1, 2):
        timed("operator_apply",
              lambda: ch.accumulate_operator(dt * a[i], 0, chi * dt * z[i], 1))
        ch.exchange()
        timed("explicit_operator", lambda: ch.nonlinear(1, False))
        timed("state_update", lambda: ch.add_scaled(chi * dt * g[i], 1, 0))
        timed("implicit_solve", lambda: ch.invert_operator(-dt * b[i], 0))
    return {k: round(v, 3) for k, v in acc.items()}


def percentiles(ms):
    ms = np.asarray(ms, dtype=float)
    return {"p50": round(float(np.percentile(ms, 50)), 3),
            "p99": round(float(np.percentile(ms, 99)), 3),
            "mean": round(float(ms.mean()), 3), "count": int(ms.size)}


def solve_workload(steps, warmup, cpu=True):
    """The batched FP64 implicit solve (S=5, n=96, N=480, KL=KU=34, 1152
    (kx, kz) systems, one right hand side each): assemble + permute + ZGBTRF
    + ZGBTRS + un-permute per pass, inputs resident in HBM."""
    import torch
    import suzerain_amd as sa
    from suzerain_amd.workload import implicit_solve_inputs

    dev = torch.device("cuda:0")
    inp = implicit_solve_inputs()
    op = sa.BsplineOp.channel(inp["n"], inp["k"], inp["htdelta"])
    S, n = inp["S"], inp["n"]
    kl = max(op.ku[:3])
    ku = max(op.kl[:3])
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    N, nsys = S * n, len(inp["kx"])
    ldab = 2 * KL + KU + 1

    kx = torch.from_numpy(inp["kx"]).to(dev)
    kz = torch.from_numpy(inp["kz"]).to(dev)
    cij = torch.from_numpy(inp["cij"]).to(dev)
    uij = torch.from_numpy(inp["uij"]).to(dev)
    rhs = torch.from_numpy(inp["rhs"]).to(dev)
    lu = torch.zeros((nsys, N, ldab), dtype=torch.complex128, device=dev)
    pb = torch.zeros((nsys, 1, N), dtype=torch.complex128, device=dev)
    x = torch.zeros((nsys, N), dtype=torch.complex128, device=dev)
    ipiv = torch.zeros((nsys, N), dtype=torch.int32, device=dev)
    info = torch.zeros((nsys,), dtype=torch.int32, device=dev)
    op.device_ptr()

    stream = torch.cuda.current_stream()
    ev, pev = [], []

    def step(record):
        if record:
            p0 = torch.cuda.Event(enable_timing=True)
            p0.record(stream)
        sa.model_assemble(op, S, kx, kz, cij, uij, inp["phidt"], inp["nu"], lu,
                          KL, KU)
        sa.aPxpby("N", S, n, 1, rhs, 0, pb.view(nsys, N))
        if record:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        sa.gbtrf(lu, KL, KU, ipiv, info)
        if record:
            e1.record(stream)
            ev.append((e0, e1))
        sa.gbtrs("N", lu, KL, KU, ipiv, pb)
        sa.aPxpby("T", S, n, 1, pb.view(nsys, N), 0, x)
        if record:
            p1 = torch.cuda.Event(enable_timing=True)
            p1.record(stream)
            pev.append((p0, p1))

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    trf_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    out = {
        "metric": "implicit wavenumber solves/s",
        "value": round(nsys * steps / elapsed, 1),
        "unit": "solves/s",
        "steps": steps, "warmup": warmup,
        "ms_per_step": round(elapsed * 1e3 / steps, 4),
        "latency_ms_per_batch": percentiles([a.elapsed_time(b) for a, b in pev]),
        "dtype": "complex128 (f64)",
        "config": {"workload": "batched FP64 implicit solve: S=5 x n=96 "
                               "(k=8, 90 stretched breakpoints), N=480, "
                               "KL=KU=34, 1152 (kx,kz) systems, 1 RHS",
                   "systems": nsys, "N": N, "KL": KL, "KU": KU},
        "roofline": lu_roofline(trf_ms, nsys, N, KL, KU,
                                kernel=lu_kernel_name(KL, KU)),
    }
    if cpu:
        threads, nproc = host_threads()
        rate, dt = cpu_baseline(inp, op, 1152, threads)
        rate1, dt1 = cpu_baseline(inp, op, 96, 1)
        out["cpu_baseline"] = {
            "value": round(rate, 1), "unit": "solves/s", "cores": threads,
            "nproc": nproc, "kind": "port",
            "value_1_thread": round(rate1, 1),
            "sample": f"all 1152 systems on {threads} threads ({dt:.1f} s) "
                      f"and 96 evenly spaced systems on 1 thread ({dt1:.1f} "
                      f"s): pack + permute in the oracle's C restatement, "
                      f"LAPACK ZGBTRF + ZGBTRS from scipy's OpenBLAS"}
    return out


def pmc_traffic(kernel_regex, child_args, copy_mib=256, timeout=150):
    """HBM bytes per dispatch of the kernels matching kernel_regex, from two
    rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch)
    over a child run of this script, each scaled by the same counter's
    reading for a plain device copy of copy_mib MiB.  Returns a dict, or
    None with the reason when rocprofv3 is unavailable or fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    me = os.path.abspath(__file__)
    tmp = tempfile.mkdtemp(prefix="sz_pmc_")

    def one(counter, args, regex):
        d = os.path.join(tmp, f"{counter}_{len(os.listdir(tmp))}")
        cmd = ["rocprofv3", "--pmc", counter, "-d", d, "-o", "p",
               "--output-format", "csv"]
        if regex:
            cmd += ["--kernel-include-regex", regex]
        cmd += ["--", sys.executable, "-u", me, *args]
        try:
            subprocess.run(cmd, timeout=timeout, capture_output=True,
                           cwd=ROOT)
        except subprocess.TimeoutExpired:
            return None
        vals = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"),
                           recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") == counter:
                        vals.append((row.get("Kernel_Name", ""),
                                     float(row["Counter_Value"])))
        return vals

    try:
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            k = one(counter, child_args, kernel_regex)
            c = one(counter, ["--copy-calibration", str(copy_mib)],
                    "copyBuffer")
            if not k or not c:
                return None, f"no {counter} readings"
            kib = float(np.mean([v for _, v in k]))
            cal = float(np.mean([v for _, v in c]))
            moved = (copy_mib << 20)        # a copy reads and writes this
            factor = moved / (cal * 1024.0)
            res[counter] = {"kib_per_dispatch": round(kib, 1),
                            "dispatches": len(k),
                            "copy_kib_per_dispatch": round(cal, 1),
                            "bytes_per_kib": round(factor * 1024.0, 2),
                            "bytes": kib * 1024.0 * factor}
        total = res["FETCH_SIZE"]["bytes"] + res["WRITE_SIZE"]["bytes"]
        return {"bytes": total, "passes": res}, None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def pmc_child(kind):
    """The short run a counter pass profiles: one step of the workload."""
    import torch
    if kind == "rk3":
        from suzerain_amd.channel import Channel, laminar_state
        ch = Channel(**RK3_CONFIG)
        ch.set_state(torch.from_numpy(laminar_state(ch)).to("cuda:0"))
        ch.step()
    torch.cuda.synchronize()


def main_rk3(args):
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    ch = Channel(**RK3_CONFIG)
    W0 = laminar_state(ch)
    ch.set_state(torch.from_numpy(W0).to("cuda:0"))
    for _ in range(args.warmup):
        ch.step()
    torch.cuda.synchronize()
    ch.factor_timing(enable=1)
    ch.substep_timing(enable=1)
    t0 = time.perf_counter()
    dts = []
    for _ in range(args.steps):
        dts.append(ch.step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fms, fcount = ch.factor_timing(enable=0)
    sub_ms = ch.substep_timing(enable=0)
    stages = stage_breakdown(ch)
    substeps = 3 * args.steps
    N = 5 * ch.Ny
    KL = KU = 5 * (7) - 1
    out = {
        "metric": "RK substeps/s",
        "value": round(substeps / elapsed, 3),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: isothermal laminar channel + seeded Fourier "
                "perturbations (amplitude 1e-2 |k|^-2, |k| <= 8)",
        "config": {"workload": "SMR91 RK3 step (3 substeps), compressible "
                               "channel 96x96x96, dealiased 144x144, k=8, "
                               "htdelta 3, Re 3000, Ma 1.5, rhome_xyz",
                   "implicit_systems_per_substep": ch.nsys,
                   "N": N, "KL": KL, "KU": KU,
                   "dt_last": dts[-1],
                   "parallelism": "single GPU"},
        "latency_ms_per_substep": percentiles(sub_ms),
        "implicit_solves_per_s": round(ch.nsys * substeps / elapsed, 1),
        "stages_ms_per_step": stages,
        "roofline": lu_roofline(fms / max(fcount, 1), ch.nsys, N, KL, KU,
                                rhs=True),
    }
    out["roofline"]["launches_timed"] = fcount
    kname = lu_kernel_name(KL, KU, fused_ut=True)
    out["roofline"]["kernel"] = kname
    # the factorization's binding resource (issue and latency of its step
    # chain, not bytes); bound stays the HBM roofline it is priced against
    out["roofline"]["limiter"] = "VALU/LDS issue and per-step latency"
    del ch
    torch.cuda.empty_cache()
    out["config1_solve"] = solve_workload(max(10, args.steps // 2), 3,
                                          cpu=not args.no_cpu)
    if not args.no_traffic:
        t, why = pmc_traffic("k_gbtrf_wave", ["--pmc-child", "rk3"])
        out["roofline"]["traffic"] = t["bytes"] if t else None
        if t:
            out["roofline"]["traffic_passes"] = t["passes"]
        else:
            out["roofline"]["traffic_note"] = why
    if not args.no_cpu:
        threads, nproc = host_threads()
        ch = Channel(**RK3_CONFIG)
        rate, el = cpu_baseline_rk3(ch, W0, threads)
        del ch
        out["cpu_baseline"] = {
            "value": round(rate, 5), "unit": "substeps/s", "cores": threads,
            "nproc": nproc, "kind": "port",
            "sample": f"one full SMR91 step (3 substeps) at the full 96^3 "
                      f"size, {el:.1f} s: explicit operator in numpy/"
                      f"pocketfft, operator apply in the oracle's C, "
                      f"4560 implicit solves per substep with LAPACK "
                      f"ZGBTRF/ZGBTRS (scipy's OpenBLAS), per-wavenumber "
                      f"work on {threads} threads"}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): this "
              "build has no multi-GPU path. The implicit solves would shard by "
              "wavenumber with no communication; the explicit stage would add an "
              "all-to-all pencil exchange per substep.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.pmc_child:
        pmc_child(args.pmc_child)
        return
    if args.workload == "rk3":
        main_rk3(args)
        return
    out = solve_workload(args.steps, args.warmup, cpu=not args.no_cpu)
    out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None,
                "data": "synthetic (seeded operator and right hand sides)"})
    out["config"]["parallelism"] = "single GPU"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
