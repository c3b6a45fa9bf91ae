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
    p.add_argument("--workload", choices=("rk3", "solve", "precision", "ensemble",
                                          "serving", "physics"), default="rk3")
    p.add_argument("--steps", type=int, default=40)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes")
    p.add_argument("--pmc-child", default=None, help=argparse.SUPPRESS)
    p.add_argument("--no-extra", action="store_true",
                   help="rk3 only: skip the precision, ensemble and serving keys")
    p.add_argument("--prec-sizes", default="96,192",
                   help="cube sizes of the precision-form comparison")
    p.add_argument("--prec-steps", type=int, default=4)
    p.add_argument("--members", type=int, default=48,
                   help="ensemble members (the uncertainty sweep's first ones)")
    p.add_argument("--ens-steps", type=int, default=3)
    p.add_argument("--requests", type=int, default=16,
                   help="serving trace length")
    p.add_argument("--rate", type=float, default=2.0,
                   help="serving arrivals per second")
    p.add_argument("--sat-rate", type=float, default=16.0,
                   help="serving: a second run of the same requests at this "
                        "rate, high enough to keep both servers busy (0: none)")
    p.add_argument("--req-steps", default="10,60",
                   help="serving: steps per request ~ U{lo..hi}")
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
    """The short run a counter pass profiles: one step of the workload, or
    (t16:NSYS:N) two launches of the single precision transposed solve."""
    import torch
    if kind.startswith("t16:"):
        _, nsys, N = kind.split(":")
        t16_roofline(int(nsys), int(N), 34, reps=2)
    elif kind == "rk3":
        from suzerain_amd.channel import Channel, laminar_state
        ch = Channel(**RK3_CONFIG)
        ch.set_state(torch.from_numpy(laminar_state(ch)).to("cuda:0"))
        ch.step()
    torch.cuda.synchronize()


PREC_SPEC = "zcgbsvx,reuse=true,siter=3,diter=5,tolsc=0.1"


def field_errors(got, want):
    """Per-field max |got - want| / max |want| over the five conserved
    fields (rho E, rho u, rho v, rho w, rho)."""
    return [float(np.abs(got[f] - want[f]).max() / np.abs(want[f]).max())
            for f in range(5)]


def t16_roofline(nsys, N, KL, reps=5):
    """The mixed precision path's dominant kernel, the single precision
    transposed band solve (sixteen lanes per system), launched as the
    precision form launches it: grouped, on factors in the split layout
    (each system's U columns, then its multiplier columns, contiguous), on
    nsys systems of order N with KL = KU.  One launch reads the U and L rows
    of every column (N (2 KL + KL + 1) complex floats) and the pivots, and
    reads and writes the right hand side.  Timed with HIP events on the
    stream it runs on (torch's current stream) over `reps` launches; the
    factors are random with a dominant diagonal and identity pivots (the
    kernel's work does not depend on the values)."""
    import ctypes as C
    import torch
    from suzerain_amd._lib import load
    lib = load()
    f = lib.sz_internal_gbtrs_grouped
    vp, i32 = C.c_void_p, C.c_int
    f.argtypes = [i32, C.c_char, i32, i32, vp, i32, vp, vp, vp, vp, i32, i32, vp]
    f.restype = i32
    dev = torch.device("cuda:0")
    ldab = 3 * KL + 1
    U = 2 * KL + 1
    lu = torch.randn((nsys, N * ldab), dtype=torch.complex64, device=dev)
    lu.view(nsys, N * ldab)[:, :N * U].view(nsys, N, U)[:, :, 2 * KL] += 4.0 * (2 * KL + 1)
    ipiv = torch.arange(1, N + 1, dtype=torch.int32, device=dev).repeat(nsys, 1)
    b = torch.randn((nsys, N), dtype=torch.complex64, device=dev)
    gdt = np.dtype([("n", "<i4"), ("layout", "<i4"), ("ab", "<u8"), ("ipiv", "<u8"),
                    ("b", "<u8")])
    g = np.zeros(nsys, dtype=gdt)
    g["n"], g["layout"] = N, 1
    g["ab"] = lu.data_ptr() + np.arange(nsys, dtype=np.uint64) * (N * ldab * 8)
    g["ipiv"] = ipiv.data_ptr() + np.arange(nsys, dtype=np.uint64) * (N * 4)
    g["b"] = b.data_ptr() + np.arange(nsys, dtype=np.uint64) * (N * 8)
    gs = torch.from_numpy(g.view(np.uint8)).to(dev)
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)

    def launch():
        rc = f(1, b"T", KL, KL, None, ldab, None, None, None, gs.data_ptr(), nsys, N, sp)
        if rc:
            raise RuntimeError(f"grouped transposed solve failed ({rc})")
    launch()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        launch()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    bytes_sys = N * (3 * KL + 1) * 8 + N * 4 + 2 * N * 8
    achieved = bytes_sys * nsys / (ms * 1e-3) / 1e9
    del lu, ipiv, b, gs
    torch.cuda.empty_cache()
    return {"bound": "hbm", "kernel": f"k_gbtrs_t16<float, {KL}, {2 * KL}, true, false>",
            "layout": "split (U columns, then multiplier columns)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": round(ms, 4), "bytes_per_launch": bytes_sys * nsys,
            "systems": nsys}


def precision_workload(size, steps, dt):
    """The precision form on the channel substep at size^3 (dealiased
    3/2 size, k = 8, the rk3 workload's physics): the plain FP64 path
    (zgbsv every substep) and the FP32-refined path (float factors reused
    across steps per substep slot, FP64 residual sweeps) advance the same
    initial state with the same fixed time step; one untimed step first
    (it holds the refined path's factorizations), then `steps` timed."""
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    cfg = dict(Nx=size, Ny=size, Nz=size, dNx=3 * size // 2, dNz=3 * size // 2,
               k=8, htdelta=3.0)
    out = {"size": [size, size, size], "dealiased": [3 * size // 2] * 2,
           "spec": PREC_SPEC, "dt": dt, "steps_timed": steps}
    states, rates = {}, {}
    for name, spec in (("fp64", "zgbsv"), ("refined", PREC_SPEC)):
        ch = Channel(**cfg, solver=spec)
        if name == "fp64":
            W0 = torch.from_numpy(laminar_state(ch)).to("cuda:0")
            out["systems"], out["N"] = ch.nsys, 5 * size
        ch.set_state(W0)
        got = ch.step(max_dt=dt)
        assert got == dt, (got, dt)
        first = ch.get_state(0).cpu().numpy()
        torch.cuda.synchronize()
        ch.substep_timing(enable=1)
        t0 = time.perf_counter()
        for _ in range(steps):
            assert ch.step(max_dt=dt) == dt
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        sub = ch.substep_timing(enable=0)
        rates[name] = el / steps
        states[name] = (first, ch.get_state(0).cpu().numpy())
        if name == "refined":
            si = ch.solver_stats("siter")
            out["siter_last_substep"] = {"min": int(si.min()), "mean": round(float(si.mean()), 3),
                                         "max": int(si.max())}
            out["restarted_or_double_last_substep"] = list(ch.solver_counts())[1:]
            out["latency_ms_per_substep"] = percentiles(sub)
        else:
            out["fp64_latency_ms_per_substep"] = percentiles(sub)
        del ch
        torch.cuda.empty_cache()
    out["fp64_ms_per_step"] = round(rates["fp64"] * 1e3, 3)
    out["ms_per_step"] = round(rates["refined"] * 1e3, 3)
    out["substeps_per_s"] = round(3 / rates["refined"], 3)
    out["fp64_substeps_per_s"] = round(3 / rates["fp64"], 3)
    out["implicit_solves_per_s"] = round(3 * out["systems"] / rates["refined"], 1)
    out["speedup_vs_fp64"] = round(rates["fp64"] / rates["refined"], 3)
    out["err_per_field_first_step"] = ["%.2e" % e for e in
                                       field_errors(states["refined"][0], states["fp64"][0])]
    out["err_per_field_after_all_steps"] = ["%.2e" % e for e in
                                            field_errors(states["refined"][1], states["fp64"][1])]
    out["roofline"] = t16_roofline(out["systems"], out["N"], 34)
    return out


def ensemble_workload(count, steps):
    """The batched uneven ensemble: `count` members of the seeded uncertainty
    sweep advance together (explicit stages on the autotuned streams, all
    members' implicit solves in the grouped FP32-refined solver, the step
    replayed from a captured hipGraph); against it, the plain FP64 drop-in
    steps the same members one by one.  Both start from the same states and
    take the same number of steps (2 untimed, then `steps` timed)."""
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    from suzerain_amd.ensemble import Ensemble, uncertainty_sweep, default_dt
    members = uncertainty_sweep(count)
    W = []
    for c in members:
        ch = Channel(**c, solver="none")
        W.append(laminar_state(ch, seed=7 + c["Ny"]))
        del ch
    ens = Ensemble(members, graph=True)
    for m in range(count):
        ens.set_state(m, torch.from_numpy(W[m]).to("cuda:0"))
    ens.step(1)                          # factorizations and the first capture
    first = [ens.get_state(m).cpu().numpy() for m in range(count)]
    tuning = ens.autotune(1)
    ens.step(1)                          # captured again with the tuned choice
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat = []
    for _ in range(steps):
        t1 = time.perf_counter()
        ens.step(1)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    el = time.perf_counter() - t0
    ens.check()
    nsys = sum(((c["Nx"] - 1) // 2 + 1) * (2 * ((c["Nz"] - 1) // 2) + 1) for c in members)
    out = {"members": count, "systems": nsys, "resident_gb": round(ens.bytes() / 1e9, 1),
           "spec": PREC_SPEC, "graph": True, "tuning": tuning, "steps_timed": steps,
           "ms_per_step": round(el / steps * 1e3, 2),
           "member_substeps_per_s": round(3 * count * steps / el, 2),
           "step_latency_ms": percentiles(lat)}
    finals = [ens.get_state(m).cpu().numpy() for m in range(count)]
    del ens
    torch.cuda.empty_cache()
    fp64 = 0.0
    errs, errs1 = np.zeros(5), np.zeros(5)
    for m, c in enumerate(members):
        dt = default_dt(c)
        ch = Channel(**c, solver="zgbsv")
        ch.set_state(torch.from_numpy(W[m]).to("cuda:0"))
        ch.step(max_dt=dt)
        errs1 = np.maximum(errs1, field_errors(first[m], ch.get_state(0).cpu().numpy()))
        ch.step(max_dt=dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ch.step(max_dt=dt)
        torch.cuda.synchronize()
        fp64 += time.perf_counter() - t0
        errs = np.maximum(errs, field_errors(finals[m], ch.get_state(0).cpu().numpy()))
        del ch
        torch.cuda.empty_cache()
    out["fp64_ms_per_step"] = round(fp64 / steps * 1e3, 2)
    out["fp64_member_substeps_per_s"] = round(3 * count * steps / fp64, 2)
    out["speedup_vs_fp64"] = round(fp64 / el, 3)
    out["err_per_field_first_step"] = ["%.2e" % e for e in errs1]
    out["err_per_field_after_all_steps"] = ["%.2e" % e for e in errs]
    out["err_note"] = ("max over members of max|refined - fp64| / max|fp64| per field: "
                       "after one step from the same state, and after all "
                       f"{steps + 2} steps (the two trajectories drift apart "
                       "by the refined solve's residual floor each step)")
    return out


def serving_requests(nreq, rate, steps_range, seed=20261204, large_frac=0.05,
                     large=(192, 192, 192)):
    """Open-loop Poisson arrivals at `rate` per second: each request is a
    channel drawn from the uncertainty sweep's distribution (or, with
    probability large_frac, the large case) with a step count ~ U(steps_range)
    and the sweep's initial-condition recipe."""
    from suzerain_amd.serving import poisson_trace
    from suzerain_amd.ensemble import uncertainty_sweep
    from suzerain_amd.channel import DEFAULT_SCENARIO

    def make(i, rng):
        if rng.random() < large_frac:
            cfg = dict(Nx=large[0], Ny=large[1], Nz=large[2], k=8, htdelta=3.0,
                       scenario=dict(DEFAULT_SCENARIO))
        else:
            cfg = uncertainty_sweep(1, seed=int(rng.integers(1 << 30)))[0]
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        return cfg, steps, _initial_state(cfg)
    return poisson_trace(rate, nreq, make, seed=seed)


def _initial_state(cfg):
    """A request's initial state, made when the trace is made (a request
    arrives with its state): the server only uploads it."""
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    ch = Channel(**cfg, solver="none")
    W = laminar_state(ch, seed=7 + cfg["Ny"])
    del ch
    torch.cuda.empty_cache()

    def init():
        return torch.from_numpy(W).to("cuda:0")
    return init


def serving_workload(nreq, rate, steps_range, budget=240e9, sat_rate=0.0):
    """Continuous batching of the trace within the HBM budget (FP32-refined,
    captured steps), then the plain FP64 drop-in serving the same trace
    (each request its own channel, live ones stepped round robin); final
    states compared request by request.  With sat_rate, the same requests
    again arriving at that rate, which keeps both servers busy: their
    sustained substeps/s."""
    out = serving_run(nreq, rate, steps_range, budget)
    out.update({"requests": nreq, "rate_per_s": rate, "budget_gb": budget / 1e9,
                "spec": PREC_SPEC})
    if sat_rate > 0:
        sat = serving_run(nreq, sat_rate, steps_range, budget)
        sat["rate_per_s"] = sat_rate
        out["saturated"] = sat
    return out


def serving_run(nreq, rate, steps_range, budget):
    import torch
    from suzerain_amd.serving import Server, Fp64Server
    out = {"steps_range": list(steps_range)}
    reqs = serving_requests(nreq, rate, steps_range)
    out["large_requests"] = sum(r.cfg["Nx"] == 192 for r in reqs)
    out["member_steps"] = sum(r.steps for r in reqs)
    rep = Server(budget_bytes=budget, solver=PREC_SPEC, graph=True).run(reqs)
    torch.cuda.empty_cache()
    ref = serving_requests(nreq, rate, steps_range)
    rep64 = Fp64Server(budget_bytes=budget).run(ref)
    torch.cuda.empty_cache()
    errs, per_step = np.zeros(5), np.zeros(5)
    for a, b in zip(sorted(reqs, key=lambda r: r.rid), sorted(ref, key=lambda r: r.rid)):
        e = np.asarray(field_errors(a.result, b.result))
        errs = np.maximum(errs, e)
        per_step = np.maximum(per_step, e / a.steps)
    out.update({"refined": rep, "fp64": rep64,
                "speedup_substeps_per_s": round(rep["substeps_per_s"] / rep64["substeps_per_s"], 3),
                "err_per_field_final_max_over_requests": ["%.2e" % e for e in errs],
                "err_per_field_per_step_max_over_requests": ["%.2e" % e for e in per_step],
                "err_note": f"final states after each request's {steps_range[0]}.."
                            f"{steps_range[1]} steps; per step: that error divided "
                            "by the request's step count"})
    return out


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
    if not args.no_extra:
        out.update(extra_keys(args))
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


PREC_DT = {96: 0.02, 192: 0.01}


def mms_workload(Ns=(12, 18, 24, 36), dt=5e-4):
    """The manufactured-solution convergence gate (jobscripts/
    mms_case.perfect.in: nsctpl isothermal channel, alpha = 1, k = 6,
    Nx = Nz = 2 Ny / 3, one step of 5e-4 from the exact state): L2 error
    per field against the exact solution for the FP64 and the mixed
    precision solves, and the observed orders between successive Ny."""
    from suzerain_amd.channel import Channel
    from suzerain_amd.mms import Manufactured
    out = {"Ny": list(Ns), "dt": dt, "k": 6}
    for solver in ("zgbsv", PREC_SPEC):
        errs = []
        for N in Ns:
            m = Manufactured.isothermal_channel().set(alpha=1.0)
            Nxz = 2 * N // 3
            sc = {q: getattr(m.s, q) for q in ("alpha", "beta", "gamma", "Ma", "Pr", "Re")}
            ch = Channel(Nx=Nxz, Ny=N, Nz=Nxz, dNx=3 * Nxz // 2, dNz=3 * Nxz // 2, k=6,
                         htdelta=0.0, Lx=m.s.Lx, Ly=m.s.Ly, Lz=m.s.Lz, scenario=sc,
                         solver=solver)
            ch.set_manufactured(m)
            ch.set_time(0.0)
            ch.manufactured_state(0.0, which=0)
            ch.step(max_dt=dt)
            ch.manufactured_state(ch.time, which=1)
            ch.add_scaled(-1.0, 1, 0)
            mean, fluct = ch.field_L2(0)
            errs.append(np.hypot(mean, fluct))
            del ch
        errs = np.array(errs)
        orders = [np.log(errs[i] / errs[i + 1]) / np.log(Ns[i + 1] / Ns[i])
                  for i in range(len(Ns) - 1)]
        key = "fp64" if solver == "zgbsv" else "mixed_precision"
        out[key] = {"l2_error_per_field": [[float(f"{v:.4e}") for v in e] for e in errs],
                    "observed_order_per_field": [[round(float(v), 3) for v in o]
                                                 for o in orders],
                    "min_order": round(float(np.min(orders)), 3)}
    out["max_order_difference"] = float(f"""{np.max(np.abs(
        np.array(out['fp64']['observed_order_per_field'])
        - np.array(out['mixed_precision']['observed_order_per_field']))):.3e}""")
    return out


def constrained_workload(steps=5, dt=0.01):
    """The 96^3 FP64 step with the channel driver's bulk constraints (rho
    and rho u held at the laminar state's bulk values, main_advance.cpp:
    259-272) against the unconstrained step: per-step time and the bulk
    deviation after the steps."""
    import torch
    from suzerain_amd.channel import Channel, laminar_state
    cfg = dict(Nx=96, Ny=96, Nz=96, dNx=144, dNz=144, k=8, htdelta=3.0)
    out = {"steps": steps, "dt": dt}
    W0 = None
    for name, bulk in (("unconstrained", None), ("constrained", {"rho": 1.0, "rho_u": 1.0})):
        ch = Channel(**cfg, bulk=bulk)
        if W0 is None:
            W0 = torch.from_numpy(laminar_state(ch)).to("cuda:0")
        ch.set_state(W0)
        ch.step(max_dt=dt)                      # untimed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ch.step(max_dt=dt)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        r = {"ms_per_step": round(ms, 3)}
        if bulk:
            w = ch.bulk_weights()
            s = ch.get_state(0).cpu().numpy()
            r["bulk_deviation"] = {"rho": float(f"{abs(np.dot(w, s[4, 0, 0].real) - 1):.2e}"),
                                   "rho_u": float(f"{abs(np.dot(w, s[1, 0, 0].real) - 1):.2e}")}
            r["multipliers_last_substep"] = [float(f"{v:.4e}") for v in ch.constraint_multipliers()]
        out[name] = r
        del ch
        torch.cuda.empty_cache()
    out["overhead"] = round(out["constrained"]["ms_per_step"]
                            / out["unconstrained"]["ms_per_step"] - 1, 4)
    return out


def extra_keys(args):
    """The precision form (96^3 and 192^3), the 48-member ensemble and the
    serving trace, each against the plain FP64 drop-in on the same work."""
    out = {}
    for size in [int(x) for x in args.prec_sizes.split(",") if x]:
        r = precision_workload(size, args.prec_steps, PREC_DT.get(size, 0.01))
        if not args.no_traffic and size == min(int(x) for x in args.prec_sizes.split(",") if x):
            t, why = pmc_traffic("k_gbtrs_t16", ["--pmc-child",
                                                 f"t16:{r['systems']}:{r['N']}"])
            r["roofline"]["traffic"] = t["bytes"] if t else None
            if t:
                r["roofline"]["traffic_passes"] = t["passes"]
            else:
                r["roofline"]["traffic_note"] = why
        out[f"precision_{size}"] = r
    if args.members:
        out["ensemble"] = ensemble_workload(args.members, args.ens_steps)
    if args.requests:
        lo, hi = (int(x) for x in args.req_steps.split(","))
        out["serving"] = serving_workload(args.requests, args.rate, (lo, hi),
                                          sat_rate=args.sat_rate)
    if getattr(args, "physics", True):
        out["mms"] = mms_workload()
        out["bulk_constraints"] = constrained_workload()
    return out


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
    if args.workload in ("precision", "ensemble", "serving", "physics"):
        keep = {"precision": ("members", "requests"),
                "ensemble": ("prec_sizes", "requests"),
                "serving": ("prec_sizes", "members"),
                "physics": ("prec_sizes", "members", "requests")}[args.workload]
        for k in keep:
            setattr(args, k, "" if k == "prec_sizes" else 0)
        args.physics = args.workload == "physics"
        print(json.dumps(extra_keys(args)), flush=True)
        return
    out = solve_workload(args.steps, args.warmup, cpu=not args.no_cpu)
    out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None,
                "data": "synthetic (seeded operator and right hand sides)"})
    out["config"]["parallelism"] = "single GPU"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
