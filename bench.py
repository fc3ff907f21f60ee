"""Throughput of QGN's ARK43_QGLeith time stepper on one MI355X.

Workload: expstrat_1024x16 (1024^2 x 16 layers, exponential stratification,
surface-intensified shear, quadratic + Ekman drag, ARK43_QGLeith adaptive).
A step is one accepted TimeStep (rejected attempts included in the time).
Prints one JSON line; ``value`` is simulated days per wall-second of the
mixed-precision path through the graph-captured attempts (qgn_batch_step),
with the per-call qgn_time_step path and the FP64 path measured on the same
GPU from the same spun-up state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes of each kernel group of one attempt, counting only the
# kept (unpadded) modes: rs = bytes of one complex of the pass's arithmetic
# type; the stage value is read as the projection kernel's FP32 copy (rs
# bytes) on the mixed path and as the FP64 state (16 bytes) on the FP64 path.
# ncmp = nz*Ky*Kx compact modes, nmix = nz*ny*Kx mixed values.
def kernel_bytes(kernel, g, rs, qrs=None):
    """Bytes per launch (row_jac, col_inv, col_rhs, project, update) or per
    attempt of all launches (stage: the stage kernels of an attempt; errnorm:
    the column and row passes of the error field)."""
    qrs = rs if qrs is None else qrs
    nz, ny, Kx, Ky = g["nz"], g["ny"], g["Kx"], g["Ky"]
    ncmp = nz * Ky * Kx
    nmix = nz * ny * Kx
    drag = 2 * ny * Kx * rs if g["drag"] else 0
    if kernel == "row_jac":     # read u, v, q_x, q_y; write J (+ |u|u, |u|v on layer nz)
        return 5 * nmix * rs + drag
    if kernel == "col_inv":     # read q, psi; write u, v, q_x, q_y
        return (qrs + rs) * ncmp + 4 * nmix * rs
    if kernel == "col_rhs":     # read J, q, psi; write N (+ read |u|u, |u|v for layer nz)
        return nmix * rs + (qrs + 2 * rs) * ncmp + drag
    if kernel == "project":     # mixed: the MFMA projection of every GetRHS (read the FP32 stage
        # value, write psi); FP64: the projection of q (read q, write psi)
        return ncmp * 2 * rs if rs == 8 else ncmp * (16 + rs)
    if kernel == "stage":       # mixed: NS = 0..5, read q and NS stage RHS, write the stage value
        # rounded to FP32; FP64: NS = 1..5 fused with the projection, read q and NS stage RHS,
        # write psi and the FP64 stage value
        if rs == 8:
            return sum(ncmp * (16 + ns * rs + rs) for ns in range(0, 6))
        return sum(ncmp * (16 + ns * rs + rs + 16) for ns in range(1, 6))
    if kernel == "update":      # read q and 6 stage RHS; write the candidate (FP64) and the error field
        # (float2 on the mixed path)
        return ncmp * (16 + 6 * rs + 16 + rs)
    if kernel == "errnorm":     # y-inverse of E (read compact, write mixed) and row max (read mixed),
        # in FP32 on the mixed path
        return ncmp * rs + 2 * nmix * rs
    raise ValueError(kernel)


def projection_flops(g):
    """BF16 MFMA flops of one projection launch (k_proj32_mfma16, nz <= 16):
    per tile of 32 points, 4 waves x 2 products x 3 MFMAs of 16x16x16."""
    plane = g["Ky"] * g["Kxp"]
    return -(-plane // 32) * 4 * 2 * 3 * 2 * 16 * 16 * 16


def profile_figures(workload, kernel):
    """(HBM bytes per launch, its commit, rocprof average us) of `kernel` from
    the committed summaries profiles/traffic_<workload>.txt, or Nones."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None, None, None
    for line in open(path):
        f = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
        if f.get("kernel") == kernel:
            avg = float(f["avg_us"]) if "avg_us" in f else None
            return int(f["bytes_per_launch"]), f.get("commit"), avg
    return None, None, None


def geom(p):
    ny = p.ny
    ya, yb = ny // 3, 2 * (ny // 3) + 1
    Kx = p.nx // 3
    return dict(nx=p.nx, ny=ny, nz=p.nz, Kx=Kx, Kxp=-(-Kx // 16) * 16, Ky=ya + (ny - yb), drag=p.C_d != 0.0)


def _model(wl, precision):
    from qgn_fortran_amd import Batch
    from qgn_fortran_amd.model import vertical_modes
    p = wl.params
    p.precision = precision
    H, _L, Modes, EVals, qy = vertical_modes(wl.H, wl.S, wl.uBar, p.Htot)
    # a batch of one member: qgn_time_step and the graph path on one context
    return Batch(p, [dict(H=H, Modes=Modes, EVals=EVals, uBar=wl.uBar, qyBar=qy)])


def run_gpu(wl, precision, q_spun, t0, dt0, steps, warmup, prof_kernel):
    """K accepted steps through qgn_time_step (one read-back per attempt)
    with the roofline kernel timed by HIP events on the library's stream,
    then K accepted steps from the same state through the graph-captured
    attempts of qgn_batch_step (no host round trip inside the timed run)."""
    import torch
    m = _model(wl, precision)
    m.set_state(q_spun)
    t, dt = t0, dt0
    for _ in range(warmup):
        t, dt = m.time_step(t, dt)
    q_w, t_w, dt_w = m.get_state(), t, dt
    st0 = m.stats()
    m.prof_enable("all")
    torch.cuda.synchronize()
    tw0 = time.perf_counter()
    ts = t
    for _ in range(steps):
        t, dt = m.time_step(t, dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - tw0
    per = m.prof_read_all()
    nlaunch, kms = per[prof_kernel]
    m.prof_enable(-1)
    st1 = m.stats()
    res = dict(
        precision=precision, wall_s=wall, sim_days=(t - ts) / 86400.0,
        days_per_s=(t - ts) / 86400.0 / wall, ms_per_step=1e3 * wall / steps,
        mean_dt=(t - ts) / steps, rejects=st1["rejects"] - st0["rejects"],
        rhs_evals=st1["rhs_evals"] - st0["rhs_evals"], kernel=prof_kernel,
        kernel_launches=nlaunch, kernel_ms=kms / max(nlaunch, 1), t_end=t, dt_end=dt,
        kernels={k: dict(launches=n, ms=ms) for k, (n, ms) in per.items()},
    )
    res["ms_per_rhs"] = 1e3 * wall / max(res["rhs_evals"], 1)
    # graph path from the same post-warm-up state
    m.set_state(q_w)
    m.set_time(t_w, dt_w)
    m.step(warmup)                      # captures the graph; untimed
    info = m.times()
    n0, tg0 = int(info["steps"][0]), float(info["t"][0])
    torch.cuda.synchronize()
    tw0 = time.perf_counter()
    done = 0
    attempts = 0
    while done < steps:                 # rejected attempts add to the timed work
        m.step(steps - done)
        attempts += steps - done
        info = m.times()
        done = int(info["steps"][0]) - n0
    wall_g = time.perf_counter() - tw0
    tg = float(info["t"][0])
    res["graph"] = dict(wall_s=wall_g, steps=done, attempts=attempts, sim_days=(tg - tg0) / 86400.0,
                        days_per_s=(tg - tg0) / 86400.0 / wall_g, ms_per_step=1e3 * wall_g / done,
                        mean_dt=(tg - tg0) / done)
    q_end = m.get_state()
    m.close()
    return res, q_end


def spin_up(wl, steps, precision="mixed"):
    """`steps` accepted mixed-precision steps from the seeded initial state
    (graph-captured attempts, 500 at a time)."""
    from qgn_fortran_amd import initial
    p = wl.params
    m = _model(wl, precision)
    q0 = initial.random_phase_pv(p.nx, p.ny, p.nz, wl.k_peak, wl.rms, wl.seed, p.Lx, p.Ly)
    m.initialize(q0)
    m.set_time(0.0, wl.dt0)
    tw = time.perf_counter()
    done = 0
    while done < steps:
        m.step(min(500, steps - done))
        info = m.times()
        done = int(info["steps"][0])
        print(f"# spin-up {done}/{steps}: t = {info['t'][0] / 86400:.2f} days, dt = {info['dt'][0]:.1f} s, "
              f"{time.perf_counter() - tw:.1f} s", file=sys.stderr, flush=True)
    q = m.get_state()
    m.close()
    return q, float(info["t"][0]), float(info["dt"][0])


def _cpu_model_name():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, q_spun, t0, dt0, budget_s=30.0):
    """The reference algorithm on the host cores, from the same spun-up
    state: the C++/OpenMP FP64 restatement with MKL DFTI transforms
    (oracle/qgn_cpu.cpp; MKL is the reference's FFT library), as many
    accepted ARK43_QGLeith steps as fit in ~budget_s/2, at least one.
    Falls back to the numpy/scipy oracle where the restatement or MKL is
    missing."""
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd.model import vertical_modes
    p = wl.params
    cores = int(os.environ.get("OMP_NUM_THREADS", min(os.cpu_count() or 1, 16)))
    H, _L, Modes, EVals, qy = vertical_modes(wl.H, wl.S, wl.uBar, p.Htot)
    po = orc.Params(nx=p.nx, ny=p.ny, nz=p.nz, Lx=p.Lx, Ly=p.Ly, Htot=p.Htot, beta=p.beta, A2=p.A2, A8=p.A8,
                    r0=p.r0, C_d=p.C_d, TOL=p.TOL, is_explicit=p.is_explicit, dt_max=p.dt_max)
    what = "C++/OpenMP FP64 restatement with MKL DFTI (oracle/qgn_cpu.cpp)"
    try:
        from oracle import qgn_cpu
        cm = qgn_cpu.Model(po, H, Modes, EVals, wl.uBar, qy, threads=cores)
        cores = cm.threads
        step = lambda q, t, dt: cm.ark43_step(q, t, dt, leith=True)
    except (RuntimeError, OSError, ImportError) as e:
        om = orc.Model(po, H, Modes, EVals, wl.uBar, qy, workers=cores)
        stp = orc.ARK43(om, leith=True)
        step = stp.step
        what = f"FP64 numpy/scipy oracle (the C++ restatement did not load: {e})"
    q, t, dt = q_spun, t0, dt0
    q, t, dt = step(q, t, dt)           # first step: page faults, FFT plans, thread start-up
    t1 = t
    tw = time.perf_counter()
    n = 0
    while True:
        q, t, dt = step(q, t, dt)
        n += 1
        if time.perf_counter() - tw > budget_s * 0.5 or n >= 40:
            break
    wall = time.perf_counter() - tw
    return dict(value=(t - t1) / 86400.0 / wall, unit="simulated days/s", cores=cores, kind="port",
                sample=f"{n} accepted ARK43_QGLeith steps of the same spun-up 1024^2x16 state after one untimed step, "
                       f"{what}, {wall:.1f} s, {1e3 * wall / n:.0f} ms/step",
                cpu_model=_cpu_model_name(), nproc=os.cpu_count())


def run_cnn(steps, warmup):
    """The closure network's BF16 MFMA kernels at the closure's shapes: train
    steps on 10 images of 256^2 x 16 channels (the training batch) and
    forward passes on 4 images (the 4-member coarse ensemble), with HIP-event
    timings and the dense-MFMA flops of each."""
    import torch
    from qgn_fortran_amd.closure import _he_init
    from qgn_fortran_amd.cnn import CNN, HIDDEN, to_bf16_bits
    cin, cout, H, W = 16, 8, 256, 256
    rng = np.random.default_rng(0)
    net = CNN(cin, cout, _he_init(cin, cout, 1))
    dims = [(cin, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, cout)]
    macs_px = sum(o * 9 * i for i, o in dims)        # per pixel, forward
    out = {}
    for name, n, train in (("train_step", 10, True), ("forward", 4, False)):
        x = torch.from_numpy(to_bf16_bits(rng.standard_normal((n, H, W, cin)))).cuda()
        t = torch.from_numpy(rng.standard_normal((n, H, W, cout)).astype(np.float32)).cuda()
        y = torch.empty((n, H, W, cout), device="cuda", dtype=torch.float32)
        call = (lambda: net.train_step(x, t, y, 1e-4)) if train else (lambda: net.forward(x, y))
        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            call()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        # forward 2 flops per MAC; a train step adds the data and weight gradients (2x forward, the
        # first layer's data gradient skipped)
        flops = 2.0 * n * H * W * macs_px * (3.0 if train else 1.0)
        out[name] = dict(images=n, ms=ms, tflops=flops / (ms * 1e-3) / 1e12,
                         frac_bf16_dense_peak=flops / (ms * 1e-3) / 2.5e15)
    net.close()
    return out


def run_ensemble(precision, attempts, spinup):
    """The 96-member ensemble (32 members of 8 layers on each of 128^2,
    256^2, 384^2, ARK43_QGLeith): `spinup` attempts from the seeded states,
    then `attempts` timed attempts of every member through chained graph
    launches.  Member-days per second sums the simulated time every member
    advanced."""
    import torch
    from qgn_fortran_amd import Batch, Ensemble, configs
    groups = configs.ensemble96(precision=precision)
    batches = []
    for g in groups:
        b = Batch(g.params, g.members)
        b.initialize(g.q0)
        b.set_time(0.0, g.dt0)
        batches.append(b)
    ens = Ensemble(batches)
    ens.step(spinup)
    t0s = [b.times()["t"].copy() for b in batches]
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ens.step(attempts)
    infos = [b.times() for b in batches]
    wall = time.perf_counter() - w0
    sizes = {}
    total = 0.0
    for g, t0, i in zip(groups, t0s, infos):
        d = float(np.sum(i["t"] - t0)) / 86400.0
        total += d
        sizes[f"{g.params.nx}^2x{g.params.nz} x{len(g.members)}"] = dict(
            member_days=d, rejects=int(i["rejects"].sum()), mean_dt=float(np.mean(i["dt"])))
    ens.close()
    for b in batches:
        b.close()
    return dict(precision=precision, members=96, attempts=attempts, spinup_attempts=spinup, wall_s=wall,
                ms_per_attempt=1e3 * wall / attempts, member_days_per_s=total / wall, sizes=sizes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--spinup", type=int, default=2000)
    ap.add_argument("--workload", default="expstrat_1024x16")
    ap.add_argument("--kernel", default="row_jac", help="kernel timed for the roofline figure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--cnn", action="store_true",
                    help="time the closure network's MFMA kernels instead of the stepper")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 96-member ensemble figures")
    ap.add_argument("--ensemble-only", action="store_true", help="only the 96-member ensemble figures")
    ap.add_argument("--ensemble-attempts", type=int, default=40)
    ap.add_argument("--ensemble-spinup", type=int, default=100)
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a 1 GiB read + 1 GiB write multiply three times (PMC byte-counter reference)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (ensemble members or kx/y slabs "
              "would shard across GPUs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from qgn_fortran_amd import configs

    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        sys.exit(1)
    if a.calib_copy:
        # reference kernel for the PMC byte counters: an elementwise multiply
        # that reads 1 GiB and writes 1 GiB (torch's MulFunctor kernel)
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if a.cnn:
        print(json.dumps({"cnn": run_cnn(a.steps, a.warmup)}), flush=True)
        return
    if a.ensemble_only:
        print(json.dumps({"ensemble": {p: run_ensemble(p, a.ensemble_attempts, a.ensemble_spinup)
                                       for p in ("mixed", "fp64")}}), flush=True)
        return
    wl = configs.WORKLOADS[a.workload]()
    q, t0, dt0 = spin_up(wl, a.spinup)
    mixed, _ = run_gpu(configs.WORKLOADS[a.workload](), "mixed", q, t0, dt0, a.steps, a.warmup, a.kernel)
    fp64, _ = run_gpu(configs.WORKLOADS[a.workload](), "fp64", q, t0, dt0, a.steps, a.warmup, a.kernel)
    g = geom(wl.params)
    steps_timed = a.steps
    passes = {}
    for prec, res, rs in (("mixed", mixed, 8), ("fp64", fp64, 16)):
        attempts = steps_timed + res["rejects"]
        out = {}
        for k, v in res["kernels"].items():
            if v["launches"] == 0:
                continue
            per_attempt = k in ("stage", "errnorm")      # the byte model counts all launches of an attempt
            nb = kernel_bytes(k, g, rs)
            sec = v["ms"] * 1e-3 / (attempts if per_attempt else v["launches"])
            out[k] = dict(launches=v["launches"], avg_us=round(1e6 * v["ms"] * 1e-3 / v["launches"], 2),
                          bytes=nb, gbs=round(nb / sec / 1e9, 1), frac_hbm_peak=round(nb / sec / 8e12, 4))
        passes[prec] = out
    # MFMA utilisation of the split-BF16 projection (mixed; nz <= 16): the
    # stage/projection kernels' BF16 flops over their summed time
    kp = mixed["kernels"]
    proj_ms = kp["project"]["ms"]
    proj_launches = kp["project"]["launches"]
    mfma = None
    if wl.params.nz <= 16 and proj_ms > 0:
        fl = projection_flops(g) * proj_launches
        mfma = dict(kernel="k_proj32_mfma16", bf16_flops_per_launch=projection_flops(g),
                    tflops=round(fl / (proj_ms * 1e-3) / 1e12, 2),
                    frac_bf16_dense_peak=round(fl / (proj_ms * 1e-3) / 2.5e15, 5))
    rs = 8  # mixed path: FP32 complex intermediates
    kb = kernel_bytes(a.kernel, g, rs)
    achieved = kb / (mixed["kernel_ms"] * 1e-3) / 1e9
    traffic, traffic_commit, prof_avg = profile_figures(a.workload, a.kernel)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s",
                frac=round(achieved / 8000.0, 4), traffic=traffic, kernel=a.kernel,
                bytes_per_launch=kb, avg_launch_ms=mixed["kernel_ms"],
                timing="HIP events around each launch on the library's stream, qgn_time_step path",
                profile_avg_ms=None if prof_avg is None else prof_avg * 1e-3,
                profile_frac=None if prof_avg is None else round(kb / (prof_avg * 1e-6) / 8e12, 4),
                traffic_commit=traffic_commit)
    ensemble = None
    if not a.no_ensemble:
        ensemble = {p: run_ensemble(p, a.ensemble_attempts, a.ensemble_spinup) for p in ("mixed", "fp64")}
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(configs.WORKLOADS[a.workload](), q, t0, dt0, a.cpu_budget)
    out = {
        "metric": "simulated days/wall-s, 1024^2x16-layer ARK43_QGLeith, mixed precision",
        "value": round(mixed["graph"]["days_per_s"], 4),
        "unit": "simulated days/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(mixed["graph"]["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 FFT/Jacobian, f64 state/stages/error",
        "data": "synthetic (seeded random-phase PV, spun up on the GPU)",
        "config": {"workload": a.workload, "nx": wl.params.nx, "ny": wl.params.ny, "nz": wl.params.nz,
                   "scheme": "ARK43_QGLeith", "spinup_steps": a.spinup, "parallelism": "single GPU"},
        "mixed": mixed, "fp64": fp64,
        "mixed_vs_fp64": round(mixed["graph"]["days_per_s"] / fp64["graph"]["days_per_s"], 3),
        "passes": passes, "projection_mfma": mfma,
        "roofline": roof, "ensemble": ensemble, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
