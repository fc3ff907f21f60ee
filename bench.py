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

# Algorithmic HBM bytes per launch of the RHS passes, counting only the kept
# (unpadded) modes: rs = bytes of one complex of the pass's arithmetic type;
# the stage value q is read as the projection kernel's FP32 copy (rs bytes)
# on the mixed path and as the FP64 state (16 bytes) on the FP64 path.
# ncmp = nz*Ky*Kx compact modes, nmix = nz*ny*Kx mixed values.
def kernel_bytes(kernel, g, rs):
    nz, ny, Kx, Ky = g["nz"], g["ny"], g["Kx"], g["Ky"]
    ncmp = nz * Ky * Kx
    nmix = nz * ny * Kx
    drag = 2 * ny * Kx * rs if g["drag"] else 0
    if kernel == "row_jac":     # read u, v, q_x, q_y; write J (+ |u|u, |u|v on layer nz)
        return 5 * nmix * rs + drag
    if kernel == "col_inv":     # read q, psi; write u, v, q_x, q_y
        return 2 * ncmp * rs + 4 * nmix * rs
    if kernel == "col_rhs":     # read J, q, psi; write N (+ read |u|u, |u|v for layer nz)
        return nmix * rs + 3 * ncmp * rs + drag
    raise ValueError(kernel)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/traffic_<workload>.txt), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
        if f.get("kernel") == kernel:
            return int(f["bytes_per_launch"])
    return None


def geom(p):
    ny = p.ny
    ya, yb = ny // 3, 2 * (ny // 3) + 1
    return dict(nx=p.nx, ny=ny, nz=p.nz, Kx=p.nx // 3, Ky=ya + (ny - yb), drag=p.C_d != 0.0)


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
    m.prof_enable(prof_kernel)
    torch.cuda.synchronize()
    tw0 = time.perf_counter()
    ts = t
    for _ in range(steps):
        t, dt = m.time_step(t, dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - tw0
    nlaunch, kms = m.prof_read()
    m.prof_enable(-1)
    st1 = m.stats()
    res = dict(
        precision=precision, wall_s=wall, sim_days=(t - ts) / 86400.0,
        days_per_s=(t - ts) / 86400.0 / wall, ms_per_step=1e3 * wall / steps,
        mean_dt=(t - ts) / steps, rejects=st1["rejects"] - st0["rejects"],
        rhs_evals=st1["rhs_evals"] - st0["rhs_evals"], kernel=prof_kernel,
        kernel_launches=nlaunch, kernel_ms=kms / max(nlaunch, 1), t_end=t, dt_end=dt,
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


def cpu_baseline(wl, q_spun, t0, dt0, budget_s=30.0):
    """The FP64 oracle (scipy pocketfft, multithreaded FFTs) on the host: as
    many accepted ARK43_QGLeith steps as fit in ~budget_s, at least one."""
    from oracle import qgn_oracle as orc
    from qgn_fortran_amd.model import vertical_modes
    p = wl.params
    cores = min(os.cpu_count() or 1, 16)
    H, _L, Modes, EVals, qy = vertical_modes(wl.H, wl.S, wl.uBar, p.Htot)
    po = orc.Params(nx=p.nx, ny=p.ny, nz=p.nz, Lx=p.Lx, Ly=p.Ly, Htot=p.Htot, beta=p.beta, A2=p.A2, A8=p.A8,
                    r0=p.r0, C_d=p.C_d, TOL=p.TOL, is_explicit=p.is_explicit, dt_max=p.dt_max)
    om = orc.Model(po, H, Modes, EVals, wl.uBar, qy, workers=cores)
    stp = orc.ARK43(om, leith=True)
    q, t, dt = q_spun, t0, dt0
    tw = time.perf_counter()
    n = 0
    while True:
        q, t, dt = stp.step(q, t, dt)
        n += 1
        if time.perf_counter() - tw > budget_s * 0.5 or n >= 20:
            break
    wall = time.perf_counter() - tw
    return dict(value=(t - t0) / 86400.0 / wall, unit="simulated days/s", cores=cores, kind="port",
                sample=f"{n} accepted ARK43_QGLeith step(s) of the same 1024^2x16 state, FP64 numpy/scipy "
                       f"oracle, {wall:.1f} s")


def run_cnn(steps, warmup):
    """The closure network's BF16 MFMA kernels at the closure's shapes: train
    steps on 10 images of 256^2 x 16 channels (the training batch) and
    forward passes on 4 images (the 4-member coarse ensemble), with HIP-event
    timings and the dense-MFMA flops of each."""
    import torch
    from oracle import cnn_oracle as co
    from qgn_fortran_amd.cnn import CNN, to_bf16_bits
    cin, cout, H, W = 16, 8, 256, 256
    rng = np.random.default_rng(0)
    net = CNN(cin, cout, co.init_params(cin, cout, 1))
    macs_px = sum(o * 9 * i for i, o in co.shapes(cin, cout))        # per pixel, forward
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
    wl = configs.WORKLOADS[a.workload]()
    q, t0, dt0 = spin_up(wl, a.spinup)
    mixed, _ = run_gpu(configs.WORKLOADS[a.workload](), "mixed", q, t0, dt0, a.steps, a.warmup, a.kernel)
    fp64, _ = run_gpu(configs.WORKLOADS[a.workload](), "fp64", q, t0, dt0, a.steps, a.warmup, a.kernel)
    g = geom(wl.params)
    rs = 8  # mixed path: FP32 complex intermediates
    kb = kernel_bytes(a.kernel, g, rs)
    achieved = kb / (mixed["kernel_ms"] * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=8000.0, unit="GB/s",
                frac=round(achieved / 8000.0, 4), traffic=measured_traffic(a.workload, a.kernel), kernel=a.kernel,
                bytes_per_launch=kb, avg_launch_ms=mixed["kernel_ms"])
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
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
