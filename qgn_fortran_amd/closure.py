"""Learned subgrid PV closure for coarse grids, end to end on the GPU.

The reference's only closure is the QG-Leith hyperviscosity
(ARK43_QGLeith.f90:62-79).  This module trains the network of cnn.py to
stand in for it on a coarse grid and judges it against the high-resolution
model that produced its training data:

1. data: a batch of high-resolution members (one seed each) is spun up and
   sampled.  Each snapshot q_H is filtered spectrally and truncated to the
   coarse grid, q_c = T q_H, with T = truncation x the Gaussian
   exp(-k^2 D^2 / 24) of width D = `filter_cells` coarse cells (0: the
   sharp truncation alone).  The subgrid forcing is
   S = T RHS_H(q_H) - RHS_c(q_c): what the coarse right-hand side misses of
   the filtered high-resolution one.
   Inputs are the physical q_c and psi_c per layer; the target is S
   physical per layer.
2. training on the GPU (hand-written BF16 MFMA forward/backward, Adam), with
   the last `holdout` members (seeds) held out.
3. offline skill on the held-out members: R^2 = 1 - |S_pred - S|^2 / |S|^2,
   next to no closure (R^2 = 0) and to the QG-Leith forcing -A8L k^8 q_c.
4. online: coarse runs from a coarse-grained snapshot with the network, with
   QG-Leith (also the fallback if a network run fails) and with neither.
   Their time-mean KE spectra are scored against the time-mean KE spectrum
   of the filtered high-resolution run over the same window.

Run it as ``python -m qgn_fortran_amd.closure`` (full size: 1024^2 -> 256^2,
8 layers, 4 seeds x 50 snapshots) or with smaller sizes for tests.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import background, initial
from .cnn import CNN, to_bf16_bits
from .model import Batch, Dfti, Params, vertical_modes


@dataclasses.dataclass
class Setup:
    n_hi: int = 1024
    n_lo: int = 256
    nz: int = 8
    seeds: tuple = (11, 12, 13, 14, 15)
    holdout: int = 2         # the last `holdout` members are not trained on
    snapshots: int = 50
    every: int = 25          # accepted high-resolution steps between snapshots
    spinup: int = 3000       # accepted high-resolution steps before the first snapshot
    k_peak: float = 16.0
    rms: float = 1.0e-6
    precision: str = "mixed"
    epochs: int = 2000
    batch: int = 10
    lr: float = 1.0e-3
    seed: int = 5
    filter_cells: float = 2.0    # Gaussian filter width in coarse cells (0: sharp truncation)

    def params(self, n, scheme):
        # the exponential stratification of the expstrat workloads, 8 layers
        return Params(nx=n, ny=n, nz=self.nz, Lx=1.0e6, Ly=1.0e6, Htot=4000.0, beta=1.6e-11, A2=0.0, A8=0.0,
                      r0=1.0e-7, C_d=1.0e-5, TOL=1.0e-9, is_explicit=False, scheme=scheme,
                      precision=self.precision)


def structure(nz):
    H, S, u = background.exponential(nz, 4000.0, 1.0e-4, 5.0e-3, 800.0, 0.1)
    Hs, _L, Modes, EVals, qy = vertical_modes(H, S, u, 4000.0)
    return dict(H=Hs, Modes=Modes, EVals=EVals, uBar=u, qyBar=qy)


def keep_mask(n):
    """DeAlias's kept modes (QGN_Module.f90:194-204) on an n x n grid, as a
    boolean (n, n/2+1) array; (0,0) dropped."""
    m = np.zeros((n, n // 2 + 1), dtype=bool)
    rows = np.r_[0:n // 3, 2 * (n // 3) + 1:n]
    m[np.ix_(rows, np.arange(n // 3))] = True
    m[0, 0] = False
    return m


def truncate(spec, n_hi, n_lo, filter_cells=0.0, L=1.0e6):
    """Spectral truncation of full half spectra (.., n_hi, n_hi/2+1) to the
    coarse grid, scaled for the unnormalised transforms (q = c2r(q_hat)/n^2):
    the coarse field samples the same band-limited function.  With
    filter_cells > 0 the kept modes are first multiplied by the Gaussian
    exp(-k^2 D^2 / 24), D = filter_cells * L / n_lo."""
    h = n_lo // 2
    out = np.zeros(spec.shape[:-2] + (n_lo, h + 1), dtype=np.complex128)
    out[..., :h, :h] = spec[..., :h, :h]
    out[..., h + 1:, :h] = spec[..., n_hi - h + 1:, :h]
    out *= (n_lo / n_hi) ** 2
    if filter_cells > 0:
        D = filter_cells * L / n_lo
        out *= np.exp(-keep_k2(n_lo, L) * D * D / 24.0)
    return out


def generate(st: Setup, log=print):
    """Training pairs from the high-resolution model.  Returns a dict with
    X (N, n_lo, n_lo, 2 nz) float32 (q then psi), Y (N, n_lo, n_lo, nz)
    float32 (S), member index and time per image, the coarse-grained spectral
    states and the filtered KE spectra."""
    nz, nh, nl = st.nz, st.n_hi, st.n_lo
    vs = structure(nz)
    M = len(st.seeds)
    hi = Batch(st.params(nh, "ARK43_QGLeith"), [vs] * M)
    lo = Batch(st.params(nl, "ARK43_QGLeith"), [vs] * M)
    dft = Dfti(nl, nl, M * nz)
    q0 = np.concatenate([initial.random_phase_pv(nh, nh, nz, st.k_peak, st.rms, s, 1.0e6, 1.0e6) for s in st.seeds])
    hi.initialize(q0)
    hi.set_time(0.0, 750.0)
    tw = time.perf_counter()
    done = 0
    while done < st.spinup:
        hi.step(min(500, st.spinup - done))
        info = hi.times()
        done = int(info["steps"].min())
        log(f"# spin-up {done}/{st.spinup} steps, t = {info['t'].mean() / 86400:.1f} days, "
            f"{time.perf_counter() - tw:.1f} s")
    mask = keep_mask(nl)
    X, Y, member, tsnap, qcs, ke_ref = [], [], [], [], [], []
    nn = float(nl * nl)
    wall_hi = 0.0
    for k in range(st.snapshots):
        if k:
            target = done + st.every
            t0 = time.perf_counter()
            while done < target:
                hi.step(target - done)
                info = hi.times()
                done = int(info["steps"].min())
            wall_hi += time.perf_counter() - t0
        info = hi.times()
        qh = hi.get_state()
        rh = hi.get_rhs(qh)
        qc = truncate(qh, nh, nl, st.filter_cells) * mask
        S = truncate(rh, nh, nl, st.filter_cells) * mask - lo.get_rhs(qc)
        psi = lo.get_psi(qc)
        qp = dft.backward(qc) / nn
        pp = dft.backward(psi) / nn
        sp = dft.backward(S) / nn
        lo.set_state(qc)
        ke, _ = lo.spectra()
        for m in range(M):
            sl = slice(m * nz, (m + 1) * nz)
            X.append(np.concatenate([qp[sl], pp[sl]]).transpose(1, 2, 0).astype(np.float32))
            Y.append(sp[sl].transpose(1, 2, 0).astype(np.float32))
            member.append(m)
            tsnap.append(float(info["t"][m]))
            ke_ref.append(ke[sl])
        qcs.append(qc)
        if (k + 1) % 10 == 0:
            log(f"# snapshot {k + 1}/{st.snapshots}, t = {info['t'].mean() / 86400:.1f} days")
    steps_hi = st.every * (st.snapshots - 1)
    out = dict(X=np.stack(X), Y=np.stack(Y), member=np.array(member), t=np.array(tsnap), qc=qcs,
               ke_ref=np.stack(ke_ref), vs=vs, wall_hi=wall_hi, steps_hi=steps_hi,
               days_hi=float(np.mean(info["t"] - np.array(tsnap[:M])) / 86400.0))
    hi.close()
    lo.close()
    dft.close()
    return out


def _he_init(cin, cout, seed, last_scale=0.1):
    """He-normal weights (std sqrt(2/(9 ci))), zero biases, the last layer
    scaled by last_scale; same parameter layout as include/qgn.h describes."""
    rng = np.random.default_rng(seed)
    dims = [(cin, 32), (32, 32), (32, 32), (32, cout)]
    parts = []
    for l, (i, o) in enumerate(dims):
        W = rng.standard_normal((o, 9, i)) * np.sqrt(2.0 / (9 * i))
        if l == len(dims) - 1:
            W *= last_scale
        parts += [W.ravel(), np.zeros(o)]
    return np.concatenate(parts).astype(np.float32)


def train_network(st: Setup, data, log=print):
    import torch
    nz = st.nz
    M = int(data["member"].max()) + 1
    tr = data["member"] < M - st.holdout
    X, Y = data["X"], data["Y"]
    s_in = 1.0 / X[tr].reshape(-1, 2 * nz).std(axis=0)
    s_out = Y[tr].reshape(-1, nz).std(axis=0)
    xd = torch.from_numpy(to_bf16_bits(X * s_in)).cuda()
    yd = torch.from_numpy((Y / s_out).astype(np.float32)).cuda()
    idx = np.flatnonzero(tr)
    net = CNN(2 * nz, nz, _he_init(2 * nz, nz, st.seed))
    rng = np.random.default_rng(st.seed)
    out = torch.empty((st.batch,) + Y.shape[1:3] + (nz,), device="cuda", dtype=torch.float32)
    hist = []
    tw = time.perf_counter()
    nsteps = 0
    for ep in range(st.epochs):
        lr = st.lr if ep < st.epochs // 2 else (st.lr / 4 if ep < 5 * st.epochs // 6 else st.lr / 16)
        perm = rng.permutation(idx)
        losses = []
        for b in range(0, len(perm) - st.batch + 1, st.batch):
            sel = torch.from_numpy(np.sort(perm[b:b + st.batch])).cuda()
            losses.append(net.train_step(xd.index_select(0, sel), yd.index_select(0, sel), out, lr))
            nsteps += 1
        hist.append(float(np.mean(losses)))
        if (ep + 1) % max(1, st.epochs // 10) == 0:
            log(f"# epoch {ep + 1}/{st.epochs}: loss {hist[-1]:.4f}")
    torch.cuda.synchronize()
    return net, s_in, s_out, dict(loss=hist, steps=nsteps, wall_s=time.perf_counter() - tw)


def offline_skill(st: Setup, data, net, s_in, s_out):
    """R^2 of the network, of no closure and of the QG-Leith forcing on the
    held-out members (together, and per member)."""
    import torch
    nz, nl = st.nz, st.n_lo
    M = int(data["member"].max()) + 1
    held = list(range(M - st.holdout, M))
    out = torch.empty((1, nl, nl, nz), device="cuda", dtype=torch.float32)
    num = {m: np.zeros(nz) for m in held}
    den = {m: np.zeros(nz) for m in held}
    num_l = {m: np.zeros(nz) for m in held}
    lo = Batch(st.params(nl, "ARK43_QGLeith"), [data["vs"]])
    dft = Dfti(nl, nl, nz)
    k2 = keep_k2(nl)
    for g in np.flatnonzero(data["member"] >= M - st.holdout):
        m, snap = int(data["member"][g]), int(g) // M      # images are stored snapshot by snapshot
        X, Y = data["X"][g], data["Y"][g]
        xd = torch.from_numpy(to_bf16_bits(X[None] * s_in)).cuda()
        net.forward(xd, out)
        pred = out.cpu().numpy()[0] * s_out
        num[m] += np.sum((pred - Y) ** 2, axis=(0, 1))
        den[m] += np.sum(Y ** 2, axis=(0, 1))
        # QG-Leith forcing of the same coarse state: -A8L k^8 q_c
        qc = data["qc"][snap][m * nz:(m + 1) * nz]
        lo.set_state(qc)
        a8 = lo.leith()
        fl = dft.backward(-a8[:, None, None] * k2[None] ** 4 * qc) / float(nl * nl)
        num_l[m] += np.sum((fl.transpose(1, 2, 0) - Y) ** 2, axis=(0, 1))
    lo.close()
    dft.close()
    tn, td, tl = sum(num.values()), sum(den.values()), sum(num_l.values())
    return dict(r2_network=float(1.0 - tn.sum() / td.sum()), r2_no_closure=0.0,
                r2_qg_leith=float(1.0 - tl.sum() / td.sum()),
                r2_network_per_layer=[float(v) for v in 1.0 - tn / td],
                r2_qg_leith_per_layer=[float(v) for v in 1.0 - tl / td],
                r2_network_per_member={str(st.seeds[m]): float(1.0 - num[m].sum() / den[m].sum()) for m in held},
                r2_qg_leith_per_member={str(st.seeds[m]): float(1.0 - num_l[m].sum() / den[m].sum()) for m in held},
                test_images=int(np.sum(data["member"] >= M - st.holdout)))


def keep_k2(n, L=1.0e6):
    kx = 2 * np.pi / L * np.arange(n // 2 + 1)
    ky = 2 * np.pi / L * np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    return kx[None, :] ** 2 + ky[:, None] ** 2


def _spectrum_errors(kb, rb):
    """Errors of a depth-summed KE spectrum kb against the reference rb, per
    wavenumber band."""
    rel = (kb - rb) / rb
    return dict(
        ke_total_rel_err=float(kb.sum() / rb.sum() - 1.0),
        ke_first_band_rel_err=float(rel[0]),
        # every band weighted alike (the small scales dominate) ...
        ke_spectrum_rms_rel_err=float(np.sqrt(np.mean(rel ** 2))),
        ke_spectrum_max_rel_err=float(np.max(np.abs(rel))),
        # ... and weighted by energy (the energetic scales dominate)
        ke_spectrum_l2_rel_err=float(np.linalg.norm(kb - rb) / np.linalg.norm(rb)),
        ke_spectrum_rel_per_band=[float(v) for v in rel])


def online(st: Setup, data, net, s_in, s_out, log=print):
    """Coarse runs of every member, as one batch, from their first
    coarse-grained snapshots over the snapshot window: network closure
    (QG-Leith fallback), the network on top of QG-Leith, QG-Leith, and
    neither.  Time-mean KE spectra are scored against the filtered
    high-resolution ones, for the held-out member and for the mean over all
    members (the coarse runs leave the high-resolution trajectories within
    days, so the comparison is statistical, and the largest scales, a few
    modes each, need the member mean); the cost is per member and simulated
    day."""
    nz, nl = st.nz, st.n_lo
    M = int(data["member"].max()) + 1
    held = list(range(M - st.holdout, M))
    te = np.flatnonzero(data["member"] == M - 1)
    t0, t1 = data["t"][te[0]], data["t"][te[-1]]
    band = slice(1, nl // 3)
    ref_all = np.stack([data["ke_ref"][data["member"] == m].mean(axis=0) for m in range(M)])[:, :, band].sum(axis=1)
    ref_held = ref_all[held].mean(axis=0)
    q_start = data["qc"][0]
    t_start = data["t"][:M]
    res = {}
    for method in ("network", "network_leith", "qg_leith", "none"):
        scheme = "ARK43_QGLeith" if method in ("qg_leith", "network_leith") else "ARK43"
        m = Batch(st.params(nl, scheme), [data["vs"]] * M)
        if method.startswith("network"):
            m.set_closure(net, s_in, s_out)
        m.set_state(q_start)
        m.set_time(t_start, 750.0)
        kes, fallback, failed = [], False, False
        tw = time.perf_counter()
        steps = 0
        while True:
            m.step(20)
            try:
                info = m.times()
            except Exception as e:   # non-finite or too many rejections
                failed = True
                log(f"# {method}: {e}")
                break
            steps = int(info["steps"][M - 1])
            ke, _ = m.spectra()
            if not np.all(np.isfinite(ke)) or ke.sum() > 100 * M * ref_held.sum():
                failed = True
                break
            if info["t"].min() >= t0 + 0.1 * (t1 - t0):
                kes.append(ke.reshape(M, nz, -1)[:, :, band].sum(axis=1))   # (member, band)
            if info["t"].min() >= t1:
                break
        wall = time.perf_counter() - tw
        days = float(np.mean(info["t"] - t_start)) / 86400.0 if not failed else float("nan")
        m.close()
        if failed and method == "network":
            fallback = True
        if failed or not kes:
            res[method] = dict(failed=True, fallback_to_qg_leith=fallback)
            continue
        ke = np.mean(kes, axis=0)
        # the top-level errors: the mean spectrum of the held-out members
        res[method] = dict(failed=False, steps=steps, wall_s=wall, sim_days=days,
                           wall_s_per_member_day=wall / days / M,
                           **_spectrum_errors(ke[held].mean(axis=0), ref_held))
        res[method]["per_held_out_member"] = {str(st.seeds[h]): _spectrum_errors(ke[h], ref_all[h]) for h in held}
        res[method]["all_members"] = _spectrum_errors(ke.mean(axis=0), ref_all.mean(axis=0))
        log(f"# {method}: " + str({k: v for k, v in res[method].items() if k != "ke_spectrum_rel_per_band"}))
    if res["network"].get("failed"):
        res["network_effective"] = dict(res["qg_leith"], fallback_to_qg_leith=True)
    return res


def run(st: Setup, log=print):
    t0 = time.perf_counter()
    data = generate(st, log)
    gen_s = time.perf_counter() - t0
    net, s_in, s_out, th = train_network(st, data, log)
    off = offline_skill(st, data, net, s_in, s_out)
    log(f"# offline: {off}")
    on = online(st, data, net, s_in, s_out, log)
    M = int(data["member"].max()) + 1
    hi_per_day = data["wall_hi"] / max(data["days_hi"], 1e-30) / M    # per member and simulated day
    out = dict(setup=dataclasses.asdict(st), generate_s=gen_s, images=int(len(data["X"])),
               train=dict(loss_first=th["loss"][0], loss_last=th["loss"][-1], steps=th["steps"],
                          wall_s=th["wall_s"]),
               offline=off, online=on, hi_wall_s_per_member_day=hi_per_day)
    for k in ("network", "network_leith", "qg_leith"):
        if not on.get(k, {}).get("failed", True):
            out[f"speedup_{k}_vs_hi"] = hi_per_day / on[k]["wall_s_per_member_day"]
    net.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-hi", type=int, default=1024)
    ap.add_argument("--n-lo", type=int, default=256)
    ap.add_argument("--snapshots", type=int, default=50)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--spinup", type=int, default=3000)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=1.0e-3)
    ap.add_argument("--filter-cells", type=float, default=2.0)
    ap.add_argument("--seeds", default="11,12,13,14,15")
    ap.add_argument("--holdout", type=int, default=2)
    ap.add_argument("--init-seed", type=int, default=5, help="seed of the initial weights and of the batch order")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():     # device memory for training comes from torch
        raise SystemExit("no GPU")
    st = Setup(n_hi=a.n_hi, n_lo=a.n_lo, snapshots=a.snapshots, every=a.every, spinup=a.spinup, epochs=a.epochs,
               lr=a.lr, filter_cells=a.filter_cells, seeds=tuple(int(x) for x in a.seeds.split(",")),
               holdout=a.holdout, seed=a.init_seed)
    res = run(st, log=lambda s: print(s, file=sys.stderr, flush=True))
    line = json.dumps(res)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
