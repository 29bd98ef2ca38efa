"""The learned-closure workflow on one GPU (no counterpart in the reference).

1. A fine DNS (FP32-mixed step) is started from a laminar profile with a
   large-scale divergence-free-after-projection perturbation and run until
   turbulent.
2. Every ``snap_every`` steps its state is box filtered onto the coarse grid
   (``pll_closure_samples``): the filtered field's features and the
   least-squares eddy-viscosity labels of the box subgrid stress go to a
   device dataset, and the filtered field itself is sampled into the coarse
   observer's statistics (the reference: filtered-DNS mean profile and stresses).
3. The network is trained on the GPU by the library's own MFMA forward and
   backward kernels with Adam, one snapshot per step; the last snapshots are
   held out.
4. Two coarse runs start from the filtered DNS state: without a model, and
   with the closure attached (its eddy viscosity evaluated once per cell for
   each RK stage).  Their mean profile, Reynolds stresses and wall shear are
   compared with the filtered DNS.

Device memory for the dataset comes from torch (plumbing only); every
computation is a libpll.so kernel.  Run as
``python -m powerllel_channel_sycl_amd.closure_pipeline`` for the Re_tau~395
case (fine 384x256x192, coarse 96x64x64); it prints one JSON line.
"""
import json
import math
import sys
import time
from dataclasses import asdict, dataclass

import numpy as np

from . import Channel, Closure, mesh_by_func

STAT_U, STAT_V, STAT_W, STAT_U2, STAT_V2, STAT_W2, STAT_UV, STAT_UW = 0, 1, 2, 3, 4, 5, 6, 7


@dataclass
class Config:
    # fine DNS: 2pi x pi x 2, tanh stretch 1.5, Re_b = 13750 (re = 6875)
    nx: int = 384
    ny: int = 256
    nz: int = 192
    lx: float = 2 * math.pi
    ly: float = math.pi
    lz: float = 2.0
    stretch: float = 1.5
    re: float = 6875.0
    dt: float = 0.004
    re_tau: float = 395.0          # nominal, sets the feature scales
    # coarse grid: factors 4 x 4 x 3
    fx: int = 4
    fy: int = 4
    fz: int = 3
    coarse_dt: float = 0.008
    coarse_precision: str = "fp64"
    # schedule (fine steps / coarse steps)
    amplitude: float = 0.3         # large-scale start perturbation, in bulk units
    seed: int = 1
    spin_steps: int = 7500         # fine steps before sampling (30 time units)
    sample_steps: int = 10000      # fine steps of the sampling window (40 time units)
    snap_every: int = 50           # fine steps between snapshots
    holdout: float = 0.2           # last fraction of snapshots kept for validation
    epochs: int = 3
    lr: float = 1e-3
    net_seed: int = 7
    les_spin_steps: int = 2500     # coarse steps before the coarse statistics
    les_steps: int = 10000         # coarse steps of the statistics window
    les_sample_every: int = 25


def say(msg):
    print(f"[closure] {msg}", flush=True)


def seed_turbulence(ch, amplitude, seed, lx, ly, zc):
    """Add a random large-scale perturbation (12 Fourier modes, |k| <= 4 in x
    and y, per component, shape 1 - (z-1)^2, rms ``amplitude``) to u, v, w on
    the host and refresh the halos; the first step's projection removes its
    divergence."""
    nx, ny = ch.nx, ch.ny
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 1.0) * lx / nx
    y = (np.arange(ny) + 0.5) * ly / ny
    X, Y = np.meshgrid(x, y, indexing="xy")
    shape = 1.0 - (zc - 1.0) ** 2
    for f in "uvw":
        a = ch.download(f)
        pert = np.zeros((ny, nx))
        for _ in range(12):
            kx, ky = rng.integers(1, 5), rng.integers(1, 5)
            pert += np.cos(2 * np.pi * kx * X / lx + rng.uniform(0, 2 * np.pi)) * \
                np.cos(2 * np.pi * ky * Y / ly + rng.uniform(0, 2 * np.pi))
        pert /= np.sqrt(np.mean(pert ** 2))
        a[1:-1, 1:-1, 1:-1] += amplitude * shape[:, None, None] * pert[None]
        ch.upload(f, a)
    ch.updateBoundVel()


def plane_stats(ch, nsamples):
    """Mean profile and Reynolds stresses (per cell-centre plane) of a
    context's accumulated statistics."""
    n = float(nsamples) * ch.nx * ch.ny
    m = {k: ch.stat_plane_sums(i) / n for k, i in
         (("u", STAT_U), ("v", STAT_V), ("w", STAT_W), ("u2", STAT_U2), ("v2", STAT_V2),
          ("w2", STAT_W2), ("uw", STAT_UW))}
    return dict(U=m["u"], uu=m["u2"] - m["u"] ** 2, vv=m["v2"] - m["v"] ** 2,
                ww=m["w2"] - m["w"] ** 2, uw=m["uw"] - m["u"] * m["w"])


def wall_re_tau(U, dzf1, re):
    """Re_tau of the mean wall stress nu U(1) / (dzf(1)/2), both walls."""
    tau = 0.5 * (U[0] + U[-1]) / (0.5 * dzf1) / re
    return math.sqrt(max(tau, 0.0)) * re


def profile_errors(ref, got):
    """Relative L2 errors (over the z planes) of each statistic."""
    return {k: float(np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])) for k in ref}


def run(cfg=Config(), log=say, save=None):
    """The workflow; `save` names a file for the trained closure."""
    import torch
    t_start = time.time()
    c = cfg
    ncx, ncy, ncz = c.nx // c.fx, c.ny // c.fy, c.nz // c.fz
    dzf_f = mesh_by_func(1, 0, c.stretch, c.nz, c.lz)
    dzf_c = mesh_by_func(1, 0, c.stretch, ncz, c.lz)
    zc = np.cumsum(np.concatenate([[0.5 * dzf_f[1]], 0.5 * (dzf_f[1:c.nz] + dzf_f[2:c.nz + 1])]))
    utau = c.re_tau / c.re
    scales = (1.0 / (c.re * utau * utau), utau * c.re / 100.0, 1.0 / c.re)
    out = dict(config=asdict(c), scales=dict(gscale=scales[0], yscale=scales[1], nuscale=scales[2]))

    fine = Channel(c.nx, c.ny, c.nz, c.lx, c.ly, c.lz, c.re, c.dt, dzf_f, precision="fp32_mixed")
    obs = Channel(ncx, ncy, ncz, c.lx, c.ly, c.lz, c.re, c.coarse_dt, dzf_c,
                  precision=c.coarse_precision)
    net = Closure(seed=c.net_seed)
    net.set_scales(*scales)
    try:
        # 1. fine DNS spin-up
        fine.initFlow("poi", with_noise=False)
        seed_turbulence(fine, c.amplitude, c.seed, c.lx, c.ly, zc)
        chunk = 1500
        for s in range(0, c.spin_steps, chunk):
            fine.step(min(chunk, c.spin_steps - s))
            u = fine.download("u")[1:-1, 1:-1, 1:-1]
            U = u.mean(axis=(1, 2))
            log(f"fine spin-up t={(s + min(chunk, c.spin_steps - s)) * c.dt:6.1f}  "
                f"Re_tau={wall_re_tau(U, dzf_f[1], c.re):6.1f}  urms(mid)="
                f"{float(np.sqrt(((u[c.nz // 2] - U[c.nz // 2]) ** 2).mean())):.4f}")
        # 2. snapshots: samples + observer statistics; fine wall statistics
        nsnap = c.sample_steps // c.snap_every
        ncell = ncx * ncy * ncz
        x = torch.empty((nsnap, ncell, 10), dtype=torch.float32, device="cuda:0")
        y = torch.empty((nsnap, ncell), dtype=torch.float32, device="cuda:0")
        obs.stat_enable((1, 1, 1, 1, 1, 1, 0, 1, 0, 0, 0))
        fine.stat_enable((1,) + (0,) * 10)
        t0 = time.time()
        for s in range(nsnap):
            fine.step(c.snap_every)
            net.samples(fine, obs, x[s].data_ptr(), y[s].data_ptr())
            obs.stat_accumulate()
            fine.stat_accumulate()
            if (s + 1) % max(1, nsnap // 5) == 0:
                log(f"snapshot {s + 1}/{nsnap} ({time.time() - t0:.1f} s)")
        ref = plane_stats(obs, nsnap)
        Uf = fine.stat_plane_sums(STAT_U) / (nsnap * c.nx * c.ny)
        out["dns"] = dict(re_tau=wall_re_tau(Uf, dzf_f[1], c.re), snapshots=nsnap,
                          time_units=c.sample_steps * c.dt)
        yl = y.cpu().numpy()
        out["labels"] = dict(mean=float(yl.mean()), zero_fraction=float((yl == 0).mean()),
                             capped_fraction=float((yl >= 50.0).mean()))
        log(f"DNS Re_tau={out['dns']['re_tau']:.1f}; label mean {yl.mean():.3f}, "
            f"zero {out['labels']['zero_fraction']:.2f}")
        # 3. training (one snapshot per Adam step), validation on held-out snapshots
        ntrain = max(1, int(round(nsnap * (1 - c.holdout))))
        val = list(range(ntrain, nsnap))

        def loss_on(idx):
            if not idx:
                return float("nan")
            return float(np.mean([net.gradient(x[s].data_ptr(), y[s].data_ptr(), ncell)[1]
                                  for s in idx]))
        base = float(np.mean([np.var(yl[s]) for s in val])) if val else 1.0
        out["training"] = dict(val_loss_before=loss_on(val))
        rng = np.random.default_rng(c.net_seed)
        t0 = time.time()
        for ep in range(c.epochs):
            for s in rng.permutation(ntrain):
                net.train_step(x[int(s)].data_ptr(), y[int(s)].data_ptr(), ncell, c.lr)
            net.synchronize()
            log(f"epoch {ep + 1}/{c.epochs}: held-out mse {loss_on(val):.4f} "
                f"(label variance {base:.4f}; {time.time() - t0:.1f} s)")
        out["training"].update(val_loss=loss_on(val), label_variance=base,
                               train_loss=loss_on(list(range(min(ntrain, 8)))),
                               adam_steps=c.epochs * ntrain, samples_per_step=ncell,
                               seconds=time.time() - t0)
        if save:
            net.save(save)
            log(f"closure saved to {save}")
        del x, y
        torch.cuda.empty_cache()
        # 4. coarse runs from the filtered DNS state
        runs = {}
        for name, closure in (("no_model", None), ("closure", net)):
            ch = Channel(ncx, ncy, ncz, c.lx, c.ly, c.lz, c.re, c.coarse_dt, dzf_c,
                         precision=c.coarse_precision)
            try:
                ch.filter_from(fine)
                if closure is not None:
                    ch.attach_closure(closure)
                t0 = time.time()
                ch.step(c.les_spin_steps)
                ch.stat_enable((1, 1, 1, 1, 1, 1, 0, 1, 0, 0, 0))
                nsamp = c.les_steps // c.les_sample_every
                for _ in range(nsamp):
                    ch.step(c.les_sample_every)
                    ch.stat_accumulate()
                ch.synchronize()
                st = plane_stats(ch, nsamp)
                ok = ch.checkNaN("u")
                nut = ch.closure_nut() if closure is not None else None
                runs[name] = dict(stats=st, seconds=time.time() - t0, finite=ok,
                                  re_tau=wall_re_tau(st["U"], dzf_c[1], c.re),
                                  nut_over_nu_mean=(float(nut.mean() * c.re) if nut is not None
                                                    else 0.0))
                if closure is not None:
                    ch.attach_closure(None)
            finally:
                ch.close()
            log(f"coarse {name}: Re_tau {runs[name]['re_tau']:.1f} "
                f"({runs[name]['seconds']:.1f} s)")
        # 5. errors against the filtered DNS
        out["coarse"] = {}
        for name, r in runs.items():
            e = profile_errors(ref, r["stats"])
            e["re_tau"] = abs(r["re_tau"] - out["dns"]["re_tau"]) / out["dns"]["re_tau"]
            out["coarse"][name] = dict(errors=e, re_tau=r["re_tau"], finite=r["finite"],
                                       seconds=r["seconds"],
                                       nut_over_nu_mean=r["nut_over_nu_mean"])
        out["profiles"] = dict(dns={k: v.tolist() for k, v in ref.items()},
                               **{n: {k: v.tolist() for k, v in r["stats"].items()}
                                  for n, r in runs.items()})
        out["seconds"] = time.time() - t_start
        return out
    finally:
        net.close()
        obs.close()
        fine.close()


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=None, help="write the full result (with profiles) as JSON")
    ap.add_argument("--save", default=None, help="save the trained closure (pll_closure_save)")
    for f, v in asdict(Config()).items():
        ap.add_argument("--" + f.replace("_", "-"), type=type(v), default=v)
    a = vars(ap.parse_args(argv))
    path = a.pop("out")
    save = a.pop("save")
    res = run(Config(**a), save=save)
    if path:
        with open(path, "w") as fh:
            json.dump(res, fh)
    summary = {k: res[k] for k in ("dns", "labels", "training", "coarse", "seconds")}
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
