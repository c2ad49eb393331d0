"""Learned proposal, host side: collect training rows from plain-MDAS runs
over seeded scenes, train the network on the GPU, and compare learned and
plain sampling by image error against an FP64 high-spp QMC ground truth.

Every step runs on the device through libmdas (collection mode of mdas_run,
mdas_mlp_train, scoring mode of mdas_run, mdas_render_qmc); numpy only
computes the image metrics.
"""
import math
import time

import numpy as np

from . import _lib, hip
from .kdgrid import KDGrid, Mlp, Scene
from .metrics import mse, rel_mse

SCENE_MBDOF = _lib.SCENE_SPHERES_MBDOF


def extra_bits(W, H, dim, b, spp0=4):
    """extraImgBits e giving about spp0 initial samples per pixel:
    4 * 2^(b*dim) * 4^e ~= spp0 * W * H."""
    return max(0, int(round(math.log(spp0 * W * H / (4.0 * 2 ** (b * dim)), 4))))


def make_grid(W, H, spp, dim, b=1, e=None, alpha=0.25, cand=4):
    e = extra_bits(W, H, dim, b) if e is None else e
    target = int(spp * W * H)
    cap = int(target * 1.5) + 4 * (1 << (b * dim + 2 * e)) + 1024
    return KDGrid(dim, cap, cand, b, e, alpha, float(W), float(H)), target


def collect(seeds, W, H, spp, dim=5, kind=SCENE_MBDOF, b=1, cap_rows=1 << 24, image_label=False,
            log=print):
    """Rows (device X [n,16], Y [n]) from plain-MDAS runs of the seeded scenes
    (labels: relative error reduction, or the image change if image_label)."""
    X = hip.DeviceBuffer(cap_rows * 64)
    Y = hip.DeviceBuffer(cap_rows * 4)
    n = 0
    for s in seeds:
        if n >= cap_rows:
            break
        g, target = make_grid(W, H, spp, dim, b)
        sc = Scene(kind, int(s), dim, float(W), float(H))
        mode = _lib.LEARN_COLLECT | (_lib.LEARN_IMAGE_LABEL if image_label else 0)
        g.set_learned(mode, None, X.ptr + n * 64, Y.ptr + n * 4, cap_rows - n)
        g.Run(sc, target)
        n += min(g.learned_rows(), cap_rows - n)
        g.close()
        sc.close()
    log(f"[learned] collected {n} rows from {len(seeds)} scenes ({W}x{H}, {spp} spp, D={dim})")
    return X, Y, n


def train(X, Y, n, epochs=4, batch=65536, lr=1e-3, seed=0, log=print):
    m = Mlp(seed)
    m.fit_norm(X, n)
    t0 = time.perf_counter()
    losses = m.train(X, Y, n, epochs, batch, lr, seed)
    log(f"[learned] trained {epochs} epochs x {math.ceil(n / batch)} steps in "
        f"{time.perf_counter() - t0:.2f} s; epoch losses {np.round(losses, 4).tolist()}")
    return m, losses


def heldout_loss(m, X, Y, n):
    """(network MSE, constant-predictor MSE) on device rows."""
    Xh = X.download(np.float32, n * 16).reshape(n, 16)
    Yh = Y.download(np.float32, n)
    y = m.forward(Xh)
    return float(np.mean((y - Yh) ** 2)), float(np.var(Yh))


def render(scene, W, H, spp, dim, mlp=None, b=1, image_label=False, alpha=0.25, kappa=1.0,
           kappa_y=None):
    """MDAS image (H, W, 4), samples, sampling + reconstruction ms (device;
    the second of two identical runs) for one scene.  kappa_y (optional) separates the
    network's exponent from kappa."""
    g, target = make_grid(W, H, spp, dim, b, alpha=alpha)
    if mlp is not None:
        g.set_learned(_lib.LEARN_SCORE | (_lib.LEARN_IMAGE_LABEL if image_label else 0), mlp)
        if kappa_y is not None:
            g.set_learned_exponents(0.0 if image_label else kappa, kappa_y)
        elif kappa != 1.0:
            g.set_learned_exponent(kappa)
    # the first run and reconstruction size and touch the grid's buffers (the
    # reconstruction's keys are allocated at its first call); the second
    # ones are timed
    g.Run(scene, target)
    g.integrate_image(W, H)
    g.set_timing(True)
    g.Run(scene, target)
    img = g.integrate_image(W, H)
    pt = g.phase_times()
    ms = pt["initial_ms"] + pt["errors_ms"] + pt["adaptive_ms"] + pt["recon_ms"]
    n = g.GetNumberOfSamples()
    g.close()
    return img, n, ms, pt["integrand_ms"]


def ground_truth(scene, W, H, spp=4096):
    return scene.render_qmc(W, H, spp, fp64=True)


def tune_kappa(mlp, seeds, W, H, spp, kappas, dim=5, kind=SCENE_MBDOF, gt_spp=4096,
               image_label=False, alpha=0.25, log=print):
    """Score exponent with the lowest mean image MSE on validation scenes
    (seeds disjoint from the evaluation)."""
    scenes = [Scene(kind, int(s), dim, float(W), float(H)) for s in seeds]
    gts = [ground_truth(sc, W, H, gt_spp) for sc in scenes]
    best, table = None, {}
    for k in kappas:
        e = float(np.mean([mse(render(sc, W, H, spp, dim, mlp, image_label=image_label,
                                      alpha=alpha, kappa=k)[0], gt)
                           for sc, gt in zip(scenes, gts)]))
        table[k] = e
        if best is None or e < table[best]:
            best = k
    for sc in scenes:
        sc.close()
    log(f"[learned] validation mse by kappa {table}; chose {best}")
    return best, table


def compare(mlp, seeds, W, H, spps, dim=5, kind=SCENE_MBDOF, gt_spp=4096, image_label=False,
            kappa=1.0, alpha=0.25, log=print):
    """Per seed and spp: image MSE / relMSE against the FP64 ground truth,
    sample counts and times of plain sampling (alpha), learned sampling
    (score exponent kappa) and plain sampling with the learned form's
    effective exponent kappa * alpha (the same flattening without the
    network)."""
    rows = []
    arms = (("plain", None, alpha), ("learned", mlp, alpha), ("plain_same_exponent", None, kappa * alpha))
    for s in seeds:
        sc = Scene(kind, int(s), dim, float(W), float(H))
        gt = ground_truth(sc, W, H, gt_spp)
        for spp in spps:
            r = dict(seed=int(s), spp=spp)
            for name, net, a in arms:
                img, n, ms, ims = render(sc, W, H, spp, dim, net, image_label=image_label,
                                         alpha=a, kappa=kappa)
                r[name] = dict(mse=mse(img, gt), rel_mse=rel_mse(img, gt), samples=n, ms=ms,
                               integrand_ms=ims)
            log(f"[learned] seed {s} spp {spp}: mse plain {r['plain']['mse']:.4e} learned "
                f"{r['learned']['mse']:.4e} plain(alpha*kappa) {r['plain_same_exponent']['mse']:.4e}")
            rows.append(r)
        sc.close()
    return rows


def _curve(rows, name, key, min_spp):
    """Mean error, samples and time (sampling + integrand ms) of one arm over
    seeds, per swept spp >= min_spp."""
    spps = sorted({r["spp"] for r in rows if r["spp"] >= min_spp})
    sel = [[r[name] for r in rows if r["spp"] == s] for s in spps]
    e = np.array([np.mean([a[key] for a in x]) for x in sel])
    n = np.array([np.mean([a["samples"] for a in x]) for x in sel])
    t = np.array([np.mean([a["ms"] + a["integrand_ms"] for a in x]) for x in sel])
    return spps, n, e, t


def _loglog(xs, ys, x):
    """log-log interpolation of y at x along (xs, ys) (xs increasing); None
    outside the range."""
    lx = np.log(xs)
    if not (lx[0] <= math.log(x) <= lx[-1]):
        return None
    return float(np.exp(np.interp(math.log(x), lx, np.log(ys))))


def samples_to_equal(rows, ref_spp, key="mse", arm="learned", min_spp=0):
    """Samples (and time) an arm needs to reach the plain method's mean error
    at ref_spp, relative to plain's there (log-log interpolation of the mean
    error over seeds).  Only swept points with spp >= min_spp enter (points
    below it, e.g. the initial grid alone, are the same for every arm).  If
    the arm is already below that error at its first point, that point is
    returned as an upper bound; None if its curve never gets there."""
    spps, n_p, e_p, t_p = _curve(rows, "plain", key, min_spp)
    _, n_l, e_l, t_l = _curve(rows, arm, key, min_spp)
    k = spps.index(ref_spp)
    target = e_p[k]
    le, ln = np.log(e_l), np.log(n_l)
    for j in range(len(spps) - 1):
        a, b = le[j], le[j + 1]
        if (a - np.log(target)) * (b - np.log(target)) <= 0 and a != b:
            w = (np.log(target) - a) / (b - a)
            n_eq = math.exp(ln[j] + w * (ln[j + 1] - ln[j]))
            t_eq = math.exp(np.log(t_l[j]) + w * (np.log(t_l[j + 1]) - np.log(t_l[j])))
            return dict(sample_ratio=n_eq / n_p[k], time_ratio=t_eq / t_p[k],
                        plain_err=float(target), plain_samples=float(n_p[k]))
    if le[0] <= np.log(target):
        return dict(sample_ratio=float(n_l[0] / n_p[k]), time_ratio=float(t_l[0] / t_p[k]),
                    plain_err=float(target), plain_samples=float(n_p[k]), bound="upper")
    return None


def error_at_equal_samples(rows, arm, ref_arm, key="mse", min_spp=0):
    """The arm's mean error over the reference arm's at the reference arm's
    sample count, per swept spp >= min_spp (the arm's error interpolated
    log-log in its samples; points outside the arm's sampled range are
    left out), and their geometric mean.  Sampling passes overshoot their
    target by different amounts per arm, so equal targets are not equal
    samples; this compares at equal samples."""
    spps, n_r, e_r, _ = _curve(rows, ref_arm, key, min_spp)
    _, n_a, e_a, _ = _curve(rows, arm, key, min_spp)
    order = np.argsort(n_a)
    ratios = {}
    for s, n, e in zip(spps, n_r, e_r):
        v = _loglog(n_a[order], e_a[order], n)
        if v is not None:
            ratios[str(s)] = v / float(e)
    gm = float(np.exp(np.mean(np.log(list(ratios.values()))))) if ratios else None
    return dict(ratios=ratios, geomean=gm)