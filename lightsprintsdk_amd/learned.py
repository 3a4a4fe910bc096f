"""Learned gather: an MLP surrogate of the 1024-ray final gather.

Pipeline per scene (all on the GPU):
  cheap gather  -- the plain gather at low quality (16 hemisphere rays, 2 shadow
                   rounds per light), hemisphere and light parts kept apart
  features      -- log1p(parts/s) of the texel and window means over its lightmap
                   neighbours, position in the scene box, normal, albedo, spreads (32)
  MLP           -- 32-64-64-64-64-4, BF16 weights/activations on MFMA, FP32 accumulate;
                   predicts log1p(E/s) and its own RMS log error
  decide        -- keep the prediction where predicted error <= tau, queue the rest
  fallback      -- the plain 1024-ray gather on the queued texels
  filters       -- the updateLightmaps filter chain, as after the plain gather
Training data: texels of seeded box-room scenes (10x10x4 m room, 20 boxes, ~50k
triangles, 21 lightmaps of 256x256) baked by the plain gather.
The plain 1024-ray bake stays the ground truth: RMSE is measured against it, on the
filtered lightmaps both pipelines write.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import native as N
from . import scenes as S
from .bake import SceneBake, GatherParams, TEXEL_DTYPE
from .collider import DeviceBuffer

CHEAP_QUALITY = 16
CHEAP_LIGHT_ROUNDS = 0        # 0: the reference's quality/10+1 rounds (2 at quality 16)
BOUND_SAMPLE = 4              # the error bound uses every 4th texel's second cheap gather
LAMBDA = 0.5
NFEAT = 32
FEATURE_VERSION = 3      # 32 features, outputs relative to the baseline below


def base_log(X: np.ndarray) -> np.ndarray:
    """log1p(E/s) of the baseline: 5x5 hemisphere mean + the texel's own light part
    (features 9-11 and 3-5, stored as log1p(./s))."""
    X = np.asarray(X, np.float64)
    return np.log1p(np.expm1(X[:, 9:12]) + np.expm1(X[:, 3:6]))


def rmse_rel(a: np.ndarray, b: np.ndarray) -> float:
    """Lightmap RMSE relative to the reference: sqrt(mean((a-b)^2)) / sqrt(mean(b^2))
    over the RGB of every covered texel."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


class Surrogate:
    """The MLP on the GPU (lsg_mlp_*)."""

    def __init__(self, max_batch: int = 16384, seed: int = 0):
        self.max_batch = max_batch
        h = ctypes.c_void_p()
        N.check(N.lib().lsg_mlp_create(max_batch, ctypes.byref(h)), "lsg_mlp_create")
        self._h = h.value
        self.nparams = int(N.lib().lsg_mlp_num_params())
        self.set_params(self.init_params(seed))

    @staticmethod
    def layer_shapes():
        return [(NFEAT, 64), (64, 64), (64, 64), (64, 64), (64, 16)]

    def init_params(self, seed=0):
        rng = np.random.default_rng(seed)
        out = []
        for k, (i, o) in enumerate(self.layer_shapes()):
            w = rng.normal(size=(i, o)) * np.sqrt(2.0 / i)
            if k == 4:
                w *= 0.1
            out += [w.reshape(-1), np.zeros(o)]
        return np.concatenate(out).astype(np.float32)

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert len(p) == self.nparams
        N.check(N.lib().lsg_mlp_set_params(self._h, p.ctypes.data_as(ctypes.c_void_p)), "set_params")

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        N.check(N.lib().lsg_mlp_get_params(self._h, p.ctypes.data_as(ctypes.c_void_p)), "get_params")
        return p

    def get_grads(self):
        g = np.zeros(self.nparams, np.float32)
        N.check(N.lib().lsg_mlp_get_grads(self._h, g.ctypes.data_as(ctypes.c_void_p)), "get_grads")
        return g

    def step(self, dx, dt, n, lr, lam=LAMBDA, want_loss=False, offset_rows=0):
        loss = ctypes.c_float()
        N.check(N.lib().lsg_mlp_train_step(self._h, dx.ptr + 4 * NFEAT * offset_rows, dt.ptr + 12 * offset_rows, n, lam, lr,
                                           ctypes.byref(loss) if want_loss else None, None), "train_step")
        return loss.value if want_loss else None

    def forward(self, dx_ptr, n, dy_ptr, stream=None):
        N.check(N.lib().lsg_mlp_forward(self._h, dx_ptr, n, dy_ptr, stream), "mlp_forward")

    def train(self, X: np.ndarray, T: np.ndarray, epochs: int = 4, batch: int = 16384, lr: float = 2e-3,
              seed: int = 1, log=None):
        """Adam over shuffled mini-batches; returns per-epoch mean loss."""
        rng = np.random.default_rng(seed)
        n = len(X)
        perm = rng.permutation(n)
        dx = DeviceBuffer.from_array(np.ascontiguousarray(X[perm], np.float32))
        dt = DeviceBuffer.from_array(np.ascontiguousarray(T[perm], np.float32))
        batch = min(batch, self.max_batch)
        nsteps = max(1, n // batch)
        total = epochs * nsteps
        hist = []
        k = 0
        for ep in range(epochs):
            losses = []
            for s in range(nsteps):
                frac = k / max(total - 1, 1)
                cur = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * frac)))      # cosine decay
                want = (s % 32 == 0) or s == nsteps - 1
                l = self.step(dx, dt, batch, cur, want_loss=want, offset_rows=s * batch)
                if want:
                    losses.append(l)
                k += 1
            hist.append(float(np.mean(losses)))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f}")
        return hist

    def predict(self, X: np.ndarray) -> np.ndarray:
        dx = DeviceBuffer.from_array(np.ascontiguousarray(X, np.float32))
        dy = DeviceBuffer(16 * max(len(X), 1))
        self.forward(dx.ptr, len(X), dy.ptr)
        N.check(N.lib().lsg_synchronize(), "sync")
        return dy.download(np.float32, (len(X), 4))

    def close(self):
        if getattr(self, "_h", None):
            N.lib().lsg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedScene:
    """One box-room scene prepared for the learned gather: the plain bake plan,
    its device buffers, the cheap-gather parameters and the albedo table."""

    def __init__(self, scene: S.BakeScene, quality: int = 1024, seed: int = 1):
        self.scene = scene
        self.sb = SceneBake(scene, GatherParams(quality=quality, ray_length_min=S.minimal_safe_distance(scene.verts),
                                                seed=seed))
        self.sb.first_gather()
        p = self.sb.params
        self.cheap = GatherParams(quality=CHEAP_QUALITY, hemisphere=p.hemisphere, ray_length_min=p.ray_length_min,
                                  locality=p.locality, rug_distance=p.rug_distance, light_multiplier=p.light_multiplier,
                                  seed=p.seed ^ 0x2545F491, light_rounds=CHEAP_LIGHT_ROUNDS,
                                  environment_multiplier=p.environment_multiplier)
        n = len(self.sb.texels)
        self.n = n
        self.d_parts = DeviceBuffer(32 * max(n, 1))
        self.d_cheap_out = DeviceBuffer(16 * max(self.sb.nout, 1))
        self.d_x = DeviceBuffer(4 * NFEAT * max(n, 1))
        self.d_work = DeviceBuffer(48 * max(self.sb.nout, 1))
        self.d_cheap2 = DeviceBuffer(16 * max(self.sb.nout, 1))
        self.d_hist = DeviceBuffer(8 * 513)
        # the bound's texel sample and the filter plan are set up here, not on first use
        self.d_tex_sample = DeviceBuffer.from_array(np.ascontiguousarray(self.sb.texels[::BOUND_SAMPLE]))
        self.sb._filter_plan()
        self.d_y = DeviceBuffer(16 * max(n, 1))
        self.d_fb = DeviceBuffer(16 * max(n, 1))
        self.d_nfb = DeviceBuffer(16)
        self.d_albedo = DeviceBuffer.from_array(np.ascontiguousarray(scene.albedo, np.float32))
        self.scale = 1.0

    def plain(self, stream=None, filtering=True):
        """The plain 1024-ray gather of every texel into sb.d_out (the ground truth), then the
        filter chain into sb.d_rgbf."""
        self.sb.clear()
        self.sb.final_gather(stream)
        if filtering:
            self.sb.filter(stream)

    def plain_rgb(self):
        """Gathered RGBA of every texel (before the filters when plain/learned ran with
        filtering=False; after smoothing otherwise)."""
        N.check(N.lib().lsg_synchronize(), "sync")
        out = self.sb.d_out.download(np.float32, (max(self.sb.nout, 1), 4))
        return out[self.sb.texels["out_index"]]

    def final_rgb(self):
        """The filtered RGBF lightmap value of every texel (what updateLightmaps writes)."""
        N.check(N.lib().lsg_synchronize(), "sync")
        out = self.sb.d_rgbf.download(np.float32, (max(self.sb.nout, 1), 3))
        return out[self.sb.texels["out_index"]]

    def cheap_and_features(self, stream=None):
        L = N.lib()
        p = self.cheap.c()
        N.check(L.lsg_gather_parts_device(self.sb.gs._h, self.sb.d_tex.ptr, self.n, self.sb.d_sub.ptr, ctypes.byref(p),
                                          self.d_cheap_out.ptr, self.d_parts.ptr, stream), "gather_parts")
        sc = ctypes.c_float()
        N.check(L.lsg_learned_scale(self.d_parts.ptr, self.n, ctypes.byref(sc)), "scale")
        self.scale = float(sc.value)
        N.check(L.lsg_learned_features(self.sb.gs._h, self.sb.d_tex.ptr, self.n, self.sb.d_sub.ptr, self.d_parts.ptr,
                                       self.d_albedo.ptr, self.scale, self.sb.maps, len(self.sb.slots), self.sb.nout,
                                       self.d_work.ptr, self.d_x.ptr, stream), "features")

    def features(self):
        N.check(N.lib().lsg_synchronize(), "sync")
        return self.d_x.download(np.float32, (self.n, NFEAT))

    def targets(self, plain_rgb, X):
        """Training targets: log1p(E/s) of the plain gather minus the window baseline of
        the features X."""
        return (np.log1p(np.maximum(plain_rgb[:, :3], 0) / self.scale) - base_log(X)).astype(np.float32)

    def predict_rgb(self, net: "Surrogate"):
        """The network's irradiance for every texel (no fallback), from the current features."""
        X = self.features()
        return self.scale * np.expm1(base_log(X) + net.predict(X)[:, :3])

    def close(self):
        self.sb.gs.close()

    def cheap_only(self, stream=None):
        """The cheap gather alone written as the lightmap (no network), then the filters:
        the method the surrogate improves on."""
        self.sb.clear()
        p = self.cheap.c()
        N.check(N.lib().lsg_gather_device(self.sb.gs._h, self.sb.d_tex.ptr, self.n, self.sb.d_sub.ptr, len(self.sb.subs),
                                          ctypes.byref(p), self.sb.d_out.ptr, stream), "cheap gather")
        self.sb.filter(stream)

    def error_bound(self, net: Surrogate, target: float, stream=None):
        """Per-scene bound tau for a target relative RMSE of the kept texels, estimated
        from a second, independently seeded cheap gather (lsg_learned_bound).  Needs the
        features and network outputs of the current pass."""
        L = N.lib()
        p = self.cheap.c()
        p.seed = (self.cheap.seed * 2654435761 + 0x3C6EF372) & 0xFFFFFFFF
        ns = (self.n + BOUND_SAMPLE - 1) // BOUND_SAMPLE
        N.check(L.lsg_gather_device(self.sb.gs._h, self.d_tex_sample.ptr, ns, self.sb.d_sub.ptr, len(self.sb.subs),
                                    ctypes.byref(p), self.d_cheap2.ptr, stream), "second cheap gather")
        tau = ctypes.c_float(); est = ctypes.c_float()
        N.check(L.lsg_learned_bound(self.d_y.ptr, self.d_x.ptr, self.sb.d_tex.ptr, self.d_parts.ptr, self.d_cheap2.ptr,
                                    self.n, BOUND_SAMPLE, self.scale, target, self.d_hist.ptr, ctypes.byref(tau),
                                    ctypes.byref(est), stream), "bound")
        self.last_est_rmse = float(est.value)
        return float(tau.value)

    def learned(self, net: Surrogate, tau: float | None = None, target: float | None = None, stream=None,
                filtering=True):
        """Cheap gather + features + MLP + decision + plain gather of the fallback texels,
        writing every texel into sb.d_out, then the filter chain into sb.d_rgbf.  The
        bound is a fixed tau, or with target the per-scene bound of error_bound.
        Returns the fallback count."""
        L = N.lib()
        self.sb.clear()
        self.cheap_and_features(stream)
        net.forward(self.d_x.ptr, self.n, self.d_y.ptr, stream)
        if target is not None:
            tau = self.error_bound(net, target, stream)
        self.last_tau = tau
        N.check(L.lsg_memset(self.d_nfb.ptr, 0, 4), "memset")
        N.check(L.lsg_learned_decide(self.d_y.ptr, self.d_x.ptr, self.sb.d_tex.ptr, self.d_parts.ptr, self.n, self.scale, tau,
                                     self.sb.d_out.ptr, self.d_fb.ptr, self.d_nfb.ptr, stream), "decide")
        N.check(L.lsg_synchronize(), "sync")
        nfb = int(self.d_nfb.download(np.uint32, (1,))[0])
        if nfb:
            p = self.sb.params.c()
            N.check(L.lsg_gather_device(self.sb.gs._h, self.d_fb.ptr, nfb, self.sb.d_sub.ptr, len(self.sb.subs),
                                        ctypes.byref(p), self.sb.d_out.ptr, stream), "fallback gather")
        if filtering:
            self.sb.filter(stream)
        return nfb


def box_room_family(seed: int, lightmap_size: int = 256, target_tris: int = 50_000) -> S.BakeScene:
    """A box-room scene: the closed 10x10x4 m room with 20 boxes from seed."""
    return S.box_room(seed=seed, n_boxes=20, target_tris=target_tris, lightmap_size=lightmap_size)


def city_family(seed: int) -> S.BakeScene:
    """A city-and-terrain scene at 1/16 of the 10M-triangle one's area, same densities: a 512 m terrain
    (25 cm texels) and 312 boxes in 3 city lightmaps of 2048^2, sun from the sweep."""
    step = int(np.random.default_rng(seed).integers(0, 60))      # any direction of the sweep
    return S.city_terrain(seed=seed, n_boxes=312, grid_vertices=513, terrain_tiles=1, lightmap_size=2048, city_maps=3,
                          box_tris=100_000, sun_direction=S.sun_sweep(step))


FAMILIES = {
    "box_room": dict(make=box_room_family, train=range(1000, 1020), cal=(1500, 1501), heldout=range(2000, 2005),
                     weights="surrogate_box_room.npz", sample=None),
    "city": dict(make=city_family, train=range(5000, 5020), cal=(5500, 5501, 5502), heldout=range(6000, 6003),
                 weights="surrogate_city.npz", sample=1_000_000),
}


def _make(family, seed, **kw):
    return FAMILIES[family]["make"](seed, **kw)


def build_dataset(seeds, augment=1, family="box_room", sample=None, log=None, **kw):
    """(features, targets) of the reliable texels of every scene in seeds: features of
    `augment` independently seeded cheap gathers, targets log1p(E/s) of the plain
    1024-ray gather before the filters (relative to the window baseline).  sample: at
    most this many texels per scene and cheap gather, drawn uniformly."""
    Xs, Ts = [], []
    rng = np.random.default_rng(12345)
    for sd in seeds:
        ls = LearnedScene(_make(family, sd, **kw), seed=sd)
        ls.plain(filtering=False)
        ref = ls.plain_rgb()
        ok = np.nonzero(ref[:, 3] > 0)[0]
        base = ls.cheap.seed
        for a in range(augment):
            sel = ok if (sample is None or len(ok) <= sample) else np.sort(rng.choice(ok, sample, replace=False))
            ls.cheap.seed = base + 0x632BE5AB * a & 0xFFFFFFFF
            ls.cheap_and_features()
            X = ls.features()[sel]
            Xs.append(X)
            Ts.append(ls.targets(ref[sel], X))
        if log:
            log(f"scene {sd}: {len(ok)} texels, {len(sel)} x {augment} taken, scale {ls.scale:.4f}")
        ls.close()
        del ls
    return np.concatenate(Xs), np.concatenate(Ts)


def prepare_eval(seed, family="box_room", **kw) -> LearnedScene:
    """A scene with its plain 1024-ray filtered lightmaps kept as the reference."""
    ls = LearnedScene(_make(family, seed, **kw), seed=seed)
    ls.plain()
    ls.ref_final = ls.final_rgb()
    ls.ref_raw = ls.plain_rgb()
    return ls


def evaluate(ls: LearnedScene, net: Surrogate, tau: float | None = None, target: float | None = None):
    """(relative RMSE of the filtered learned+fallback lightmaps against the plain bake,
    fallback fraction)."""
    nfb = ls.learned(net, tau=tau, target=target)
    return rmse_rel(ls.final_rgb(), ls.ref_final), nfb / max(ls.n, 1)


TAU_CANDIDATES = (0.01, 0.0125, 0.015, 0.0175, 0.02, 0.025, 0.03, 0.035, 0.04, 0.045, 0.05, 0.06, 0.07, 0.08,
                  0.1, 0.12, 0.15, 0.2, 0.3)


def calibrate_tau(net: Surrogate, prepared, target_rmse: float = 0.018, candidates=TAU_CANDIDATES):
    """Largest fixed bound tau on the predicted error whose learned+fallback bake keeps
    the filtered-lightmap RMSE of every calibration scene at or below target_rmse."""
    best = candidates[0]
    for tau in candidates:
        worst = max(evaluate(ls, net, tau=tau)[0] for ls in prepared)
        if worst <= target_rmse:
            best = tau
        else:
            break
    return best


TARGET_CANDIDATES = (0.012, 0.014, 0.016, 0.018, 0.02, 0.022, 0.024, 0.026, 0.028, 0.03, 0.033, 0.036, 0.04, 0.045, 0.05)


def calibrate_target(net: Surrogate, prepared, target_rmse: float = 0.018, candidates=TARGET_CANDIDATES):
    """Largest per-scene estimate target (lsg_learned_bound) whose learned+fallback
    bake keeps the filtered-lightmap RMSE of every calibration scene at or below
    target_rmse."""
    best = candidates[0]
    for t in candidates:
        worst = max(evaluate(ls, net, target=t)[0] for ls in prepared)
        if worst <= target_rmse:
            best = t
        else:
            break
    return best


# ---- trained surrogate for the box-room scene family ---------------------------------------

_PKG = __file__.rsplit("/", 1)[0]
WEIGHTS = _PKG + "/" + FAMILIES["box_room"]["weights"]
TRAIN_SEEDS = FAMILIES["box_room"]["train"]       # 20 training scenes
CAL_SEEDS = FAMILIES["box_room"]["cal"]           # bound calibration
HELDOUT_SEEDS = FAMILIES["box_room"]["heldout"]   # 5 held-out scenes


def weights_path(family: str = "box_room") -> str:
    return _PKG + "/" + FAMILIES[family]["weights"]


def save_weights(path, net: Surrogate, bound: dict, meta: dict):
    """bound: {"target": per-scene estimate target, "tau": calibrated fixed bound}."""
    import json
    np.savez(path, params=net.get_params(), target=np.float32(bound["target"]), tau=np.float32(bound["tau"]),
             nfeat=np.int32(NFEAT), version=np.int32(FEATURE_VERSION), meta=np.array(json.dumps(meta)))


def load_weights(path=WEIGHTS, max_batch: int = 65536):
    """(Surrogate with the stored parameters, bound dict, meta) or None when absent or
    trained for other features."""
    import json
    import os
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    if int(z["nfeat"]) != NFEAT or "version" not in z.files or int(z["version"]) != FEATURE_VERSION \
            or "target" not in z.files:
        return None
    net = Surrogate(max_batch=max_batch)
    net.set_params(z["params"])
    return net, dict(target=float(z["target"]), tau=float(z["tau"])), json.loads(str(z["meta"]))


def train_family(family="box_room", augment=2, epochs=12, batch=8192, lr=2e-3, target_rmse=0.017, log=print):
    """Train the surrogate on the family's 20 training scenes; calibrate on its
    calibration scenes both the per-scene estimate target and, for comparison, a fixed
    tau.  Returns (net, bound, meta)."""
    F = FAMILIES[family]
    t0 = time.perf_counter()
    X, T = build_dataset(F["train"], augment=augment, family=family, sample=F["sample"], log=log)
    t1 = time.perf_counter()
    net = Surrogate(max_batch=65536, seed=0)
    hist = net.train(X, T, epochs=epochs, batch=batch, lr=lr, log=log)
    t2 = time.perf_counter()
    n = len(X)
    del X, T
    cal = [prepare_eval(s, family) for s in F["cal"]]
    bound = dict(target=calibrate_target(net, cal, target_rmse), tau=calibrate_tau(net, cal, target_rmse))
    for ls in cal:
        ls.close()
    meta = dict(family=family, train_seeds=[F["train"].start, F["train"].stop], cal_seeds=list(F["cal"]),
                augment=augment, sample=F["sample"], epochs=epochs, batch=batch, lr=lr, samples=int(n), loss=hist,
                target_rmse=target_rmse, dataset_s=round(t1 - t0, 2), train_s=round(t2 - t1, 2))
    log(f"trained on {n} texels in {t2 - t1:.1f}s, bound {bound}")
    return net, bound, meta


def train_box_room(**kw):
    return train_family("box_room", **kw)


def _sync():
    N.check(N.lib().lsg_synchronize(), "sync")


def heldout_report(net: Surrogate, bound: dict, seeds=HELDOUT_SEEDS, reps: int = 3, adaptive: bool = True, log=print,
                   family="box_room"):
    """Plain vs learned+fallback bake on each held-out scene: RMSE of the filtered
    lightmaps against the plain bake, fallback fraction and the time of both.  The
    bound is per scene (bound["target"]) or, with adaptive=False, the fixed tau."""
    kw = dict(target=bound["target"]) if adaptive else dict(tau=bound["tau"])
    rows = []
    for sd in seeds:
        ls = prepare_eval(sd, family)
        ls.cheap_only()
        r_cheap = rmse_rel(ls.final_rgb(), ls.ref_final)
        r_net, _ = evaluate(ls, net, tau=float("inf"))
        tp, tl = [], []
        for _ in range(reps):
            _sync(); a = time.perf_counter(); ls.plain(); _sync(); b = time.perf_counter()
            nfb = ls.learned(net, **kw); _sync(); c = time.perf_counter()
            tp.append(b - a); tl.append(c - b)
        r, f = rmse_rel(ls.final_rgb(), ls.ref_final), nfb / max(ls.n, 1)
        row = dict(seed=sd, texels=ls.n, rmse=r, rmse_cheap=r_cheap, rmse_network=r_net, fallback=f, tau=ls.last_tau,
                   plain_ms=1e3 * float(np.median(tp)), learned_ms=1e3 * float(np.median(tl)))
        rows.append(row)
        log(f"held-out {sd}: {ls.n} texels, rmse {r:.4f} (cheap {r_cheap:.4f}, network {r_net:.4f}), fallback {f:.3f} "
            f"(tau {ls.last_tau:.4g}), plain {row['plain_ms']:.1f} ms, learned {row['learned_ms']:.1f} ms, "
            f"x{row['plain_ms'] / row['learned_ms']:.2f}")
        ls.close()
    return rows


def _main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train / evaluate the learned-gather surrogate (box-room scene family)")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--family", default="box_room", choices=sorted(FAMILIES))
    ap.add_argument("--weights", default=None)
    ap.add_argument("--augment", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--fixed", action="store_true", help="report with the fixed tau instead of the per-scene bound")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    log = lambda m: print(f"[{time.perf_counter() - t0:7.1f}s] {m}", flush=True)
    a.weights = a.weights or weights_path(a.family)
    if a.cmd == "train":
        net, bound, meta = train_family(a.family, augment=a.augment, epochs=a.epochs, log=log)
        save_weights(a.weights, net, bound, meta)
        log(f"saved {a.weights}")
    else:
        got = load_weights(a.weights)
        if got is None:
            raise SystemExit(f"no weights at {a.weights}")
        net, bound, meta = got
    heldout_report(net, bound, seeds=FAMILIES[a.family]["heldout"], adaptive=not a.fixed, log=log, family=a.family,
                   reps=3 if a.family == "box_room" else 1)


if __name__ == "__main__":
    _main()
