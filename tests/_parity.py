"""Helpers that put the HIP path's structure-of-arrays outputs next to the
oracle's reference-layout (CLRay / Reservoir) arrays."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle as O
from smallproject_amd import _lib, configs

# FP32 tolerance for float fields when a comparison is not strict:
# |a - b| <= RTOL*|b| + ATOL on >= 99.9% of pixels
RTOL, ATOL, PASS_FRACTION = 1e-5, 1e-7, 0.999

FACE_OF = {(1, 0, 0): 0, (0, 1, 0): 1, (0, 0, 1): 2, (-1, 0, 0): 3, (0, -1, 0): 4, (0, 0, -1): 5, (0, 0, 0): 7}


def expected_hitword(hits: np.ndarray) -> np.ndarray:
    side = hits["side"].astype(np.uint32)
    ds = np.take_along_axis(hits["D"], side.reshape(-1, 1).astype(np.int64), axis=1).reshape(-1)
    return hits["voxel"] | (side << 24) | ((ds > 0).astype(np.uint32) << 26)


def nlight_face(res: np.ndarray) -> np.ndarray:
    nl = res["Nlight"].astype(np.int64)
    out = np.full(len(res), 255, np.uint32)
    for k, v in FACE_OF.items():
        out[np.all(nl == np.array(k), axis=1)] = v
    return out


def split_reservoirs(res: np.ndarray):
    a = np.stack([res["wsum"], res["W"], res["pHat"], res["invPdf"]], axis=1)
    b = np.concatenate([res["pos"], res["lightIndex"].view(np.float32).reshape(-1, 1)], axis=1)
    m = res["M"] | (nlight_face(res) << 24)
    return a.astype(np.float32), b.astype(np.float32), m.astype(np.uint32)


def close(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return (np.abs(a - b) <= RTOL * np.abs(b) + ATOL) | ((a == b) | (np.isnan(a) & np.isnan(b)))


class Comparison:
    def __init__(self):
        self.lines = []
        self.ok = True

    def exact(self, name: str, got: np.ndarray, want: np.ndarray, need: float = 1.0):
        eq = np.asarray(got).view(np.uint32) == np.asarray(want).view(np.uint32) if np.asarray(got).dtype == np.float32 \
            else np.asarray(got) == np.asarray(want)
        frac = float(np.mean(eq)) if eq.size else 1.0
        self.lines.append(f"{name}: bit-exact {frac:.6f} ({int((~eq).sum())} differ)")
        if frac < need:
            self.ok = False
        return eq

    def tol(self, name: str, got: np.ndarray, want: np.ndarray, need: float = PASS_FRACTION):
        ok = close(got, want)
        frac = float(np.mean(ok)) if ok.size else 1.0
        exact = float(np.mean(np.asarray(got, np.float32).view(np.uint32) == np.asarray(want, np.float32).view(np.uint32)))
        self.lines.append(f"{name}: within tol {frac:.6f}, bit-exact {exact:.6f}")
        if frac < need:
            self.ok = False
        return ok

    def report(self) -> str:
        return "\n".join(self.lines)


def params_for_oracle(p) -> O.Params:
    return O.Params.from_buffer_copy(bytes(p))


class OracleMirror:
    """The oracle's copy of one World's frames: same world, lights, emission
    LUT and configuration, rendered with the params the GPU frame used."""

    def __init__(self, w, sc: configs.SceneConfig):
        W, H = w.cfg.width, w.cfg.height
        self.learned = sc.source == _lib.SOURCE_LEARNED
        self.view = O.View(w.grid(), w.bricks())
        lut = np.array(list(w.cfg.emit_lut), np.float32)
        ocfg = O.make_config(W, H, sc.point, gfxexp=1, quirks=sc.quirks,
                             source=O.SOURCE_SLOTS if self.learned else sc.source, lut=lut,
                             spatial_mis=bool(sc.flags & _lib.FLAG_MIS_SPATIAL))
        self.frames = O.Frames(ocfg, self.view, w.lights())
        self.nlights = len(w.lights())

    def render(self, p, r):
        """The frame with params p; r (the GPU renderer that rendered it)
        supplies the learned source's candidate slots."""
        if self.learned:
            self.frames.set_slots(*r.candidates(), r.cluster_tables(self.nlights))
        self.frames.render(params_for_oracle(p))

    def compare(self, tag: str, r, strict: bool = True) -> "Comparison":
        of = self.frames
        cmp = Comparison()
        hw, d, s = r.hits()
        oh = of.hits()
        cmp.exact(f"{tag} hitword", hw, expected_hitword(oh))
        cmp.exact(f"{tag} dist", d, oh["dist"])
        cmp.exact(f"{tag} seed", s, oh["seed"])
        for which, ores in ((1, of.initial()), (0, of.final())):
            a, b, m = r.reservoirs(which)
            ea, eb, em = split_reservoirs(ores)
            kind = "initial" if which == 1 else "final"
            cmp.exact(f"{tag} {kind} M", m & 0xFFFFFF, em & 0xFFFFFF)
            cmp.exact(f"{tag} {kind} light", b[:, 3].view(np.uint32), eb[:, 3].view(np.uint32))
            cmp.exact(f"{tag} {kind} face", m >> 24, em >> 24)
            for k, nm in enumerate(("wsum", "W", "pHat", "invPdf")):
                (cmp.exact if strict else cmp.tol)(f"{tag} {kind} {nm}", a[:, k], ea[:, k])
            (cmp.exact if strict else cmp.tol)(f"{tag} {kind} pos", b[:, :3], eb[:, :3])
        fr = r.frame()
        (cmp.exact if strict else cmp.tol)(f"{tag} radiance", fr[:, :3], of.frame()[:, :3])
        return cmp


def run_config(sc: configs.SceneConfig, frames: int, width: int | None = None, height: int | None = None,
               quirks: int | None = None, source: int | None = None, check_every: bool = True, setup=None,
               flags: int | None = None, before_frame=None, strict: bool = True, params=None):
    """Render `frames` frames of a configuration on the GPU and in the oracle
    with identical per-frame params; returns a Comparison per checked frame.

    With the learned source the oracle takes the GPU's candidate slots (light
    index and source pdf per pixel and slot) as its candidate lists, so every
    stage downstream of the proposal is compared exactly.  `setup(world,
    renderer)` runs once after the world is on the device;
    `before_frame(f, world, renderer, view, frames)` before each frame (the
    oracle's world view and frame state); `params(f, p)` may change frame
    f's render params, which both sides then take.  `strict` (the default) compares
    the float fields bit for bit too; strict=False accepts them within the
    FP32 tolerance above."""
    import dataclasses
    if quirks is not None or source is not None or flags is not None:
        sc = dataclasses.replace(sc, quirks=sc.quirks if quirks is None else quirks,
                                 source=sc.source if source is None else source,
                                 flags=sc.flags if flags is None else flags)
    w = configs.build_world(sc, width, height)
    w.force_sync()
    r = w.renderer
    if setup is not None:
        setup(w, r)
    mirror = OracleMirror(w, sc)
    results = []
    for f in range(frames):
        if before_frame is not None:
            before_frame(f, w, r, mirror.view, mirror.frames)
        pos, tgt = sc.camera(f)
        w.look_at(pos, tgt)
        p = w.prepare_frame()
        w.finish_frame()
        if params is not None:
            params(f, p)
        r.render(p)
        mirror.render(p, r)
        if not (check_every or f == frames - 1):
            continue
        results.append(mirror.compare(f"f{f}", r, strict))
    return results
