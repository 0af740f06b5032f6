"""Seeded scene batches for the batched path: 64 scenes of uneven size, band
count and library, a quarter of them snapshot-mosaic frames demosaiced on the
device.

Scene i (seed s): a 2048x1088 u16 4x4 snapshot mosaic demosaiced to
272x512x16 when i % 4 == 3, else height U{128..2048}, width U{128..1024},
bands from {16, 113, 224}; library of U{1..32} targets. Pixels come from the
seeded generator (include/hsi_synth.h) with seed s + i and 0.5% anomalies;
the mosaic frames are that generator's frame, quantised to 16 bits and laid
out in the 4x4 filter pattern (band b at mosaic position b) before the
device demosaics it back.
"""
from __future__ import annotations

import numpy as np


def uneven_scene_specs(seed, count=64):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        t = int(rng.integers(1, 33))
        if i % 4 == 3:
            h, w, b, mosaic = 272, 512, 16, True
        else:
            h, w = int(rng.integers(128, 2049)), int(rng.integers(128, 1025))
            b, mosaic = int(rng.choice([16, 113, 224])), False
        specs.append(dict(index=i, height=h, width=w, bands=b, targets=t, mosaic=mosaic, seed=seed + i))
    return specs


def make_scene(H, torch, ctx, spec, threads=16, outputs=True):
    """The scene's device cube (bands, plane stride) float32, its library and
    (with outputs) empty device maps for all four detectors."""
    h, w, b, t = spec["height"], spec["width"], spec["bands"], spec["targets"]
    cfg = H.synth_config(h, w, b, min(t, 8), seed=spec["seed"], anomaly_fraction=0.005)
    lib = H.synth_library(H.synth_config(h, w, b, t, seed=spec["seed"]))
    stride = H.plane_stride(h, w)
    if spec["mosaic"]:
        host = H.synth_frame(cfg, 0, threads=threads)[:, :h * w].reshape(b, h, w)
        raw = np.zeros((h * 4, w * 4), np.uint16)
        for band in range(b):
            oy, ox = divmod(band, 4)
            raw[oy::4, ox::4] = np.clip(np.round(host[band] * 65535.0), 0, 65535).astype(np.uint16)
        raw_d = torch.from_numpy(raw.view(np.int16)).cuda()
        cube = torch.zeros((b, stride), dtype=torch.float32, device="cuda")
        H.demosaic(ctx, raw_d, H.RAW_U16, h * 4, w * 4, 4, b, cube)
        torch.cuda.synchronize()
        del raw_d
    else:
        cube = torch.from_numpy(H.synth_frame(cfg, 0, threads=threads)).cuda()
    scene = dict(cube=cube, height=h, width=w, library=lib)
    if outputs:
        p = h * w
        f32 = dict(dtype=torch.float32, device="cuda")
        i32 = dict(dtype=torch.int32, device="cuda")
        scene.update(sam_score=torch.empty(p, **f32), sam_target=torch.empty(p, **i32),
                     cem_score=torch.empty(p, **f32), cem_target=torch.empty(p, **i32),
                     rx=torch.empty(p, **f32), ace_score=torch.empty(p, **f32),
                     ace_target=torch.empty(p, **i32))
    return scene
