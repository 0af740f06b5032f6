"""Helpers shared by the GPU parity tests (tolerances, device buffers)."""
import numpy as np

# Parity bars. Scores: |gpu - oracle| <= REL * max(|oracle|, FLOOR), with a
# floor for the unit-scaled ACE and CEM responses, which can sit at or
# cross zero. Angles: absolute radians.
REL = 1e-5
FLOOR = {"rx": 1.0, "ace": 1.0, "cem": 1.0}
ANGLE_TOL = 1e-6


def rel_err(got, want, floor):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.abs(got - want) / np.maximum(np.abs(want), floor)


def device_buffers(torch, pixels, cap=None):
    dev = "cuda"
    out = {
        "score": torch.full((pixels,), float("nan"), dtype=torch.float32, device=dev),
        "target": torch.full((pixels,), -7, dtype=torch.int32, device=dev),
        "aux": torch.full((pixels,), float("nan"), dtype=torch.float32, device=dev),
    }
    if cap is not None:
        out["dets"] = torch.zeros((max(cap, 1), 4), dtype=torch.int32, device=dev)
        out["count"] = torch.zeros(1, dtype=torch.int32, device=dev)
    return out


def upload_golden_stats(ctx, g):
    """Install the reference's own DetectorStatistics (float) on the device."""
    b = g["cube"].shape[0]
    ctx.stats_upload(b, g["st_mean"], g["st_whitener"], g["st_target_white"],
                     g["st_target_norm2"], g["st_cem_weight"])


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def near_cut(score, h, w, thr, nms, tol):
    """Pixels whose score is within tol of the threshold or of a neighbour's
    score inside the NMS window: where a `tol` relative score difference may
    flip whether the pixel is listed."""
    sc = score.reshape(h, w).astype(np.float64)
    near = np.abs(sc - thr) <= tol * max(abs(thr), 1.0)
    pad = np.pad(sc, nms, constant_values=np.nan)
    for dy in range(-nms, nms + 1):
        for dx in range(-nms, nms + 1):
            if dy or dx:
                nb = pad[nms + dy:nms + dy + h, nms + dx:nms + dx + w]
                near |= np.abs(nb - sc) <= tol * np.maximum(np.abs(sc), 1.0)
    return near.reshape(-1)
