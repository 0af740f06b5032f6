"""The learned pre-screen for RX/ACE, driven through the C ABI.

A small network (B -> 64 -> 32 -> 1) is trained on the GPU by the library's
own forward/backward/Adam kernels to predict which pixels the plain RX/ACE
detectors will put over threshold. At run time one pass over the frame
(hsi_mlp_screen) computes the logits, lists the pixels at or above a
conservative cut and writes their spectra to rows of a buffer; only those
candidates are rescored exactly
(hsi_score_candidate_spectra on those rows: with precision PLAIN their
scores are bit-identical to the plain full-frame path; with MIXED they are
the precision form's, FP16 hi/lo matrix cores), and the usual threshold +
NMS runs on the result. The plain full-frame detectors are
the judge (training labels, recall) and the fallback (a candidate list that
overflows its capacity makes the device score the whole frame).

Labels: a pixel is positive when the plain RX score reaches the chi-squared
threshold of ten times `rx_false_alarm_rate`, or the plain ACE score reaches
`label_slack` times `ace_threshold`: positives include pixels just short of
the detection thresholds, which keeps the cut conservative. Inputs are centred on the frame's scene mean. The default
false-alarm rate (1e-6 per pixel) makes RX detections the anomalous pixels:
at 1e-3 about one background pixel in a thousand crosses the threshold by
chance, an excursion spread over all whitened components that only the full
quadratic form can see and no pre-screen can anticipate.
"""
from __future__ import annotations

import numpy as np

from . import MIXED, PLAIN, Context, DetectionParams, Detector, MlpTrainParams, rx_threshold


class Prescreen:
    def __init__(self, ctx: Context, bands: int, rx_false_alarm_rate=1e-6, ace_threshold=0.5,
                 nms_radius=2, seed=1, capacity=1 << 20, label_slack=0.7, precision=PLAIN):
        import torch
        if precision not in (PLAIN, MIXED):
            raise ValueError("precision is PLAIN or MIXED")
        self.precision = precision
        self.torch = torch
        self.ctx = ctx
        self.bands = bands
        self.rx_thr = float(rx_threshold(bands, rx_false_alarm_rate))
        self.ace_thr = float(ace_threshold)
        # Training positives also include pixels just short of a threshold:
        # ACE at `label_slack` of its threshold, RX at the chi-squared
        # threshold of a ten times higher false-alarm rate (a fixed fraction
        # of the RX threshold would admit ordinary noise excursions).
        self.label_ace = float(label_slack) * self.ace_thr
        self.label_rx = float(rx_threshold(bands, min(0.5, 10.0 * rx_false_alarm_rate)))
        self.nms = nms_radius
        self.capacity = capacity
        self.rng = np.random.default_rng(seed)
        self.cut = None
        ctx.mlp_init(bands, seed)
        self._x, self._y = [], []
        self._maps = {}

    # -- plain path (judge) ---------------------------------------------------
    def _buffers(self, pixels):
        t = self.torch
        if self._maps.get("pixels") != pixels:
            f32 = dict(dtype=t.float32, device="cuda")
            i32 = dict(dtype=t.int32, device="cuda")
            self._maps = {"pixels": pixels, "ace": t.empty(pixels, **f32), "rx": t.empty(pixels, **f32),
                          "target": t.empty(pixels, **i32), "logit": t.empty(pixels, **f32),
                          "cand": t.empty((self.capacity, 4), **i32), "count": t.empty(1, **i32),
                          "spectra": t.empty((self.capacity, (self.bands + 31) // 32 * 32), **f32),
                          "dets": t.empty((1 << 16, 4), **i32), "ndet": t.empty(1, **i32)}
        return self._maps

    def statistics(self, cube, h, w, loading=1e-6, precision=PLAIN):
        self.ctx.stats_accumulate(cube, h, w, self.bands, precision=precision)
        self.ctx.stats_build(loading)
        self.ctx.stats_check()

    def plain(self, cube, h, w):
        """Full-frame ACE with the RX side output (statistics must be built)."""
        m = self._buffers(h * w)
        self.ctx.score(Detector.ACE, cube, h, w, self.bands, m["ace"], m["target"], m["rx"])
        return m

    def labels(self, cube, h, w):
        m = self.plain(cube, h, w)
        self.torch.cuda.synchronize()
        rx = m["rx"].cpu().numpy()
        ace = m["ace"].cpu().numpy()
        return (rx >= self.label_rx) | (ace >= self.label_ace)

    # -- training ----------------------------------------------------------------
    def add_training_frame(self, cube, h, w, negatives_per_positive=16, max_positives=4096):
        """Statistics, plain labels, and a gathered sample set from one frame
        (at most `max_positives` positives, sampled, and
        `negatives_per_positive` times as many negatives)."""
        t = self.torch
        self.statistics(cube, h, w)
        pos = self.labels(cube, h, w)
        pidx = np.flatnonzero(pos)
        if pidx.size > max_positives:
            pidx = np.sort(self.rng.choice(pidx, max_positives, replace=False))
        nidx = np.flatnonzero(~pos)
        k = min(nidx.size, max(negatives_per_positive * max(pidx.size, 1), 4096))
        nsel = self.rng.choice(nidx, k, replace=False)
        idx = np.concatenate([pidx, nsel]).astype(np.int32)
        lab = np.concatenate([np.ones(pidx.size, np.float32), np.zeros(nsel.size, np.float32)])
        di = t.from_numpy(idx).cuda()
        x = t.empty((self.bands, idx.size), dtype=t.float32, device="cuda")
        self.ctx.gather_pixels(cube, h, w, self.bands, di, x, center=True)
        self._x.append(x)
        self._y.append(lab)
        return int(pidx.size), int(nsel.size)

    def train(self, epochs=60, batch=1024, learning_rate=2e-3, pos_weight=8.0, max_steps=8000):
        """Adam over shuffled batches; epochs are cut so the run stays within
        `max_steps` updates."""
        t = self.torch
        n = sum(x.shape[1] for x in self._x)
        stride = (n + 7) // 8 * 8
        X = t.zeros((self.bands, stride), dtype=t.float32, device="cuda")
        o = 0
        for x in self._x:
            X[:, o:o + x.shape[1]] = x
            o += x.shape[1]
        Y = np.concatenate(self._y)
        self._train_x, self._train_y, self._n = X, Y, n
        xb = t.empty((self.bands, batch), dtype=t.float32, device="cuda")
        params = MlpTrainParams(learning_rate=learning_rate, pos_weight=pos_weight)
        loss = t.zeros(1, dtype=t.float32, device="cuda")
        history = []
        epochs = min(epochs, max(1, max_steps // max(1, n // batch))) if epochs else 0
        for epoch in range(epochs):
            perm = self.rng.permutation(n).astype(np.int32)
            for s0 in range(0, n - batch + 1, batch):
                sel = perm[s0:s0 + batch]
                self.ctx.gather_pixels(X, 1, stride, self.bands, t.from_numpy(sel).cuda(), xb, center=False)
                yb = t.from_numpy(Y[sel]).cuda()
                self.ctx.mlp_train_step(xb, yb, params, loss)
            t.cuda.synchronize()
            history.append(float(loss.item()))
        return history

    def calibrate(self, margin=1.0, quantile=1e-4):
        """Cut = the `quantile` of the training positives' logits (the lowest
        one for small sets), less a margin."""
        t = self.torch
        X, Y, n = self._train_x, self._train_y, self._n
        logit = t.empty(X.shape[1], dtype=t.float32, device="cuda")
        self.ctx.mlp_infer(X, 1, X.shape[1], self.bands, logit, center=False)
        t.cuda.synchronize()
        lg = logit.cpu().numpy()[:n]
        self.cut = float(np.quantile(lg[Y > 0.5], quantile, method="lower") - margin)
        return self.cut

    # -- model file (read by hsi_detect --prescreen) ---------------------------------
    MAGIC = b"HSIMLP01"

    def save(self, path):
        """Write the trained network and its cut: "HSIMLP01", int32 bands,
        float32 cut, uint32 count, float32 weights[count] (little-endian)."""
        if self.cut is None:
            raise ValueError("calibrate the cut before saving")
        w = np.ascontiguousarray(self.ctx.mlp_get_weights(), dtype="<f4")
        with open(path, "wb") as f:
            f.write(self.MAGIC)
            f.write(np.array([self.bands], "<i4").tobytes())
            f.write(np.array([self.cut], "<f4").tobytes())
            f.write(np.array([w.size], "<u4").tobytes())
            f.write(w.tobytes())

    def load(self, path):
        """Install a network and cut written by save()."""
        with open(path, "rb") as f:
            data = f.read()
        if data[:8] != self.MAGIC:
            raise ValueError(f"{path}: not a pre-screen model")
        bands = int(np.frombuffer(data, "<i4", 1, 8)[0])
        cut = float(np.frombuffer(data, "<f4", 1, 12)[0])
        n = int(np.frombuffer(data, "<u4", 1, 16)[0])
        if bands != self.bands:
            raise ValueError(f"{path}: a network for {bands} bands, not {self.bands}")
        w = np.frombuffer(data, "<f4", n, 20).astype(np.float32)
        self.ctx.mlp_set_weights(bands, w)
        self.cut = cut

    # -- run time --------------------------------------------------------------------
    def screen(self, cube, h, w):
        """Learned path (statistics must be built): one screening pass
        (candidates at the cut and their spectra), exact rescoring of the
        candidates; returns the score buffers."""
        m = self._buffers(h * w)
        self.ctx.mlp_screen(cube, h, w, self.bands, self.cut, m["cand"], m["count"], m["spectra"])
        self.ctx.score_candidate_spectra(Detector.ACE, cube, h, w, self.bands, m["cand"], m["count"],
                                         m["spectra"], m["ace"], m["target"], m["rx"],
                                         precision=self.precision)
        return m

    def screen_fused(self, cube, h, w, shift=None, loading=1e-6):
        """The learned path from the frame alone: the mixed-form statistics
        and the screen in one pass over the cube (the network sees the pixel
        minus `shift`, the sampled mean when None), the build, and the
        rescoring of the candidate rows; returns the score buffers."""
        m = self._buffers(h * w)
        self.ctx.stats_accumulate_screen(cube, h, w, self.bands, self.cut, m["cand"], m["count"], m["spectra"],
                                         shift=shift)
        self.ctx.stats_build(loading)
        self.ctx.score_candidate_spectra(Detector.ACE, cube, h, w, self.bands, m["cand"], m["count"],
                                         m["spectra"], m["ace"], m["target"], m["rx"],
                                         precision=self.precision)
        return m

    def screen_two_pass(self, cube, h, w):
        """The same result in separate steps: the logit map, hsi_detect at the
        cut (raster order), rescoring that gathers the candidates from the
        cube."""
        m = self._buffers(h * w)
        self.ctx.mlp_infer(cube, h, w, self.bands, m["logit"], center=True)
        self.ctx.detect(m["logit"], None, h, w, DetectionParams(self.cut, 0, self.capacity, True),
                        m["cand"], m["count"])
        self.ctx.score_candidates(Detector.ACE, cube, h, w, self.bands, m["cand"], m["count"], m["ace"],
                                  m["target"], m["rx"])
        return m

    def detections(self, m, h, w):
        """(rx, ace) detection sets {(x, y)} from score buffers, threshold + NMS."""
        t = self.torch
        out = []
        for score, target, thr in ((m["rx"], None, self.rx_thr), (m["ace"], m["target"], self.ace_thr)):
            self.ctx.detect(score, target, h, w, DetectionParams(thr, self.nms, m["dets"].shape[0], True),
                            m["dets"], m["ndet"])
            t.cuda.synchronize()
            k = min(int(m["ndet"].item()), m["dets"].shape[0])
            out.append(set(map(tuple, m["dets"][:k, :2].cpu().numpy().tolist())))
        return out

    def evaluate(self, cube, h, w, fused=False):
        """Recall of the learned path against the plain path on one frame
        (statistics built here). With precision MIXED the learned path also
        builds its statistics in the mixed form (the plain path keeps the
        plain form), and the result holds the rescored candidates' largest
        error against the plain scores: RX relative, ACE absolute. fused:
        the learned path is screen_fused (statistics and screen in one
        pass)."""
        self.statistics(cube, h, w)
        m = self.plain(cube, h, w)
        plain = self.detections(m, h, w)
        ref = None
        if self.precision == MIXED or fused:
            ref = (m["rx"].clone(), m["ace"].clone())
        if fused:
            m = self.screen_fused(cube, h, w)
        else:
            if self.precision == MIXED:
                self.statistics(cube, h, w, precision=MIXED)
            m = self.screen(cube, h, w)
        self.torch.cuda.synchronize()
        flagged = int(m["count"].item())
        learned = self.detections(m, h, w)
        res = {"candidates": flagged, "candidate_fraction": flagged / (h * w)}
        if ref is not None and 0 < flagged <= self.capacity:
            t = self.torch
            rec = m["cand"][:flagged].long()
            px = rec[:, 1] * w + rec[:, 0]
            rx_p, rx_l = ref[0][px].double(), m["rx"][px].double()
            res["rx_max_rel_err"] = float((t.abs(rx_l - rx_p) / t.clamp(t.abs(rx_p), min=1e-30)).max())
            res["ace_max_abs_err"] = float(t.abs(m["ace"][px].double() - ref[1][px].double()).max())
        for name, want, got in (("rx", plain[0], learned[0]), ("ace", plain[1], learned[1])):
            hit = len(want & got)
            res[name] = {"plain": len(want), "learned": len(got), "recall": hit / max(len(want), 1),
                         "false_skips": len(want) - hit, "extra": len(got - want)}
        return res
