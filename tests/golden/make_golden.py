#!/usr/bin/env python3
"""Write the golden fixtures from the compiled reference CPU scorer.

Run in a container that holds the reference checkout (default
/root/reference, override with HSI_REF_ROOT):

    make -C oracle            # builds oracle/_ref/libhsiref.so from the reference sources
    python tests/golden/make_golden.py

Inputs are synthetic scenes drawn here with numpy (a Dirichlet mix of smooth
endmember spectra under a per-pixel gain, Gaussian noise clipped at zero, and
targets planted at high abundance), plus hand-built edge cases. Outputs come
from the reference's own functions (sam_best_cpu, sam_cpu, detect_cpu,
accumulate_moments_cpu, build_statistics, rx_cpu, ace_best_cpu, cem_best_cpu,
rx_threshold_for_false_alarm_rate, cholesky, triangular_inverse_lower,
cholesky_solve, cholesky_inverse, quadratic_form, moments_from_shifted,
mean_spectrum) called through oracle/ref_shim.cpp. Each fixture is an .npz
holding inputs and expected outputs; tests load them with
numpy.load(allow_pickle=False).
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libhsiref.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def load_ref():
    if not os.path.exists(REF_LIB):
        sys.exit(f"{REF_LIB} missing: run `make -C oracle` with the reference checkout present")
    lib = C.CDLL(REF_LIB)
    lib.ref_library_norms.argtypes = [f32p, C.c_int, C.c_int, f32p]
    lib.ref_sam_best.argtypes = [f32p, C.c_int, C.c_int, C.c_int, f32p, f32p, C.c_int, f32p, i32p]
    lib.ref_sam_all.argtypes = [f32p, C.c_int, C.c_int, C.c_int, f32p, f32p, C.c_int, f32p]
    lib.ref_detect.restype = C.c_int64
    lib.ref_detect.argtypes = [f32p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                               i32p, f32p, C.c_int64]
    lib.ref_moments.argtypes = [f32p, C.c_int, C.c_int, C.c_int, f64p, f64p, f64p]
    lib.ref_build_statistics.restype = C.c_int
    lib.ref_build_statistics.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                         C.c_int, C.c_double, f32p, f32p, f32p, f32p, f32p,
                                         C.c_char_p, C.c_int]
    lib.ref_rx.argtypes = [f32p, C.c_int, C.c_int, C.c_int, f32p, f32p, f32p]
    lib.ref_ace_best.argtypes = [f32p, C.c_int, C.c_int, C.c_int, f32p, f32p, f32p, f32p, f32p,
                                 C.c_int, f32p, i32p, f32p]
    lib.ref_cem_best.argtypes = [f32p, C.c_int, C.c_int, C.c_int, f32p, f32p, f32p, f32p, f32p,
                                 C.c_int, f32p, i32p]
    lib.ref_rx_threshold.restype = C.c_double
    lib.ref_rx_threshold.argtypes = [C.c_int, C.c_double]
    lib.ref_cholesky.restype = C.c_int
    lib.ref_cholesky.argtypes = [f64p, C.c_int]
    lib.ref_triangular_inverse_lower.argtypes = [f64p, C.c_int, f64p]
    lib.ref_cholesky_solve.argtypes = [f64p, C.c_int, f64p, f64p]
    lib.ref_cholesky_inverse.argtypes = [f64p, C.c_int, f64p]
    lib.ref_quadratic_form.restype = C.c_double
    lib.ref_quadratic_form.argtypes = [f64p, f64p, C.c_int]
    lib.ref_moments_from_shifted.argtypes = [f64p, f64p, C.c_double, f64p, C.c_int, f64p, f64p]
    lib.ref_mean_spectrum.restype = C.c_int
    lib.ref_mean_spectrum.argtypes = [f32p, C.c_int, C.c_int, C.c_int,
                                      np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS"), f32p]
    return lib


# ---------------------------------------------------------------------------
# Scene drawing (fixture inputs only).

def endmembers(rng, count, bands):
    lam = np.linspace(0.0, 1.0, bands)
    out = []
    for _ in range(count):
        s = rng.uniform(0.1, 0.5) + rng.uniform(-0.3, 0.3) * lam
        for _ in range(6):
            c, w, h = rng.uniform(0, 1), rng.uniform(0.03, 0.2), rng.uniform(-0.3, 0.6)
            s = s + h * np.exp(-0.5 * ((lam - c) / w) ** 2)
        s = 0.05 + 0.85 * (s - s.min()) / (s.max() - s.min())
        out.append(s)
    return np.asarray(out)


def scene(seed, h, w, bands, n_end=8, n_targets=4, plant=0.02):
    rng = np.random.default_rng(seed)
    e = endmembers(rng, n_end, bands)
    p = h * w
    mix = rng.dirichlet(np.ones(n_end), size=p) @ e
    u = rng.uniform(size=p)
    for t in range(n_targets):
        sel = (u >= t * plant) & (u < (t + 1) * plant)
        a = rng.uniform(0.6, 1.0, size=int(sel.sum()))[:, None]
        mix[sel] = a * e[t % n_end] + (1 - a) * mix[sel]
    gain = rng.uniform(0.3, 1.2, size=p)[:, None]
    x = np.clip(gain * mix + rng.normal(0.0, 0.005, size=(p, bands)), 0.0, None)
    cube = np.ascontiguousarray(x.T.reshape(bands, h, w), dtype=np.float32)
    lib = np.ascontiguousarray(e[np.arange(n_targets) % n_end], dtype=np.float32)
    return cube, lib


# ---------------------------------------------------------------------------

def run_all(ref, cube, library, loading, thresholds):
    """Every reference output for one cube and library."""
    b, h, w = cube.shape
    t = library.shape[0]
    p = h * w
    out = {"cube": cube, "library": library, "loading": np.float64(loading)}
    norms = np.empty(t, np.float32)
    ref.ref_library_norms(library, t, b, norms)
    out["norms"] = norms

    score = np.empty(p, np.float32)
    target = np.empty(p, np.int32)
    ref.ref_sam_best(cube, h, w, b, library, norms, t, score, target)
    out["sam_score"], out["sam_target"] = score.copy(), target.copy()
    allv = np.empty(t * p, np.float32)
    ref.ref_sam_all(cube, h, w, b, library, norms, t, allv)
    out["sam_all"] = allv.reshape(t, p)

    cnt = np.zeros(1, np.float64)
    s1 = np.zeros(b, np.float64)
    s2 = np.zeros(b * b, np.float64)
    ref.ref_moments(cube, h, w, b, cnt, s1, s2)
    out["m_count"], out["m_sum"], out["m_outer"] = cnt[0], s1, s2.reshape(b, b)

    mean = np.zeros(b, np.float32)
    white = np.zeros(b * b, np.float32)
    tw = np.zeros(t * b, np.float32)
    tn2 = np.zeros(t, np.float32)
    cw = np.zeros(t * b, np.float32)
    err = C.create_string_buffer(512)
    ok = ref.ref_build_statistics(cube, h, w, b, library.ctypes.data, t, b, loading, mean, white,
                                  tw, tn2, cw, err, 512)
    if not ok:
        raise RuntimeError(err.value.decode())
    out.update(st_mean=mean, st_whitener=white.reshape(b, b), st_target_white=tw.reshape(t, b),
               st_target_norm2=tn2, st_cem_weight=cw.reshape(t, b))

    rx = np.empty(p, np.float32)
    ref.ref_rx(cube, h, w, b, mean, white, rx)
    out["rx_score"] = rx
    ace = np.empty(p, np.float32)
    acet = np.empty(p, np.int32)
    acerx = np.empty(p, np.float32)
    ref.ref_ace_best(cube, h, w, b, mean, white, tw, tn2, cw, t, ace, acet, acerx)
    out.update(ace_score=ace, ace_target=acet, ace_rx=acerx)
    cem = np.empty(p, np.float32)
    cemt = np.empty(p, np.int32)
    ref.ref_cem_best(cube, h, w, b, mean, white, tw, tn2, cw, t, cem, cemt)
    out.update(cem_score=cem, cem_target=cemt)

    # Detection lists per detector, NMS off and on, at the given thresholds.
    maps = {"sam": (out["sam_score"], out["sam_target"], 0),
            "rx": (rx, None, 1), "ace": (ace, acet, 1), "cem": (cem, cemt, 1)}
    for name, (sc, tg, higher) in maps.items():
        thr = np.float32(thresholds[name](sc))
        out[f"{name}_threshold"] = thr
        for nms in (0, 2):
            xyt = np.zeros(3 * p, np.int32)
            ds = np.zeros(p, np.float32)
            n = ref.ref_detect(sc, None if tg is None else tg.ctypes.data, h, w, float(thr), nms,
                               higher, xyt, ds, p)
            out[f"{name}_det{nms}_xyt"] = xyt[: 3 * n].reshape(n, 3)
            out[f"{name}_det{nms}_score"] = ds[:n]
    return out


def quantile(q):
    return lambda s: np.quantile(s.astype(np.float64), q)


def edge_cases(ref):
    """Hand-built inputs for the cases the reference tests pin by property."""
    out = {}
    # SAM: zero pixel, orthogonal pixel, pixel equal to a target, a target
    # duplicated (tie -> first wins), a zero-norm target, a scaled copy, and a
    # pixel with negative entries.
    bands = 6
    cube = np.zeros((bands, 1, 8), np.float32)
    lib = np.array([[1, 0, 1, 0, 1, 0],
                    [0.2, 0.4, 0.6, 0.8, 1.0, 1.2],
                    [0.2, 0.4, 0.6, 0.8, 1.0, 1.2],
                    [0, 0, 0, 0, 0, 0],
                    [0.6, 0.5, 0.4, 0.3, 0.2, 0.1]], np.float32)
    pix = np.array([[0, 0, 0, 0, 0, 0],
                    [0, 1, 0, 1, 0, 1],
                    [0.1, 0.2, 0.3, 0.4, 0.5, 0.6],
                    [1, 0, 1, 0, 1, 0],
                    [0.6, 0.5, 0.4, 0.3, 0.2, 0.1],
                    [-0.1, 0.2, -0.3, 0.4, -0.5, 0.6],
                    [3e-20, 0, 0, 0, 0, 0],
                    [250, 500, 750, 1000, 1250, 1500]], np.float32)
    cube[:, 0, :] = pix.T
    norms = np.empty(lib.shape[0], np.float32)
    ref.ref_library_norms(lib, lib.shape[0], bands, norms)
    score = np.empty(8, np.float32)
    target = np.empty(8, np.int32)
    ref.ref_sam_best(cube, 1, 8, bands, lib, norms, lib.shape[0], score, target)
    allv = np.empty(lib.shape[0] * 8, np.float32)
    ref.ref_sam_all(cube, 1, 8, bands, lib, norms, lib.shape[0], allv)
    out.update(edge_cube=cube, edge_library=lib, edge_norms=norms, edge_sam_score=score,
               edge_sam_target=target, edge_sam_all=allv.reshape(lib.shape[0], 8))

    # Detect: the reference's hand-built maps (plateau tie, blob, polarity).
    maps = []
    m = np.full((9, 9), 1.0, np.float32); m[3:6, 3:6] = 0.05; m[4, 4] = 0.01
    maps.append(("blob", m, 0.1, 2, 0))
    m = np.full((5, 5), 1.0, np.float32); m[1:3, 1:3] = 0.02
    maps.append(("plateau", m, 0.1, 2, 0))
    m = np.full((7, 7), 0.5, np.float32); m[2, 2] = 0.01; m[5, 5] = 0.99
    maps.append(("pol_low", m, 0.1, 1, 0))
    maps.append(("pol_high", m, 0.9, 1, 1))
    m = np.full((9, 9), 0.5, np.float32); m[3:6, 3:6] = 0.8; m[3, 3] = 0.6; m[5, 5] = 0.95
    maps.append(("extremum", m, 0.55, 3, 1))
    rng = np.random.default_rng(7)
    m = rng.integers(0, 6, size=(23, 31)).astype(np.float32) / 5.0  # many exact ties
    maps.append(("ties_low", m, 0.4, 2, 0))
    maps.append(("ties_high", m, 0.6, 1, 1))
    for name, mp, thr, nms, higher in maps:
        h, w = mp.shape
        tg = (np.arange(h * w, dtype=np.int32) % 3).astype(np.int32)
        xyt = np.zeros(3 * h * w, np.int32)
        ds = np.zeros(h * w, np.float32)
        n = ref.ref_detect(np.ascontiguousarray(mp.reshape(-1)), tg.ctypes.data, h, w, thr, nms,
                           higher, xyt, ds, h * w)
        out[f"det_{name}_map"] = mp
        out[f"det_{name}_args"] = np.array([thr, nms, higher], np.float64)
        out[f"det_{name}_xyt"] = xyt[: 3 * n].reshape(n, 3)
        out[f"det_{name}_score"] = ds[:n]

    # Linear algebra on random SPD matrices, and chi-squared thresholds.
    rng = np.random.default_rng(11)
    for n in (1, 2, 5, 17, 40):
        bm = rng.normal(size=(n, n))
        a = bm @ bm.T + n * np.eye(n)
        l = np.ascontiguousarray(a.copy())
        assert ref.ref_cholesky(l, n) == 1
        inv = np.empty((n, n))
        ref.ref_triangular_inverse_lower(l, n, inv)
        rhs = rng.normal(size=n)
        x = np.empty(n)
        ref.ref_cholesky_solve(l, n, rhs, x)
        out.update({f"la{n}_a": a, f"la{n}_l": l, f"la{n}_inv": inv, f"la{n}_rhs": rhs,
                    f"la{n}_x": x})
    bad = np.array([[1.0, 1.0], [1.0, 1.0]])
    out["la_singular_ok"] = np.int32(ref.ref_cholesky(np.ascontiguousarray(bad), 2))
    rates = np.array([0.5, 0.1, 1e-3, 1e-6])
    bands_list = np.array([2, 16, 50, 113, 224], np.int32)
    thr = np.array([[ref.ref_rx_threshold(int(b), float(r)) for r in rates] for b in bands_list])
    out.update(chi2_rates=rates, chi2_bands=bands_list, chi2_threshold=thr)
    return out


def helper_cases(ref):
    """mean_spectrum, moments_from_shifted, cholesky_inverse and quadratic_form."""
    out = {}
    rng = np.random.default_rng(23)
    # mean_spectrum over masks of a 9x11x7 cube: a sparse mask, a full mask,
    # a single pixel, and an empty mask (the reference throws).
    cube = np.ascontiguousarray(rng.normal(0.4, 0.2, size=(7, 9, 11)).astype(np.float32))
    masks = np.zeros((4, 99), np.uint8)
    masks[0] = rng.random(99) < 0.3
    masks[1] = 1
    masks[2, 57] = 1
    means = np.zeros((4, 7), np.float32)
    oks = np.zeros(4, np.int32)
    for i in range(4):
        oks[i] = ref.ref_mean_spectrum(cube, 9, 11, 7, np.ascontiguousarray(masks[i]), means[i])
    out.update(ms_cube=cube, ms_masks=masks, ms_mean=means, ms_ok=oks)
    # moments_from_shifted: the shifted moments of a small cube about a
    # reference near its mean (computed here in float64), rebuilt by the
    # reference; and a zero count.
    bands = 9
    x = rng.normal(0.5, 0.1, size=(bands, 200))
    r = (x.mean(axis=1) + rng.normal(0, 0.01, bands)).astype(np.float64)
    d = x - r[:, None]
    s1 = d.sum(axis=1)
    s2 = np.tril(d @ d.T)
    s2[np.triu_indices(bands, 1)] = 7.0  # the upper triangle is ignored
    osum = np.zeros(bands)
    oouter = np.zeros(bands * bands)
    ref.ref_moments_from_shifted(s1, np.ascontiguousarray(s2), 200.0, r, bands, osum, oouter)
    zsum = np.zeros(bands)
    zouter = np.zeros(bands * bands)
    ref.ref_moments_from_shifted(s1, np.ascontiguousarray(s2), 0.0, r, bands, zsum, zouter)
    out.update(mfs_s1=s1, mfs_s2=s2, mfs_ref=r, mfs_sum=osum, mfs_outer=oouter.reshape(bands, bands),
               mfs_zero_sum=zsum, mfs_zero_outer=zouter.reshape(bands, bands))
    # cholesky_inverse and quadratic_form on SPD matrices.
    for n in (1, 3, 12, 33):
        bm = rng.normal(size=(n, n))
        a = bm @ bm.T + n * np.eye(n)
        l = np.ascontiguousarray(a.copy())
        assert ref.ref_cholesky(l, n) == 1
        inv = np.empty((n, n))
        ref.ref_cholesky_inverse(l, n, inv)
        y = rng.normal(size=n)
        q = ref.ref_quadratic_form(np.ascontiguousarray(a), y, n)
        out.update({f"ci{n}_a": a, f"ci{n}_l": l, f"ci{n}_inv": inv, f"ci{n}_y": y,
                    f"ci{n}_q": np.float64(q)})
    return out


def near_singular_case(ref):
    """A scene whose covariance the reference factorises although one pivot is
    below 4*256*eps of its diagonal: band 1 is band 0 moved by one float ulp on
    part of the pixels, so the two are collinear to about 1e-15."""
    rng = np.random.default_rng(77)
    h, w, b = 16, 16, 6
    tol = 4.0 * 256 * np.finfo(np.float64).eps
    for _ in range(100):
        base = rng.normal(0.0, 0.2, size=(b, h * w)).astype(np.float32)
        sel = rng.random(h * w) < rng.uniform(0.05, 0.9)
        up = rng.random(int(sel.sum())) < 0.5
        base[1] = base[0]
        base[1, sel] = np.nextafter(base[0, sel], np.where(up, np.inf, -np.inf).astype(np.float32))
        xd = base.astype(np.float64)
        cov = xd @ xd.T / (h * w) - np.outer(xd.mean(axis=1), xd.mean(axis=1))
        piv = np.linalg.eigvalsh(cov).min() / np.diag(cov).max()
        if not (0.005 * tol < piv < 0.1 * tol):
            continue
        cube = np.ascontiguousarray(base.reshape(b, h, w))
        out = {"nsing_cube": cube}
        for key, loading in (("nsing_l0", 0.0), ("nsing_l3", 1e-3)):
            mean = np.zeros(b, np.float32)
            white = np.zeros(b * b, np.float32)
            z1 = np.zeros(1, np.float32)
            err = C.create_string_buffer(512)
            ok = ref.ref_build_statistics(cube, h, w, b, None, 0, b, loading, mean, white, z1, z1,
                                          z1, err, 512)
            assert ok == 1, err.value
            out[f"{key}_whitener"] = white.reshape(b, b)
            out[f"{key}_mean"] = mean
        return out
    raise RuntimeError("no near-singular scene found")


def singular_case(ref):
    """Duplicated bands: build_statistics must fail at loading 0 and name --loading."""
    rng = np.random.default_rng(5)
    shared = rng.normal(0.5, 0.1, size=100).astype(np.float32)
    other = rng.normal(0.5, 0.1, size=100).astype(np.float32)
    cube = np.ascontiguousarray(np.stack([shared, shared, other, other]).reshape(4, 10, 10))
    out = {"sing_cube": cube}
    for loading in (0.0, 1e-3):
        mean = np.zeros(4, np.float32)
        white = np.zeros(16, np.float32)
        z1 = np.zeros(1, np.float32)
        err = C.create_string_buffer(512)
        ok = ref.ref_build_statistics(cube, 10, 10, 4, None, 0, 4, loading, mean, white, z1, z1,
                                      z1, err, 512)
        key = "sing_l0" if loading == 0.0 else "sing_l3"
        out[f"{key}_ok"] = np.int32(ok)
        out[f"{key}_error"] = np.frombuffer(err.value, np.uint8)
        out[f"{key}_whitener"] = white.reshape(4, 4)
        out[f"{key}_mean"] = mean
    return out


def main():
    ref = load_ref()
    th = {"sam": lambda s: 0.10, "rx": quantile(0.98), "ace": quantile(0.98),
          "cem": quantile(0.98)}
    cases = {
        # small, 16 bands, 3 targets
        "scene_24x32x16_t3": (scene(101, 24, 32, 16, n_end=6, n_targets=3, plant=0.03), 1e-6),
        # a crop of the plain-method scene: 113 bands, 4 targets
        "scene_32x40x113_t4": (scene(202, 32, 40, 113, n_end=8, n_targets=4, plant=0.01), 1e-6),
        # 224 bands, 16 targets (more bands than most tiles; fewer pixels than bands squared)
        "scene_20x20x224_t16": (scene(303, 20, 20, 224, n_end=8, n_targets=16, plant=0.01), 1e-5),
        # 8 targets at a width that is not a multiple of 8 (ragged plane)
        "scene_13x37x113_t8": (scene(404, 13, 37, 113, n_end=8, n_targets=8, plant=0.02), 1e-6),
    }
    for name, ((cube, lib), loading) in cases.items():
        out = run_all(ref, cube, lib, loading, th)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")
    out = edge_cases(ref)
    out.update(singular_case(ref))
    out.update(helper_cases(ref))
    out.update(near_singular_case(ref))
    path = os.path.join(HERE, "edge_cases.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")


if __name__ == "__main__":
    main()
