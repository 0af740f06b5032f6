"""The C++ host layer: hsi_score and hsi_detect (the reference's two CLIs,
score_main.cpp and main.cpp) over the Pipeline and the file readers.

CPU tests exercise argument handling and the ENVI / CSV readers' error paths,
all of which run before any device call. GPU tests score golden scenes
through hsi_score (ENVI + CSV in, raw f32/i32 dumps out, compared with the
reference's own outputs) and run the frame pipeline, eager and as replayed
hipGraphs, against the Python path on the same seeded frames.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, SCENES, golden

BIN = os.path.join(PKG, "bin")


def run(args, timeout=120):
    return subprocess.run([os.path.join(BIN, args[0])] + list(args[1:]), capture_output=True, text=True,
                          timeout=timeout)


def write_envi(path_hdr, cube, interleave="bsq", dtype=np.float32, data_type=4, scale=None):
    """cube: (bands, H, W) float32 -> ENVI files; integer types are stored as
    round(value / scale)."""
    b, h, w = cube.shape
    if interleave == "bsq":
        arr = cube
    elif interleave == "bip":
        arr = np.transpose(cube, (1, 2, 0))
    else:  # bil: line, band, sample
        arr = np.transpose(cube, (1, 0, 2))
    if scale is not None:
        arr = np.round(arr / scale)
    arr = np.ascontiguousarray(arr).astype(dtype)
    with open(path_hdr, "w") as f:
        f.write(f"ENVI\nsamples = {w}\nlines = {h}\nbands = {b}\nheader offset = 0\n"
                f"file type = ENVI Standard\ndata type = {data_type}\ninterleave = {interleave}\n"
                f"byte order = 0\n")
    arr.tofile(path_hdr[:-4] + ".img")


def write_csv(path, library, names=None):
    with open(path, "w") as f:
        f.write("# test library\n")
        for i, row in enumerate(library):
            name = names[i] if names else f"t{i}"
            f.write(name + "," + ",".join(repr(float(v)) for v in row) + "\n")


# ---- CPU --------------------------------------------------------------------

@pytest.mark.parametrize("args,needle", [
    (["hsi_detect", "--bogus=1"], "unknown option --bogus"),
    (["hsi_detect", "stray"], "unexpected argument"),
    (["hsi_detect", "--variant=fast"], "--variant must be"),
    (["hsi_detect", "--memory=zerocopy"], "--memory=copy"),
    (["hsi_detect", "--memory=pinned"], "--memory must be"),
    (["hsi_detect", "--source=hot", "--hot-dir=x"], "unknown option --hot-dir"),
    (["hsi_detect", "--source=hot"], "hsi_demosaic"),
    (["hsi_score", "--cube=x.hdr", "--frobnicate"], "unknown option --frobnicate"),
])
def test_cli_refuses_unknown_flags_and_unsupported_modes(args, needle):
    """The reference's flags are honoured or refused with exit 2, never
    silently ignored (main.cpp:181-188 refuses bad --variant/--memory values)."""
    r = run(args)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert needle in r.stderr


def test_cli_refuses_an_invalid_rx_false_alarm_rate():
    r = run(["hsi_detect", "--detector=rx", "--width=16", "--height=16", "--rx-false-alarm=2"])
    assert r.returncode == 1
    assert "false-alarm rate must lie in (0, 1)" in r.stderr


def test_cli_help_and_usage():
    for tool in ("hsi_score", "hsi_detect"):
        r = run([tool, "--help"])
        assert r.returncode == 0, r.stderr
        assert "--threshold" in r.stdout and "--nms" in r.stdout
    r = run(["hsi_score"])
    assert r.returncode == 2  # missing --cube / --library prints usage
    r = run(["hsi_score", "--cube=x.hdr", "--library=y.csv", "--detector=bogus"])
    assert r.returncode == 2 and "--detector" in r.stderr


def test_score_reports_file_errors_before_touching_the_device(tmp_path):
    cube = np.ones((3, 2, 4), np.float32)
    hdr = str(tmp_path / "c.hdr")
    write_envi(hdr, cube)
    r = run(["hsi_score", f"--cube={tmp_path / 'missing.hdr'}", "--library=x.csv"])
    assert r.returncode == 1 and "cannot open ENVI header" in r.stderr
    bad = tmp_path / "bad.csv"
    bad.write_text("a,1,2,zz\n")
    r = run(["hsi_score", f"--cube={hdr}", f"--library={bad}"])
    assert r.returncode == 1 and "'zz' is not a number" in r.stderr
    ragged = tmp_path / "ragged.csv"
    ragged.write_text("a,1,2,3\nb,1,2\n")
    r = run(["hsi_score", f"--cube={hdr}", f"--library={ragged}"])
    assert r.returncode == 1 and "has 2 bands, expected 3" in r.stderr
    wrong = tmp_path / "wrong.csv"
    wrong.write_text("wavelength,400,500,600,700\na,1,2,3,4\n")
    r = run(["hsi_score", f"--cube={hdr}", f"--library={wrong}"])
    assert r.returncode == 1 and "4 bands but the cube has 3" in r.stderr
    r = run(["hsi_score", f"--cube={hdr}", f"--library={tmp_path / 'none.csv'}"])
    assert r.returncode == 1 and "cannot open spectral library" in r.stderr
    short = tmp_path / "s.hdr"
    short.write_text("ENVI\nsamples = 4\nlines = 2\nbands = 30\ndata type = 4\ninterleave = bsq\n")
    (tmp_path / "s.img").write_bytes(b"\0" * 16)
    r = run(["hsi_score", f"--cube={short}", "--detector=rx"])
    assert r.returncode == 1 and "header promises" in r.stderr
    r = run(["hsi_score", f"--cube={hdr}", "--library=x.csv", "--target=nobody"])
    assert r.returncode == 1


@pytest.mark.parametrize("args,needle", [
    (["hsi_detect", "--batch", "--precision=mixed"], "plain precision"),
    (["hsi_detect", "--batch", "--variant=half"], "plain precision"),
    (["hsi_detect", "--batch", "--prefill"], "--prefill does not apply to --batch"),
    (["hsi_detect", "--batch", "--frames=1000"], "1..256 synthetic frames"),
    (["hsi_detect", "--autotune"], "--autotune applies to --batch"),
])
def test_batch_mode_refuses_what_it_does_not_do(args, needle):
    r = run(args)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert needle in r.stderr


def test_detect_rejects_bad_options():
    r = run(["hsi_detect", "--detector=sam,foo"])
    assert r.returncode == 2 and "--detector" in r.stderr
    r = run(["hsi_detect", "--precision=half"])
    assert r.returncode == 2
    r = run(["hsi_detect", "--source=camera"])
    assert r.returncode == 2 and "unknown --source" in r.stderr


# ---- GPU --------------------------------------------------------------------

def score_cli(tmp_path, g, detector, extra=(), cube=None, hdr_kw=None):
    cube = g["cube"] if cube is None else cube
    hdr = str(tmp_path / "cube.hdr")
    write_envi(hdr, cube, **(hdr_kw or {}))
    csv = str(tmp_path / "lib.csv")
    write_csv(csv, g["library"])
    s, t = str(tmp_path / "score.f32"), str(tmp_path / "target.i32")
    print(f"hsi_score --detector={detector} on {cube.shape}", flush=True)
    r = run(["hsi_score", f"--cube={hdr}", f"--library={csv}", f"--detector={detector}",
             f"--dump-score={s}", f"--dump-target={t}", "--print=3"] + list(extra))
    assert r.returncode == 0, r.stderr
    return np.fromfile(s, np.float32), np.fromfile(t, np.int32), r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", SCENES)
def test_hsi_score_sam_matches_reference_dump(tmp_path, name):
    g = golden(name)
    thr = repr(float(g["sam_threshold"]))
    s, t, out = score_cli(tmp_path, g, "sam", extra=[f"--threshold={thr}", "--nms=2"])
    assert np.abs(s.astype(np.float64) - g["sam_score"]).max() <= 1e-6
    assert np.array_equal(t, g["sam_target"])
    assert f" {int(g['sam_det2_score'].size)} detections" in out


@pytest.mark.gpu
@pytest.mark.parametrize("det", ["rx", "ace", "cem"])
def test_hsi_score_statistics_detectors_match_reference(tmp_path, det):
    g = golden("scene_32x40x113_t4")
    s, t, out = score_cli(tmp_path, g, det)
    want = g[f"{det}_score"].astype(np.float64)
    err = np.abs(s - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() <= 1e-5, err.max()
    if det != "rx":
        agree = (t == g[f"{det}_target"]).mean()
        assert agree > 0.999


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [("bip", np.float32, 4, None), ("bil", np.float32, 4, None),
                                    ("bsq", np.uint16, 12, 1 / 65535.0), ("bip", np.int16, 2, 1 / 32767.0)])
def test_hsi_score_reads_every_envi_layout(tmp_path, oracle_lib, layout):
    interleave, dtype, code, scale = layout
    g = golden("scene_24x32x16_t3")
    cube = g["cube"]
    # the value the reader should reconstruct, in float32 like the reference
    if scale is None:
        want_cube = cube
    else:
        q = np.round(cube / scale).astype(dtype)
        full = {np.uint16: 65535.0, np.int16: 32767.0}[dtype]
        want_cube = q.astype(np.float32) * (np.float32(1.0) / np.float32(full))
    s, t, _ = score_cli(tmp_path, g, "sam", cube=cube,
                        hdr_kw=dict(interleave=interleave, dtype=dtype, data_type=code, scale=scale))
    ws, wt = oracle_lib.sam_best(want_cube, g["library"])
    assert np.abs(s.astype(np.float64) - ws).max() <= 1e-6
    assert (t == wt).mean() > 0.999


def detect_cli(tmp_path, extra):
    csvp = str(tmp_path / f"frames_{len(os.listdir(tmp_path))}.csv")
    args = ["hsi_detect", "--width=160", "--height=96", "--bands=113", "--targets=8", "--seed=77",
            "--frames=6", "--warmup=2", "--streams=2", "--detector=sam,rx,ace,cem", f"--csv={csvp}",
            "--json", "--quiet", "--threshold=0.1", "--print-detections=5"] + list(extra)
    print(" ".join(args), flush=True)
    r = run(args, timeout=180)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    summary = json.loads(lines[-1])
    rows = np.genfromtxt(csvp, delimiter=",", names=True)
    return summary, rows, lines[:-1]


@pytest.mark.gpu
def test_hsi_detect_graph_replay_equals_eager_and_python(tmp_path):
    import torch
    import hsi_mi355x as H
    g_sum, g_rows, g_lines = detect_cli(tmp_path, [])
    e_sum, e_rows, e_lines = detect_cli(tmp_path, ["--no-graph"])
    assert g_sum["graph"] and not e_sum["graph"]
    assert g_sum["frames"] == 6 and g_sum["statistics_ok"]
    for col in ("found_sam", "found_rx", "found_ace", "found_cem", "frame"):
        assert np.array_equal(g_rows[col], e_rows[col]), col
    assert g_lines == e_lines  # last frame's first detections, every detector
    # frame indices 2..7 of the same seeded stream through the Python path
    cfg = H.synth_config(96, 160, 113, 8, seed=77)
    lib = H.synth_library(cfg)
    ctx = H.Context(0)
    ctx.set_library(lib)
    p = 96 * 160
    s = torch.empty(p, dtype=torch.float32, device="cuda")
    t = torch.empty(p, dtype=torch.int32, device="cuda")
    out = torch.empty((4096, 4), dtype=torch.int32, device="cuda")
    cnt = torch.empty(1, dtype=torch.int32, device="cuda")
    for row in g_rows:
        k = int(row["frame"])
        cube = torch.from_numpy(H.synth_frame(cfg, k)).cuda()
        ctx.score(H.Detector.SAM, cube, 96, 160, 113, s, t)
        ctx.detect(s, t, 96, 160, H.DetectionParams(0.1, 2, 4096, False), out, cnt)
        torch.cuda.synchronize()
        assert int(cnt.item()) == int(row["found_sam"]), k
        ctx.stats_accumulate(cube, 96, 160, 113)
        ctx.stats_build(1e-6)
        ctx.score(H.Detector.CEM, cube, 96, 160, 113, s, t)
        # CEM's own threshold: the pipeline default of 0.5
        ctx.detect(s, t, 96, 160, H.DetectionParams(0.5, 2, 4096, True), out, cnt)
        torch.cuda.synchronize()
        assert int(cnt.item()) == int(row["found_cem"]), k
    ctx.close()


@pytest.mark.gpu
def test_hsi_detect_prefill_and_mixed(tmp_path):
    summ, rows, _ = detect_cli(tmp_path, ["--prefill", "--precision=mixed"])
    assert summ["frames"] == 6 and summ["statistics_ok"]
    # a prefilled ring of 2 slots replays frames 0 and 1
    assert np.array_equal(rows["found_sam"][0::2], rows["found_sam"][0::2][:1].repeat(3))


def read_dumped_lists(path):
    rows = np.genfromtxt(path, delimiter=",", names=True, dtype=None, encoding=None)
    rows = np.atleast_1d(rows)
    out = {}
    for r in rows:
        out.setdefault((int(r["frame"]), str(r["detector"])), []).append(
            (int(r["x"]), int(r["y"]), int(r["target"]), float(r["score"])))
    return out


def oracle_frame_lists(H, oracle_lib, cfg, frame, thresholds, nms, half=False):
    """The oracle's four lists for one seeded frame: statistics from the FP32
    frame; with half=True the scorers read the frame rounded to FP16 (the
    pipeline's --variant=half)."""
    h, w, b = cfg.height, cfg.width, cfg.bands
    p = h * w
    lib = H.synth_library(cfg)
    host = H.synth_frame(cfg, frame)[:, :p]
    st = oracle_lib.statistics_of_cube(np.ascontiguousarray(host).reshape(b, h, w), lib, 1e-6)
    assert st.valid, st.error
    cube = host.astype(np.float16).astype(np.float32) if half else host
    cube = np.ascontiguousarray(cube).reshape(b, h, w)
    scores = {"sam": oracle_lib.sam_best(cube, lib),
              "rx": (oracle_lib.rx(cube, st), None),
              "ace": oracle_lib.ace_best(cube, st)[:2],
              "cem": oracle_lib.cem_best(cube, st)}
    out = {}
    for det, (sc, tg) in scores.items():
        higher = det != "sam"
        thr = thresholds[det]
        dets, found = oracle_lib.detect(sc, tg, h, w, thr, nms, higher)
        out[det] = (sc, [(int(d["x"]), int(d["y"]), int(d["target"]), float(d["score"])) for d in dets], found)
    return out


def compare_pipeline_lists(got_lists, want, h, w, nms, frames, tol):
    from gpu_util import near_cut
    report = {}
    for k in frames:
        for det, (sc, wl, wfound) in want[k].items():
            gl = got_lists.get((k, det), [])
            gset = {(y * w + x, t) for x, y, t, _ in gl}
            wset = {(y * w + x, t) for x, y, t, _ in wl}
            if det == "sam":
                # SAM scores are the reference's to the bit: identical lists
                assert gset == wset, (k, det, sorted(gset ^ wset)[:5])
                assert [g[:3] for g in gl] == [v[:3] for v in wl], (k, det)
                continue
            near = near_cut(sc, h, w, thr_of(det, want, k), nms, tol)
            off = [(q, t) for q, t in gset ^ wset if not near[q]]
            assert not off, (k, det, off[:10])
            report[(k, det)] = (len(gset), len(wset), len(gset ^ wset))
    return report


def thr_of(det, want, k):
    return want["_thr"][det]


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["optimized", "half"])
def test_hsi_detect_lists_match_the_oracle_per_detector(tmp_path, oracle_lib, variant):
    """hsi_detect --detector=sam,rx,ace,cem over four seeded frames (two
    eager, two replayed from captured graphs per slot), each detector with its
    own threshold (RX from a false-alarm rate): every frame's four lists equal
    detect_cpu on the oracle's scores - SAM element for element, RX/ACE/CEM as
    (pixel, target) sets apart from pixels within 1e-5 of the threshold or of
    an NMS neighbour. With --variant=half the oracle scores the FP16-rounded
    frame (statistics from the FP32 frame, as the pipeline does)."""
    import hsi_mi355x as H
    h, w, b, t, seed = 96, 160, 113, 8, 77
    rate, ace_thr, cem_thr = 1e-3, 0.3, 0.5
    dump = str(tmp_path / "lists.csv")
    args = ["hsi_detect", f"--width={w}", f"--height={h}", f"--bands={b}", f"--targets={t}", f"--seed={seed}",
            "--frames=4", "--warmup=0", "--streams=2", "--detector=sam,rx,ace,cem", "--threshold=0.1",
            f"--rx-false-alarm={rate}", f"--ace-threshold={ace_thr}", f"--cem-threshold={cem_thr}",
            "--max-detections=20000", f"--variant={variant}", f"--dump-detections={dump}", "--quiet"]
    print(" ".join(args), flush=True)
    r = run(args, timeout=180)
    assert r.returncode == 0, r.stderr
    got = read_dumped_lists(dump)
    cfg = H.synth_config(h, w, b, t, seed=seed)
    thresholds = {"sam": 0.1, "rx": float(np.float32(H.rx_threshold(b, rate))), "ace": ace_thr, "cem": cem_thr}
    want = {k: oracle_frame_lists(H, oracle_lib, cfg, k, thresholds, 2, half=variant == "half")
            for k in range(4)}
    want["_thr"] = thresholds
    report = compare_pipeline_lists(got, want, h, w, 2, range(4), 1e-5)
    for (k, det), (ng, nw, nd) in sorted(report.items()):
        print(f"frame {k} {det}: {ng} device / {nw} oracle, {nd} differ at near-threshold or near-tie pixels")
        assert nw > 0


# ---- learned pre-screen from the CLI ----------------------------------------------

def write_model(path, bands, cut, weights, magic=b"HSIMLP01"):
    with open(path, "wb") as f:
        f.write(magic)
        f.write(np.array([bands], "<i4").tobytes())
        f.write(np.array([cut], "<f4").tobytes())
        f.write(np.array([len(weights)], "<u4").tobytes())
        f.write(np.asarray(weights, "<f4").tobytes())


def mlp_count(bands):
    return 64 * bands + 64 + 32 * 64 + 32 + 32 + 1


def test_hsi_detect_checks_the_prescreen_model_before_the_device(tmp_path):
    """A missing, foreign, truncated or mis-sized model file, or one whose
    cut or weights are not finite, is reported with exit 1 before any frame
    is produced; --prescreen with no file is a usage error."""
    base = ["hsi_detect", "--detector=rx", "--width=16", "--height=16", "--bands=16"]
    r = run(base + ["--prescreen="])
    assert r.returncode == 2 and "needs a model file" in r.stderr
    r = run(base + [f"--prescreen={tmp_path / 'none.bin'}"])
    assert r.returncode == 1 and "cannot open pre-screen model" in r.stderr
    foreign = tmp_path / "foreign.bin"
    write_model(foreign, 16, 0.0, np.zeros(mlp_count(16)), magic=b"NOTAMODL")
    r = run(base + [f"--prescreen={foreign}"])
    assert r.returncode == 1 and "bad magic" in r.stderr
    sized = tmp_path / "sized.bin"
    write_model(sized, 16, 0.0, np.zeros(mlp_count(16) - 1))
    r = run(base + [f"--prescreen={sized}"])
    assert r.returncode == 1 and f"(expected {mlp_count(16)})" in r.stderr
    short = tmp_path / "short.bin"
    write_model(short, 16, 0.0, np.zeros(mlp_count(16)))
    short.write_bytes(short.read_bytes()[:-8])
    r = run(base + [f"--prescreen={short}"])
    assert r.returncode == 1 and "truncated" in r.stderr
    nan_cut = tmp_path / "nan.bin"
    write_model(nan_cut, 16, np.nan, np.zeros(mlp_count(16)))
    r = run(base + [f"--prescreen={nan_cut}"])
    assert r.returncode == 1 and "cut is not finite" in r.stderr
    w = np.zeros(mlp_count(16))
    w[3] = np.inf
    bad_w = tmp_path / "badw.bin"
    write_model(bad_w, 16, 0.0, w)
    r = run(base + [f"--prescreen={bad_w}"])
    assert r.returncode == 1 and "non-finite weight" in r.stderr


@pytest.mark.gpu
def test_hsi_detect_prescreen_lists_equal_the_python_learned_path(tmp_path):
    """A network trained from Python (hsi_mi355x.prescreen) and saved, run by
    hsi_detect --prescreen --detector=rx,ace --precision=mixed over four
    seeded frames (two eager, two from captured graphs): each frame's RX and
    ACE lists equal the Python learned path's (the same one-pass statistics
    and screening, the build, precision-form rescoring of the candidate rows,
    detect) element for element; against the plain full-frame lists of the
    same frames the recall is printed and must reach 0.99."""
    import torch
    import hsi_mi355x as H
    from hsi_mi355x.prescreen import Prescreen
    h, w, b, t, seed, rate = 192, 256, 113, 8, 91, 1e-6
    cfg = H.synth_config(h, w, b, t, seed=seed)
    lib = H.synth_library(cfg)
    ctx = H.Context(0)
    ctx.set_library(lib)
    ps = Prescreen(ctx, b, rx_false_alarm_rate=rate, seed=4, capacity=max(h * w // 4, 4096), precision=H.MIXED)
    for k in (40, 41, 42, 43):
        ps.add_training_frame(torch.from_numpy(H.synth_frame(cfg, k)).cuda(), h, w)
    ps.train(epochs=40)
    ps.calibrate()
    model = str(tmp_path / "net.bin")
    ps.save(model)
    common = ["hsi_detect", f"--width={w}", f"--height={h}", f"--bands={b}", f"--targets={t}", f"--seed={seed}",
              "--frames=4", "--warmup=0", "--streams=2", "--detector=rx,ace", f"--rx-false-alarm={rate}",
              "--ace-threshold=0.5", "--max-detections=20000", "--quiet"]
    dump_l, dump_p = str(tmp_path / "learned.csv"), str(tmp_path / "plain.csv")
    r = run(common + ["--precision=mixed", f"--prescreen={model}", f"--dump-detections={dump_l}"], timeout=180)
    assert r.returncode == 0, r.stderr
    r = run(common + [f"--dump-detections={dump_p}"], timeout=180)
    assert r.returncode == 0, r.stderr
    got, plain = read_dumped_lists(dump_l), read_dumped_lists(dump_p)
    # the Python learned path, frame by frame
    again = Prescreen(ctx, b, rx_false_alarm_rate=rate, capacity=max(h * w // 4, 4096), precision=H.MIXED)
    again.load(model)
    thr = {"rx": float(np.float32(H.rx_threshold(b, rate))), "ace": 0.5}
    dets = torch.empty((20000, 4), dtype=torch.int32, device="cuda")
    nd = torch.empty(1, dtype=torch.int32, device="cuda")
    for k in range(4):
        cube = torch.from_numpy(H.synth_frame(cfg, k)).cuda()
        m = again.screen_fused(cube, h, w)
        for det, score, target in (("rx", m["rx"], None), ("ace", m["ace"], m["target"])):
            ctx.detect(score, target, h, w, H.DetectionParams(thr[det], 2, 20000, True), dets, nd)
            torch.cuda.synchronize()
            n = int(nd.item())
            want = [tuple(int(v) for v in row[:3]) for row in dets[:n].cpu().numpy()]
            have = [(x, y, tg if det == "ace" else -1) for (x, y, tg, _) in got.get((k, det), [])]
            if det == "rx":
                want = [(x, y, -1) for (x, y, _) in want]
            assert have == want, (k, det, len(have), len(want))
            ref = {(x, y) for (x, y, _, _) in plain.get((k, det), [])}
            hit = len(ref & {(x, y) for (x, y, _) in have})
            print(f"frame {k} {det}: learned {len(have)}, plain {len(ref)}, recall {hit / max(len(ref), 1):.4f}",
                  flush=True)
            assert len(ref) == 0 or hit / len(ref) >= 0.99
    ctx.close()


@pytest.mark.gpu
def test_hsi_detect_batch_lists_equal_the_stream_per_scene(tmp_path):
    """hsi_detect --batch over three ENVI scenes of different sizes (SAM, RX,
    ACE, CEM; with --autotune, graph replay): each scene's lists equal those
    hsi_detect gives the same cube as a one-frame stream, element for
    element; an RX-only batch mixes band counts (16 and 113)."""
    import hsi_mi355x as H
    shapes = [(40, 56), (64, 32), (24, 90)]
    cfg = H.synth_config(64, 96, 113, 6, seed=13)
    lib = H.synth_library(cfg)
    libcsv = str(tmp_path / "lib.csv")
    write_csv(libcsv, lib)
    hdrs = []
    for i, (hh, ww) in enumerate(shapes):
        full = H.synth_frame(cfg, i, stride=64 * 96).reshape(113, 64, 96)
        cube = np.ascontiguousarray(full[:, :hh, :ww])
        hdr = str(tmp_path / f"s{i}.hdr")
        write_envi(hdr, cube)
        hdrs.append(hdr)
    common = [f"--library={libcsv}", "--detector=sam,rx,ace,cem", "--threshold=0.1", "--rx-false-alarm=1e-3",
              "--ace-threshold=0.3", "--max-detections=20000", "--quiet"]
    dump_b = str(tmp_path / "batch.csv")
    r = run(["hsi_detect", "--source=envi", f"--envi={','.join(hdrs)}", "--batch", "--autotune", "--json",
             f"--dump-detections={dump_b}"] + common, timeout=180)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["scenes"] == 3 and summary["statistics_ok"] and summary["graph"]
    got = read_dumped_lists(dump_b)
    for i, hdr in enumerate(hdrs):
        dump_s = str(tmp_path / f"stream{i}.csv")
        r = run(["hsi_detect", "--source=envi", f"--envi={hdr}", "--frames=1", "--warmup=0", "--streams=1",
                 f"--dump-detections={dump_s}"] + common, timeout=180)
        assert r.returncode == 0, r.stderr
        want = read_dumped_lists(dump_s)
        for det in ("sam", "rx", "ace", "cem"):
            g, wnt = got.get((i, det), []), want.get((0, det), [])
            print(f"scene {i} {det}: batch {len(g)}, stream {len(wnt)}", flush=True)
            assert g == wnt, (i, det)
            assert len(wnt) > 0 or det == "sam"
    # RX only: scenes of 16 and 113 bands in one batch, no library
    c16 = H.synth_frame(H.synth_config(30, 50, 16, 2, seed=5), 0, stride=30 * 50).reshape(16, 30, 50)
    h16 = str(tmp_path / "b16.hdr")
    write_envi(h16, np.ascontiguousarray(c16))
    dump_m = str(tmp_path / "mixed_bands.csv")
    r = run(["hsi_detect", "--source=envi", f"--envi={h16},{hdrs[0]}", "--batch", "--detector=rx",
             "--rx-false-alarm=1e-2", "--quiet", f"--dump-detections={dump_m}"], timeout=180)
    assert r.returncode == 0, r.stderr
    mixed = read_dumped_lists(dump_m)
    assert len(mixed.get((0, "rx"), [])) > 0 and len(mixed.get((1, "rx"), [])) > 0
