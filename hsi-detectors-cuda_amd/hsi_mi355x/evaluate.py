"""Score a dumped score map against a class mask on the GPU.

The counterpart of the reference's scripts/eval_detection.py, with the same
arguments and report; the AUC, the operating-point counts and the per-class
table come from hsi_evaluate on the device:

    python -m hsi_mi355x.evaluate /tmp/F_1.f32 F_1_gt.u8 --width=696 --height=520

Scores are raw little-endian float32 (hsi_detect --dump-score), the mask raw
uint8 (labels above 8 are unannotated and left out). SAM angles are ranked
lower-is-better; pass --higher-is-better for RX/ACE/CEM maps.
"""
import argparse
import pathlib
import sys

import numpy as np

CLASS_NAMES = {0: "background", 1: "blood", 2: "ketchup", 3: "artificial_blood", 4: "beetroot_juice",
               5: "poster_paint", 6: "tomato_concentrate", 7: "acrylic_paint", 8: "uncertain_blood"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("score_map", type=pathlib.Path, help="raw float32 score map")
    ap.add_argument("ground_truth", type=pathlib.Path, help="raw uint8 class mask")
    ap.add_argument("--width", type=int, required=True)
    ap.add_argument("--height", type=int, required=True)
    ap.add_argument("--target", type=int, default=1, help="positive class [1=blood]")
    ap.add_argument("--threshold", type=float, default=0.10, help="operating point (radians for SAM)")
    ap.add_argument("--higher-is-better", action="store_true", help="RX/ACE/CEM score maps")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)

    import torch
    import hsi_mi355x as H

    pixels = args.width * args.height
    score = np.fromfile(args.score_map, dtype="<f4")
    truth = np.fromfile(args.ground_truth, dtype=np.uint8)
    for name, arr in (("score map", score), ("ground truth", truth)):
        if arr.size != pixels:
            raise SystemExit(f"{name} has {arr.size} entries, expected {pixels} ({args.height}x{args.width})")
    if not (truth == args.target).any():
        raise SystemExit(f"class {args.target} ({CLASS_NAMES.get(args.target)}) is absent from "
                         f"{args.ground_truth.name}")
    dev = f"cuda:{args.device}"
    ctx = H.Context(args.device)
    r = ctx.evaluate(torch.from_numpy(score).to(dev), torch.from_numpy(truth).to(dev), args.target, 8,
                     args.higher_is_better, args.threshold)
    ctx.close()
    tp, fp, fn, tn = r["tp"], r["fp"], r["fn"], r["tn"]
    precision = tp / (tp + fp) if tp + fp else float("nan")
    recall = tp / (tp + fn) if tp + fn else float("nan")
    f1 = 2 * precision * recall / (precision + recall) if precision + recall else float("nan")
    name = CLASS_NAMES.get(args.target, str(args.target))
    print(f"target class      {args.target} ({name})")
    print(f"annotated pixels  {r['positives'] + r['negatives']} of {pixels}")
    print(f"positives         {r['positives']}")
    print(f"\nROC AUC           {r['auc']:.4f}")
    print(f"\nat threshold {args.threshold:.3f}")
    print(f"  precision       {precision:.4f}")
    print(f"  recall          {recall:.4f}")
    print(f"  f1              {f1:.4f}")
    print(f"  tp {tp}  fp {fp}  fn {fn}  tn {tn}")
    print("\nmean score by class")
    target_mean = r["classes"][args.target][1]
    for label, (count, mean, mn) in sorted(r["classes"].items()):
        gap = mean - target_mean
        marker = "  <- target" if label == args.target else f"  gap {gap:+.4f}"
        print(f"  {CLASS_NAMES.get(label, str(label)):>20} n={count:7d}  mean {mean:.4f}  min {mn:.4f}{marker}")
    return 1 if np.isnan(r["auc"]) else 0


if __name__ == "__main__":
    sys.exit(main())
