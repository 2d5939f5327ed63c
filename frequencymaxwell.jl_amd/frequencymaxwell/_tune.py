"""Tile-width autotuner of the specialised sweep kernels.

The x passes move P lines per tile, the y and z passes P columns; P sets the
registers and LDS of a block and therefore its occupancy.  The tuner picks P
per (pass, axis length, precision, pipelines) from measurements:

  build    compile one variant library per candidate width set into --libdir
           (CBS_TUNE_XP / CBS_TUNE_YP / CBS_TUNE_ZP force one width for every
           length; the LDS budget still caps it)
  measure  (on the GPU) time the five sweeps of every specialised length with
           each variant, one child process per library, and write JSON
  write    keep the fastest width of each (pass, length, precision, S) and
           write csrc/cbs_tuned.h and a markdown summary

    python -m frequencymaxwell._tune build --libdir DIR
    python -m frequencymaxwell._tune measure --libdir DIR --out tune.json
    python -m frequencymaxwell._tune write --json tune.json --summary tuning.md

After ``write``, rebuild the library (``_build.py``).
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
TUNED_H = CSRC / "cbs_tuned.h"

# (x lines, y columns, z columns) forced in one variant build each
CANDIDATES = [(1, 4, 2), (2, 8, 4), (4, 16, 8), (8, 32, 16)]
LENGTHS = (32, 64, 96, 128, 160, 192, 224, 256, 288, 320)
CROSS = 160  # the two other axes of a tuning grid
KERNELS = ("xfwd", "y_fwd", "z", "y_inv", "xinv", "x_fused")


def variant_name(c) -> str:
    return f"libcbs_x{c[0]}_y{c[1]}_z{c[2]}.so"


def cmd_build(libdir: Path):
    from . import _build

    libdir.mkdir(parents=True, exist_ok=True)
    for c in CANDIDATES:
        defs = (f"CBS_TUNE_XP={c[0]}", f"CBS_TUNE_YP={c[1]}", f"CBS_TUNE_ZP={c[2]}")
        out = _build.build(defines=defs, out=libdir / variant_name(c))
        print(f"built {out}", flush=True)


def _grid(axis: int, L: int):
    g = [CROSS, CROSS, CROSS]
    g[axis] = L
    return tuple(g)


def child_measure(out: Path, reps: int):
    """Runs with CBS_LIBRARY set to one variant: time the sweeps of every length."""
    from . import engine

    res = 50e-9
    k_bg = 2 * np.pi * 1.33 / 500e-9
    eps = 0.3 * k_bg ** 2
    rows = []
    for L in LENGTHS:
        for prec in ("fp32", "fp64"):
            for avg in (False, True):
                for axis, ks in ((0, (0, 4, 5)), (1, (1, 3)), (2, (2,))):
                    shape = _grid(axis, L)
                    spec = engine.GreenSpec(shape=shape, resolution=(res,) * 3,
                                            shifts=(0.25,) * 3 if avg else (0.0,) * 3,
                                            use_averaging=avg, k_square=k_bg ** 2 + 1j * eps,
                                            eps_imag=eps)
                    V = np.full(shape, -1j * eps, dtype=np.complex128, order="F")
                    S = np.ones(shape + (3,), dtype=np.complex128, order="F")
                    p = engine.CBSProblem(V, spec, precision=prec)
                    p.set_source(S)
                    ms, _ = p.time_sweeps(prec, reps=reps)
                    p.close()
                    rows.append({"L": L, "prec": prec, "S": 2 if avg else 1, "axis": axis,
                                 "ms": sum(ms[k] for k in ks), "ms_k": list(ms)})
        print(f"  length {L} done", flush=True)
    out.write_text(json.dumps(rows))


def cmd_measure(libdir: Path, out: Path, reps: int):
    allrows = {}
    for c in CANDIDATES:
        lib = libdir / variant_name(c)
        tmp = out.with_suffix(f".{c[0]}_{c[1]}_{c[2]}.json")
        env = dict(os.environ, CBS_LIBRARY=str(lib))
        print(f"variant {c}", flush=True)
        t = time.time()
        subprocess.run([sys.executable, "-m", "frequencymaxwell._tune", "child", "--out", str(tmp),
                        "--reps", str(reps)], env=env, check=True, cwd=str(PKG.parent))
        print(f"variant {c}: {time.time() - t:.1f} s", flush=True)
        allrows["x%d_y%d_z%d" % c] = json.loads(tmp.read_text())
    out.write_text(json.dumps({"candidates": CANDIDATES, "rows": allrows}, indent=0))


def objective(r, passid: int) -> float:
    """Cost of one width for table pass `passid`: 0 the x sweeps (x forward +
    x inverse of a Krylov operator application), 3 the fused x pass of a Born
    sweep, 1 y forward + inverse, 2 the z pass."""
    k = r.get("ms_k")
    if k is None:
        return r["ms"]
    return {0: k[0] + k[4], 3: k[5], 1: k[1] + k[3], 2: k[2]}[passid]


def cmd_write(js: Path, summary: Path | None):
    data = json.loads(js.read_text())
    cands = [tuple(c) for c in data["candidates"]]
    rows = data["rows"]
    # best[(pass, L, eb, S)] = (width, ms); y and z keyed by S = 0 (sum over S)
    times = {}
    for c in cands:
        for r in rows["x%d_y%d_z%d" % c]:
            eb = 8 if r["prec"] == "fp32" else 16
            axis = r["axis"]
            w = c[axis]
            # the x axis sets two tables: the x sweeps (0) and the fused x pass (3)
            for passid in ((0, 3) if axis == 0 and "ms_k" in r else (axis,)):
                S = r["S"] if passid in (0, 3) else 0
                key = (passid, r["L"], eb, S)
                times.setdefault(key, {}).setdefault(w, 0.0)
                times[key][w] += objective(r, passid)
    entries, lines = [], []
    for key in sorted(times):
        tw = times[key]
        best = min(tw, key=tw.get)
        entries.append((*key, best))
        lines.append((key, best, tw))
    body = ",\n".join(f"    {{{p}, {L}, {eb}, {S}, {w}}}" for p, L, eb, S, w in entries)
    TUNED_H.write_text(
        "// Tile widths measured on MI355X by frequencymaxwell/_tune.py (generated;\n"
        "// an empty table leaves the heuristic widths in force).\n"
        "// {pass (0 x lines, 1 y columns, 2 z columns), length, bytes per complex,\n"
        "//  pipelines (x pass only), width}\n"
        "#pragma once\n\nnamespace cbs {\n\n"
        f"constexpr TileTune kTileTune[] = {{\n{body}}};\n"
        f"constexpr int kTileTuneN = {len(entries)};\n\n}}  // namespace cbs\n")
    print(f"wrote {TUNED_H} ({len(entries)} entries)")
    if summary:
        names = {0: "x (xfwd+xinv)", 3: "x fused pass", 1: "y (fwd+inv)", 2: "z"}
        out = ["| pass | length | precision | S | chosen width | ms per width |", "|---|---|---|---|---|---|"]
        for (p, L, eb, S), best, tw in lines:
            per = ", ".join(f"{w}: {tw[w]:.4f}" for w in sorted(tw))
            out.append(f"| {names[p]} | {L} | {'fp32' if eb == 8 else 'fp64'} | "
                       f"{S if p in (0, 3) else '1+2'} | {best} | {per} |")
        summary.write_text("\n".join(out) + "\n")
        print(f"wrote {summary}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build")
    b.add_argument("--libdir", type=Path, required=True)
    m = sub.add_parser("measure")
    m.add_argument("--libdir", type=Path, required=True)
    m.add_argument("--out", type=Path, required=True)
    m.add_argument("--reps", type=int, default=20)
    c = sub.add_parser("child")
    c.add_argument("--out", type=Path, required=True)
    c.add_argument("--reps", type=int, default=20)
    w = sub.add_parser("write")
    w.add_argument("--json", type=Path, required=True)
    w.add_argument("--summary", type=Path)
    a = ap.parse_args(argv)
    if a.cmd == "build":
        cmd_build(a.libdir)
    elif a.cmd == "measure":
        cmd_measure(a.libdir, a.out, a.reps)
    elif a.cmd == "child":
        child_measure(a.out, a.reps)
    else:
        cmd_write(a.json, a.summary)


if __name__ == "__main__":
    main()
