"""Write the golden vectors: run the compiled reference on every parity case.

Usage (in a container that has /root/reference):
    make -C oracle ref                 # builds oracle/_ref/run from the reference sources
    python tests/golden/make_golden.py [case ...]

For each case the reference executable runs in input mode with ``write_precision = 17`` (which
round-trips every double), and the frames after setup, after step 1 and after the last step are
kept for the variables listed in the case's ``output_flags`` (every k-th frame for the long cases of
``KEEP_EVERY``).  The step sizes of every step are
reconstructed exactly from the ``dt`` grid of the preceding frame,
``dt_step = epsilon * min(dt over the dt bounds)`` (the interior, widened by the ghost cells of open_moc sides) (evolution.cpp:62; grid.cpp:71-82).

A successful reference run ends in SIGABRT after its success message (evolution.cpp:54-56); that
exit is accepted.  Outputs: ``tests/golden/<case>.npz`` holding the input texts and the vectors.

Each case also runs once more with ``std_out_interval = 1`` (nothing else changed): the reference's
progress line of every step (PlasmaDomain::printUpdate, fileio.cpp:283-295), which carries the
modules' integer subcycle counts, goes to ``tests/golden/progress.json``.
``python tests/golden/make_golden.py --progress [case ...]`` refreshes only those lines.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from cases import CASES, KEEP_EVERY, ROOT, build_case  # noqa: E402

from spruce_amd.stateio import parse_config, read_out  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "run")


def dt_bounds(cfg, xdim, ydim):
    bc = [cfg.settings[k] for k in ("x_bound_1", "x_bound_2", "y_bound_1", "y_bound_2")]
    xl = 0 if bc[0] == "periodic" else 2
    xu = xdim - 1 if bc[1] == "periodic" else xdim - 3
    yl = 0 if bc[2] == "periodic" else 2
    yu = ydim - 1 if bc[3] == "periodic" else ydim - 3
    # open_moc sides take part in the step size (plasmadomain.cpp:155-159)
    xl -= 2 if bc[0] == "open_moc" else 0
    xu += 2 if bc[1] == "open_moc" else 0
    yl -= 2 if bc[2] == "open_moc" else 0
    yu += 2 if bc[3] == "open_moc" else 0
    return xl, xu, yl, yu


def run_case(name: str) -> str:
    st, cfg = build_case(name)
    state_text = st.to_text(17)
    config_text = cfg.to_text()
    with tempfile.TemporaryDirectory() as tmp:
        sp = os.path.join(tmp, "case.state")
        cp = os.path.join(tmp, "case.config")
        with open(sp, "w") as f:
            f.write(state_text)
        with open(cp, "w") as f:
            f.write(config_text)
        out_dir = os.path.join(tmp, "out")
        r = subprocess.run([REF_BIN, "-m", "input", "-o", out_dir, "-s", sp, "-c", cp],
                           capture_output=True, text=True, timeout=600, cwd=tmp)
        ok = r.returncode in (-6, 134) and "successfully reached" in r.stderr
        # with max_iterations = 0 the run writes the set-up frame and returns normally
        ok = ok or (r.returncode == 0 and parse_config(config_text).settings["max_iterations"] == "0")
        if not ok:
            raise RuntimeError("reference failed on %s (rc=%d):\n%s\n%s" % (name, r.returncode, r.stdout[-2000:],
                                                                            r.stderr[-2000:]))
        out = read_out(os.path.join(out_dir, "mhd.out"))
    parsed = parse_config(config_text)
    eps = float(parsed.settings["epsilon"])
    xl, xu, yl, yu = dt_bounds(parsed, st.xdim, st.ydim)
    nframes = len(out.frames)
    names = list(parsed.output_flags)
    dts = []
    for k in range(nframes - 1):
        g = out.frames[k]["dt"][xl:xu + 1, yl:yu + 1]
        g = g[~np.isnan(g)]
        dts.append(eps * float(g.min()))
    if name in KEEP_EVERY:
        keep = sorted(set(range(0, nframes, KEEP_EVERY[name])) | {nframes - 1})
    else:
        keep = sorted({0, min(1, nframes - 1), nframes - 1})
    data = np.stack([np.stack([out.frames[k][v] for v in names]) for k in keep])
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path,
                        state=np.frombuffer(state_text.encode(), dtype=np.uint8),
                        config=np.frombuffer(config_text.encode(), dtype=np.uint8),
                        names=np.array(names), frames=np.array(keep), data=data,
                        dt_steps=np.array(dts), bounds=np.array([xl, xu, yl, yu]))
    return path


PROGRESS = os.path.join(HERE, "progress.json")


def run_progress(name: str):
    """The reference's per-step progress lines of a case (the case's config with std_out_interval = 1)."""
    st, cfg = build_case(name)
    cfg.settings["std_out_interval"] = "1"
    with tempfile.TemporaryDirectory() as tmp:
        sp = os.path.join(tmp, "case.state")
        cp = os.path.join(tmp, "case.config")
        with open(sp, "w") as f:
            f.write(st.to_text(17))
        with open(cp, "w") as f:
            f.write(cfg.to_text())
        r = subprocess.run([REF_BIN, "-m", "input", "-o", os.path.join(tmp, "out"), "-s", sp, "-c", cp],
                           capture_output=True, text=True, timeout=600, cwd=tmp)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("Iter: ")]
    steps = int(cfg.settings["max_iterations"])
    if len(lines) != steps:
        raise RuntimeError("%s: %d progress lines for %d steps (rc=%d)" % (name, len(lines), steps, r.returncode))
    return lines


def main(argv):
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -C oracle ref")
    progress_only = "--progress" in argv
    names = [a for a in argv[1:] if not a.startswith("--")] or sorted(CASES)
    prog = json.load(open(PROGRESS)) if os.path.exists(PROGRESS) else {}
    for n in names:
        if not progress_only:
            p = run_case(n)
            print("%-28s -> %s (%d bytes)" % (n, os.path.relpath(p, ROOT), os.path.getsize(p)))
        prog[n] = run_progress(n)
        print("%-28s progress: %d lines" % (n, len(prog[n])))
    with open(PROGRESS, "w") as f:
        json.dump(dict(sorted(prog.items())), f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv)
