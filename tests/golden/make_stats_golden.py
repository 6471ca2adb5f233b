"""Writes tests/golden/stats_<case>.json: the `statistics` block the serial
reference prints after every step (src/nhwave.F:3119-3208, formats 101/102:
E12.4), read from its log.txt, for the device statistics test.

The reference build is oracle/_ref/nhwave_ref (oracle/build_ref.sh); the
scenario inputs are the generators named below, with SCREEN_INTV = 0 so
that `statistics` runs after every step (src/nhwave.F:270-275).

Run in the build container:
    python tests/golden/make_stats_golden.py
"""
from __future__ import annotations

import json
import re
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from nhwave_amd import scenario as S  # noqa: E402
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent

CASES = {
    "beach_hydro": ("beach_runup", dict(steps=40, non_hydro=False)),
    "bumps48_nh": ("bumps_beach", dict(n=48, kglob=4, dx=200.0, steps=12, tol=1e-13)),
    "solitary512_nh": ("solitary_channel", dict(steps=20, tol=1e-10)),
}

NUM = r"[-+]?\d*\.\d+E[-+]\d+"


def parse_log(text: str) -> list[dict]:
    """The statistics blocks of a log.txt, in order."""
    lines = text.splitlines()
    out = []
    for q, line in enumerate(lines):
        if "STATISTICS" not in line:
            continue
        tdt = re.findall(NUM, lines[q + 2])
        vals = re.findall(NUM, lines[q + 4])
        out.append(dict(time=float(tdt[0]), dt=float(tdt[1]), mass_volume=float(vals[0]),
                        energy=float(vals[1]), max_eta=float(vals[2]), min_eta=float(vals[3]),
                        max_u=float(vals[4]), max_v=float(vals[5]), max_w=float(vals[6])))
    return out


def main() -> None:
    for name, (gen, kw) in CASES.items():
        s = getattr(S, gen)(**kw)
        with tempfile.TemporaryDirectory(prefix="nhw_stats_") as d:
            S.write_reference_inputs(s, Path(d))
            inp = Path(d) / "input.txt"
            inp.write_text(re.sub(r"(?m)^SCREEN_INTV = .*$", "SCREEN_INTV = 0.0", inp.read_text()))
            res = subprocess.run([str(O.REF_BIN)], cwd=d, capture_output=True, text=True, timeout=600)
            if res.returncode != 0 or "Normal Termination" not in res.stdout:
                raise RuntimeError(f"{name}: reference failed: {res.stdout[-2000:]}{res.stderr[-2000:]}")
            blocks = parse_log((Path(d) / "log.txt").read_text())
        if len(blocks) != s.steps:
            raise RuntimeError(f"{name}: {len(blocks)} statistics blocks for {s.steps} steps")
        rec = dict(generator=gen, kwargs=kw, format="E12.4 as printed by the reference", steps=blocks)
        (OUT / f"stats_{name}.json").write_text(json.dumps(rec, indent=1))
        print(name, len(blocks), "steps")


if __name__ == "__main__":
    main()
