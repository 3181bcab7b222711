"""The C++ facade (arcwarden_amd/include/pic, the reference's arc:: API) and the
run_deck runner built on it.  CPU: every facade header compiles on its own
with a plain C++20 compiler.  GPU: the facade check program and run_deck on a
small deck."""
import csv
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PIC = ROOT / "arcwarden_amd" / "include" / "pic"
DECKS = ROOT / "tests" / "golden" / "decks"


@pytest.mark.parametrize("header", sorted(p.name for p in PIC.glob("*.hpp")))
def test_facade_header_compiles_alone(tmp_path, header):
    src = tmp_path / "tu.cpp"
    src.write_text(f'#include "pic/{header}"\nint main() {{ return 0; }}\n')
    subprocess.run(["g++", "-std=c++20", "-Wall", "-Werror", "-fsyntax-only",
                    f"-I{ROOT / 'include'}", f"-I{ROOT / 'arcwarden_amd' / 'include'}", str(src)],
                   check=True)


def _run(cmd, cwd, timeout=120):
    print(">>", " ".join(map(str, cmd)), flush=True)
    r = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    return r


@pytest.mark.gpu
def test_facade_check_program(gpu, tmp_path):
    exe = ROOT / "tests" / "cpp" / "facade_check"
    assert exe.exists(), "built by arcwarden_amd/build.py"
    r = _run([exe, DECKS / "two_stream.ini"], tmp_path)
    assert r.returncode == 0 and "PASSED" in r.stdout
    rows = list(csv.reader(open(tmp_path / "facade_energy.csv")))
    assert rows[0] == ["step", "time", "ke", "ee", "total", "charge", "max_e"]
    assert len(rows) >= 9


@pytest.mark.gpu
@pytest.mark.parametrize("em", [False, True])
def test_run_deck(gpu, tmp_path, em):
    exe = ROOT / "arcwarden_amd" / "tools" / "run_deck"
    assert exe.exists(), "built by arcwarden_amd/build.py"
    deck = tmp_path / "small.ini"
    deck.write_text("[grid]\nnx = 64\nny = 8\nLx = 10.2633\nLy = 1.0\n"
                    "[time]\ndt = 0.05\nnsteps = 100\ndump_every = 25\n"
                    "[plasma]\nnoisy = true\nseed = 20260627\n"
                    "[species beam_plus]\ndensity = 0.5\nppc = 128\nuth = 0.1 0.1 0.1\n"
                    "ufl = 1 0 0\n"
                    "[species beam_minus]\ndensity = 0.5\nppc = 128\nuth = 0.1 0.1 0.1\n"
                    "ufl = -1 0 0\n")
    out = tmp_path / "out"
    r = _run([exe, deck, out] + (["--em", "--c=20"] if em else []), tmp_path)
    assert r.returncode == 0
    assert ("Darwin EM" if em else "electrostatic") in r.stdout
    man = list(csv.DictReader(open(out / "manifest.csv")))
    assert [int(m["step"]) for m in man] == [0, 25, 50, 75, 100]
    frame = list(csv.DictReader(open(out / man[-1]["file"])))
    assert len(frame) == min(64 * 8 * 256, 100000)  # run_deck samples at most 100000
    assert all(0.0 <= float(f["x"]) < 10.2633 + 1e-4 for f in frame[:1000])
    energy = list(csv.DictReader(open(out / "energy.csv")))
    assert [int(e["step"]) for e in energy] == [0, 25, 50, 75]
    q = [float(e["charge"]) for e in energy]
    assert max(q) - min(q) <= 1e-6 * abs(q[0])
