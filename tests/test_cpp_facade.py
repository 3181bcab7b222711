"""The C++ facade (arcwarden_amd/include/pic, the reference's arc:: API) and the
runner built on it (arcsim, and its run_deck command line).  CPU: every
facade header compiles on its own with a plain C++20 compiler.  GPU: the
facade check program, run_deck on a small deck, and arcsim on electrostatic,
Darwin and Yee decks with every diagnostics module, each output file read back
in the reference's layout (include/pic/diag/*.hpp, include/pic/run_meta.hpp,
tools/arcsim.cu)."""
import csv
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PIC = ROOT / "arcwarden_amd" / "include" / "pic"
DECKS = ROOT / "tests" / "golden" / "decks"


@pytest.mark.parametrize("header", sorted(str(p.relative_to(PIC)) for p in PIC.rglob("*.hpp")))
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
    assert ("darwin model" if em else "electrostatic model") in r.stdout
    man = list(csv.DictReader(open(out / "manifest.csv")))
    assert [int(m["step"]) for m in man] == [0, 25, 50, 75, 100]
    frame = list(csv.DictReader(open(out / man[-1]["file"])))
    assert len(frame) == min(64 * 8 * 256, 100000)  # run_deck samples at most 100000
    assert all(0.0 <= float(f["x"]) < 10.2633 + 1e-4 for f in frame[:1000])
    energy = list(csv.DictReader(open(out / "energy.csv")))
    assert [int(e["step"]) for e in energy] == [0, 25, 50, 75]
    q = [float(e["charge"]) for e in energy]
    assert max(q) - min(q) <= 1e-6 * abs(q[0])


def _meta(path):
    out = {}
    for line in open(path):
        k, v = line.split(None, 1)
        out[k] = v.strip()
    return out


_DIAG_DECK = """[grid]
nx = {nx}
ny = {ny}
Lx = {Lx}
Ly = {Ly}
[time]
dt = {dt}
nsteps = {nsteps}
dump_every = {dump}
[plasma]
noisy = true
seed = 20260627
[field]
model = {model}
c = 10.0
[background]
wce = 0.5
theta_deg = 30
[pump]
enable = true
mode = 3
w0 = 0.15
amp = 20.0
trmp = 1.0
toff = 1000
[diagnostics]
enable = spectrum kt phase_video snapshot em_energy phase_frames
kt_stride = 4
n_frames = 5
band_lo = 2
band_hi = 6
prefix = run
[species electrons]
density = 1.0
ppc = {ppc}
uth = 0.05 0.05 0.05
ufl = 0 0 0
"""


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["electrostatic", "darwin", "yee"])
def test_arcsim_every_diagnostic(gpu, tmp_path, model):
    exe = ROOT / "arcwarden_amd" / "tools" / "arcsim"
    assert exe.exists(), "built by arcwarden_amd/build.py"
    nx, ny, ppc, nsteps = 64, 8, 64, 40
    dt = 0.02 if model == "yee" else 0.05
    deck = tmp_path / "diag.ini"
    deck.write_text(_DIAG_DECK.format(nx=nx, ny=ny, Lx=nx * 0.5, Ly=ny * 0.5, dt=dt,
                                      nsteps=nsteps, dump=10, model=model, ppc=ppc))
    out = tmp_path / "out"
    r = _run([exe, deck, out, "--nsteps=40"], tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"{model} model" in r.stdout
    n = nx * ny * ppc
    # provenance
    meta = {}
    for line in open(out / "RUN_META.txt"):
        k, v = line.split("=", 1)
        meta[k.strip()] = v.strip()
    assert set(meta) == {"timestamp", "git_hash", "deck", "S_d", "eps_L", "cmdline"}
    assert meta["S_d"] == "1" and meta["eps_L"] == "0" and "--nsteps=40" in meta["cmdline"]
    assert (out / "deck_snapshot.ini").read_bytes() == deck.read_bytes()
    # phase frames and manifest (dump_every = 10)
    man = list(csv.DictReader(open(out / "manifest.csv")))
    assert [int(m["step"]) for m in man] == [0, 10, 20, 30, 40]
    assert [m["file"] for m in man] == [f"frame_{k:04d}.csv" for k in range(5)]
    frame = list(csv.DictReader(open(out / man[-1]["file"])))
    assert len(frame) == min(n, 100000) and list(frame[0]) == ["x", "vx"]
    # written with the stream's default 6 significant digits, as the reference
    # writes them: x just below Lx prints as Lx
    assert all(0.0 <= float(f["x"]) <= nx * 0.5 for f in frame)
    # k-t dump: rows of nx floats every kt_stride = 4 steps, steps 0..40
    km = _meta(out / "run_kt.meta")
    assert int(km["nx"]) == nx
    assert float(km["dt_sample"]) == pytest.approx(4 * dt)
    assert int(km["M"]) == 3 and int(km["band_lo"]) == 2 and int(km["band_hi"]) == 6
    kt = np.fromfile(out / "run_kt.bin", np.float32)
    if model == "electrostatic":  # no E_L grid on the electrostatic path
        assert kt.size == 0 and int(km["nsamp"]) == 0
    else:
        assert kt.size == 11 * nx and np.isfinite(kt).all() and int(km["nsamp"]) == 11
    # phase video: n_frames = 5 -> every 8 steps, 6 frames
    vm = _meta(out / "run_vid.meta")
    assert (int(vm["nframe"]), int(vm["nxb"]), int(vm["nvb"])) == (6, 256, 200)
    assert float(vm["vlo"]) == -6 and float(vm["vhi"]) == 6
    vp = np.fromfile(out / "run_vid_phase.bin", np.float32).reshape(6, 200, 256)
    assert (vp >= 0).all() and (vp == np.round(vp)).all()
    assert (vp.sum(axis=(1, 2)) > 0.999 * n).all() and (vp.sum(axis=(1, 2)) <= n).all()
    vf = np.fromfile(out / "run_vid_field.bin", np.float32)
    assert vf.size == 6 * 3 * nx
    # snapshot at the end: fields.csv, fhist.bin over every particle
    rows = list(csv.reader(open(out / "run_fields.csv")))
    assert rows[0] == ["x", "Epar", "Eperp", "Ey", "dBy", "dBmag"] and len(rows) == nx + 1
    fh = np.fromfile(out / "run_fhist.bin", np.float32)
    assert fh.size == 480 * 512 and 0.999 * n < fh.sum() <= n
    fm = _meta(out / "run_fhist.meta")
    assert (int(fm["nxb"]), int(fm["nvb"])) == (512, 480)
    assert float(fm["tsnap"]) == pytest.approx(nsteps * dt)
    # EM energies (every nsteps/40 = 1 step) on the Darwin and Yee paths
    em = list(csv.DictReader(open(out / "run_emenergy.csv")))
    if model == "electrostatic":
        assert em == []
    else:
        assert list(em[0]) == ["step", "time", "E_total", "B", "E_T"] and len(em) == 41
        assert all(float(e["B"]) > 0 for e in em)
    energy = out / "energy.csv"
    if model == "yee":
        mx = list(csv.DictReader(open(out / "run_maxwell.csv")))
        assert list(mx[0]) == ["step", "time", "WE", "WB", "divB_rms", "gauss_drift_rms"]
        assert [int(m["step"]) for m in mx] == [0, 10, 20, 30, 40]
        assert max(float(m["divB_rms"]) for m in mx) < 1e-5
        assert max(float(m["gauss_drift_rms"]) for m in mx) < 1e-4
    else:
        rows = list(csv.DictReader(open(energy)))
        assert [int(e["step"]) for e in rows] == [0, 10, 20, 30]
