"""CPU checks of bench.py's contract that need no GPU: a request for more
than one GPU exits with a message before anything touches a device, and the
measured HBM traffic it reports comes from the committed counter summary,
whose figure agrees with the counter values and calibration written in it."""
import importlib.util
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_traffic_matches_the_counter_summary():
    text = (ROOT / "profiles" / "push_deposit_es_512x512_ppc256_pmc.txt").read_text()
    fetch = float(re.search(r"FETCH_SIZE [\d,-]+ KiB/dispatch -> x([\d.]+) = ([\d,]+) KiB",
                            text).group(2).replace(",", ""))
    write = float(re.search(r"WRITE_SIZE [\d,-]+ KiB/dispatch -> x([\d.]+) = ([\d,]+) KiB",
                            text).group(2).replace(",", ""))
    traffic = _bench().traffic_from_profiles()
    assert traffic == pytest.approx((fetch + write) * 1024, rel=1e-3)
    # within 10% of the algorithmic 32 B x 67,108,864 particles
    assert traffic == pytest.approx(32 * 67_108_864, rel=0.1)
