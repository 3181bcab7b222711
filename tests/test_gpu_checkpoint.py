"""Restart from a checkpoint file: an electrostatic and a Darwin run (CIC and
Esirkepov currents) saved after some steps and restored into a new
Simulation continue exactly as the uninterrupted run -- rho, every field
grid and the particle multiset bit for bit, the energies to summation
order -- and a mismatched simulation refuses the file."""
import numpy as np
import pytest

from arcwarden_amd import _lib as L
from arcwarden_amd.sim import SimConfig, Simulation

pytestmark = pytest.mark.gpu


def _multiset(st):
    cols = [st[k] for k in ("x", "y", "ux", "uy", "uz")]
    order = np.lexsort(cols[::-1])
    return np.stack([c[order] for c in cols])


def _es():
    g = L.grid(64, 48, 32.0, 24.0)
    rp = L.run_params(dt=0.1, qm=-1.0, eps0=1.0, noisy_load=1, rng_seed=5, dump_every=3,
                      B0=(0.0, 0.0, 0.4))
    sp = [L.species("plus", 0.5, 16, (0.02,) * 3, (0.2, 0, 0)),
          L.species("minus", 0.5, 16, (0.02,) * 3, (-0.2, 0, 0))]
    return g, rp, sp, SimConfig(has_b0=True, sort_every=7)


def _darwin(current):
    g = L.grid(64, 32, 32.0, 16.0)
    rp = L.run_params(dt=0.05, qm=-1.0, eps0=1.0, c=5.0, n0=1.0, noisy_load=1, rng_seed=6,
                      B0=(0.4, 0.0, 0.23), dump_every=3)
    rp.pump = 1
    rp.pump_ex, rp.pump_ey, rp.pump_ez = 1e-3, 2e-3, -2e-3
    rp.pump_k0, rp.pump_w0, rp.pump_trmp, rp.pump_toff = 2 * np.pi * 3 / 32.0, 0.15, 1.0, 100.0
    sp = [L.species("e", 1.0, 32, (0.1,) * 3, (0, 0, 0.05))]
    return g, rp, sp, SimConfig(has_b0=True, field_model=L.AW_MODEL_DARWIN, sort_every=5,
                                current_deposit=current)


@pytest.mark.parametrize("kind", ["es", "darwin_cic", "darwin_esirkepov"])
def test_restart_continues_bit_for_bit(gpu, tmp_path, kind):
    g, rp, sp, cfg = (_es() if kind == "es" else
                      _darwin(L.AW_CURRENT_CIC if kind == "darwin_cic" else
                              L.AW_CURRENT_ESIRKEPOV))
    n_save, n_total = 9, 17
    a = Simulation(g, rp, sp, cfg)
    a.init()
    for n in range(n_save):
        a.step(n)
    path = tmp_path / "run.arcw"
    a.save_checkpoint(str(path), n_save, deck="[grid]\nnx = 64\n")
    for n in range(n_save, n_total):
        a.step(n)
    b = Simulation(g, rp, sp, cfg)
    b.restore_checkpoint(str(path))
    for n in range(n_save, n_total):
        b.step(n)
    fa, fb = a.fields(), b.fields()
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), k
    sa, sb = a.sources_all(), b.sources_all()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(_multiset(a.particles()), _multiset(b.particles()))
    ha = [h for h in a.history() if h["step"] >= n_save]
    hb = b.history()
    assert [h["step"] for h in ha] == [h["step"] for h in hb]
    for x, y in zip(ha, hb):
        assert x["ee"] == y["ee"] and x["charge"] == y["charge"]
        assert x["ke"] == pytest.approx(y["ke"], rel=1e-12)
    # the header carries the step, time and seed
    from arcwarden_amd import _lib as LL
    import ctypes as C
    ck = C.c_void_p()
    LL.call("aw_checkpoint_open", str(path).encode(), C.byref(ck))
    h = LL.AwCheckpointHeader()
    LL.call("aw_checkpoint_info", ck, C.byref(h), None, None, None)
    LL.call("aw_checkpoint_close", ck)
    assert h.step == n_save and h.time == pytest.approx(n_save * rp.dt) and h.rng_seed == rp.rng_seed
    a.close()
    b.close()


def test_restore_refuses_a_mismatched_run(gpu, tmp_path):
    g, rp, sp, cfg = _es()
    a = Simulation(g, rp, sp, cfg)
    a.init()
    a.step(0)
    path = tmp_path / "es.arcw"
    a.save_checkpoint(str(path), 1)
    a.close()
    sp2 = [L.species("plus", 0.5, 8, (0.02,) * 3, (0.2, 0, 0))]
    b = Simulation(g, rp, sp2, cfg)
    with pytest.raises(L.AwError, match="does not match"):
        b.restore_checkpoint(str(path))
    b.close()
    g2, rp2, sp3, cfg2 = _darwin(L.AW_CURRENT_CIC)
    c = Simulation(g2, rp2, sp3, cfg2)
    with pytest.raises(L.AwError):
        c.restore_checkpoint(str(path))
    c.close()
