"""Runs driven by a UDF deck: the deck's parameters and particles step the
same as a run built directly from the same values, and a restart deck
written from a run (resume = CONTINUE, Save_Restart_udf's record,
resume.cxx:10-74) resumes it: the restored run follows the uninterrupted
one to 1e-12 of each quantity's scale (atomic scatters reorder sums)."""
import os

import numpy as np
import pytest

from oracle import kref
from kapsel3_amd import udf

from test_gpu_restart import _close

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECK = os.path.join(HERE, "golden", "udf", "sedimentation.udf")
SHEAR = os.path.join(HERE, "golden", "udf", "shear.udf")


def test_deck_run_and_udf_restart(gpu, tmp_path):
    print("udf deck 32^3 x 4 (motor + passive), restart through a UDF deck", flush=True)
    dk = udf.read_deck(DECK)
    parts = udf.deck_particles(dk)
    a = gpu.Run(dk.params, dk.nparticles)
    a.set_particles(parts)
    a.init_fluid()
    assert not udf.restore_deck_state(a, dk)      # resume NEW
    a.step(4)
    out = tmp_path / "restart.udf"
    udf.save_restart_udf(a, DECK, out, include=os.path.join(HERE, "golden", "udf", "define.udf"))
    a.step(5)
    dk2 = udf.read_deck(out)
    b = gpu.Run(dk2.params, dk2.nparticles)
    assert udf.restore_deck_state(b, dk2)
    assert b.get_time().ts == 4
    b.step(5)
    _close(b, a, 1e-12)
    pa = a.get_particles()
    assert np.isfinite(pa["v"]).all() and np.abs(pa["v"]).max() > 0
    a.close()
    b.close()


def test_sheared_deck_and_udf_restart(gpu, tmp_path):
    """A Lees-Edwards deck (the strain passes a remap on the way) restarted
    through a UDF deck: the record carries degree_oblique
    (resume.cxx:67-73); the resumed process starts Shear_rate_eff at 0 as
    the reference's does, so the uninterrupted run is given the same."""
    print("udf sheared deck 32^3 x 3, restart through a UDF deck", flush=True)
    dk = udf.read_deck(SHEAR)
    a = gpu.Run(dk.params, dk.nparticles)
    a.set_particles(udf.deck_particles(dk))
    a.init_fluid()
    a.set_shear(*dk.shear)
    a.set_shear_state(0.999, dk.shear[0])
    a.step(4)
    g4 = a.shear_state()[0]
    assert 0.0 < g4 < 0.01                        # remapped
    out = tmp_path / "restart.udf"
    udf.save_restart_udf(a, SHEAR, out, include=os.path.join(HERE, "golden", "udf", "define.udf"))
    a.set_shear_state(g4, 0.0)
    a.step(5)
    dk2 = udf.read_deck(out)
    b = gpu.Run(dk2.params, dk2.nparticles)
    assert udf.restore_deck_state(b, dk2)
    assert b.shear_state() == (g4, 0.0)
    b.step(5)
    _close(b, a, 1e-12)
    assert a.shear_state() == b.shear_state()
    a.close()
    b.close()


def test_udf_restart_deck_resumes_gpu_and_oracle_alike(gpu, tmp_path):
    """One restart deck (written from a GPU run after 6 steps) restored into
    a fresh GPU run and into the oracle (Force_restore_parameters,
    resume.cxx:507-587): 10 more steps agree to 1e-10, AB2 history and the
    noise counter included."""
    from test_gpu_step import cmp_particles, rel
    print("udf restart deck -> GPU and oracle, 10 steps", flush=True)
    dk = udf.read_deck(DECK)
    a = gpu.Run(dk.params, dk.nparticles)
    a.set_particles(udf.deck_particles(dk))
    a.init_fluid()
    a.step(6)
    out = tmp_path / "restart.udf"
    udf.save_restart_udf(a, DECK, out, include=os.path.join(HERE, "golden", "udf", "define.udf"))
    a.close()
    dk2 = udf.read_deck(out)
    g = gpu.Run(dk2.params, dk2.nparticles)
    o = kref.Run(dk2.params, dk2.nparticles)
    assert udf.restore_deck_state(g, dk2) and udf.restore_deck_state(o, dk2)
    p0 = o.get_particles()
    g.step(10)
    o.step(10)
    assert g.get_time().ts == o.get_time().ts == 16
    cmp_particles(g.get_particles(), o.get_particles(), 1e-10, p0=p0, a=4.0)
    zg, zo = g.get_fluid(), o.get_fluid()
    assert rel(zg[0], zo[0]) < 1e-10 and rel(zg[1], zo[1]) < 1e-10
    g.close()
