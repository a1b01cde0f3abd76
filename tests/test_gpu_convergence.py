"""4096-spp means: the device estimate and an independent oracle estimate of
the same pixel integrals agree within 3 sigma of their Monte Carlo standard
error.

The device renders seed 0 and the oracle seed 1, so the two are independent
estimates (with the same seed they share every random number and agree to
1e-4, which says nothing about the expectation).  Each side renders 16
batches of 256 consecutive samples of a 4096-sample sequence; the spread of
the batch means gives each side's standard error per pixel.  With 15
degrees of freedom a 3-sigma bound holds for about 99.1 % of pixels when the
expectations agree, so the per-pixel bar is 98.5 %, and the frame mean is
held to 3 sigma of its own error.  A biased estimator (a missing transport
term, a wrong pdf) moves whole regions and fails both."""
import sys

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import scenes
from psycles_amd.render import RendererBackend

pytestmark = pytest.mark.gpu

TOTAL = 4096
BATCHES = 16


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def batch_means_device(snap, settings):
    be = RendererBackend(0)
    comp = be.compile_scene(snap)
    assert comp.ok()
    sess = be.create_session(comp.scene, settings)
    per = TOTAL // BATCHES
    means, previous = [], None
    for b in range(BATCHES):
        sess.render_device(b * per, per, TOTAL)
        comb = sess.read_film()[0][..., :3].astype(np.float64)
        chunk = comb if previous is None else comb - previous
        previous = comb
        means.append(chunk / per)
    return np.stack(means)


def batch_means_oracle(snap, settings):
    osc = ob.OracleScene(snap)
    per = TOTAL // BATCHES
    means = []
    for b in range(BATCHES):
        film, _ = osc.render(settings, b * per, per, TOTAL)
        means.append(film.combined[..., :3].astype(np.float64) / per)
    return np.stack(means)


@pytest.mark.parametrize("name", ["cornell", "cornell_lamps"])
def test_4096_spp_means_within_3_sigma(name):
    snap = scenes.cornell_box() if name == "cornell" else scenes.cornell_lamps()
    w, h = 32, 20
    log(f"case: {name} {w}x{h} at {TOTAL} spp, device seed 0 vs oracle seed 1")
    s_dev = scenes.cornell_settings(w, h)
    s_ora = scenes.cornell_settings(w, h)
    s_ora.seed = 1
    dev = batch_means_device(snap, s_dev)
    ora = batch_means_oracle(snap, s_ora)
    m_dev, m_ora = dev.mean(0), ora.mean(0)
    se2 = dev.var(0, ddof=1) / BATCHES + ora.var(0, ddof=1) / BATCHES
    z = np.abs(m_dev - m_ora) / np.sqrt(np.maximum(se2, 1e-30))
    inside = np.all((z <= 3.0) | (se2 < 1e-20), axis=-1)
    frac = float(inside.mean())
    fd, fo = dev.mean(axis=(1, 2)), ora.mean(axis=(1, 2))
    frame_se = np.sqrt(fd.var(0, ddof=1) / BATCHES + fo.var(0, ddof=1) / BATCHES)
    frame_z = np.abs(fd.mean(0) - fo.mean(0)) / frame_se
    log(f"  pixels within 3 sigma: {frac:.4f}; frame mean device {fd.mean(0)}, oracle {fo.mean(0)}, "
        f"z {frame_z}")
    assert frac >= 0.985
    assert np.all(frame_z <= 3.0)
