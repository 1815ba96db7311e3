"""OGIP files without astropy (xpsi/Instrument.py:388-537, xpsi/Data.py:398-533).

The reader is checked on the NICER effective-area file the reference ships
(examples/examples_modeling_tutorial/model_data/nicer_20170601v005.arf,
copied into tests/golden by tests/golden/make_ogip.py): its bins abut, its
ENERGY column is the bin midpoint, its layout accounts for every byte of
the file, and its area peaks near 1.5 keV as NICER's does.  Response
matrices with variable-length rows, spectra and event lists are written by
tests/fits_writer.py and read back into the reference's dense layouts."""
import os

import numpy as np
import pytest

from tests.fits_writer import bintable, write
from xpsi_amd import ogip
from xpsi_amd.data import Data
from xpsi_amd.instrument import Instrument

ARF = os.path.join(os.path.dirname(__file__), "golden", "nicer_20170601v005.arf")


def test_reads_the_reference_arf():
    hdus = ogip.read_fits(ARF)
    assert [h["name"] for h in hdus] == ["PRIMARY", "SPECRESP"]
    a = ogip.hdu(hdus, "SPECRESP")
    h, c = a["header"], a["columns"]
    assert h["INSTRUME"] == "XTI" and h["NAXIS2"] == 3451 and h["NAXIS1"] == 32
    assert a["data_start"] == 5 * 2880
    assert a["data_start"] + ((a["data_bytes"] + 2879) // 2880) * 2880 == os.path.getsize(ARF)
    lo, hi = c["ENERG_LO"], c["ENERG_HI"]
    assert len(lo) == len(hi) == len(c["SPECRESP"]) == 3451
    assert np.array_equal(hi[:-1], lo[1:])
    assert np.array_equal(c["ENERGY"], (lo + hi) / np.float32(2.0))
    assert lo[0] == np.float32(0.1) and hi[0] == np.float32(0.105) and hi[-1] == np.float32(20.0)
    k = int(np.argmax(c["SPECRESP"]))
    assert 1.4 < lo[k] < 1.7 and 1700.0 < c["SPECRESP"][k] < 1950.0
    assert np.all(c["TRANS_DET_QE"] >= 0.0) and np.all(c["TRANS_DET_QE"] <= 1.0)


def _rmf(path, n_in, tlmin=0, detchans=80, seed=0):
    """A response with one or two channel groups per input row (some rows
    empty), its dense form returned alongside."""
    rng = np.random.default_rng(seed)
    elo = np.linspace(0.1, 0.1 + 0.005 * n_in, n_in, endpoint=False).astype(np.float32)
    ehi = (elo + np.float32(0.005)).astype(np.float32)
    dense = np.zeros((detchans, n_in))
    ngrp, fch, nch, mats = [], [], [], []
    for i in range(n_in):
        if i % 17 == 5:
            ngrp.append(0); fch.append([]); nch.append([]); mats.append([])
            continue
        groups = [(tlmin + (i * 3) % 40, 6)] + ([(tlmin + 50 + i % 10, 3)] if i % 2 else [])
        row = []
        for f, n in groups:
            v = rng.uniform(0.0, 1.0, n).astype(np.float32)
            dense[f - tlmin:f - tlmin + n, i] += v
            row.extend(v.tolist())
        ngrp.append(len(groups)); fch.append([g[0] for g in groups]); nch.append([g[1] for g in groups])
        mats.append(row)
    matrix = bintable("MATRIX", [("ENERG_LO", ">f4", elo), ("ENERG_HI", ">f4", ehi),
                                 ("N_GRP", ">i2", np.array(ngrp)), ("F_CHAN", ">i4", fch),
                                 ("N_CHAN", ">i4", nch), ("MATRIX", ">f4", mats)],
                      extra=[("INSTRUME", "XTI"), ("DETCHANS", detchans), ("TLMIN4", tlmin),
                             ("TLMAX4", tlmin + detchans - 1)])
    chans = np.arange(tlmin, tlmin + detchans)
    emin = (0.1 + 0.01 * np.arange(detchans)).astype(np.float32)
    ebounds = bintable("EBOUNDS", [("CHANNEL", ">i4", chans), ("E_MIN", ">f4", emin),
                                   ("E_MAX", ">f4", emin + np.float32(0.01))])
    write(path, [matrix, ebounds])
    return dense, elo, ehi, emin


def test_response_from_rmf_and_the_reference_arf(tmp_path):
    arf = ogip.hdu(ogip.read_fits(ARF), "SPECRESP")["columns"]["SPECRESP"]
    n_in = 300
    dense, elo, ehi, emin = _rmf(str(tmp_path / "r.rmf"), n_in)
    inst = Instrument.from_ogip_fits(str(tmp_path / "r.rmf"), ARF)
    rsp = dense * np.asarray(arf[:n_in], dtype=np.float64)
    keep_c = ~np.all(rsp == 0, axis=1)
    keep_i = ~np.all(rsp == 0, axis=0)
    np.testing.assert_array_equal(inst.matrix, rsp[keep_c][:, keep_i])
    inputs = np.arange(1, n_in + 1)[keep_i]
    np.testing.assert_array_equal(inst.energy_edges, np.append(elo[inputs - 1], ehi[inputs[-1] - 1]))
    assert inst.name == "XTI" and inst.matrix.shape[0] == keep_c.sum()
    # channel and input windows, and an RSP alone (no ARF)
    inst2 = Instrument.from_ogip_fits(str(tmp_path / "r.rmf"), None, min_channel=5, max_channel=60,
                                      min_input=10, max_input=200)
    sub = dense[5:61][:, 9:200]
    np.testing.assert_array_equal(inst2.matrix, sub[~np.all(sub == 0, axis=1)][:, ~np.all(sub == 0, axis=0)])


def test_pha_and_event_list(tmp_path):
    chans = np.arange(20, 60)
    counts = np.arange(40) * 3 + 1
    sp = bintable("SPECTRUM", [("CHANNEL", ">i4", chans), ("COUNTS", ">i4", counts)],
                  extra=[("EXPOSURE", 1234.5), ("BACKSCAL", 0.5), ("ANCRFILE", "a.arf"), ("RESPFILE", "r.rmf"),
                         ("HDUCLAS2", "TOTAL"), ("BACKFILE", "b.pha")])
    write(str(tmp_path / "s.pha"), [sp])
    d = Data.from_pha(str(tmp_path / "s.pha"), channels=np.arange(25, 35))
    np.testing.assert_array_equal(d.counts[:, 0], counts[5:15])
    assert d.exposure_time == 1234.5 and d.backfile == "b.pha" and d.last == 9 and d.counts.shape == (10, 1)
    with pytest.raises(ValueError):
        Data.from_pha(str(tmp_path / "s.pha"), channels=np.arange(55, 65))
    rng = np.random.default_rng(4)
    pi = rng.integers(30, 50, 5000)
    phase = rng.random(5000)
    ev = bintable("EVENTS", [("PI", ">i4", pi), ("PULSE_PHASE", ">f8", phase)], extra=[("EXPOSURE", 1e4)])
    write(str(tmp_path / "e.evt"), [ev])
    d = Data.from_evt(str(tmp_path / "e.evt"), n_phases=16, phase_shift=0.25)
    ref, _, _ = np.histogram2d(pi, (phase + 0.25) % 1.0, bins=[np.arange(30, 51) - 0.5, np.linspace(0, 1, 17)])
    np.testing.assert_array_equal(d.counts, ref)
    assert d.exposure_time == 1e4 and d.counts.sum() == 5000
