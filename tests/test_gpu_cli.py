"""The vdifFILE2dedisp_PIPELINE program end to end on the GPU, against the
oracle replay of the reference main loop (windows, overlap carry, the
dumped window's filterbank and dedispersed series, SIGPROC header)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from cli_support import BUILD, reference_windows, write_vdif

pytestmark = pytest.mark.gpu
PIPE = os.path.join(BUILD, "vdifFILE2dedisp_PIPELINE")


def test_pipeline_cli_matches_oracle(gpu, orc, tmp_path):
    vfrb = gpu
    nchan, nave, nb, window = 128, 8, 5, 0.05
    nfft = 2 * nchan
    p = vfrb.params_derive(nfft, nave, nb, window, fullband=True)
    dms = np.array([0.0, 0.5, 1.0, 1.5, 2.0], np.float32)
    with open(tmp_path / "DM101.txt", "w") as f:
        for d in dms:
            f.write(f"{d:f}\n")
    off, maxd, _ = vfrb.delay_table(dms, p.nchanout, p.fch1, p.chanbw, p.tchan)
    ntime = p.maxnt - maxd
    # enough payload for two windows
    rows = p.maxnt + (p.maxnt - (maxd // nb) * nb) + nb
    nblocks = rows * nave + 8
    spec_pulse = [(1.0, 0.03, 0.001, vfrb.amp_for_snr(25, nave, p.nchanout))]
    write_vdif(tmp_path / "in.vdif", vfrb, nfft, nblocks, seed=4, pulses=spec_pulse)
    print(f"[case] CLI maxnt={p.maxnt} max_delay={maxd} ntime={ntime}", flush=True)
    r = subprocess.run([PIPE, "-d", "-nchan", str(nchan), "-nave", str(nave), "-nb", str(nb), "-i", "in.vdif",
                        "-o", "run", "--fullband", "--window", str(window), "--dump-window", "1",
                        "--cands", "run.cands"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + open(tmp_path / "run.log").read()
    assert "TOTAL TIME" in r.stdout
    data = open(tmp_path / "in.vdif", "rb").read()
    wins = reference_windows(data, nfft, nave, nb, p.nskip, p.nchanout, p.maxnt, maxd, orc)
    assert len(wins) >= 2
    bp = orc.bandpass_f64(wins[1])
    fil = np.fromfile(tmp_path / "run.nohead.fil", np.float32).reshape(p.maxnt, p.nchanout)
    assert np.max(np.abs(fil - bp)) / bp.std() < 1e-3
    ref = orc.dedisperse_f64(bp, off, ntime)
    dd = np.fromfile(tmp_path / "OUTDD.bin", np.float32).reshape(len(dms), 1 + ntime)
    assert np.array_equal(dd[:, 0], dms)
    assert np.max(np.abs(dd[:, 1:] - ref)) / ref.std() < 1e-3
    head = open(tmp_path / "run.head", "rb").read()
    assert head.startswith(struct.pack("<i", 12) + b"HEADER_START") and head.endswith(b"HEADER_END")
    assert struct.pack("<d", p.fch1) in head and struct.pack("<d", p.tchan) in head
    cands = open(tmp_path / "run.cands").read().splitlines()
    assert cands[0].startswith("#")
