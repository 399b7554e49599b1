"""Command-line behaviour of the drop-in programs that needs no GPU: usage
and exit codes of the reference parsers (PIPELINE:241-245,1245-1288)."""
import os
import subprocess

import pytest

from cli_support import BUILD

PIPE = os.path.join(BUILD, "vdifFILE2dedisp_PIPELINE")


def run(args, cwd):
    return subprocess.run([PIPE] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(PIPE), reason="CLI not built")
def test_pipeline_usage_and_exit_codes(tmp_path):
    r = run([], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run(["-d", "-bogus"], tmp_path)
    assert r.returncode == 1 and "Error! Unknown argument: -bogus" in r.stderr
    r = run(["-nave", "4"], tmp_path)   # no -d: shapes stay unset
    assert r.returncode == 1
