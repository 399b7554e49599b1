"""Command-line behaviour of the drop-in programs that needs no GPU: usage
and exit codes of the reference parsers (PIPELINE:241-245,1245-1288;
TIMESERIES:1196-1258; STATS:811-834)."""
import os
import subprocess

import pytest

from cli_support import BUILD

PIPE = os.path.join(BUILD, "vdifFILE2dedisp_PIPELINE")
TS = os.path.join(BUILD, "vdifDADA2timeseries")
ST = os.path.join(BUILD, "vdifDADA2stats")


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


def run_prog(prog, args, cwd):
    return subprocess.run([prog] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(TS), reason="CLI not built")
def test_timeseries_usage_and_exit_codes(tmp_path):
    r = run_prog(TS, [], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run_prog(TS, ["-d", "-bogus"], tmp_path)
    assert r.returncode == 1 and "Error! Unknown argument: -bogus" in r.stderr
    r = run_prog(TS, ["-h"], tmp_path)   # in the usage text, but not parsed by the reference
    assert r.returncode == 1 and "Unknown argument: -h" in r.stderr


@pytest.mark.skipif(not os.path.exists(ST), reason="CLI not built")
def test_stats_usage_and_exit_codes(tmp_path):
    r = run_prog(ST, ["-h"], tmp_path)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run_prog(ST, ["-bogus"], tmp_path)
    assert r.returncode == 1 and "Error! Unknown argument: -bogus" in r.stderr
