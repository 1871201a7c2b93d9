"""The command-line driver (spatial_cell_sim, main.cu's run over the C ABI):
argument and input errors, without a GPU."""
import os
import subprocess

import pytest

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spatial-cell_amd", "spatial_cell_sim")


def run(args, cwd):
    return subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is missing: build it with spatial-cell_amd/build.py")


def test_unknown_argument_prints_usage(tmp_path):
    r = run(["--bogus"], tmp_path)
    assert r.returncode == 2 and "usage:" in r.stderr


def test_missing_documents_are_reported(tmp_path):
    r = run([], tmp_path)                                  # no ./config.json here
    assert r.returncode == 1 and "cannot read ./config.json" in r.stderr
    (tmp_path / "config.json").write_text("{}")
    r = run([], tmp_path)
    assert r.returncode == 1 and "cannot read ./setup.json" in r.stderr
    (tmp_path / "setup.json").write_text("{not json")
    r = run([], tmp_path)
    assert r.returncode == 1 and "setup.json" in r.stderr
