"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU):
the library's host side and the DWG_HIP command line are built with the
sanitizers on their host code (`make -C dwg_amd/csrc asan`), then
tests/host/host_checks.cpp exercises the .xyz reader on malformed files, the
voxel grid on degenerate boxes, the writers and the C ABI's argument checks,
and the command line is driven through its error paths (missing, empty and
three-column inputs, options without or with malformed values).  Any
sanitizer report fails the test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dwg_amd", "csrc")
ASAN = os.path.join(CSRC, "_build", "asan")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
           HIP_VISIBLE_DEVICES="")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-j", str(min(16, os.cpu_count() or 4)), "-C", CSRC, "asan"], check=True,
                   capture_output=True)
    return ASAN


def _clean(r):
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text and "LeakSanitizer" not in text, text


def test_host_checks(built, tmp_path):
    r = subprocess.run([os.path.join(built, "host_checks"), str(tmp_path)], capture_output=True, text=True, env=ENV,
                       timeout=120)
    _clean(r)
    assert r.returncode == 0 and "host checks ok" in r.stdout, r.stdout + r.stderr


def test_cli_error_paths(built, tmp_path):
    exe = os.path.join(built, "DWG_HIP")
    (tmp_path / "empty.xyz").write_text("")
    (tmp_path / "three.xyz").write_text("0 0 0\n1 0 0\n0 1 0\n0 0 1\n")
    base = ["--in_path", str(tmp_path), "--out_path", str(tmp_path)]
    cases = [
        (base + ["--in_name", "missing.xyz"], 2, "cannot read point cloud"),
        (base + ["--in_name", "empty.xyz"], 2, "cannot read point cloud"),
        (base + ["--in_name", "three.xyz", "--grid_reso"], 1, "has no value"),
        (base + ["--in_name", "three.xyz", "--grid_reso", "many"], 1, "bad value"),
        (base + ["--in_name", "three.xyz", "--total_iter", "9999999999999"], 1, "bad value"),
        # a readable input gets as far as the device, which this machine may lack
        (base + ["--in_name", "three.xyz"], None, "point cloud cnt = 4"),
    ]
    for args, rc, text in cases:
        r = subprocess.run([exe] + args, capture_output=True, text=True, env=ENV, timeout=120)
        _clean(r)
        assert text in r.stdout + r.stderr, (args, r.stdout, r.stderr)
        if rc is not None:
            assert r.returncode == rc, (args, r.returncode, r.stderr)
