"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code (CPU).

The oracle (oracle/md_oracle.c) is built with oracle/sanitize_oracle.c as its
driver; the library's host side (every source built with host-side
sanitizers, csrc/host_checks.cpp as the driver) runs the entry points that
need no GPU: the pair function, the host initialisers, md_frame_load on
well-formed and malformed files, and md_create's argument checks.  A finding
aborts the binary, so a clean exit with the driver's last line is the pass.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def _run(cmd, **kw):
    env = dict(os.environ, **SAN_ENV, **kw.pop("env", {}))
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=env, **kw)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_asan_ubsan(tmp_path):
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", f"SAN_DIR={tmp_path}"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    r = _run([str(tmp_path / "oracle_san")])
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "oracle checks passed"


def test_host_runtime_under_asan_ubsan(tmp_path):
    from moldyn_amd.build import HIPCC, build_host_checks
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        pytest.skip("needs hipcc")
    exe = build_host_checks(str(tmp_path / "build"))
    work = tmp_path / "run"
    work.mkdir()
    # HIP_VISIBLE_DEVICES hides any GPU: md_create must end in the device check
    r = _run([exe, str(work)], env={"MD_HOST_CHECKS_NO_DEVICE": "1", "HIP_VISIBLE_DEVICES": "-1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "host checks passed"
