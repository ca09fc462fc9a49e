"""CPU sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer) of
the FP64 oracle and of the library's host code (tests/sanitize/Makefile),
each run on its driver: the oracle's fixed scenarios must run clean and give
the same bytes as the unsanitised build, and the host-only entry points
(presets, ghost maps, QT, the reference-format readers on good and bad
files) must run clean and pass their checks."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")

pytestmark = pytest.mark.skipif(not (shutil.which("gcc") and os.path.exists("/opt/rocm/bin/hipcc")),
                                reason="needs gcc and hipcc")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    out = tmp_path_factory.mktemp("sanitize")
    r = subprocess.run(["make", "-s", "-C", SAN, f"OUT={out}", "-j8"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return out


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "abort_on_error=0:halt_on_error=1:detect_leaks=1:exitcode=66"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=67"
    return env


def test_oracle_clean_and_unchanged(built):
    outs = []
    for exe in ("oracle_checks_asan", "oracle_checks_plain"):
        dst = os.path.join(built, exe + ".bin")
        r = subprocess.run([os.path.join(built, exe), dst], capture_output=True, text=True, env=_env(), timeout=300)
        assert r.returncode == 0, (exe, r.returncode, r.stdout[-2000:], r.stderr[-4000:])
        assert "oracle checks ok" in r.stdout
        outs.append(open(dst, "rb").read())
    assert len(outs[0]) > 100000
    assert outs[0] == outs[1], "the sanitised oracle computes different bytes"


def test_host_entry_points_clean(built):
    scratch = os.path.join(built, "scratch")
    os.makedirs(scratch, exist_ok=True)
    r = subprocess.run([os.path.join(built, "host_checks"), scratch], capture_output=True, text=True, env=_env(),
                       timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-4000:])
    for part in ("presets ok", "ghost maps ok", "qt ok", "readers ok", "host checks ok"):
        assert part in r.stdout, r.stdout
