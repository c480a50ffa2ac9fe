"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
driver oracle/sanitize_check.c runs every oracle entry point the parity
tests use (SVD, the eight materials' return maps and stresses, substeps with
every kind of grid post-process, restart of F, rollouts) in FP64 and FP32
builds and must finish without a sanitizer report; a self-test build with a
read planted one past an array must be stopped by one of the sanitizers."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def _make(tmp_path, target):
    return subprocess.run(["make", "-s", "-C", ORACLE, f"SANDIR={tmp_path}", target], capture_output=True,
                          text=True, timeout=300)


@pytest.fixture(scope="module")
def have_sanitizers(tmp_path_factory):
    if shutil.which("make") is None or shutil.which(os.environ.get("CC", "cc")) is None:
        pytest.skip("no C toolchain")
    d = tmp_path_factory.mktemp("san_probe")
    src = d / "probe.c"
    src.write_text("int main(void) { return 0; }\n")
    r = subprocess.run([os.environ.get("CC", "cc"), "-fsanitize=address,undefined", "-o", str(d / "probe"),
                        str(src)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("the C compiler has no sanitizer runtime")
    return True


def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path, have_sanitizers):
    r = _make(tmp_path, "sanitize")
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.count("0 failing calls") == 2
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_sanitizer_catches_a_planted_out_of_bounds_read(tmp_path, have_sanitizers):
    r = _make(tmp_path, f"{tmp_path}/sanitize_selftest")
    assert r.returncode == 0, r.stderr[-4000:]
    run = subprocess.run([str(tmp_path / "sanitize_selftest")], capture_output=True, text=True, timeout=120)
    assert run.returncode != 0
    # UBSan's object-size check or ASan's heap-buffer-overflow, whichever fires first
    assert ("runtime error" in run.stderr and "insufficient space" in run.stderr) or \
        "heap-buffer-overflow" in run.stderr, run.stderr[-2000:]
