"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
`make -C oracle sanitize` builds oracle/selftest.c with the oracle sources
(-fsanitize=address,undefined, no recovery) and runs every entry point on
small synthetic inputs; any finding fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_sanitizers(tmp_path):
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("no gcc / make")
    probe = tmp_path / "probe.c"
    probe.write_text("int main(void) { return 0; }\n")
    if subprocess.run(["gcc", "-fsanitize=address,undefined", str(probe), "-o",
                       str(tmp_path / "probe")], capture_output=True).returncode != 0:
        pytest.skip("gcc cannot link the sanitizer runtimes here")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "oracle selftest passed" in r.stdout
