"""The engine's device self-tests, the counterpart of the reference's unitTest* argv modes
(PrivacyAmplification.cu:299-737, 1645-1652): the correction term, DC zeroing, the spectral product
and real-FFT splits, the bit unpack with its popcount (binIn[i] = i, A000788) and the rounding /
XOR / byte-order epilogue (v from the bits of i/32, corr 1.9, key rest 0xAAAAAAAA), each run by this
engine's own device code on the reference test's closed-form inputs.  Through the C ABI and through
pa_server's argv modes (exit 0, or 100 on a failure, as the reference).  Runs on an MI355X."""
import pathlib
import subprocess

import pytest

import privacyamplification_amd as P

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("name", P.SELFTESTS)
def test_selftest_through_the_c_abi(engine, name):
    print("[case] selftest", name, flush=True)
    ok, report = engine.selftest(name)
    assert ok, report
    assert "passed" in report


def test_selftest_unknown_name_is_refused(engine):
    with pytest.raises(P.EngineError) as ei:
        engine.selftest("NoSuchTest")
    assert ei.value.status == -1


@pytest.mark.parametrize("name", ["CalculateCorrectionFloat", "BinInt2float", "ToBinaryArray"])
def test_pa_server_unit_test_modes(tmp_path, name):
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    (tmp_path / "config.yaml").write_text("show_zeromq_status: false\n")
    print("[case] pa_server unitTest" + name, flush=True)
    r = subprocess.run([str(exe), "unitTest" + name], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"Started {name} Unit Test" in r.stdout and f"Completed {name} Unit Test" in r.stdout
