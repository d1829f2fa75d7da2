"""Host-side pieces without a GPU: the TPIE subset behind nni.h, and the
nni.h API library's exports."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tpie_subset(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "tpie_test")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "tpie subset ok" in r.stdout


def test_nni_api_library_exports_the_reference_functions():
    lib = ctypes.CDLL(os.path.join(ROOT, "terranni_amd", "_lib", "libterranni_nni.so"))
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", lib._name], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("nni::init(nni::Settings const&)", "nni::cleanup()", "nni::printTimes()",
                "nni::setRenderbuffers(int, int, int)", "nni::clearRenderbuffers()",
                "nni::setupRun(nni::WorkPackage const&, int, int)",
                "nni::performNNI(nni::WorkPackage const&, tpie::ami::stream<nni::OutPoint>&, "
                "tpie::ami::stream<nni::Point4D>&, float, float, bool, bool)"):
        assert sym in out, sym


def test_nni_init_without_gpu_fails_cleanly(tmp_path):
    if os.path.exists("/dev/kfd"):
        import pytest
        pytest.skip("a GPU is visible")
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "nni_api_test")
    (tmp_path / "p.bin").write_bytes(b"")
    r = subprocess.run([exe, str(tmp_path / "p.bin"), str(tmp_path / "o.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 4 and "engine initialisation failed" in r.stderr
