"""Host-side pieces without a GPU: the TPIE subset behind nni.h, the nni.h
API library's exports, and the synthetic tile generator."""
import ctypes
import os
import subprocess

import pytest

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
                "tpie::ami::stream<nni::Point4D>&, float, float, bool, bool)",
                # this engine's additions
                "nni::setAccumulationMode(int)", "nni::setOrigin64(double, double)",
                "nni::lastStatus()", "nni::lastError()"):
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


def test_shape_class_separates_density_grid_and_form():
    """The tuned table's key (host only): tiles alike in geometry and density
    share a class; the density class, the grid side class and the form
    (plain / batched) each split it."""
    import terranni_amd as T
    p = T.TileParams(nrows=1000, ncols=1000, cell_size=0.5, scale=1)
    base = T.shape_class(p, 1_000_000, True)
    assert T.shape_class(p, 1_050_000, True) == base
    assert T.shape_class(p, 4_000_000, True) != base
    assert T.shape_class(p, 1_000_000, False) != base
    q = T.TileParams(nrows=2000, ncols=2000, cell_size=0.5, scale=1)
    assert T.shape_class(q, 4_000_000, True) != T.shape_class(p, 1_000_000, True)
    c = T.TileParams(nrows=1000, ncols=1000, cell_size=0.5, scale=1, algorithm=T.ALG_CONES)
    assert T.shape_class(c, 1_000_000, True) != base


@pytest.mark.timeout(60)
def test_clustered_tile_smaller_than_a_water_disc():
    """A clustered tile of a few metres can lie wholly in one water disc of
    the generator's 2 m mask; it still gets its points (land everywhere)."""
    import terranni_amd.synth as S
    t = S.make_tile(51, 200, 7, cell_size=0.25, scale=4, kind="clustered")
    assert t.n == 200


def test_tile_kernel_stage_and_label_tile_indices_in_bounds():
    """Every index the tile kernel forms into its per-wave LDS stage (window
    offsets, row prefixes, staged slots), its label tile and search list,
    the bucket offsets and the distance maps stays inside its buffer, for
    every 8x8 block of every work item of the benchmark tiles, bucket and
    cell-tile sweeps with and without the border, and 300 random
    geometries: a host emulation over the engine's own tile layout."""
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "stage_bounds_test")
    r = subprocess.run([exe, "300"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "violations 0" in r.stdout
