"""HipModelAdapter (braingrid_amd/adapter) compiles against the reference's own headers with the
reference's own flags (C++98) and links with the reference's objects and libbghip.so into
BrainGrid's driver (oracle/Makefile.ref target `hip`). Needs the reference sources; skipped
without them."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "growth_hip")
LIB = os.path.join(ROOT, "braingrid_amd", "lib", "libbghip.so")
CONFIGS = os.path.join(ROOT, "tests", "golden", "configs")

needs_ref = pytest.mark.skipif(not os.path.isdir("/root/reference") or not os.path.exists(LIB),
                               reason="needs the reference sources and a built libbghip.so")


@needs_ref
def test_adapter_links_into_the_reference_driver():
    r = subprocess.run(["make", "-s", "-j8", "-f", "oracle/Makefile.ref", "hip"], cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]
    assert os.path.exists(DRIVER)
    deps = subprocess.run(["ldd", DRIVER], stdout=subprocess.PIPE, text=True).stdout
    assert "libbghip.so => " + os.path.realpath(LIB) in deps.replace(os.path.join(ROOT, "oracle", "_ref", "..", ".."), ROOT)
    undef = subprocess.run(["nm", "-D", "--undefined-only", DRIVER], stdout=subprocess.PIPE, text=True).stdout
    for sym in ("bghip_ctx_create", "bghip_net_create", "bghip_upload_neurons", "bghip_upload_synapses",
                "bghip_upload_layout", "bghip_upload_growth", "bghip_advance", "bghip_update_connections",
                "bghip_download_spikes", "bghip_download_growth", "bghip_update_history", "bghip_net_save",
                "bghip_net_load"):
        assert sym in undef, sym
    # the reference's classes are linked in, not re-implemented: its simulator and recorder symbols
    defined = subprocess.run(["nm", "-C", "--defined-only", DRIVER], stdout=subprocess.PIPE, text=True).stdout
    for sym in ("Simulator::simulate", "XmlGrowthRecorder::compileHistories", "ConnGrowth::setupConnections",
                "HipModelAdapter::advance"):
        assert sym in defined, sym


@needs_ref
def test_driver_runs_the_reference_setup_then_needs_a_device():
    """Without a GPU the driver parses the XML through the reference's reader, runs Model::setupSim
    (the reference's draws) and stops at the engine's context creation with its error."""
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present: the GPU tests run the driver")
    if not os.path.exists(DRIVER):
        pytest.skip("growth_hip not built")
    r = subprocess.run([DRIVER, "-t", "conn-lif-ds.xml", "-o", "/tmp/_unused.xml"], cwd=CONFIGS,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 1
    assert "Done initializing neurons" in r.stderr or "Done initializing neurons" in r.stdout
    assert "bghip_ctx_create failed" in r.stderr
