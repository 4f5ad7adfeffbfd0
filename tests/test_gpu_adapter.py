"""The reference's own driver with the engine behind IModel: BrainGrid's Simulator loop, XML
parameter reader, class factory and recorders (linked from the reference's objects) drive
HipModelAdapter (braingrid_amd/adapter), which forwards IModel to libbghip.so. The XML state file
the reference's recorder writes must be byte-identical to the host model's and to the oracle's
(deterministic noise mode) for the same configuration.

oracle/_ref/growth_hip is built by `make -f oracle/Makefile.ref hip` where the reference sources are
present; the test is skipped where it is not built."""
import os
import subprocess

import numpy as np

import pytest

from braingrid_amd import host
from oracle import oracle_py as O
from tests import helpers as H
from tests import xmlout

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "growth_hip")


def _run_driver(args, cwd, timeout=300):
    r = subprocess.run([DRIVER, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stdout[-1500:], r.stderr[-1500:])
    return r


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="oracle/_ref/growth_hip is not built (needs the reference sources)")
@pytest.mark.parametrize("name", ["conn-lif-ds", "conn-izh-stdp", "conn-lif-dynstdp", "static-lif-ds", "static-izh-all"])
def test_reference_driver_on_the_engine_writes_the_same_state_file(hip_ctx, tmp_path, name):
    print(f"\n[adapter] {name}", flush=True)
    out = tmp_path / "adapter-out.xml"
    _run_driver(["-t", name + ".xml", "-o", str(out)], H.CONFIG_DIR)
    got = out.read_text()

    cfg = host.read_config(os.path.join(H.CONFIG_DIR, name + ".xml"), H.CONFIG_DIR)
    m = host.Model(hip_ctx, cfg)
    for _ in range(cfg["num_epochs"]):
        m.run_epoch()
    p = tmp_path / "host-out.xml"
    m.save_xml(str(p))
    m.close()
    assert got == p.read_text()

    ocfg = H.load_config(name)
    ora = O.OracleNet(ocfg)
    for _ in range(ocfg["num_epochs"]):
        ora.run_epoch()
    want = xmlout.oracle_growth_xml(ora, ocfg)
    ora.close()
    assert got == want


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="oracle/_ref/growth_hip is not built (needs the reference sources)")
def test_reference_driver_memory_image_resumes_the_run(hip_ctx, tmp_path):
    """BGDriver -w after 4 epochs, then -r of that image in a fresh process for 4 more
    (Core/BGDriver.cpp:88-116): the resumed run continues the first exactly. Its recorder starts at
    the resume, as the reference's does (g_simulationStep restarts at 0, Simulator::deserialize
    re-initialises the history rows from the loaded radii and rates), so its histories equal the
    second half of one uninterrupted 8-epoch run of the host model."""
    name = "conn-lif-ds"
    img = tmp_path / "mem.bin"
    a = tmp_path / "a.xml"
    b = tmp_path / "b.xml"
    _run_driver(["-t", name + ".xml", "-o", str(a), "-w", str(img)], H.CONFIG_DIR)
    raw = img.read_bytes()
    assert raw.split(b"\0", 1)[0] == b"25"        # Model::serialize writes the neuron count first
    _run_driver(["-t", name + ".xml", "-o", str(b), "-r", str(img)], H.CONFIG_DIR)

    cfg = host.read_config(os.path.join(H.CONFIG_DIR, name + ".xml"), H.CONFIG_DIR)
    E = cfg["num_epochs"]
    N = cfg["width"] * cfg["height"]
    cfg["num_epochs"] = 2 * E
    m = host.Model(hip_ctx, cfg)
    for _ in range(2 * E):
        m.run_epoch()
    radii = m.array("radiiHistory").reshape(2 * E + 1, N)
    rates = m.array("ratesHistory").reshape(2 * E + 1, N)
    burst = m.array("burstinessHist")
    spikes = m.array("spikesHistory")
    m.close()

    def matrix(text, name):
        body = text.split(f'name="{name}"')[1].split("</Matrix>")[0]
        return np.array([float(v) for v in body.split(">", 1)[1].split()], dtype=np.float32)

    got = b.read_text()
    f32 = lambda v: np.array(["%g" % float(x) for x in v], dtype=np.float64).astype(np.float32)  # noqa: E731
    assert np.array_equal(matrix(got, "radiiHistory"), f32(radii[E:].ravel()))
    assert np.array_equal(matrix(got, "ratesHistory"), f32(rates[E:].ravel()))
    nb = len(burst) // 2
    assert np.array_equal(matrix(got, "burstinessHist"), f32(burst[nb:]))
    ns = len(spikes) // 2
    assert np.array_equal(matrix(got, "spikesHistory"), f32(spikes[ns:]))
    # and the first run's file is the first half of it
    first = a.read_text()
    assert np.array_equal(matrix(first, "radiiHistory"), f32(radii[:E + 1].ravel()))
    assert np.array_equal(matrix(first, "burstinessHist"), f32(burst[:nb]))
