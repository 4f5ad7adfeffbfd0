"""Regenerates tests/golden/ref_rates_10k.npz from the compiled reference (this container only).

The workload is the first two 10 s epochs of the benchmark-scale plain method: a 100 x 100 grid of
AllLIFNeurons with the test-small parameter distributions, DynamicLayout with 90% excitatory and
10% starter neurons, AllDSSynapses, ConnGrowth (startRadius 0.4: no synapse forms in these
epochs), dt = 1e-4 s. It is run by oracle/_ref/bgref_dump -- the reference's own objects, driven
like the reference `growth` binary -- once per noise seed. The reference draws parameters and
layout from its fixed rng(1) (Utils/Global.cpp:79) and seeds only its Norm noise stream with the
configured seed (Core/SingleThreadedSpikingModel.cpp:30), so the seeds give eight noise
realisations of one network.

Stored: `counts` (uint16, seeds x epochs x neurons), every neuron's spike count per epoch, taken
from the ConnGrowth rates after each growth update (rate = count / epoch duration,
Connections/ConnGrowth.cpp:295-310); `burst` (uint32, seeds x 1 s bins), the recorder's
burstinessHist (Recorders/XmlRecorder.cpp:71-104), and `radii` (float32, seeds x neurons), the
final row of radiiHistory (Recorders/XmlGrowthRecorder.cpp:72-97), both from the recorder output
the harness writes; and `seeds`.

    python tests/golden/make_rate_fixture.py
"""
import os
import re
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle_py as O  # noqa: E402

SIDE, EPOCH_S, EPOCHS = 100, 10.0, 2
SEEDS = tuple(range(11, 19))
M, MAX_RATE = 200, 200
OUT = os.path.join(HERE, "ref_rates_10k.npz")


def config_xml(seed: int) -> str:
    """The conn-lif-ds-dynamic parameter file at 100 x 100, startRadius 0.4 and the given seed."""
    with open(os.path.join(HERE, "configs", "conn-lif-ds-dynamic.xml")) as f:
        x = f.read()
    reps = [
        ('<x name="x">8</x><y name="y">8</y>', f'<x name="x">{SIDE}</x><y name="y">{SIDE}</y>'),
        ('<Tsim name="Tsim">0.5</Tsim><numSims name="numSims">3</numSims>',
         f'<Tsim name="Tsim">{EPOCH_S}</Tsim><numSims name="numSims">{EPOCHS}</numSims>'),
        ('<maxFiringRate name="maxFiringRate">200</maxFiringRate><maxSynapsesPerNeuron name="maxSynapsesPerNeuron">40'
         '</maxSynapsesPerNeuron>',
         f'<maxFiringRate name="maxFiringRate">{MAX_RATE}</maxFiringRate><maxSynapsesPerNeuron '
         f'name="maxSynapsesPerNeuron">{M}</maxSynapsesPerNeuron>'),
        ('<Seed name="Seed"><value name="value">7</value></Seed>', f'<Seed name="Seed"><value name="value">{seed}'
                                                                   '</value></Seed>'),
        ('<startRadius name="startRadius">1.2</startRadius>', '<startRadius name="startRadius">0.4</startRadius>'),
    ]
    for a, b in reps:
        assert a in x, a
        x = x.replace(a, b)
    return x


def xml_matrix(text: str, name: str) -> np.ndarray:
    """The values of <Matrix name="name" ...> in the recorder's XML output."""
    m = re.search(r'<Matrix name="%s"[^>]*rows="(\d+)" columns="(\d+)"[^>]*>(.*?)</Matrix>' % name, text, re.S)
    rows, cols = int(m.group(1)), int(m.group(2))
    return np.array(m.group(3).split(), np.float64).reshape(rows, cols)


def run_seed(seed: int):
    exe = os.path.join(O.REF_DIR, "bgref_dump")
    with tempfile.TemporaryDirectory() as td:
        cfg = os.path.join(td, "rates.xml")
        with open(cfg, "w") as f:
            f.write(config_xml(seed))
        dump = os.path.join(td, "dump.bin")
        subprocess.run([exe, cfg, dump, "--dynamic-layout", "0.9", "0.1", "--no-history"], cwd=td, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        recs = O.read_dump(dump)
        with open(dump + ".xml") as f:
            xml = f.read()
    rows = [d["rates"] for tag, ep, d in recs if tag == "grown"]
    assert len(rows) == EPOCHS
    counts = np.rint(np.array(rows, np.float64) * EPOCH_S)
    assert counts.max() < 65536
    burst = xml_matrix(xml, "burstinessHist")[0]
    radii = xml_matrix(xml, "radiiHistory")[-1]
    assert burst.shape == (int(EPOCHS * EPOCH_S),) and radii.shape == (SIDE * SIDE,)
    return counts.astype(np.uint16), np.rint(burst).astype(np.uint32), radii.astype(np.float32)


def main():
    if not os.path.exists(os.path.join(O.REF_DIR, "bgref_dump")):
        sys.exit("oracle/_ref/bgref_dump missing: make -f oracle/Makefile.ref")
    with ThreadPoolExecutor(len(SEEDS)) as ex:
        res = list(ex.map(run_seed, SEEDS))
    counts = np.array([r[0] for r in res])
    burst = np.array([r[1] for r in res])
    radii = np.array([r[2] for r in res])
    np.savez_compressed(OUT, counts=counts, burst=burst, radii=radii, seeds=np.array(SEEDS, np.int32))
    print(OUT, counts.shape, "spikes per seed", counts.sum(axis=(1, 2)), "burst totals", burst.sum(axis=1),
          os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
