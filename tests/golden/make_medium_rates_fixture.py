"""Regenerates tests/golden/ref_rates_medium.npz from the compiled reference (this container only).

The workload is the reference's own configfiles/test-medium.xml as committed under
tests/golden/ref/configfiles: a 30 x 30 grid of AllLIFNeurons (FixedLayout, 9 starter and 18
inhibitory neurons from its NList files), AllDSSynapses, ConnGrowth from startRadius 0.4, twenty
100 s epochs, dt = 1e-4 s. Synapses form once the growing neurites overlap (epoch 12 or so), and
from then on the firing is shaped by them. It is run by oracle/_ref/growth -- the reference's CPU
simulator built from its sources -- once per noise seed: the reference draws parameters from its
fixed rng(1) (Utils/Global.cpp:79) and seeds only its Norm noise stream with the configured seed
(Core/SingleThreadedSpikingModel.cpp:30), so the seeds give noise realisations of one network.

Stored, from each run's recorder output (Recorders/XmlGrowthRecorder.cpp:104-154):
`counts` (uint16, seeds x epochs x neurons), every neuron's spike count per epoch, from
ratesHistory rows 1.. (rate = count / epoch duration, Connections/ConnGrowth.cpp:295-310);
`burst` (uint32, seeds x 1 s bins), burstinessHist; `radii` (float32, seeds x epochs+1 x neurons),
radiiHistory; and `seeds`.

    python tests/golden/make_medium_rates_fixture.py [--workers K]
"""
import argparse
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle_py as O  # noqa: E402

CONFIG = os.path.join(HERE, "ref", "configfiles", "test-medium.xml")
NLIST = os.path.join(HERE, "ref", "configfiles", "NList")
SEEDS = tuple(range(11, 19))
SIDE, EPOCH_S, EPOCHS = 30, 100.0, 20
OUT = os.path.join(HERE, "ref_rates_medium.npz")


def config_xml(seed: int) -> str:
    with open(CONFIG) as f:
        x = f.read()
    a = '<Seed name="Seed">\n      <value name="value">1</value>\n    </Seed>'
    assert a in x
    return x.replace(a, f'<Seed name="Seed">\n      <value name="value">{seed}</value>\n    </Seed>')


def xml_matrix(text: str, name: str) -> np.ndarray:
    """The values of <Matrix name="name" ...> in the recorder's XML output."""
    m = re.search(r'<Matrix name="%s"[^>]*rows="(\d+)" columns="(\d+)"[^>]*>(.*?)</Matrix>' % name, text, re.S)
    rows, cols = int(m.group(1)), int(m.group(2))
    return np.array(m.group(3).split(), np.float64).reshape(rows, cols)


def run_seed(seed: int):
    exe = os.path.join(O.REF_DIR, "growth")
    t0 = time.time()
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "configfiles"))
        shutil.copytree(NLIST, os.path.join(td, "configfiles", "NList"))
        cfg = os.path.join(td, "medium.xml")
        with open(cfg, "w") as f:
            f.write(config_xml(seed))
        out = os.path.join(td, "out.xml")
        subprocess.run([exe, "-t", cfg, "-o", out], cwd=td, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        with open(out) as f:
            xml = f.read()
    N = SIDE * SIDE
    rates = xml_matrix(xml, "ratesHistory")
    radii = xml_matrix(xml, "radiiHistory")
    burst = xml_matrix(xml, "burstinessHist")[0]
    assert rates.shape == (EPOCHS + 1, N) and radii.shape == (EPOCHS + 1, N)
    assert burst.shape == (int(EPOCHS * EPOCH_S),)
    counts = np.rint(rates[1:] * EPOCH_S)
    assert np.allclose(counts / EPOCH_S, rates[1:], rtol=1e-5) and counts.max() < 65536
    print(f"seed {seed}: {int(counts.sum())} spikes, {time.time() - t0:.0f} s", flush=True)
    return counts.astype(np.uint16), np.rint(burst).astype(np.uint32), radii.astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=len(SEEDS))
    args = ap.parse_args()
    if not os.path.exists(os.path.join(O.REF_DIR, "growth")):
        sys.exit("oracle/_ref/growth missing: make -f oracle/Makefile.ref")
    with ThreadPoolExecutor(args.workers) as ex:
        res = list(ex.map(run_seed, SEEDS))
    counts = np.array([r[0] for r in res])
    burst = np.array([r[1] for r in res])
    radii = np.array([r[2] for r in res])
    np.savez_compressed(OUT, counts=counts, burst=burst, radii=radii, seeds=np.array(SEEDS, np.int32))
    print(OUT, counts.shape, "spikes per seed", counts.sum(axis=(1, 2)), os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
