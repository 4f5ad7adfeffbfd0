"""Regenerates tests/golden/ref_<config>.npz from the compiled reference (this container only).

For every validation config it runs oracle/_ref/bgref_dump -- the reference's own objects built
from /root/reference by oracle/Makefile.ref, driven exactly like the reference `growth` binary --
and stores the model state after setup, after each epoch's steps and after each growth update:
neuron membrane/refractory/spike bookkeeping, every synapse slot, the index maps, rates and radii.
tests/test_oracle_vs_reference.py compares the oracle against these vectors bit for bit, so the
oracle stays pinned on machines where the reference is absent.

    python tests/golden/make_ref_fixtures.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle_py as O  # noqa: E402
from make_configs import INPUT_CONFIGS, STATIC_CONFIGS  # noqa: E402

CONFIGS = {
    # name: (config path relative to base dir, base dir, extra harness args)
    "test-tiny": ("configfiles/test-tiny.xml", os.path.join(HERE, "ref"), []),
    "test-small": ("configfiles/test-small.xml", os.path.join(HERE, "ref"), []),
}
for _n in ("conn-lif-ds", "conn-lif-ds-dense", "conn-lif-spiking", "conn-lif-stdp", "conn-lif-dynstdp",
           "conn-izh-ds", "conn-izh-stdp", "conn-lif-ds-dynamic"):
    CONFIGS[_n] = (_n + ".xml", os.path.join(HERE, "configs"),
                   ["--dynamic-layout", "0.9", "0.1"] if "dynamic" in _n else [])
for _n in STATIC_CONFIGS:
    CONFIGS[_n] = (_n + ".xml", os.path.join(HERE, "configs"), [])
for _n, (_kw, _stim) in INPUT_CONFIGS.items():
    CONFIGS[_n] = (_n + ".xml", os.path.join(HERE, "configs"), ["--stim", _stim])

KEEP = ("Vm", "Vthresh", "Vreset", "nStepsInRefr", "spikeCount", "spikeCountOffset", "spike_history", "u",
        "in_use", "sourceNeuronIndex", "destNeuronIndex", "W", "psr", "type", "synapse_counts", "decay", "total_delay",
        "delayQueue", "delayIdx", "r", "u_syn", "lastSpike", "delayQueuePost", "delayIdxPost", "radii", "rates",
        "incomingSynapseIndexMap", "incomingSynapseBegin", "incomingSynapseCount", "outgoingSynapseIndexMap",
        "outgoingSynapseBegin", "outgoingSynapseCount", "xloc", "yloc", "neuron_type_map", "starter_map", "Iinject",
        "Inoise", "Vinit", "Trefract", "I0", "C1", "C2", "Aconst", "Bconst", "Cconst", "Dconst", "C3",
        "in_nISIs", "in_psr", "in_r", "in_u", "in_W", "in_lastSpike", "in_delayQueue", "in_delayIdx", "in_values",
        "in_shift", "in_steps_in_cycle", "summation_map")


def main():
    exe = os.path.join(O.REF_DIR, "bgref_dump")
    if not os.path.exists(exe):
        sys.exit("oracle/_ref/bgref_dump missing: make -f oracle/Makefile.ref")
    only = sys.argv[1:]
    for name, (cfg, base, extra) in CONFIGS.items():
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as td:
            dump = os.path.join(td, "dump.bin")
            subprocess.run([exe, cfg, dump] + extra, cwd=base, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            recs = O.read_dump(dump)
        out = {}
        for i, (tag, ep, d) in enumerate(recs):
            out[f"{i:03d}__marker"] = np.array([ep], np.int32)
            out[f"{i:03d}__tag"] = np.frombuffer(tag.encode().ljust(16, b" "), np.uint8)
            for k, v in d.items():
                if k not in KEEP:
                    continue
                if k == "spike_history" and v.size > 2_000_000:
                    continue
                out[f"{i:03d}_{k}"] = v
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(path, **out)
        print(name, len(recs), "records", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
