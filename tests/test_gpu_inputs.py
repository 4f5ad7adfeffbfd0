"""GPU parity of the stimulus inputs (SInputPoisson with its DS input synapse and Philox interval
draws, SInputRegular with asynchronous and wave shifts) against the oracle in deterministic mode:
every neuron, synapse and input-state array bit for bit after every epoch and growth update."""
import numpy as np
import pytest

from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(H.INPUT_STIM))
def test_input_parity(hip_ctx, name):
    print(f"\n[input parity {name}]", flush=True)
    cfg = H.load_config(name)
    ora = O.OracleNet(cfg)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    steps = H.steps_per_epoch(cfg)
    try:
        bad = H.compare_input(ora, net, cfg, "upload")
        assert not bad, bad
        for ep in range(cfg["num_epochs"]):
            ora.advance(steps)
            net.advance(steps)
            bad = H.compare_state(ora, net, cfg, f"advanced{ep}") + H.compare_input(ora, net, cfg, f"advanced{ep}")
            assert not bad, bad
            ora.update_connections()
            net.update_connections()
            ora.update_history()
            net.update_history()
            bad = H.compare_state(ora, net, cfg, f"grown{ep}")
            assert not bad, bad
        assert net.error_bits() == 0
        assert ora.array("spike_history").size and (ora.array("radii") > 0).all()
    finally:
        net.close()
        ora.close()


def test_poisson_input_resumes_mid_run(hip_ctx):
    """Upload of a Poisson input state taken from the oracle after 777 steps (rings at a phase
    that is not a multiple of 32) continues bit for bit."""
    name = "input-izh-poisson"
    cfg = H.load_config(name)
    ora = O.OracleNet(cfg)
    ora.advance(777)
    net = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    try:
        assert not H.compare_input(ora, net, cfg, "upload")
        for c in (5, 300, 1000):
            ora.advance(c)
            net.advance(c)
            bad = H.compare_state(ora, net, cfg, f"s{ora.step}") + H.compare_input(ora, net, cfg, f"s{ora.step}")
            assert not bad, bad
        assert int(np.sum(ora.array("in_lastSpike") != np.iinfo(np.uint64).max)) > 0
    finally:
        net.close()
        ora.close()
