"""bghip_update_weights_static, the entry point that allocates no device memory (its copy of the
caller's matrix goes into the library's static per-GPU scratch array): through the C ABI, its results
equal those of bghip_update_weights (which allocates that copy) bit for bit, on two calls in a row,
the second on a smaller network (a smaller matrix in the same scratch), and on the oracle's.
Matrices larger than the scratch are refused."""
import numpy as np
import pytest

from braingrid_amd import bghip as B
from oracle import oracle_py as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _net_pair(hip_ctx, side, M, radius, seed):
    cfg = O.default_config()
    cfg.update(width=side, height=side, max_synapses_per_neuron=M, layout="DynamicLayout", frac_exc=0.8,
               frac_starter=0.1, noise_mode=1, math_mode=1, sum_mode=1, epoch_duration=0.05, num_epochs=2,
               seed=seed)
    ora = O.OracleNet(cfg)
    ora.set_radii(radius)
    a = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    b = H.gpu_net_from_oracle(hip_ctx, ora, cfg)
    ora.update_connections()
    a.update_connections()
    b.update_connections()
    return cfg, ora, a, b


def _matrix(ora, N, M, rng):
    use = ora.array("in_use").astype(bool)
    slots = np.flatnonzero(use)
    src = ora.array("sourceNeuronIndex")[slots].astype(np.int64)
    dst = slots // M
    # about 5 new synapses per destination, well inside the rows' free slots
    Wm = rng.uniform(-1.0, 1.0, (N, N)).astype(np.float32)
    Wm[rng.random((N, N)) >= 10.0 / N] = 0.0
    Wm[src, dst] = rng.choice(np.array([-0.5, 0.0, 0.25, 3.0], np.float32), slots.size)
    return Wm


def _same(a, b):
    sa, sb = a.download_synapses(), b.download_synapses()
    bad = [f for f in sa if not np.array_equal(np.asarray(sa[f]).view(np.uint8), np.asarray(sb[f]).view(np.uint8))]
    ma, mb = a.download_index_map(), b.download_index_map()
    bad += [f for f in ma if not np.array_equal(ma[f], mb[f])]
    return bad


def test_static_weight_update_equals_allocating_one(hip_ctx):
    rng = np.random.default_rng(3)
    # first call: 40 x 40 = 1,600 neurons (a 10 MB matrix); second: 20 x 20 = 400 neurons
    for side, M, radius, seed in ((40, 60, 1.6, 5), (20, 40, 1.3, 6)):
        print(f"[no_alloc] {side}x{side}, M={M}", flush=True)
        N = side * side
        cfg, ora, a, b = _net_pair(hip_ctx, side, M, radius, seed)
        try:
            for rnd in range(2):
                Wm = _matrix(ora, N, M, rng)
                ora.update_weights(Wm)
                sa = a.update_weights(Wm)
                sb = b.update_weights_static(Wm)
                print(f"[no_alloc] round {rnd}: {sb.created} created, {sb.erased} erased", flush=True)
                assert (sa.created, sa.erased, sa.total_synapses) == (sb.created, sb.erased, sb.total_synapses)
                assert sb.created > 0 and sb.erased > 0
                bad = _same(a, b)
                bad += H.compare_state(ora, b, cfg, f"static-{side}-{rnd}")
                assert not bad, bad[:10]
                ora.advance(200)
                a.advance(200)
                b.advance(200)
            bad = _same(a, b)
            assert not bad, bad[:10]
        finally:
            a.close()
            b.close()
            ora.close()


def test_static_weight_update_refuses_matrices_beyond_its_scratch(hip_ctx):
    from braingrid_amd import model as MD
    desc = MD.net_desc(N=2049, M=4, neuron_model=B.NEURON_LIF, synapse_model=B.SYN_DS, max_spikes=10)
    net = B.Net(hip_ctx, desc)
    try:
        N = 2049
        net.upload_layout(np.zeros(N, np.float32), np.arange(N, dtype=np.float32), np.full(N, 2, np.int32))
        with pytest.raises(B.BgHipError, match="static scratch"):
            net.update_weights_static(np.zeros((N, N), np.float32))
    finally:
        net.close()
