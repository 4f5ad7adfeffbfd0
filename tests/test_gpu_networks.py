"""GPU parity of the static Izhikevich + DynamicSTDP network (per-synapse delays of 11..201 steps
on a 256-bit delay ring, STDP learning, Tsodyks-Markram PSR) against the oracle, bit-exact in the
deterministic-noise mode. The device network is built by braingrid_amd.networks (the product
builder) and the oracle network by the oracle's own addSynapse/createSynapse chain."""
import pytest

from braingrid_amd import networks as NW
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,k,seed,chunks", [(200, 20, 7, (250, 750, 1000, 1000)), (1000, 100, 5, (600, 600)),
                                           (14400, 100, 9, (600, 600))])   # the largest sweep network: 1.44M synapses
def test_izh_dynstdp_static_parity(hip_ctx, N, k, seed, chunks):
    print(f"\n[izh_dynstdp_static N={N} k={k} seed={seed}]", flush=True)
    neu, syn, meta = NW.izh_stdp_arrays(N, seed, k=k)
    ora, cfg = H.izh_static_oracle(neu, syn, meta, seed)
    net, meta2 = NW.izh_stdp(hip_ctx, N, seed, k=k)
    try:
        assert meta2["M"] == meta["M"]
        bad = H.compare_state(ora, net, cfg, "upload")
        assert not bad, bad
        for c in chunks:
            ora.advance(c)
            net.advance(c)
            bad = H.compare_state(ora, net, cfg, f"step{ora.step}")
            assert not bad, bad
        assert net.error_bits() == 0
        assert ora.array("spikeCount").sum() > 0
    finally:
        net.close()
        ora.close()


def test_izh_dynstdp_multi_epoch_full_rings(hip_ctx):
    """Several epochs with a short spike-history ring, so that the rings fill and wrap and the
    pre-spike learning walks cross stale entries of earlier epochs: the device's early exit from
    those walks must leave every weight exactly as the oracle's full walk does."""
    N, k, seed = 300, 30, 21
    print(f"\n[izh_dynstdp_multi_epoch N={N} k={k} seed={seed}]", flush=True)
    epoch_s, max_rate = 0.5, 80
    neu, syn, meta = NW.izh_stdp_arrays(N, seed, k=k, max_rate=max_rate, epoch_s=epoch_s)
    ora, cfg = H.izh_static_oracle(neu, syn, meta, seed, epoch_s=epoch_s, max_rate=max_rate)
    net, _ = NW.izh_stdp(hip_ctx, N, seed, k=k, max_rate=max_rate, epoch_s=epoch_s)
    steps = int(round(epoch_s / 1e-4))
    try:
        wrapped = False
        for ep in range(5):
            ora.advance(steps)
            net.advance(steps)
            wrapped |= bool((ora.array("spikeCount") + ora.array("spikeCountOffset") >= meta["max_spikes"]).any())
            bad = H.compare_state(ora, net, cfg, f"epoch{ep}")
            assert not bad, bad
            ora.update_history()
            net.update_history()
        assert wrapped, "no history ring wrapped; the test does not reach the stale-entry walks"
        assert net.error_bits() == 0
    finally:
        net.close()
        ora.close()


def test_device_built_network_equals_uploaded_slots(hip_ctx):
    """The product builder creates the synapses on the device from their list
    (bghip_add_synapses_ex); the same network uploaded slot by slot (bghip_upload_synapses of
    izh_stdp_arrays) has identical slot arrays, and out-of-range per-synapse delays are refused."""
    import numpy as np
    from braingrid_amd import bghip as B
    N, k, seed = 400, 40, 13
    neu, syn, meta = NW.izh_stdp_arrays(N, seed, k=k)
    a, _ = NW.izh_stdp(hip_ctx, N, seed, k=k)
    b = B.Net(hip_ctx, NW.izh_stdp_desc(meta, seed))
    try:
        b.upload_neurons(neu)
        b.upload_synapses(syn, 0)
        sa, sb = a.download_synapses(), b.download_synapses()
        # slot fields are defined for synapses in use (the reference's per-slot arrays hold
        # uninitialised memory until createSynapse/resetSynapse writes a slot)
        use = np.asarray(sb["in_use"]).astype(bool)
        assert np.array_equal(np.asarray(sa["in_use"]), np.asarray(sb["in_use"])) and use.sum() == meta["S_active"]
        bad = []
        for f in sb:
            x, y = np.asarray(sa[f]), np.asarray(sb[f])
            if x.shape == use.shape:
                x, y = x[use], y[use]
            if not np.array_equal(x.view(np.uint8), y.view(np.uint8)):
                bad.append(f)
        assert not bad, bad
        assert a.total_synapses == b.total_synapses == meta["S_active"]
        one = np.zeros(1, np.int32)
        for d in (0, 256):
            with pytest.raises(B.BgHipError):
                a.add_synapses(one, one + 1, np.ones(1, np.float32), total_delay=np.full(1, d, np.int32))
    finally:
        a.close()
        b.close()


def test_stdp_parameters_differing_within_a_type(hip_ctx):
    """Synapses of one type with different STDP parameters (taupos, Apos and Wex of every 7th
    synapse scaled): no per-type table describes them, so the device must read each synapse's own
    values, and every state array equals the oracle's over 2,000 steps. The same network with the
    builder's values (one set per type) is covered by test_izh_dynstdp_static_parity."""
    import numpy as np
    from braingrid_amd import bghip as B
    N, k, seed = 300, 30, 23
    print(f"\n[stdp_params_per_synapse N={N} k={k} seed={seed}]", flush=True)
    neu, syn, meta = NW.izh_stdp_arrays(N, seed, k=k)
    slots = meta["slot"][::7]
    for f, x in (("taupos", 1.25), ("Apos", 0.8), ("Wex", 1.1)):
        syn[f] = syn[f].copy()
        syn[f][slots] = (syn[f][slots] * np.float32(x)).astype(np.float32)
    ora, cfg = H.izh_static_oracle(neu, syn, meta, seed)
    for f in ("taupos", "Apos"):
        ora.view(f)[slots] = syn[f][slots]
    net = B.Net(hip_ctx, NW.izh_stdp_desc(meta, seed))
    try:
        net.upload_neurons(neu)
        net.upload_synapses(syn, 0)
        bad = H.compare_state(ora, net, cfg, "upload")
        assert not bad, bad
        for c in (700, 1300):
            ora.advance(c)
            net.advance(c)
            bad = H.compare_state(ora, net, cfg, f"step{ora.step}")
            assert not bad, bad
        assert net.error_bits() == 0
        assert ora.array("spikeCount").sum() > 0
    finally:
        net.close()
        ora.close()
