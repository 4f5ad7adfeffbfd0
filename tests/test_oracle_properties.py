"""Oracle pieces that the reference fixtures leave open, pinned by mathematics:
closed-form LIF trajectories and spike times, the circle-lens overlap area, the delay-ring
delivery time, the Tsodyks-Markram recurrences, index-map construction, first-free-slot synapse
creation, and the agreement of the wave64 summation order with the sequential one."""
import math

import numpy as np
import pytest

from oracle import oracle_py as O

F32 = np.float32


def _single_lif(vthresh, inoise=(0.0, 0.0), steps=0):
    cfg = O.default_config()
    cfg.update(width=1, height=1, max_synapses_per_neuron=1, Inoise=inoise, Vthresh=(vthresh, vthresh),
               inh_list=[], active_list=[], conn="None", epoch_duration=1.0, num_epochs=1)
    net = O.OracleNet(cfg)
    if steps:
        net.advance(steps)
    return net


def test_lif_subthreshold_closed_form():
    net = _single_lif(1.0)
    C1, C2, I0 = (float(net.array(k)[0]) for k in ("C1", "C2", "I0"))
    v0 = float(net.array("Vm")[0])
    vinf = C2 * I0 / (1.0 - C1)
    for n in (1, 10, 100, 1000, 3000):
        net.advance(n - net.step)
        want = vinf + (v0 - vinf) * C1 ** n
        assert abs(float(net.array("Vm")[0]) - want) <= 1e-5 * abs(want)
    assert abs(C1 - math.exp(-1e-4 / 0.03)) < 1e-7 and abs(C2 - 1e6 * (1 - C1)) < 1e-2


def test_lif_spike_times_and_refractory_period():
    vth = 13.4e-3
    net = _single_lif(vth, steps=4000)
    C1, C2, I0 = (float(net.array(k)[0]) for k in ("C1", "C2", "I0"))
    vinf = C2 * I0 / (1.0 - C1)

    def steps_to_threshold(v):
        return math.ceil(math.log((vth - vinf) / (v - vinf)) / math.log(C1))

    first = steps_to_threshold(13.0e-3)
    hist = net.array("spike_history")
    spikes = sorted(int(s) for s in hist if s != np.iinfo(np.uint64).max)
    assert spikes[0] == first
    # after a spike: 30 refractory steps (int(3 ms / 0.1 ms + 0.5)), reset to Vreset = 13.5 mV which
    # is above threshold, so the neuron fires again on the first non-refractory step
    assert all(b - a == 31 for a, b in zip(spikes, spikes[1:]))


def _lens_area(d, r1, r2):
    if d >= r1 + r2:
        return 0.0
    if d + min(r1, r2) <= max(r1, r2):
        return math.pi * min(r1, r2) ** 2
    a1 = r1 * r1 * math.acos((d * d + r1 * r1 - r2 * r2) / (2 * d * r1))
    a2 = r2 * r2 * math.acos((d * d + r2 * r2 - r1 * r1) / (2 * d * r2))
    k = 0.5 * math.sqrt((-d + r1 + r2) * (d + r1 - r2) * (d - r1 + r2) * (d + r1 + r2))
    return a1 + a2 - k


@pytest.mark.parametrize("math_mode", [0, 1])
def test_overlap_area_is_the_lens_area(math_mode):
    import ctypes
    L = O.lib()
    L.bgo_overlap_area.restype = ctypes.c_float
    L.bgo_overlap_area.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int]
    rng = np.random.default_rng(11)
    for _ in range(3000):
        x2, y2 = (float(F32(v)) for v in rng.integers(0, 20, 2))
        r1, r2 = (float(F32(v)) for v in rng.uniform(0.1, 12.0, 2))
        d = math.hypot(x2, y2)
        got = L.bgo_overlap_area(0.0, 0.0, r1, x2, y2, r2, math_mode)
        want = _lens_area(d, r1, r2) if d > 0 else 0.0
        if d + min(r1, r2) <= max(r1, r2) and d > 0:
            assert abs(got - want) <= 2e-6 * want
        elif want > 1e-3:
            assert abs(got - want) <= 2e-4 * want + 1e-5, (d, r1, r2, got, want)
        elif d >= r1 + r2:
            assert got == 0.0


def _pair_net(syn="AllDSSynapses", delay=None):
    cfg = O.default_config()
    cfg.update(width=2, height=1, max_synapses_per_neuron=4, Inoise=(0.0, 0.0), conn="None", inh_list=[],
               active_list=[0], Vthresh=(1.0, 1.0), starter_Vthresh=(13.4e-3, 13.4e-3), synapse=syn,
               epoch_duration=1.0, num_epochs=1, noise_mode=1)
    if delay:
        cfg["delay_override"] = [delay] * 4
    net = O.OracleNet(cfg)
    net.add_synapse(0, 1, 1e-12)
    net.rebuild_index_map()
    return net


def test_ring_delivers_after_total_delay():
    net = _pair_net()
    slot = 1 * 4 + 0
    D = int(net.array("total_delay")[slot])
    assert D == int(F32(F32(1.5e-3) / F32(1e-4))) + 1 == 16   # EE synapse, 1.5 ms
    net.advance(3000)
    T = net.step
    spikes = sorted(int(s) for s in net.array("spike_history")[:net.cfg["max_firing_rate"]]
                    if s != np.iinfo(np.uint64).max)
    delivered = [s for s in spikes if s + D <= T - 1]
    pending = [s for s in spikes if s + D > T - 1]
    assert int(net.array("lastSpike")[slot]) == delivered[-1] + D
    ring = int(net.array("delayQueue")[slot])
    idx = int(net.array("delayIdx")[slot])
    assert idx == T % 32
    want = 0
    for s in pending:
        want |= 1 << ((idx + (s + D - T)) % 32)
    assert ring == want


def test_tsodyks_markram_recurrence():
    net = _pair_net()
    slot = 4
    net.advance(5000)
    spikes = sorted(int(s) for s in net.array("spike_history")[:net.cfg["max_firing_rate"]]
                    if s != np.iinfo(np.uint64).max)
    D = 16
    deliveries = [s + D for s in spikes if s + D <= net.step - 1]
    U, Dt, Ft = 0.5, 1.1, 0.05                          # EE (Synapses/AllDSSynapses.cpp:205-209)
    r, u = 1.0, 0.4
    for a, b in zip(deliveries, deliveries[1:]):
        isi = (b - a) * 1e-4
        r = 1 + (r * (1 - u) - 1) * math.exp(-isi / Dt)
        u = U + u * (1 - U) * math.exp(-isi / Ft)
    assert abs(float(net.array("r")[slot]) - r) < 1e-5
    assert abs(float(net.array("u_syn")[slot]) - u) < 1e-5


def test_index_maps_are_sorted_views_of_the_slots():
    cfg = O.default_config()
    cfg.update(width=6, height=6, max_synapses_per_neuron=40, start_radius=1.6, epoch_duration=0.1, num_epochs=1,
               inh_list=[3, 20], active_list=[0, 11, 30])
    net = O.OracleNet(cfg)
    net.run_epoch()
    use = net.array("in_use").astype(bool)
    src = net.array("sourceNeuronIndex")
    slots = np.nonzero(use)[0].astype(np.uint32)
    assert np.array_equal(net.array("incomingSynapseIndexMap"), slots)
    order = np.argsort(src[slots], kind="stable")
    assert np.array_equal(net.array("outgoingSynapseIndexMap"), slots[order])
    assert np.array_equal(net.array("outgoingSynapseCount"), np.bincount(src[slots], minlength=36).astype(np.uint32))


def test_growth_creates_every_overlap_in_first_free_slots_by_source_order():
    import ctypes
    L = O.lib()
    L.bgo_overlap_area.restype = ctypes.c_float
    L.bgo_overlap_area.argtypes = [ctypes.c_float] * 6 + [ctypes.c_int]
    cfg = O.default_config()
    cfg.update(width=7, height=5, max_synapses_per_neuron=40, start_radius=1.3, epoch_duration=0.05, num_epochs=1,
               inh_list=[4, 17], active_list=[0, 9])
    net = O.OracleNet(cfg)
    net.advance(500)
    net.update_connections()
    radii = net.array("radii")
    x, y = net.array("xloc"), net.array("yloc")
    use = net.array("in_use").reshape(35, 40).astype(bool)
    src = net.array("sourceNeuronIndex").reshape(35, 40)
    for d in range(35):
        want = [s for s in range(35) if s != d and L.bgo_overlap_area(x[s], y[s], radii[s], x[d], y[d], radii[d], 0) > 0]
        k = len(want)
        assert use[d, :k].all() and not use[d, k:].any()
        assert list(src[d, :k]) == want


def test_wave64_summation_stays_within_tolerance_of_sequential():
    cfgs = []
    for mode in (0, 1):
        cfg = O.default_config()
        cfg.update(width=6, height=6, max_synapses_per_neuron=40, start_radius=1.6, epoch_duration=0.05,
                   num_epochs=2, inh_list=[3, 20], active_list=[0, 11, 30], noise_mode=1, math_mode=1, sum_mode=mode)
        cfgs.append(cfg)
    a, b = O.OracleNet(cfgs[0]), O.OracleNet(cfgs[1])
    for n in (a, b):
        n.run_epoch()
        n.advance(100)
    va, vb = a.array("Vm"), b.array("Vm")
    assert np.array_equal(a.array("spike_history"), b.array("spike_history"))
    assert np.max(np.abs(va - vb) / np.maximum(np.abs(va), 1e-6)) < 1e-5


def test_poisson_intervals_follow_the_rejection_density():
    """The counter-based Poisson input draws (bg_poisson_isi_steps) against the distribution the
    reference's loop defines (Inputs/HostSInputPoisson.cpp:75-84): an exponential interval of mean
    lambda ms, accepted with probability 1 - exp(-isi^2 / 32), then rounded to steps. Checked by
    mean, by a Kolmogorov-Smirnov test against the numerically integrated CDF, and by the
    independence of neighbouring counters."""
    from scipy import integrate, stats
    L = O.lib()
    lam, dt = 100.0, 1e-4           # 10 Hz (fr_mean 10 -> 0.01 per ms -> lambda 100 ms)
    n = 40000
    isi = np.array([L.bgo_poisson_isi(1, k % 977, 1000 + k // 977, lam, dt) for k in range(n)], np.float64)
    dens = lambda x: math.exp(-x / lam) / lam * (1.0 - math.exp(-x * x / 32.0))
    Z = integrate.quad(dens, 0, np.inf)[0]
    mean_ms = integrate.quad(lambda x: x * dens(x), 0, np.inf)[0] / Z
    sd_ms = math.sqrt(integrate.quad(lambda x: x * x * dens(x), 0, np.inf)[0] / Z - mean_ms ** 2)
    steps_per_ms = 1e-3 / dt
    assert abs(isi.mean() / steps_per_ms - mean_ms) < 4 * sd_ms / math.sqrt(n)
    xs = np.sort(isi / steps_per_ms)
    grid = np.linspace(0, 1000, 2001)
    cdf_grid = np.concatenate([[0.0], np.cumsum([integrate.quad(dens, a, b)[0] for a, b in zip(grid[:-1], grid[1:])])]) / Z
    # rounding to whole steps moves a sample by at most 0.05 ms; compare against the CDF shifted both ways
    emp = np.searchsorted(xs, grid, side="right") / n
    lo = np.interp(grid - 0.05, grid, cdf_grid)
    hi = np.interp(grid + 0.05, grid, cdf_grid)
    dev = np.maximum(emp - hi, lo - emp).max()
    assert dev < stats.kstwo.ppf(0.999, n)
    # intervals inside the refractoriness are almost never accepted
    assert (xs < 1.0).mean() < 1e-3
    # neighbouring neurons and steps are uncorrelated
    a = isi[:-977]
    b = isi[977:]
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.03
    assert abs(np.corrcoef(isi[:-1], isi[1:])[0, 1]) < 0.03


def _ref_update_weights(in_use, src, W, counts, Wm, ntype, M):
    """Connections/ConnGrowth.cpp:405-461 as the reference runs it: source outer, destination inner;
    an existing synapse (s, d) is erased when W(s, d) < 0 (Synapses/AllSynapses.cpp:339-344) or gets
    W(s, d) * sign * 1e-8; a missing one with W(s, d) > 0 goes into d's first free slot
    (Synapses/AllSynapses.cpp:357-377). Sign: -1 for an inhibitory source (Layouts/Layout.cpp:129-141)."""
    N = len(counts)
    adj = F32(1.0e-8)
    for s in range(N):
        for d in range(N):
            row = range(d * M, (d + 1) * M)
            sign = F32(-1.0) if ntype[s] == 1 else F32(1.0)
            w = F32(Wm[s, d])
            connected = False
            for k in row:
                if in_use[k] and src[k] == s:
                    connected = True
                    if w < 0:
                        in_use[k] = 0
                        counts[d] -= 1
                    else:
                        W[k] = w * sign * adj
            if not connected and w > 0:
                k = next(k for k in row if not in_use[k])
                in_use[k], src[k], W[k] = 1, s, w * sign * adj
                counts[d] += 1


def test_update_weights_matches_reference_loop_order_with_erasures():
    """The oracle's dense-W weight update (destination outer) equals the reference's source-outer
    loop on a network with existing synapses: some erased, their slots reused by later sources,
    some reweighted (zero weights kept), new ones created; NaN entries neither erase nor create."""
    cfg = O.default_config()
    cfg.update(width=6, height=6, max_synapses_per_neuron=30, layout="DynamicLayout", frac_exc=0.8,
               frac_starter=0.1, noise_mode=1, math_mode=1, sum_mode=1)
    net = O.OracleNet(cfg)
    net.set_radii(0.75)
    net.update_connections()
    N, M = 36, 30
    in_use = net.array("in_use").astype(np.int64)
    src = net.array("sourceNeuronIndex").astype(np.int64)
    W = net.array("W").copy()
    counts = net.array("synapse_counts").astype(np.int64)
    ntype = net.array("neuron_type_map")
    assert in_use.sum() > 100
    rng = np.random.default_rng(5)
    Wm = rng.uniform(-1.0, 1.0, (N, N)).astype(np.float32)
    Wm[rng.random((N, N)) < 0.7] = 0.0          # keep rows under M
    existing = np.zeros((N, N), bool)
    for k in np.flatnonzero(in_use):
        existing[src[k], k // M] = True
    Wm[existing] = rng.choice(np.array([-0.5, 0.0, 0.25, np.nan], np.float32), existing.sum())
    _ref_update_weights(in_use, src, W, counts, Wm, ntype, M)
    net.update_weights(Wm)
    got_use = net.array("in_use").astype(np.int64)
    assert np.array_equal(got_use, in_use)
    live = got_use == 1
    assert np.array_equal(net.array("sourceNeuronIndex")[live], src[live])
    # NaN weights stay NaN (their sign bit depends on how the compiler multiplies by -1)
    got_w = net.array("W")[live]
    nan = np.isnan(W[live])
    assert nan.sum() > 5 and np.array_equal(np.isnan(got_w), nan)
    assert np.array_equal(got_w[~nan].view(np.uint32), W[live][~nan].view(np.uint32))
    assert np.array_equal(net.array("synapse_counts").astype(np.int64), counts)
    erased = int(((Wm < 0) & existing).sum())
    assert erased > 10 and int(((Wm > 0) & ~existing).sum()) > 10
    net.close()
