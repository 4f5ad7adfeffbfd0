// HIP kernels of the engine (gfx950). All memory-bound integer/FP32 work: no MFMA anywhere.
#pragma once
#include "bghip_device.h"
#include "bghip_launch.h"

#define BG_WAVE 64
#define BG_BLOCK 256
#define BG_KMAX 32

// Loads and stores at a 32-bit byte offset from a wave-uniform base: the address is the scalar
// base plus a zero-extended 32-bit lane offset, which global memory instructions take as is (the
// kernels that use them run only for networks whose slot arrays stay below 4 GiB).
template <typename T>
__device__ __forceinline__ T ld_off(const T *base, uint32_t byteoff)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byteoff);
}
template <typename T>
__device__ __forceinline__ void st_off(T *base, uint32_t byteoff, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byteoff) = v;
}


// ---------------------------------------------------------------------------------------------
// Per-step path (used for STDP synapses, which need every neuron's step-t spike before any
// synapse of step t): one thread per neuron. AllSpikingNeurons::advanceNeurons
// (Neurons/AllSpikingNeurons.cpp:98-147); the outgoing/incoming ring notifications become pulls in
// k_synapse_step.
template <int NK>
__device__ __forceinline__ bool neuron_step_body(const DevNet &n, int i, uint64_t t)
{
    NeuronRegs s;
    neuron_load<NK>(n, i, s);
    const float noise = bg_noise_normal(n.seed, (uint32_t)i, t);
    float sp = n.sp[i], add;
    if (n.in_kind != BG_INPUT_NONE && input_step(n, i, t, &add)) sp = sp + add;
    bool fired = neuron_advance<NK>(n, i, t, sp, noise, s);
    neuron_store<NK>(n, i, s);
    n.sp[i] = 0.0f;
    if (fired) n.fired_step[i] = t;
    return fired;
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_step(DevNet n, uint64_t t)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n.N) return;
    neuron_step_body<NK>(n, i, t);
}

// One wavefront per destination neuron: delay-ring insert (pull of sources that fired at t),
// ring test/clear, STDP learning, PSR update and decay, and the destination's summation in the
// canonical lane-strided + butterfly order. AllSynapses::advanceSynapses
// (Synapses/AllSynapses.cpp:325-331), AllSpikingSynapses::advanceSynapse (:297-319),
// AllSTDPSynapses::advanceSynapse (Synapses/AllSTDPSynapses.cpp:283-421).
template <int SK, int W32>
__device__ __forceinline__ void synapse_step_body(const DevNet &n, int d, int lane, uint64_t t)
{
    const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
    const bool dense = (cnt == hi);
    const bool post = (SK == SK_STDP || SK == SK_DYNSTDP) && n.fired_step[d] == t;
    const int R = n.R;
    const uint32_t tb = (uint32_t)(t & (uint64_t)(R - 1));
    float acc = 0.0f;
    for (uint32_t q = lane; q < cnt; q += BG_WAVE) {
        const uint64_t slot = dense ? (uint64_t)d * n.M + q : (uint64_t)n.in_map[n.in_begin[d] + q];
        const uint32_t meta = n.meta[slot];
        const int src = (int)(meta & BG_META_SRC_MASK);
        const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
        const int delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
        constexpr bool planes = SK == SK_STDP || SK == SK_DYNSTDP;   // STDP rings are byte planes
        if (n.fired_step[src] == t) {   // preSpikeHit at the firing step
            const uint32_t b = (tb + (uint32_t)delay) & (uint32_t)(R - 1);
            if (planes) {
                uint32_t m;
                uint32_t *w = sring_word(n, b, slot, m);
                if (atomicOr(w, m) & m) dev_error(n, 1u);
            } else {
                uint32_t *w = &n.ring[(uint64_t)(b >> 5) * n.S + slot];
                const uint32_t m = 1u << (b & 31);
                if (*w & m) dev_error(n, 1u);
                *w |= m;
            }
        }
        bool pre;
        if (planes) {
            uint32_t m;
            uint32_t *w = sring_word(n, tb, slot, m);
            pre = (*w & m) != 0;
            if (pre) atomicAnd(w, ~m);
        } else {
            uint32_t *w = &n.ring[(uint64_t)(tb >> 5) * n.S + slot];
            const uint32_t m = 1u << (tb & 31);
            const uint32_t word = *w;
            pre = (word & m) != 0;
            if (pre) *w = word & ~m;
        }
        float psr = n.psr[slot];
        const float decay = n.decay[type];
        if (SK == SK_STDP || SK == SK_DYNSTDP) {
            if (pre || post) psr = stdp_event<SK>(n, slot, type, t, pre, post, src, d, delay, psr, decay);
        } else if (pre) {
            psr = change_psr<SK>(n, slot, type, t, psr, decay);
        }
        psr = psr * decay;
        n.psr[slot] = psr;
        acc += psr;
    }
    acc = wave_sum(acc);
    if (lane == 0) n.sp[d] = n.sp[d] + acc;
}


// ---------------------------------------------------------------------------------------------
// Networks without synapses (growth networks before their first connection, for instance): one
// thread per neuron advances `steps` steps with neuron and stimulus-input state in registers. The
// summation point then only ever holds 0 (the sum over no synapses) plus stimulus input, exactly
// as in the per-step and fused paths.
template <int NK>
__device__ __forceinline__ void neuron_only_body(const DevNet &n, int i, uint64_t t0, int steps)
{
    NeuronRegs s;
    neuron_load<NK>(n, i, s);
    InputRegs q;
    input_load(n, i, q);
    float sp = n.sp[i];
    uint64_t last = BG_NO_SPIKE;
    for (int k = 0; k < steps; k++) {
        const uint64_t t = t0 + (uint64_t)k;
        float add;
        if (input_step_regs(n, i, t, q, &add)) sp = sp + add;
        if (neuron_advance<NK>(n, i, t, sp, bg_noise_normal(n.seed, (uint32_t)i, t), s)) last = t;
        sp = 0.0f;
    }
    neuron_store<NK>(n, i, s);
    input_store(n, i, q);
    n.sp[i] = sp;
    if (last != BG_NO_SPIKE) n.fired_step[i] = last;
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_only(DevNet n, uint64_t t0, int steps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n.N) return;
    neuron_only_body<NK>(n, i, t0, steps);
}

// ---------------------------------------------------------------------------------------------
// Per-step STDP / DynamicSTDP path. The reference's step is: every neuron advances (a spike
// sets a bit in the delay ring of each of its outgoing synapses, preSpikeHit), then every synapse
// advances (Synapses/AllSTDPSynapses.cpp:283-421). Here a step is four launches:
//   1. k_stdp_neurons: thread per neuron; then, for each neuron of the wave that fired, all 64
//      lanes set the ring bits of its outgoing synapses (outgoing map);
//   2. k_stdp_stream: wave per destination; every synapse without a spike event (no pre-spike
//      leaving its ring, destination did not fire) gets its PSR decay; a destination without
//      events finishes its summation here. Event synapses are listed (ring bit still set, PSR
//      untouched) and so is their destination;
//   3. k_stdp_event_syn: thread per event synapse: ring clear, learning walks, PSR update, decay;
//   4. k_stdp_event_sum: wave per listed destination: the canonical summation of all its PSRs.
// Spike events are rare, and their learning walks and double-precision arithmetic would set the
// register budget of a single fused kernel; split this way the streaming pass runs at full
// occupancy and the event work runs with every lane busy. Event synapses are independent of one
// another (each reads its own weight and the two neurons' spike histories, which no synapse
// writes), so pass 3 may run them in any order.

// preSpikeHit of every outgoing synapse of `src` at step t, by the whole wave.
__device__ __forceinline__ void push_spike(const DevNet &n, int src, uint64_t t, int lane)
{
    const uint32_t beg = n.out_begin[src], c = n.out_count[src];
    const uint32_t Rm = (uint32_t)n.R - 1;
    const uint32_t tb = (uint32_t)t & Rm;
    for (uint32_t q = lane; q < c; q += BG_WAVE) {
        const uint32_t slot = n.out_map[beg + q];
        const uint32_t b = (tb + (uint32_t)n.out_delay[beg + q]) & Rm;
        uint32_t m;
        uint32_t *w = sring_word(n, b, slot, m);
        if (atomicOr(w, m) & m) dev_error(n, 1u);
    }
}

// The single-network per-step kernels take the step as *n.tbase + off: a graph captured for S
// steps (offsets 0..S-1, then k_step_bump) replays for every S steps of any advance call.
static __global__ void k_step_set(uint64_t *tbase, uint64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *tbase = v;
}
static __global__ void k_step_bump(uint64_t *tbase, uint32_t by)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *tbase += by;
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_neurons(DevNet n, uint32_t off)
{
    const uint64_t t = *n.tbase + off;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (BG_WAVE - 1);
    if (i < 2 * BG_EV_SHARDS) n.ev.count[i * BG_EV_PAD] = 0;   // launches have at least 256 threads
    const bool fired = i < n.N && neuron_step_body<NK>(n, i, t);
    uint64_t fm = __ballot(fired);
    while (fm) {
        const int l = __ffsll((unsigned long long)fm) - 1;
        fm &= fm - 1;
        push_spike(n, __shfl(i, l), t, lane);
    }
}

// Streaming pass for destination d (work item w). Returns whether d has events, in which case
// its summation is left to the event passes.
// Whether a post-spike event of this synapse can change its weight: the post-spike walk
// (stdp_event) reads the pre neuron's newest spike first and stops at once when that spike is
// out of the learning window, leaving the synapse as an event-free one. fired_step holds every
// neuron's newest history entry (getSpikeHistory(i, -1)); the test is the walk's own.
__device__ __forceinline__ bool post_walk_learns(const DevNet &n, uint64_t slot, uint32_t meta, uint64_t t)
{
    const uint64_t sh = n.fired_step[meta & BG_META_SRC_MASK];
    if (sh == BG_NO_SPIKE) return false;
    const int total_delay = (int)((meta >> BG_META_DELAY_SHIFT) & 0xFFu);
    const float delta = (float)(uint64_t)(t - sh - (uint64_t)(int64_t)total_delay) * n.dt;
    const float taupos = n.stdp_typed ? n.tp[(meta >> BG_META_TYPE_SHIFT) & 3u].taupos : n.taupos[slot];
    return !(delta <= 0 || (double)delta >= 3.0 * (double)taupos);
}

// The decay of a synapse of type k from the four held in registers (no dependent load).
__device__ __forceinline__ float decay_of(const float (&dec)[4], uint32_t k)
{
    const float d0 = dec[0], d1 = dec[1], d2 = dec[2], d3 = dec[3];
    const float lo = (k & 1u) ? d1 : d0, hi = (k & 1u) ? d3 : d2;
    return (k & 2u) ? hi : lo;
}

// Classifies the synapses of ranks qa and qb = qa + 64 whose ring word, type (stype) and psr are
// given: ea / eb mark event synapses (a pre spike arrives, or the destination fired and the
// post-spike walk can change W; only then is meta read); event-free ones are decayed, stored and
// added to acc.
__device__ __forceinline__ void stdp_stream_classify(const DevNet &n, const float (&dec)[4], uint64_t t, uint32_t m,
                                                     bool post, bool va, uint64_t sa, uint32_t ra, uint32_t ma,
                                                     float pa, bool vb, uint64_t sb, uint32_t rb, uint32_t mb,
                                                     float pb, float &acc, bool &ea, bool &eb)
{
    ea = va && ((ra & m) != 0 || (post && post_walk_learns(n, sa, n.meta[sa], t)));
    eb = vb && ((rb & m) != 0 || (post && post_walk_learns(n, sb, n.meta[sb], t)));
    if (va && !ea) {
        pa = pa * decay_of(dec, ma);
        n.psr[sa] = pa;
        acc += pa;
    }
    if (vb && !eb) {
        pb = pb * decay_of(dec, mb);
        n.psr[sb] = pb;
        acc += pb;
    }
}

// Destinations d0 .. d0+nd-1 (nd <= D) of one network on one wavefront, all in event-list shard
// `shard`. The pass is bound by memory latency: a destination's chain is header -> row -> store,
// and a wave with one destination of ~100 synapses spends it waiting. Here the first 128 slots of
// all D rows are loaded together with the headers (lane j loads destination j's), on the guess
// that the rows are dense (no holes, so rank = slot offset, the usual case); a row the header
// shows to have holes is reloaded through the incoming map. The chain is thus one level shorter
// and paid once for D destinations; the wave's event synapses and event destinations are then
// listed with one reservation per list. Each destination keeps its own lane-strided accumulation
// and wave tree, so its sum is the same as with one destination per wave; rows longer than 128
// ranks continue in a loop of their own.
template <int SK, int W32, int D, int NS>
__device__ __forceinline__ void stdp_stream_dests(const DevNet &n, int d0, int nd, int lane, uint64_t t, uint32_t g0,
                                                  uint32_t shard, const EvLists &ev)
{
    const float dec[4] = {n.decay[0], n.decay[1], n.decay[2], n.decay[3]};
    const uint32_t tb = (uint32_t)(t & (uint64_t)(n.R - 1));
    const uint8_t *const rplane = (const uint8_t *)n.ring + (uint64_t)(tb >> 3) * n.S;   // this step's byte plane
    const uint32_t m = 1u << (tb & 7);
    uint32_t h_cnt = 0, h_hi = 0, h_beg = 0;
    uint64_t h_fs = BG_NO_SPIKE;
    float h_sp = 0.0f;
    if (lane < nd) {
        const int d = d0 + lane;
        h_cnt = n.syn_count[d];
        h_hi = n.slot_hi[d];
        h_beg = n.in_begin[d];
        h_fs = n.fired_step[d];
        h_sp = n.sp[d];
    }
    uint32_t cnt[D];
    const uint32_t *imap[D];
    uint64_t sa[D], sb[D];
    uint32_t ra[D], rb[D], ma[D], mb[D];
    float pa[D], pb[D];
    const uint32_t qa = (uint32_t)lane, qb = qa + BG_WAVE;
    const uint32_t M = (uint32_t)n.M;
#pragma unroll
    for (int j = 0; j < D; j++) {
        const uint64_t row = (uint64_t)(d0 + j) * n.M;
        sa[j] = row + qa; sb[j] = row + qb;
        ra[j] = rb[j] = ma[j] = mb[j] = 0u;
        pa[j] = pb[j] = 0.0f;
        if (j < nd && qa < M) { ra[j] = rplane[sa[j]]; ma[j] = n.stype[sa[j]]; pa[j] = n.psr[sa[j]]; }
        if (j < nd && qb < M) { rb[j] = rplane[sb[j]]; mb[j] = n.stype[sb[j]]; pb[j] = n.psr[sb[j]]; }
    }
    const uint64_t dense_mask = __ballot(h_cnt == h_hi), post_mask = __ballot(lane < nd && h_fs == t);
#pragma unroll
    for (int j = 0; j < D; j++) {
        cnt[j] = j < nd ? (uint32_t)__builtin_amdgcn_readlane((int)h_cnt, j) : 0u;
        const bool dense = (dense_mask >> j) & 1ull;
        imap[j] = dense ? nullptr : n.in_map + (uint32_t)__builtin_amdgcn_readlane((int)h_beg, j);
        if (!dense) {
            if (qa < cnt[j]) {
                sa[j] = (uint64_t)imap[j][qa];
                ra[j] = rplane[sa[j]]; ma[j] = n.stype[sa[j]]; pa[j] = n.psr[sa[j]];
            }
            if (qb < cnt[j]) {
                sb[j] = (uint64_t)imap[j][qb];
                rb[j] = rplane[sb[j]]; mb[j] = n.stype[sb[j]]; pb[j] = n.psr[sb[j]];
            }
        }
    }
    float acc[D];
    uint64_t ema[D], emb[D];
    uint32_t event_dests = 0;   // bit j: destination j has events (and is summed by the event pass)
#pragma unroll
    for (int j = 0; j < D; j++) {
        acc[j] = 0.0f;
        ema[j] = emb[j] = 0ull;
        if (j >= nd) continue;
        const bool post = (post_mask >> j) & 1ull;
        bool ea, eb;
        stdp_stream_classify(n, dec, t, m, post, qa < cnt[j], sa[j], ra[j], ma[j], pa[j], qb < cnt[j], sb[j], rb[j],
                             mb[j], pb[j], acc[j], ea, eb);
        ema[j] = __ballot(ea);
        emb[j] = __ballot(eb);
        bool more = false;
        if (cnt[j] > 2u * BG_WAVE) {   // long rows: the remaining ranks, listed pass by pass
            const uint32_t w = g0 + (uint32_t)j;
            const bool dense = (dense_mask >> j) & 1ull;
            const uint64_t row = (uint64_t)(d0 + j) * n.M;
            for (uint32_t q0 = 2 * BG_WAVE; q0 < cnt[j]; q0 += 2 * BG_WAVE) {
                const uint32_t xa = q0 + (uint32_t)lane, xb = xa + BG_WAVE;
                const bool va = xa < cnt[j], vb = xb < cnt[j];
                uint64_t s1 = 0, s2 = 0;
                uint32_t r1 = 0, r2 = 0, m1 = 0, m2 = 0;
                float p1 = 0.0f, p2 = 0.0f;
                if (va) {
                    s1 = dense ? row + xa : (uint64_t)imap[j][xa];
                    r1 = rplane[s1]; m1 = n.stype[s1]; p1 = n.psr[s1];
                }
                if (vb) {
                    s2 = dense ? row + xb : (uint64_t)imap[j][xb];
                    r2 = rplane[s2]; m2 = n.stype[s2]; p2 = n.psr[s2];
                }
                bool e1, e2;
                stdp_stream_classify(n, dec, t, m, post, va, s1, r1, m1, p1, vb, s2, r2, m2, p2, acc[j], e1, e2);
                const uint64_t b1 = __ballot(e1), b2 = __ballot(e2);
                if (b1 | b2) {
                    more = true;
                    const uint32_t n1 = (uint32_t)__popcll(b1);
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&ev.count[(NS + shard) * BG_EV_PAD], n1 + (uint32_t)__popcll(b2));
                    base = __shfl(base, 0) + ev.syn_off[shard];
                    const uint64_t below = (1ull << lane) - 1ull;
                    if (e1) ev.syn[base + (uint32_t)__popcll(b1 & below)] = make_uint2(w, (uint32_t)s1);
                    if (e2) ev.syn[base + n1 + (uint32_t)__popcll(b2 & below)] = make_uint2(w, (uint32_t)s2);
                }
            }
        }
        if (ema[j] | emb[j] | (uint64_t)more) event_dests |= 1u << j;
    }
    // the first-pass event synapses and the event destinations of all D rows, one reservation each
    uint32_t nsyn = 0;
#pragma unroll
    for (int j = 0; j < D; j++) nsyn += (uint32_t)(__popcll(ema[j]) + __popcll(emb[j]));
    if (event_dests) {
        uint32_t bs = 0, bd = 0;
        if (lane == 0) {
            if (nsyn) bs = atomicAdd(&ev.count[(NS + shard) * BG_EV_PAD], nsyn);
            bd = atomicAdd(&ev.count[shard * BG_EV_PAD], (uint32_t)__popc(event_dests));
        }
        bs = (uint32_t)__shfl((int)bs, 0) + ev.syn_off[shard];
        bd = (uint32_t)__shfl((int)bd, 0) + ev.dest_off[shard];
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (int j = 0; j < D; j++) {
            if (!((event_dests >> j) & 1u)) continue;
            const uint32_t w = g0 + (uint32_t)j;
            if ((ema[j] >> lane) & 1ull) ev.syn[bs + (uint32_t)__popcll(ema[j] & below)] = make_uint2(w, (uint32_t)sa[j]);
            bs += (uint32_t)__popcll(ema[j]);
            if ((emb[j] >> lane) & 1ull) ev.syn[bs + (uint32_t)__popcll(emb[j] & below)] = make_uint2(w, (uint32_t)sb[j]);
            bs += (uint32_t)__popcll(emb[j]);
            if (lane == 0) ev.dest[bd] = w;
            bd++;
        }
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
        if (j >= nd || ((event_dests >> j) & 1u)) continue;
        const float tot = wave_sum(acc[j]);
        if (lane == j) n.sp[d0 + j] = h_sp + tot;
    }
}

// destinations per wavefront of the streaming pass
#ifndef BG_STREAM_DESTS
#define BG_STREAM_DESTS 4
#endif
template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_stream(DevNet n, uint32_t off)
{
    const uint64_t t = *n.tbase + off;
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int d0 = __builtin_amdgcn_readfirstlane(
        (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE) * BG_STREAM_DESTS);
    if (d0 >= n.N) return;
    const uint32_t chunk = (uint32_t)d0 / BG_STREAM_DESTS;
    stdp_stream_dests<SK, W32, BG_STREAM_DESTS, BG_EV_SHARDS>(n, d0, min(BG_STREAM_DESTS, n.N - d0), lane, t,
                                                              (uint32_t)d0, chunk % BG_EV_SHARDS, n.ev);
}

// One event synapse (destination d, rank q): the body of the reference's advanceSynapse for a
// synapse with fPre or fPost. A single network's event pass runs stdp_event_slot, which requests
// everything before its first store (12.9 -> 9.1 us per step on a 4,624-neuron Izhikevich +
// DynamicSTDP network). The batch's event pass keeps this form, which reads descriptor fields where
// they are first needed: on the 256-network sweep the up-front form was slower there, 59 against
// 56 us per step with shards shared by several networks (every descriptor field read a divergent
// gather) and still 52.6 against 48.1 us with a shard range per member; its 182 VGPRs leave two
// wavefronts per SIMD for the batch's larger event grid.
template <int SK>
__device__ __forceinline__ void stdp_event_syn_lazy(const DevNet &n, int d, uint64_t slot, uint64_t t)
{
    const uint32_t tb = (uint32_t)(t & (uint64_t)(n.R - 1));
    uint32_t m;
    uint32_t *wp = sring_word(n, tb, slot, m);
    const bool pre = (*wp & m) != 0;
    if (pre) atomicAnd(wp, ~m);
    const bool post = n.fired_step[d] == t;
    const uint32_t meta = n.meta[slot];
    const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
    const float decay = n.decay[type];
    const float psr = stdp_event<SK>(n, slot, type, t, pre, post, (int)(meta & BG_META_SRC_MASK), d,
                                     (int)((meta >> BG_META_DELAY_SHIFT) & 0xFFu), n.psr[slot], decay);
    n.psr[slot] = psr * decay;
}

// The descriptor is read from device memory: the out-of-line learning call takes its address,
// which for a by-value kernel argument would force a private copy of it in every lane.
// Grids of the event passes: BG_EV_SHARDS * G blocks; block b works on shard b mod BG_EV_SHARDS
// together with the other G - 1 blocks of that shard.
template <int SK>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_event_syn(const DevNet *np, uint32_t off)
{
    const DevNet &n = *np;
    const uint64_t t = *n.tbase + off;
    const uint32_t s = blockIdx.x % BG_EV_SHARDS, g = blockIdx.x / BG_EV_SHARDS;
    const uint32_t G = gridDim.x / BG_EV_SHARDS;
    const uint32_t c = n.ev.count[(BG_EV_SHARDS + s) * BG_EV_PAD];
    const uint2 *list = n.ev.syn + n.ev.syn_off[s];
    for (uint32_t k = g * BG_BLOCK + threadIdx.x; k < c; k += G * BG_BLOCK) {
        const uint2 e = list[k];
        stdp_event_slot<SK>(n, (int)e.x, e.y, t);
    }
}

// The summation of a destination that had events: its psr values are final after the event pass.
// The first 128 slots are loaded together with the header on the guess that the row is dense
// (rank = slot offset), as in the streaming pass; a row with holes goes through the incoming map.
__device__ __forceinline__ void stdp_sum_body(const DevNet &n, int d, int lane)
{
    const uint64_t row = (uint64_t)d * n.M;
    const uint32_t M = (uint32_t)n.M, qa = (uint32_t)lane, qb = qa + BG_WAVE;
    const float sa = qa < M ? n.psr[row + qa] : 0.0f, sb = qb < M ? n.psr[row + qb] : 0.0f;
    const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
    const float sp0 = n.sp[d];
    float acc = 0.0f;
    if (cnt == hi) {
        if (qa < cnt) acc += sa;
        if (qb < cnt) acc += sb;
        for (uint32_t q = qb + BG_WAVE; q < cnt; q += BG_WAVE) acc += n.psr[row + q];
    } else {
        const uint32_t *imap = n.in_map + n.in_begin[d];
        for (uint32_t q = qa; q < cnt; q += BG_WAVE) acc += n.psr[imap[q]];
    }
    acc = wave_sum(acc);
    if (lane == 0) n.sp[d] = sp0 + acc;
}

// stdp_sum_body for destinations d_0 .. d_{nd-1} (nd <= D) of one network on one wavefront (lane j
// holds d_j): the D rows' first 128 slots and headers are requested together, so the chain is paid
// once for D destinations; each destination's sum has stdp_sum_body's order and tree.
template <int D>
__device__ __forceinline__ void stdp_sum_dests(const DevNet &n, int dl, int nd, int lane)
{
    const uint32_t M = (uint32_t)n.M, qa = (uint32_t)lane, qb = qa + BG_WAVE;
    uint32_t h_cnt = 0, h_hi = 0, h_beg = 0;
    float h_sp = 0.0f;
    if (lane < nd) {
        h_cnt = n.syn_count[dl];
        h_hi = n.slot_hi[dl];
        h_beg = n.in_begin[dl];
        h_sp = n.sp[dl];
    }
    float sa[D], sb[D];
    uint64_t row[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        row[j] = (uint64_t)__builtin_amdgcn_readlane(dl, j) * M;
        sa[j] = j < nd && qa < M ? n.psr[row[j] + qa] : 0.0f;
        sb[j] = j < nd && qb < M ? n.psr[row[j] + qb] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
        if (j >= nd) break;
        const uint32_t cnt = (uint32_t)__builtin_amdgcn_readlane((int)h_cnt, j);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)h_hi, j);
        float acc = 0.0f;
        if (cnt == hi) {
            if (qa < cnt) acc += sa[j];
            if (qb < cnt) acc += sb[j];
            for (uint32_t q = qb + BG_WAVE; q < cnt; q += BG_WAVE) acc += n.psr[row[j] + q];
        } else {
            const uint32_t *imap = n.in_map + (uint32_t)__builtin_amdgcn_readlane((int)h_beg, j);
            for (uint32_t q = qa; q < cnt; q += BG_WAVE) acc += n.psr[imap[q]];
        }
        acc = wave_sum(acc);
        if (lane == j) n.sp[dl] = h_sp + acc;
    }
}

static __global__ __launch_bounds__(BG_BLOCK) void k_stdp_event_sum(DevNet n)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t s = blockIdx.x % BG_EV_SHARDS, g = blockIdx.x / BG_EV_SHARDS;
    const uint32_t G = gridDim.x / BG_EV_SHARDS, wpb = BG_BLOCK / BG_WAVE;
    const uint32_t c = n.ev.count[s * BG_EV_PAD];
    const uint32_t *list = n.ev.dest + n.ev.dest_off[s];
    for (uint32_t k = g * wpb + threadIdx.x / BG_WAVE; k < c; k += G * wpb) stdp_sum_body(n, (int)list[k], lane);
}

// ---------------------------------------------------------------------------------------------
// Fused, temporally blocked step kernel for Spiking/DS synapses. A launch advances K consecutive
// steps (K <= every synapse's total_delay, so no spike fired inside the block is delivered inside
// it). One wavefront owns one destination neuron and all its incoming synapses:
//   1. pull the spikes its sources fired during the previous block into the delay rings;
//   2. run every synapse through the K steps in registers (ring test, PSR update, decay),
//      accumulating the K per-step summation points in the canonical order;
//   3. advance the neuron K steps with those inputs and publish a K-bit fire mask.
// Each synapse's meta/ring/psr are read and written once per K steps instead of once per step.
// A synapse's first delivered spike of the block has its DS update evaluated before the step
// loop, so the divergent transcendental work runs once per synapse and block rather than once
// per step; later spikes in the same block (possible only when a source's refractory period is
// shorter than the block) are handled inside the loop.
// One synapse's cold state (W, and for DS r, u, lastSpike), loaded ahead when its delivery is
// likely (valid) or on demand otherwise.
struct ColdDS {
    float W, r, u;
    uint64_t last;
    bool valid;
};

template <int SK>
__device__ __forceinline__ void cold_load(const DevNet &n, uint64_t slot, ColdDS &c)
{
    c.W = n.W[slot];
    if (SK == SK_DS) {
        c.r = n.r[slot];
        c.u = n.us[slot];
        c.last = n.last[slot];
    }
    c.valid = true;
}

// The per-type constants of a network held in registers for a wavefront: a per-lane type then
// selects among them instead of indexing an array of the descriptor (a memory access).
struct TypeRegs {
    float dec[4], U[4], D[4], F[4];
};

__device__ __forceinline__ float pick4(const float (&v)[4], int k)
{   // selects values (not array elements), so the array stays in registers
    const float v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
    const float lo = (k & 1) ? v1 : v0, hi = (k & 1) ? v3 : v2;
    return (k & 2) ? hi : lo;
}

__device__ __forceinline__ void type_regs(const DevNet &n, TypeRegs &tr)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        tr.dec[k] = n.decay[k];
        tr.U[k] = n.tU[k];
        tr.D[k] = n.tD[k];
        tr.F[k] = n.tF[k];
    }
}

// AllDSSynapses::changePSR's increment (Synapses/AllDSSynapses.cpp:227-247) on register state:
// the same operations in the same order as psr_increment, with r, u and lastSpike carried in c.
template <int SK>
__device__ __forceinline__ float psr_increment_regs(const DevNet &n, const TypeRegs &tr, uint64_t slot, int type,
                                                    uint64_t t, float decay, ColdDS &c, bool &ru_changed)
{
    if (SK == SK_DS) {
        if (c.last != BG_NO_SPIKE) {
            // steps since the last delivery, converted from 32 bits when they fit (the same value,
            // a shorter conversion than from 64 bits)
            const uint64_t gap = t - c.last;
            const float isi = ((gap >> 32) ? (float)gap : (float)(uint32_t)gap) * n.dt;
            float D, F, U;
            if (n.ds_typed) {
                D = pick4(tr.D, type);
                F = pick4(tr.F, type);
                U = pick4(tr.U, type);
            } else {
                D = n.Dd[slot];
                F = n.F[slot];
                U = n.U[slot];
            }
            c.r = 1.0f + (c.r * (1.0f - c.u) - 1.0f) * bg_expf(-isi / D);
            c.u = U + c.u * (1.0f - U) * bg_expf(-isi / F);
            ru_changed = true;
        }
        c.last = t;
        return ((c.W / decay) * c.u) * c.r;
    } else {
        return c.W / decay;
    }
}

// synapse_step_body for Spiking/DS synapses with 32-bit rings, as a software pipeline over the
// row's passes of 64 ranks: while pass p is processed, pass p+1 has its sources' last firing steps
// and (for its delivering lanes) its cold state in flight, and pass p+2 its meta, ring word and psr.
// Every load instruction is unconditional (ranks past the row's end are clamped to its last rank;
// lanes without a delivery read the cold state of the row's first slot, one shared cache line), so
// the waits before a pass cover only its own operands. Same per-slot operations, same order.
struct StepRank {
    uint32_t slot, meta, ring;
    float psr;
};
struct StepCold {
    float W, r, u;
    uint64_t last;
};

template <int SK, bool DENSE>
__device__ __forceinline__ void step_hot_load(const DevNet &n, uint32_t row, const uint32_t *imap, uint32_t q,
                                              StepRank &h)
{
    h.slot = DENSE ? row + q : imap[q];
    const uint32_t off = h.slot << 2;
    h.meta = ld_off(n.meta, off);
    h.ring = ld_off(n.ring, off);
    h.psr = ld_off(n.psr, off);
}

template <int SK>
__device__ __forceinline__ void step_cold_load(const DevNet &n, uint32_t row, const StepRank &h, uint32_t tb,
                                               StepCold &c)
{
    const uint32_t off = (((h.ring >> tb) & 1u) ? h.slot : row) << 2;
    c.W = ld_off(n.W, off);
    if (SK == SK_DS) {
        c.r = ld_off(n.r, off);
        c.u = ld_off(n.us, off);
        c.last = ld_off(n.last, off << 1);
    }
}

template <int SK, bool DENSE>
__device__ __forceinline__ void synapse_step_row(const DevNet &n, const TypeRegs &tr, int d, uint32_t cnt, uint64_t t,
                                                 float &acc)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t row = (uint32_t)d * (uint32_t)n.M;
    const uint32_t *imap = n.in_map + (DENSE ? 0 : n.in_begin[d]);
    const uint32_t tb = (uint32_t)(t & 31u);
    const uint32_t P = (cnt + BG_WAVE - 1) / BG_WAVE, last = cnt - 1;
    auto rank = [&](uint32_t pass) {
        const uint32_t q = pass * BG_WAVE + (uint32_t)lane;
        return q < cnt ? q : last;
    };
    StepRank x, y, z;
    StepCold cx, cy, cz;
    uint64_t fx, fy, fz;
    step_hot_load<SK, DENSE>(n, row, imap, rank(0), x);
    step_hot_load<SK, DENSE>(n, row, imap, rank(1), y);
    __builtin_amdgcn_sched_barrier(0);
    fx = ld_off(n.fired_step, (x.meta & BG_META_SRC_MASK) << 3);
    step_cold_load<SK>(n, row, x, tb, cx);
    auto pass = [&](uint32_t p, StepRank &cur, StepCold &ccur, uint64_t &fcur, StepRank &nxt, StepCold &cnxt,
                    uint64_t &fnxt, StepRank &far) {
        step_hot_load<SK, DENSE>(n, row, imap, rank(p + 2), far);
        fnxt = ld_off(n.fired_step, (nxt.meta & BG_META_SRC_MASK) << 3);
        step_cold_load<SK>(n, row, nxt, tb, cnxt);
        const bool valid = p * BG_WAVE + (uint32_t)lane < cnt;
        const int type = (cur.meta >> BG_META_TYPE_SHIFT) & 3;
        const uint32_t delay = (cur.meta >> BG_META_DELAY_SHIFT) & 0xFF;
        uint32_t w = cur.ring;
        bool changed = false;
        if (fcur == t) {   // preSpikeHit at the firing step
            const uint32_t m = 1u << ((tb + delay) & 31u);
            if (valid && (w & m)) dev_error(n, 1u);
            w |= m;
            changed = true;
        }
        const bool pre = (w >> tb) & 1u;
        if (pre) {
            w &= ~(1u << tb);
            changed = true;
        }
        const float decay = pick4(tr.dec, type);
        float psr = cur.psr;
        if (pre) {
            ColdDS c;
            c.W = ccur.W;
            c.r = ccur.r;
            c.u = ccur.u;
            c.last = ccur.last;
            c.valid = true;
            bool ru_changed = false;
            psr = psr + psr_increment_regs<SK>(n, tr, cur.slot, type, t, decay, c, ru_changed);
            if (valid && SK == SK_DS) {
                const uint32_t off = cur.slot << 2;
                if (ru_changed) {
                    st_off(n.r, off, c.r);
                    st_off(n.us, off, c.u);
                }
                st_off(n.last, off << 1, c.last);
            }
        }
        psr = psr * decay;
        if (valid) {
            st_off(n.psr, cur.slot << 2, psr);
            if (changed) st_off(n.ring, cur.slot << 2, w);
            acc += psr;
        }
    };
    for (uint32_t p = 0; p < P; p += 3) {
        pass(p, x, cx, fx, y, cy, fy, z);
        if (p + 1 >= P) break;
        pass(p + 1, y, cy, fy, z, cz, fz, x);
        if (p + 2 >= P) break;
        pass(p + 2, z, cz, fz, x, cx, fx, y);
    }
}

template <int SK, int W32>
__device__ __forceinline__ void synapse_step_any(const DevNet &n, int d, int lane, uint64_t t)
{
    if (W32 == 1 && (SK == SK_SPIKING || SK == SK_DS)) {
        if (n.S < (1ull << 29)) {   // 32-bit byte offsets (8-byte lastSpike included)
            TypeRegs tr;
            type_regs(n, tr);
            const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
            float acc = 0.0f;
            if (cnt) {
                if (cnt == hi) synapse_step_row<SK, true>(n, tr, d, cnt, t, acc);
                else synapse_step_row<SK, false>(n, tr, d, cnt, t, acc);
            }
            acc = wave_sum(acc);
            if (lane == 0) n.sp[d] = n.sp[d] + acc;
            return;
        }
    }
    synapse_step_body<SK, W32>(n, d, lane, t);
}

template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_synapse_step(DevNet n, uint64_t t)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int d = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (d >= n.N) return;
    synapse_step_any<SK, W32>(n, d, lane, t);
}

// A synapse through the K steps of a block: pull the spikes its source fired in the previous block
// (fire mask fm) into the ring, take the block's deliveries out of it, and run the PSR with the
// canonical per-step accumulation. The first delivery's increment is evaluated before the step
// loop, so the divergent transcendental work runs once per synapse and block. With a 32-bit ring
// the pulled spikes are one rotated copy of fm and the block's deliveries one rotated window of the
// ring word; wider rings take them bit by bit.
// Block length of the fused kernels: KMAX = BG_K_COMMON is an instantiation for exactly that many
// steps, with no guard at all; any other KMAX (8, 16, 32) takes any K <= KMAX at run time (the step
// loops are unrolled to KMAX with a guard on k < K, which the compiler turns into selects on every
// step). The 8-step one runs the shorter last blocks of a call in the registers of the 9-step one.
// the exact block length instantiated: the minimum reference delay at dt = 0.1 ms, int(0.8/0.1)+1
#define BG_K_COMMON 9
#define BG_K_EXACT(KMAX) ((KMAX) == BG_K_COMMON)

// Occupancy target of the fused kernels (waves per SIMD): bounds their registers.
#ifndef BG_FUSED_WAVES
#define BG_FUSED_WAVES 5
#endif

template <int SK, int W32, int KMAX>
__device__ __forceinline__ void fused_synapse(const DevNet &n, const TypeRegs &tr, uint64_t slot, uint32_t meta,
                                              uint32_t (&rw)[W32], float &psr, uint32_t fm, ColdDS &c, uint64_t t0,
                                              int K, uint64_t tprev0, float (&acc)[KMAX], uint32_t &hit)
{
    const uint32_t Rm = (uint32_t)n.R - 1;
    const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
    const int Kb = BG_K_EXACT(KMAX) ? KMAX : K;
    const uint32_t kmask = Kb >= 32 ? 0xFFFFFFFFu : ((1u << Kb) - 1u);
    bool changed = false;
    uint32_t hm = 0;
    if (W32 == 1) {
        if (fm) {
            // spike j of the previous block is delivered at bit (tprev0 + j + delay) mod 32
            const uint32_t delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
            const uint32_t sh = (uint32_t)((tprev0 + (uint64_t)delay) & 31u);
            const uint32_t ins = (fm << sh) | (sh ? fm >> (32u - sh) : 0u);
            if (rw[0] & ins) dev_error(n, 1u);
            rw[0] |= ins;
            changed = true;
        }
        const uint32_t b0 = (uint32_t)(t0 & 31u);
        hm = __builtin_amdgcn_alignbit(rw[0], rw[0], b0) & kmask;   // bit k = ring bit (t0 + k) mod 32
        const uint32_t clear = (kmask << b0) | (b0 ? kmask >> (32u - b0) : 0u);
        rw[0] &= ~clear;
    } else {
        if (fm) {
            const uint32_t delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
            while (fm) {
                const int b0 = __ffs(fm) - 1;
                fm &= fm - 1;
                const uint32_t b = (uint32_t)((tprev0 + (uint64_t)b0 + delay) & Rm);
                const uint32_t m = 1u << (b & 31);
#pragma unroll
                for (int w = 0; w < W32; w++)
                    if ((int)(b >> 5) == w) {
                        if (rw[w] & m) dev_error(n, 1u);
                        rw[w] |= m;
                    }
                changed = true;
            }
        }
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                const uint32_t b = (uint32_t)((t0 + (uint64_t)k) & Rm);
                const uint32_t m = 1u << (b & 31);
#pragma unroll
                for (int w = 0; w < W32; w++)
                    if ((int)(b >> 5) == w && (rw[w] & m)) {
                        rw[w] &= ~m;
                        hm |= 1u << k;
                    }
            }
        }
    }
    hit = hm;
    const float decay = pick4(tr.dec, type);
    int k1 = -1;
    float c1 = 0.0f;
    bool ru_changed = false;
    if (hm) {
        changed = true;
        if (!c.valid) cold_load<SK>(n, slot, c);
        k1 = __ffs(hm) - 1;
        c1 = psr_increment_regs<SK>(n, tr, slot, type, t0 + (uint64_t)k1, decay, c, ru_changed);
        hm &= hm - 1;
    }
    float p = psr;
    if (__ballot(hm != 0) == 0ull) {
        // no lane has a second delivery in this block (a source's refractory period is longer
        // than the block in every network so far): the step loop without the per-step test
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                p = (k == k1) ? p + c1 : p;
                p = p * decay;
                acc[k] += p;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                p = (k == k1) ? p + c1 : p;
                if ((hm >> k) & 1u)
                    p = p + psr_increment_regs<SK>(n, tr, slot, type, t0 + (uint64_t)k, decay, c, ru_changed);
                p = p * decay;
                acc[k] += p;
            }
        }
    }
    psr = p;
    if (k1 >= 0 && SK == SK_DS) {
        if (ru_changed) {
            n.r[slot] = c.r;
            n.us[slot] = c.u;
        }
        n.last[slot] = c.last;
    }
    if (changed) {
#pragma unroll
        for (int w = 0; w < W32; w++) n.ring[(uint64_t)w * n.S + slot] = rw[w];
    }
}

// The hot state of one rank in the pipeline of fused_block_body.
template <int W32>
struct HotRank {
    uint64_t slot;
    uint32_t meta;
    uint32_t rw[W32];
    float psr;
};

template <int W32>
__device__ __forceinline__ void hot_load(const DevNet &n, bool dense, uint64_t row, const uint32_t *imap, uint32_t q,
                                         uint32_t cnt, HotRank<W32> &h)
{
    h.slot = 0;
    h.meta = 0;
#pragma unroll
    for (int w = 0; w < W32; w++) h.rw[w] = 0;
    h.psr = 0.0f;
    if (q < cnt) {
        h.slot = dense ? row + q : (uint64_t)imap[q];
        h.meta = n.meta[h.slot];
#pragma unroll
        for (int w = 0; w < W32; w++) h.rw[w] = n.ring[(uint64_t)w * n.S + h.slot];
        h.psr = n.psr[h.slot];
    }
}

template <int NK, int SK, int W32, int KMAX>
__device__ __forceinline__ void fused_block_body(const DevNet &n, int d, uint64_t t0, int K, int pull,
                                                 uint64_t tprev0, int sel_in)
{
    __shared__ float s_noise[BG_BLOCK / BG_WAVE][KMAX];
    __shared__ float s_in[BG_BLOCK / BG_WAVE][KMAX];
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int wv = threadIdx.x / BG_WAVE;
    const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
    const bool dense = (cnt == hi);
    const uint32_t *mprev = n.mask[sel_in];
    const uint64_t row = (uint64_t)d * n.M;
    const uint32_t *imap = n.in_map + (dense ? 0 : n.in_begin[d]);
    // noise samples of the K steps, computed one per lane up front and read back by lane 0
    if (lane < K) s_noise[wv][lane] = bg_noise_normal(n.seed, (uint32_t)d, t0 + (uint64_t)lane);
    float acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) acc[k] = 0.0f;

    // Lane-strided ranks (the canonical summation order) in a three-stage pipeline: while rank q is
    // processed, rank q+64 has its source's fire mask and (when deliveries are frequent) its cold
    // state in flight, and rank q+128 its meta/ring/psr. Cold state is fetched ahead only while the
    // previous pass had at least 4 of 64 lanes with a delivery: then nearly every cache line of it
    // is needed anyway, and otherwise the few delivering synapses load it on demand.
    const bool do_pull = pull != 0;
    TypeRegs tr;
    type_regs(n, tr);
    HotRank<W32> x, y, z;
    ColdDS cx, cy, cz;
    uint32_t fx = 0, fy = 0, fz = 0;
    bool spec = false;
    uint32_t q = lane;
    hot_load<W32>(n, dense, row, imap, q, cnt, x);
    hot_load<W32>(n, dense, row, imap, q + BG_WAVE, cnt, y);
    cx.valid = false;
    if (do_pull && q < cnt) fx = mprev[x.meta & BG_META_SRC_MASK];
    // one pass: process `cur`, issue the fire-mask gather (and cold state) of `nxt`, whose hot state
    // has arrived, and the hot loads of `far`. The loop is unrolled by three so that the three
    // stages trade roles instead of being copied: a copy of registers whose loads are still in
    // flight would wait for them and drain the pipeline.
    auto pass = [&](HotRank<W32> &cur, ColdDS &ccur, uint32_t &fcur, HotRank<W32> &nxt, ColdDS &cnxt, uint32_t &fnxt,
                    HotRank<W32> &far) {
        hot_load<W32>(n, dense, row, imap, q + 2 * BG_WAVE, cnt, far);
        cnxt.valid = false;
        fnxt = 0;
        if (q + BG_WAVE < cnt) {
            if (do_pull) fnxt = mprev[nxt.meta & BG_META_SRC_MASK];
            if (spec) cold_load<SK>(n, nxt.slot, cnxt);
        }
        uint32_t hit = 0;
        fused_synapse<SK, W32, KMAX>(n, tr, cur.slot, cur.meta, cur.rw, cur.psr, fcur, ccur, t0, K, tprev0, acc, hit);
        n.psr[cur.slot] = cur.psr;
        spec = __popcll(__ballot(hit != 0)) >= 4;
        q += BG_WAVE;
    };
    while (true) {
        if (q >= cnt) break;
        pass(x, cx, fx, y, cy, fy, z);
        if (q >= cnt) break;
        pass(y, cy, fy, z, cz, fz, x);
        if (q >= cnt) break;
        pass(z, cz, fz, x, cx, fx, y);
    }
#pragma unroll
    for (int k = 0; k < KMAX; k++)
        if (BG_K_EXACT(KMAX) || k < K) acc[k] = wave_sum(acc[k]);

    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    if (lane == 0) {
        // stimulus input of the K steps (independent of the neuron's state), then the neuron
        uint32_t inbits = 0;
        if (n.in_kind != BG_INPUT_NONE) {
            for (int k = 0; k < K; k++) {
                float add;
                if (input_step(n, d, t0 + (uint64_t)k, &add)) {
                    s_in[wv][k] = add;
                    inbits |= 1u << k;
                }
            }
        }
        NeuronRegs s;
        neuron_load<NK>(n, d, s);
        float sp = n.sp[d];
        uint32_t mask = 0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (BG_K_EXACT(KMAX) || k < K) {
                if ((inbits >> k) & 1u) sp = sp + s_in[wv][k];
                if (neuron_advance<NK>(n, d, t0 + (uint64_t)k, sp, s_noise[wv][k], s)) mask |= 1u << k;
                sp = 0.0f + acc[k];
            }
        }
        neuron_store<NK>(n, d, s);
        n.sp[d] = sp;
        n.mask[sel_in ^ 1][d] = mask;
        if (mask) n.fired_step[d] = t0 + (uint64_t)(31 - __clz(mask));
    }
}

template <int NK, int SK, int W32, int KMAX>
__global__ __launch_bounds__(BG_BLOCK) __attribute__((amdgpu_waves_per_eu(BG_FUSED_WAVES))) void k_fused_block(
    DevNet n, uint64_t t0, int K, int pull, uint64_t tprev0, int sel_in)
{
    const int d = (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE);
    if (d >= n.N) return;
    fused_block_body<NK, SK, W32, KMAX>(n, d, t0, K, pull, tprev0, sel_in);
}

// ---------------------------------------------------------------------------------------------
// Multi-network forms. One launch advances every network of a batch: work item w belongs to
// network i with base[i] <= w < base[i+1] (found by a wave-uniform binary search), and each
// network keeps its own step counter, noise seed and fire-mask buffers, so a network's results do
// not depend on which other networks share the launch.
struct BatchNet {
    DevNet n;
    uint64_t t0;        // step of the network at the first launch of this batch call
    uint64_t tprev0;    // start step of its last block before this call
    int sel;            // mask buffer holding that block's masks
    int pull;           // 1 if those spikes are still to be inserted into the rings
};

template <int NK, int SK, int W32, int KMAX>
__global__ __launch_bounds__(BG_BLOCK) __attribute__((amdgpu_waves_per_eu(BG_FUSED_WAVES))) void k_fused_batch(
    BatchMap bm, int K, int Kfull, int launch)
{
    // wave-uniform by construction (blockDim is a multiple of 64); readfirstlane lets the
    // compiler keep the network's descriptor in scalar registers
    const uint32_t w = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (w >= bm.total) return;
    const int i = __builtin_amdgcn_readfirstlane((int)bm.nof[w]);
    const BatchNet &b = bm.nets[i];
    const uint64_t t0 = b.t0 + (uint64_t)launch * (uint64_t)Kfull;
    const int sel = b.sel ^ (launch & 1);
    const int pull = launch > 0 ? 1 : b.pull;
    const uint64_t tprev0 = launch > 0 ? t0 - (uint64_t)Kfull : b.tprev0;
    fused_block_body<NK, SK, W32, KMAX>(b.n, (int)(w - bm.base[i]), t0, K, pull, tprev0, sel);
}

// Thread-per-neuron kernels take their neurons from the per-wave table nwaves[v] = {member, first
// neuron}: every wavefront holds neurons of one network (its last wavefront may be partial), so the
// member's descriptor is wave-uniform and read through scalar loads instead of per-lane gathers.
struct NeuronLane {
    int member;     // wave-uniform
    int neuron;     // this lane's neuron in the member network
    bool valid;     // the lane has a neuron
};
__device__ __forceinline__ NeuronLane neuron_lane(const BatchMap &bm)
{
    NeuronLane r = {0, 0, false};
    const uint32_t v = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (v >= bm.nnwaves) return r;
    const uint2 e = bm.nwaves[v];
    r.member = __builtin_amdgcn_readfirstlane((int)e.x);
    r.neuron = (int)e.y + (int)(threadIdx.x & (BG_WAVE - 1));
    r.valid = r.neuron < bm.nets[r.member].n.N;
    return r;
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_batch(BatchMap bm, uint64_t step)
{
    const NeuronLane q = neuron_lane(bm);
    if (!q.valid) return;
    const BatchNet &b = bm.nets[q.member];
    neuron_step_body<NK>(b.n, q.neuron, b.t0 + step);
}

template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_synapse_batch(BatchMap bm, uint64_t step)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (w >= bm.total) return;
    const int i = __builtin_amdgcn_readfirstlane((int)bm.nof[w]);
    const BatchNet &b = bm.nets[i];
    synapse_step_any<SK, W32>(b.n, (int)(w - bm.base[i]), threadIdx.x & (BG_WAVE - 1), b.t0 + step);
}

// The per-step STDP passes over a batch; event lists hold batch work-item indices.
template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_neurons_batch(BatchMap bm, uint64_t step, EvLists ev)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (BG_WAVE - 1);
    if (w < 2 * BG_EV_SHARDS_BATCH) ev.count[w * BG_EV_PAD] = 0;   // launches have at least that many threads
    const NeuronLane q = neuron_lane(bm);
    const BatchNet &b = bm.nets[q.member];
    bool fired = false;
    if (q.valid) fired = neuron_step_body<NK>(b.n, q.neuron, b.t0 + step);
    uint64_t fm = __ballot(fired);
    while (fm) {
        const int l = __ffsll((unsigned long long)fm) - 1;
        fm &= fm - 1;
        push_spike(b.n, __shfl(q.neuron, l), b.t0 + step, lane);
    }
}

// chunks[c] = {member | shard << 16, first destination} of the c-th run of D destinations (runs
// never cross networks). Shards are assigned per run of BG_STREAM_DESTS destinations whatever D is
// (D divides BG_STREAM_DESTS), so the event-list capacities do not depend on D.
template <int SK, int W32, int D>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_stream_batch(BatchMap bm, uint64_t step, EvLists ev)
{
    const uint32_t c = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (c >= bm.nchunks) return;
    const uint2 e = bm.chunks[c];
    const int i = __builtin_amdgcn_readfirstlane((int)(e.x & 0xFFFFu)), d0 = __builtin_amdgcn_readfirstlane((int)e.y);
    const uint32_t shard = __builtin_amdgcn_readfirstlane(e.x >> 16);
    const BatchNet &b = bm.nets[i];
    stdp_stream_dests<SK, W32, D, BG_EV_SHARDS_BATCH>(b.n, d0, min(D, b.n.N - d0), threadIdx.x & (BG_WAVE - 1),
                                                      b.t0 + step, bm.base[i] + (uint32_t)d0, shard, ev);
}

template <int SK>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_event_syn_batch(BatchMap bm, uint64_t step, EvLists ev)
{
    const uint32_t s = blockIdx.x % BG_EV_SHARDS_BATCH, g = blockIdx.x / BG_EV_SHARDS_BATCH;
    const uint32_t G = gridDim.x / BG_EV_SHARDS_BATCH;
    const uint32_t c = ev.count[(BG_EV_SHARDS_BATCH + s) * BG_EV_PAD];
    const uint2 *list = ev.syn + ev.syn_off[s];
    for (uint32_t k = g * blockDim.x + threadIdx.x; k < c; k += G * blockDim.x) {
        const uint2 e = list[k];
        const int i = bm.nof[e.x];
        const BatchNet &b = bm.nets[i];
        stdp_event_syn_lazy<SK>(b.n, (int)(e.x - bm.base[i]), e.y, b.t0 + step);
    }
}

// Four listed destinations per wavefront: a shard lists one member's events unless the batch has
// more members than shards, so the four are normally one network's (else one at a time).
static __global__ __launch_bounds__(BG_BLOCK) void k_stdp_event_sum_batch(BatchMap bm, EvLists ev)
{
    constexpr int D = 4;
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t s = blockIdx.x % BG_EV_SHARDS_BATCH, g = blockIdx.x / BG_EV_SHARDS_BATCH;
    const uint32_t G = gridDim.x / BG_EV_SHARDS_BATCH, wpb = blockDim.x / BG_WAVE;
    const uint32_t c = ev.count[s * BG_EV_PAD];
    const uint32_t *list = ev.dest + ev.dest_off[s];
    for (uint32_t k0 = (g * wpb + threadIdx.x / BG_WAVE) * D; k0 < c; k0 += G * wpb * D) {
        const int nd = (int)min((uint32_t)D, c - k0);
        const uint32_t wl = lane < nd ? list[k0 + lane] : 0u;
        const int il = lane < nd ? (int)bm.nof[wl] : 0;
        const int i0 = __builtin_amdgcn_readfirstlane(il);
        if (__ballot(lane < nd && il != i0) == 0ull) {
            stdp_sum_dests<D>(bm.nets[i0].n, (int)(wl - bm.base[i0]), nd, lane);
        } else {
            for (int j = 0; j < nd; j++) {
                const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)wl, j);
                const int i = __builtin_amdgcn_readlane(il, j);
                stdp_sum_body(bm.nets[i].n, (int)(w - bm.base[i]), lane);
            }
        }
    }
}

// Advances every member's step base by `by`: the last node of a captured run of batch steps, so
// the next replay (or the launches after it) continue from there.
static __global__ void k_batch_bump(BatchNet *nets, int nn, uint32_t by)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) nets[i].t0 += by;
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_only_batch(BatchMap bm, uint64_t step, int steps)
{
    const NeuronLane q = neuron_lane(bm);
    if (!q.valid) return;
    const BatchNet &b = bm.nets[q.member];
    neuron_only_body<NK>(b.n, q.neuron, b.t0 + step, steps);
}
