// HIP kernels of the engine (gfx950). All memory-bound integer/FP32 work: no MFMA anywhere.
#pragma once
#include "bghip_device.h"

#define BG_WAVE 64
#define BG_BLOCK 256
#define BG_KMAX 32

// ---------------------------------------------------------------------------------------------
// Per-step path (used for STDP synapses, which need every neuron's step-t spike before any
// synapse of step t): one thread per neuron. AllSpikingNeurons::advanceNeurons
// (Neurons/AllSpikingNeurons.cpp:98-147); the outgoing/incoming ring notifications become pulls in
// k_synapse_step.
template <int NK>
__device__ __forceinline__ void neuron_step_body(const DevNet &n, int i, uint64_t t)
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
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_step(DevNet n, uint64_t t)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && n.ev_count) *n.ev_count = 0;   // event list of this step's STDP passes
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
        if (n.fired_step[src] == t) {   // preSpikeHit at the firing step
            uint32_t b = (tb + (uint32_t)delay) & (uint32_t)(R - 1);
            uint32_t *w = &n.ring[(uint64_t)(b >> 5) * n.S + slot];
            uint32_t m = 1u << (b & 31);
            if (*w & m) dev_error(n, 1u);
            *w |= m;
        }
        uint32_t *w = &n.ring[(uint64_t)(tb >> 5) * n.S + slot];
        const uint32_t m = 1u << (tb & 31);
        const uint32_t word = *w;
        const bool pre = (word & m) != 0;
        if (pre) *w = word & ~m;
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

template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_synapse_step(DevNet n, uint64_t t)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int d = (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE);
    if (d >= n.N) return;
    synapse_step_body<SK, W32>(n, d, lane, t);
}

// ---------------------------------------------------------------------------------------------
// Per-step STDP / DynamicSTDP synapse advance in two passes. Spike events (a pre-synaptic spike
// leaving the delay ring, or the destination's own spike) are rare, but the learning walks and
// double-precision arithmetic they need would set the register budget of the whole kernel. The
// streaming pass therefore does the delay-ring inserts, the ring test and the PSR decay of every
// event-free synapse and completes the summation of every destination without events; a
// destination with at least one event is appended to a list and left to the event pass, which
// handles its event synapses (learning, PSR update, decay) exactly as the one-pass body and then
// sums all its synapses in the canonical order. Event synapses keep their ring bit set until the
// event pass, so both passes classify them the same way.
template <int SK, int W32>
__device__ __forceinline__ bool stdp_stream_body(const DevNet &n, int d, int lane, uint64_t t)
{
    const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
    const bool dense = (cnt == hi);
    const bool post = n.fired_step[d] == t;
    const int R = n.R;
    const uint32_t tb = (uint32_t)(t & (uint64_t)(R - 1));
    const uint64_t wbase = (uint64_t)(tb >> 5) * n.S;
    const uint32_t m = 1u << (tb & 31);
    float acc = 0.0f;
    bool ev = post && cnt > 0;
    for (uint32_t q = lane; q < cnt; q += BG_WAVE) {
        const uint64_t slot = dense ? (uint64_t)d * n.M + q : (uint64_t)n.in_map[n.in_begin[d] + q];
        const uint32_t meta = n.meta[slot];
        const int src = (int)(meta & BG_META_SRC_MASK);
        if (n.fired_step[src] == t) {   // preSpikeHit at the firing step
            const int delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
            const uint32_t b = (tb + (uint32_t)delay) & (uint32_t)(R - 1);
            uint32_t *w = &n.ring[(uint64_t)(b >> 5) * n.S + slot];
            const uint32_t mb = 1u << (b & 31);
            if (*w & mb) dev_error(n, 1u);
            *w |= mb;
        }
        if (post) continue;
        if (n.ring[wbase + slot] & m) {
            ev = true;
            continue;
        }
        const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
        const float psr = n.psr[slot] * n.decay[type];
        n.psr[slot] = psr;
        acc += psr;
    }
    if (__ballot(ev) != 0) return true;
    acc = wave_sum(acc);
    if (lane == 0) n.sp[d] = n.sp[d] + acc;
    return false;
}

template <int SK, int W32>
__device__ __forceinline__ void stdp_event_body(const DevNet &n, int d, int lane, uint64_t t)
{
    const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
    const bool dense = (cnt == hi);
    const bool post = n.fired_step[d] == t;
    const int R = n.R;
    const uint32_t tb = (uint32_t)(t & (uint64_t)(R - 1));
    const uint64_t wbase = (uint64_t)(tb >> 5) * n.S;
    const uint32_t m = 1u << (tb & 31);
    float acc = 0.0f;
    for (uint32_t q = lane; q < cnt; q += BG_WAVE) {
        const uint64_t slot = dense ? (uint64_t)d * n.M + q : (uint64_t)n.in_map[n.in_begin[d] + q];
        uint32_t *w = &n.ring[wbase + slot];
        const uint32_t word = *w;
        const bool pre = (word & m) != 0;
        float psr = n.psr[slot];
        if (pre || post) {
            if (pre) *w = word & ~m;
            const uint32_t meta = n.meta[slot];
            const int src = (int)(meta & BG_META_SRC_MASK);
            const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
            const int delay = (meta >> BG_META_DELAY_SHIFT) & 0xFF;
            const float decay = n.decay[type];
            psr = stdp_event<SK>(n, slot, type, t, pre, post, src, d, delay, psr, decay);
            psr = psr * decay;
            n.psr[slot] = psr;
        }
        acc += psr;
    }
    acc = wave_sum(acc);
    if (lane == 0) n.sp[d] = n.sp[d] + acc;
}

template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_stream(DevNet n, uint64_t t)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int d = (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE);
    if (d >= n.N) return;
    if (stdp_stream_body<SK, W32>(n, d, lane, t) && lane == 0) n.ev_list[atomicAdd(n.ev_count, 1u)] = (uint32_t)d;
}

// The descriptor is read from device memory: the out-of-line learning call takes its address,
// which for a by-value kernel argument would force a private copy of it in every lane.
template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_events(const DevNet *np, uint64_t t)
{
    const DevNet &n = *np;
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t nw = gridDim.x * (BG_BLOCK / BG_WAVE);
    const uint32_t c = *n.ev_count;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / BG_WAVE; i < c; i += nw)
        stdp_event_body<SK, W32>(n, (int)n.ev_list[i], lane, t);
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
template <int SK, int W32, int KMAX>
__device__ __forceinline__ void fused_synapse(const DevNet &n, uint64_t slot, uint32_t meta, uint32_t (&rw)[W32],
                                              float &psr, uint64_t t0, int K, bool pull, uint64_t tprev0,
                                              const uint32_t *mprev, float (&acc)[KMAX])
{
    const uint32_t Rm = (uint32_t)n.R - 1;
    const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
    bool changed = false;
    if (pull) {
        uint32_t fm = mprev[meta & BG_META_SRC_MASK];
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
    uint32_t hm = 0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k < K) {
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
    const float decay = n.decay[type];
    int k1 = -1;
    float c1 = 0.0f;
    if (hm) {
        changed = true;
        k1 = __ffs(hm) - 1;
        c1 = psr_increment<SK>(n, slot, type, t0 + (uint64_t)k1, decay);
        hm &= hm - 1;
    }
    float p = psr;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k < K) {
            p = (k == k1) ? p + c1 : p;
            if ((hm >> k) & 1u) p = p + psr_increment<SK>(n, slot, type, t0 + (uint64_t)k, decay);
            p = p * decay;
            acc[k] += p;
        }
    }
    psr = p;
    if (changed) {
#pragma unroll
        for (int w = 0; w < W32; w++) n.ring[(uint64_t)w * n.S + slot] = rw[w];
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

    // lane-strided ranks (the canonical summation order), with the next rank's meta/ring/psr
    // loads issued before the current rank is processed
    uint32_t q = lane;
    uint64_t slot = 0;
    uint32_t meta = 0;
    uint32_t rw[W32];
    float psr = 0.0f;
    if (q < cnt) {
        slot = dense ? row + q : (uint64_t)imap[q];
        meta = n.meta[slot];
#pragma unroll
        for (int w = 0; w < W32; w++) rw[w] = n.ring[(uint64_t)w * n.S + slot];
        psr = n.psr[slot];
    }
    while (q < cnt) {
        const uint32_t qn = q + BG_WAVE;
        uint64_t slot_n = slot;
        uint32_t meta_n = 0;
        uint32_t rw_n[W32];
        float psr_n = 0.0f;
        if (qn < cnt) {
            slot_n = dense ? row + qn : (uint64_t)imap[qn];
            meta_n = n.meta[slot_n];
#pragma unroll
            for (int w = 0; w < W32; w++) rw_n[w] = n.ring[(uint64_t)w * n.S + slot_n];
            psr_n = n.psr[slot_n];
        }
        fused_synapse<SK, W32, KMAX>(n, slot, meta, rw, psr, t0, K, pull != 0, tprev0, mprev, acc);
        n.psr[slot] = psr;
        q = qn;
        slot = slot_n;
        meta = meta_n;
#pragma unroll
        for (int w = 0; w < W32; w++) rw[w] = rw_n[w];
        psr = psr_n;
    }
#pragma unroll
    for (int k = 0; k < KMAX; k++)
        if (k < K) acc[k] = wave_sum(acc[k]);

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
            if (k < K) {
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
__global__ __launch_bounds__(BG_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_fused_block(
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

__device__ __forceinline__ int batch_find(const uint32_t *base, int nn, uint32_t w)
{
    int lo = 0, hi = nn;   // base[lo] <= w < base[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (base[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int NK, int SK, int W32, int KMAX>
__global__ __launch_bounds__(BG_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_fused_batch(
    const BatchNet *nets, const uint32_t *wbase, int nn, int K, int Kfull, int launch)
{
    // wave-uniform by construction (blockDim is a multiple of 64); readfirstlane lets the
    // compiler keep the network's descriptor in scalar registers
    const uint32_t w = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (w >= wbase[nn]) return;
    const int i = __builtin_amdgcn_readfirstlane(batch_find(wbase, nn, w));
    const BatchNet &b = nets[i];
    const uint64_t t0 = b.t0 + (uint64_t)launch * (uint64_t)Kfull;
    const int sel = b.sel ^ (launch & 1);
    const int pull = launch > 0 ? 1 : b.pull;
    const uint64_t tprev0 = launch > 0 ? t0 - (uint64_t)Kfull : b.tprev0;
    fused_block_body<NK, SK, W32, KMAX>(b.n, (int)(w - wbase[i]), t0, K, pull, tprev0, sel);
}

template <int NK>
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_batch(const BatchNet *nets, const uint32_t *nbase, int nn,
                                                           uint64_t step, uint32_t *ev_count)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w == 0 && ev_count) *ev_count = 0;
    if (w >= nbase[nn]) return;
    const int i = batch_find(nbase, nn, w);
    const BatchNet &b = nets[i];
    neuron_step_body<NK>(b.n, (int)(w - nbase[i]), b.t0 + step);
}

template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_synapse_batch(const BatchNet *nets, const uint32_t *wbase, int nn,
                                                            uint64_t step)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (w >= wbase[nn]) return;
    const int i = __builtin_amdgcn_readfirstlane(batch_find(wbase, nn, w));
    const BatchNet &b = nets[i];
    synapse_step_body<SK, W32>(b.n, (int)(w - wbase[i]), threadIdx.x & (BG_WAVE - 1), b.t0 + step);
}

// Two-pass STDP step over a batch: the event list holds batch work-item indices.
template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_stream_batch(const BatchNet *nets, const uint32_t *wbase, int nn,
                                                                uint64_t step, uint32_t *ev_count, uint32_t *ev_list)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(
        (uint32_t)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE));
    if (w >= wbase[nn]) return;
    const int i = __builtin_amdgcn_readfirstlane(batch_find(wbase, nn, w));
    const BatchNet &b = nets[i];
    const int lane = threadIdx.x & (BG_WAVE - 1);
    if (stdp_stream_body<SK, W32>(b.n, (int)(w - wbase[i]), lane, b.t0 + step) && lane == 0)
        ev_list[atomicAdd(ev_count, 1u)] = w;
}

template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_stdp_events_batch(const BatchNet *nets, const uint32_t *wbase, int nn,
                                                                uint64_t step, const uint32_t *ev_count,
                                                                const uint32_t *ev_list)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const uint32_t nw = gridDim.x * (BG_BLOCK / BG_WAVE);
    const uint32_t c = *ev_count;
    for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) / BG_WAVE; k < c; k += nw) {
        const uint32_t w = __builtin_amdgcn_readfirstlane(ev_list[k]);
        const int i = __builtin_amdgcn_readfirstlane(batch_find(wbase, nn, w));
        const BatchNet &b = nets[i];
        stdp_event_body<SK, W32>(b.n, (int)(w - wbase[i]), lane, b.t0 + step);
    }
}
