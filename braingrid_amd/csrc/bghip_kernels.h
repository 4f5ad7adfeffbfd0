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
__global__ __launch_bounds__(BG_BLOCK) void k_neuron_step(DevNet n, uint64_t t)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n.N) return;
    NeuronRegs s;
    neuron_load<NK>(n, i, s);
    bool fired = neuron_advance<NK>(n, i, t, n.sp[i], s);
    neuron_store<NK>(n, i, s);
    n.sp[i] = 0.0f;
    if (fired) n.fired_step[i] = t;
}

// One wavefront per destination neuron: delay-ring insert (pull of sources that fired at t),
// ring test/clear, STDP learning, PSR update and decay, and the destination's summation in the
// canonical lane-strided + butterfly order. AllSynapses::advanceSynapses
// (Synapses/AllSynapses.cpp:325-331), AllSpikingSynapses::advanceSynapse (:297-319),
// AllSTDPSynapses::advanceSynapse (Synapses/AllSTDPSynapses.cpp:283-421).
template <int SK, int W32>
__global__ __launch_bounds__(BG_BLOCK) void k_synapse_step(DevNet n, uint64_t t)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int d = (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE);
    if (d >= n.N) return;
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
            if (pre || post) psr = stdp_event<SK>(n, slot, t, pre, post, src, d, delay, psr, decay);
        } else if (pre) {
            psr = change_psr<SK>(n, slot, t, psr, decay);
        }
        psr = psr * decay;
        n.psr[slot] = psr;
        acc += psr;
    }
    acc = wave_sum(acc);
    if (lane == 0) n.sp[d] = n.sp[d] + acc;
}

// ---------------------------------------------------------------------------------------------
// Fused, temporally blocked step kernel for Spiking/DS synapses. A launch advances K consecutive
// steps (K <= every synapse's total_delay, so no spike fired inside the block is delivered inside
// it). One wavefront owns one destination neuron and all its incoming synapses:
//   1. pull the spikes its sources fired during the previous block into the delay rings;
//   2. run every synapse through the K steps in registers (ring test, PSR update, decay),
//      accumulating the K per-step summation points in the canonical order;
//   3. lane 0 advances the neuron K steps with those inputs and publishes a K-bit fire mask.
// Each synapse's meta/ring/psr are read and written once per K steps instead of once per step.
template <int NK, int SK, int W32, int KMAX>
__global__ __launch_bounds__(BG_BLOCK) void k_fused_block(DevNet n, uint64_t t0, int K, int pull,
                                                          uint64_t tprev0, int sel_in)
{
    const int lane = threadIdx.x & (BG_WAVE - 1);
    const int d = (int)((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / BG_WAVE);
    if (d >= n.N) return;
    const uint32_t cnt = n.syn_count[d], hi = n.slot_hi[d];
    const bool dense = (cnt == hi);
    const uint32_t Rm = (uint32_t)n.R - 1;
    const uint32_t *mprev = n.mask[sel_in];
    float acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) acc[k] = 0.0f;

    for (uint32_t q = lane; q < cnt; q += BG_WAVE) {
        const uint64_t slot = dense ? (uint64_t)d * n.M + q : (uint64_t)n.in_map[n.in_begin[d] + q];
        const uint32_t meta = n.meta[slot];
        const int type = (meta >> BG_META_TYPE_SHIFT) & 3;
        uint32_t rw[W32];
#pragma unroll
        for (int w = 0; w < W32; w++) rw[w] = n.ring[(uint64_t)w * n.S + slot];
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
        float psr = n.psr[slot];
        const float decay = n.decay[type];
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k < K) {
                const uint64_t t = t0 + (uint64_t)k;
                const uint32_t b = (uint32_t)(t & Rm);
                const uint32_t m = 1u << (b & 31);
                bool hit = false;
#pragma unroll
                for (int w = 0; w < W32; w++)
                    if ((int)(b >> 5) == w && (rw[w] & m)) {
                        rw[w] &= ~m;
                        hit = true;
                    }
                if (hit) {
                    changed = true;
                    psr = change_psr<SK>(n, slot, t, psr, decay);
                }
                psr = psr * decay;
                acc[k] += psr;
            }
        }
        n.psr[slot] = psr;
        if (changed) {
#pragma unroll
            for (int w = 0; w < W32; w++) n.ring[(uint64_t)w * n.S + slot] = rw[w];
        }
    }
#pragma unroll
    for (int k = 0; k < KMAX; k++)
        if (k < K) acc[k] = wave_sum(acc[k]);

    if (lane == 0) {
        NeuronRegs s;
        neuron_load<NK>(n, d, s);
        float sp = n.sp[d];
        uint32_t mask = 0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k < K) {
                if (neuron_advance<NK>(n, d, t0 + (uint64_t)k, sp, s)) mask |= 1u << k;
                sp = 0.0f + acc[k];
            }
        }
        neuron_store<NK>(n, d, s);
        n.sp[d] = sp;
        n.mask[sel_in ^ 1][d] = mask;
        if (mask) n.fired_step[d] = t0 + (uint64_t)(31 - __clz(mask));
    }
}

