// Multi-network launchers (k_fused_batch / k_neuron_batch / k_synapse_batch in bghip_kernels.h).
#include "bghip_kernels.h"
#include "bghip_launch.h"
#include <algorithm>

template <int NK, int SK, int W32>
static void fused(const BatchNet *nets, const uint32_t *wbase, int nn, uint32_t waves, hipStream_t s, int kmax,
                  int K, int Kfull, int launch)
{
    unsigned blocks = (unsigned)(((uint64_t)waves * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    if (kmax <= 16)
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, 16>), dim3(blocks), dim3(BG_BLOCK), 0, s, nets, wbase, nn, K,
                           Kfull, launch);
    else
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, 32>), dim3(blocks), dim3(BG_BLOCK), 0, s, nets, wbase, nn, K,
                           Kfull, launch);
}

template <int NK, int SK>
static void fused_w(const BatchNet *nets, const uint32_t *wbase, int nn, uint32_t waves, hipStream_t s, int w32,
                    int kmax, int K, int Kfull, int launch)
{
    switch (w32) {
    case 1: fused<NK, SK, 1>(nets, wbase, nn, waves, s, kmax, K, Kfull, launch); break;
    case 2: fused<NK, SK, 2>(nets, wbase, nn, waves, s, kmax, K, Kfull, launch); break;
    case 4: fused<NK, SK, 4>(nets, wbase, nn, waves, s, kmax, K, Kfull, launch); break;
    default: fused<NK, SK, 8>(nets, wbase, nn, waves, s, kmax, K, Kfull, launch); break;
    }
}

void launch_fused_batch(const BatchNet *nets, const uint32_t *wbase, int nn, uint32_t waves, hipStream_t s, int nk,
                        int sk, int w32, int kmax, int K, int Kfull, int launch)
{
    if (nk == NK_LIF) {
        if (sk == SK_DS) fused_w<NK_LIF, SK_DS>(nets, wbase, nn, waves, s, w32, kmax, K, Kfull, launch);
        else fused_w<NK_LIF, SK_SPIKING>(nets, wbase, nn, waves, s, w32, kmax, K, Kfull, launch);
    } else {
        if (sk == SK_DS) fused_w<NK_IZH, SK_DS>(nets, wbase, nn, waves, s, w32, kmax, K, Kfull, launch);
        else fused_w<NK_IZH, SK_SPIKING>(nets, wbase, nn, waves, s, w32, kmax, K, Kfull, launch);
    }
}

template <int SK, int W32>
static void syn(const BatchNet *nets, const uint32_t *base, int nn, uint32_t waves, hipStream_t s, uint64_t step)
{
    unsigned blocks = (unsigned)(((uint64_t)waves * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL((k_synapse_batch<SK, W32>), dim3(blocks), dim3(BG_BLOCK), 0, s, nets, base, nn, step);
}
template <int SK>
static void syn_w(const BatchNet *nets, const uint32_t *base, int nn, uint32_t waves, hipStream_t s, int w32,
                  uint64_t step)
{
    switch (w32) {
    case 1: syn<SK, 1>(nets, base, nn, waves, s, step); break;
    case 2: syn<SK, 2>(nets, base, nn, waves, s, step); break;
    case 4: syn<SK, 4>(nets, base, nn, waves, s, step); break;
    default: syn<SK, 8>(nets, base, nn, waves, s, step); break;
    }
}

// Event passes: G blocks per event-list shard, more for larger batches (each shard holds about
// total / BG_EV_SHARDS destinations).
static unsigned event_grid(uint32_t total, unsigned gmax)
{
    const unsigned g = (unsigned)(((uint64_t)total + 16 * BG_EV_SHARDS - 1) / (16 * BG_EV_SHARDS));
    return BG_EV_SHARDS * std::max(1u, std::min(gmax, g));
}

template <int SK, int W32>
static void stdp(const BatchNet *nets, const uint32_t *base, int nn, uint32_t waves, hipStream_t s, uint64_t step,
                 const EvLists &ev)
{
    const uint64_t sw = ((uint64_t)waves + BG_STREAM_DESTS - 1) / BG_STREAM_DESTS;
    const unsigned blocks = (unsigned)((sw * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL((k_stdp_stream_batch<SK, W32>), dim3(blocks), dim3(BG_BLOCK), 0, s, nets, base, nn, step, ev);
    hipLaunchKernelGGL(k_stdp_event_syn_batch<SK>, dim3(event_grid(waves, 4)), dim3(BG_BLOCK), 0, s, nets, base, nn,
                       step, ev);
    hipLaunchKernelGGL(k_stdp_event_sum_batch, dim3(event_grid(waves, 8)), dim3(BG_BLOCK), 0, s, nets, base, nn, ev);
}
template <int SK>
static void stdp_w(const BatchNet *nets, const uint32_t *base, int nn, uint32_t waves, hipStream_t s, int w32,
                   uint64_t step, const EvLists &ev)
{
    switch (w32) {
    case 1: stdp<SK, 1>(nets, base, nn, waves, s, step, ev); break;
    case 2: stdp<SK, 2>(nets, base, nn, waves, s, step, ev); break;
    case 4: stdp<SK, 4>(nets, base, nn, waves, s, step, ev); break;
    default: stdp<SK, 8>(nets, base, nn, waves, s, step, ev); break;
    }
}

void launch_neuron_only_batch(const BatchNet *nets, const uint32_t *base, int nn, uint32_t total, hipStream_t s,
                              int nk, uint64_t step, int steps)
{
    const unsigned nb = (unsigned)((total + BG_BLOCK - 1) / BG_BLOCK);
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_only_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step, steps);
    else
        hipLaunchKernelGGL(k_neuron_only_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step, steps);
}

void launch_step_batch(const BatchNet *nets, const uint32_t *base, int nn, uint32_t total, hipStream_t s, int nk,
                       int sk, int w32, uint64_t step, const EvLists &ev)
{
    const unsigned nb = (unsigned)((total + BG_BLOCK - 1) / BG_BLOCK);
    if (sk == SK_STDP || sk == SK_DYNSTDP) {
        if (nk == NK_LIF)
            hipLaunchKernelGGL(k_stdp_neurons_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step, ev);
        else
            hipLaunchKernelGGL(k_stdp_neurons_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step, ev);
        if (sk == SK_STDP) stdp_w<SK_STDP>(nets, base, nn, total, s, w32, step, ev);
        else stdp_w<SK_DYNSTDP>(nets, base, nn, total, s, w32, step, ev);
        return;
    }
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step);
    else
        hipLaunchKernelGGL(k_neuron_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step);
    if (sk == SK_SPIKING) syn_w<SK_SPIKING>(nets, base, nn, total, s, w32, step);
    else syn_w<SK_DS>(nets, base, nn, total, s, w32, step);
}
