// Multi-network launchers (k_fused_batch / k_neuron_batch / k_synapse_batch / k_stdp_*_batch in
// bghip_kernels.h). Every kernel takes the batch's BatchMap by value.
#include "bghip_kernels.h"
#include "bghip_launch.h"
#include <algorithm>

static unsigned wave_blocks(uint32_t waves) { return (unsigned)(((uint64_t)waves * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK); }
static unsigned thread_blocks(uint32_t threads) { return (unsigned)(((uint64_t)threads + BG_BLOCK - 1) / BG_BLOCK); }

template <int NK, int SK, int W32>
static void fused(const BatchMap &bm, hipStream_t s, int kmax, int K, int Kfull, int launch)
{
    if (W32 == 1 && K == BG_K_COMMON)
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, BG_K_COMMON>), dim3(wave_blocks(bm.total)), dim3(BG_BLOCK), 0, s,
                           bm, K, Kfull, launch);
    else if (kmax <= 16)
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, 16>), dim3(wave_blocks(bm.total)), dim3(BG_BLOCK), 0, s, bm, K,
                           Kfull, launch);
    else
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, 32>), dim3(wave_blocks(bm.total)), dim3(BG_BLOCK), 0, s, bm, K,
                           Kfull, launch);
}

template <int NK, int SK>
static void fused_w(const BatchMap &bm, hipStream_t s, int w32, int kmax, int K, int Kfull, int launch)
{
    switch (w32) {
    case 1: fused<NK, SK, 1>(bm, s, kmax, K, Kfull, launch); break;
    case 2: fused<NK, SK, 2>(bm, s, kmax, K, Kfull, launch); break;
    case 4: fused<NK, SK, 4>(bm, s, kmax, K, Kfull, launch); break;
    default: fused<NK, SK, 8>(bm, s, kmax, K, Kfull, launch); break;
    }
}

void launch_fused_batch(const BatchMap &bm, hipStream_t s, int nk, int sk, int w32, int kmax, int K, int Kfull,
                        int launch)
{
    if (nk == NK_LIF) {
        if (sk == SK_DS) fused_w<NK_LIF, SK_DS>(bm, s, w32, kmax, K, Kfull, launch);
        else fused_w<NK_LIF, SK_SPIKING>(bm, s, w32, kmax, K, Kfull, launch);
    } else {
        if (sk == SK_DS) fused_w<NK_IZH, SK_DS>(bm, s, w32, kmax, K, Kfull, launch);
        else fused_w<NK_IZH, SK_SPIKING>(bm, s, w32, kmax, K, Kfull, launch);
    }
}

template <int SK, int W32>
static void syn(const BatchMap &bm, hipStream_t s, uint64_t step)
{
    hipLaunchKernelGGL((k_synapse_batch<SK, W32>), dim3(wave_blocks(bm.total)), dim3(BG_BLOCK), 0, s, bm, step);
}
template <int SK>
static void syn_w(const BatchMap &bm, hipStream_t s, int w32, uint64_t step)
{
    switch (w32) {
    case 1: syn<SK, 1>(bm, s, step); break;
    case 2: syn<SK, 2>(bm, s, step); break;
    case 4: syn<SK, 4>(bm, s, step); break;
    default: syn<SK, 8>(bm, s, step); break;
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
static void stdp(const BatchMap &bm, hipStream_t s, uint64_t step, const EvLists &ev)
{
    hipLaunchKernelGGL((k_stdp_stream_batch<SK, W32>), dim3(wave_blocks(bm.nchunks)), dim3(BG_BLOCK), 0, s, bm, step,
                       ev);
    hipLaunchKernelGGL(k_stdp_event_syn_batch<SK>, dim3(event_grid(bm.total, 4)), dim3(BG_BLOCK), 0, s, bm, step, ev);
    hipLaunchKernelGGL(k_stdp_event_sum_batch, dim3(event_grid(bm.total, 8)), dim3(BG_BLOCK), 0, s, bm, ev);
}
template <int SK>
static void stdp_w(const BatchMap &bm, hipStream_t s, int w32, uint64_t step, const EvLists &ev)
{
    switch (w32) {
    case 1: stdp<SK, 1>(bm, s, step, ev); break;
    case 2: stdp<SK, 2>(bm, s, step, ev); break;
    case 4: stdp<SK, 4>(bm, s, step, ev); break;
    default: stdp<SK, 8>(bm, s, step, ev); break;
    }
}

void launch_neuron_only_batch(const BatchMap &bm, hipStream_t s, int nk, uint64_t step, int steps)
{
    const unsigned nb = thread_blocks(bm.total);
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_only_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, bm, step, steps);
    else
        hipLaunchKernelGGL(k_neuron_only_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, bm, step, steps);
}

void launch_step_batch(const BatchMap &bm, hipStream_t s, int nk, int sk, int w32, uint64_t step, const EvLists &ev)
{
    const unsigned nb = thread_blocks(bm.total);
    if (sk == SK_STDP || sk == SK_DYNSTDP) {
        if (nk == NK_LIF)
            hipLaunchKernelGGL(k_stdp_neurons_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, bm, step, ev);
        else
            hipLaunchKernelGGL(k_stdp_neurons_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, bm, step, ev);
        if (sk == SK_STDP) stdp_w<SK_STDP>(bm, s, w32, step, ev);
        else stdp_w<SK_DYNSTDP>(bm, s, w32, step, ev);
        return;
    }
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, bm, step);
    else
        hipLaunchKernelGGL(k_neuron_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, bm, step);
    if (sk == SK_SPIKING) syn_w<SK_SPIKING>(bm, s, w32, step);
    else syn_w<SK_DS>(bm, s, w32, step);
}
