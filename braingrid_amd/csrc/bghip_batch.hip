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

template <int SK, int W32>
static void stdp(const BatchNet *nets, const uint32_t *base, int nn, uint32_t waves, hipStream_t s, uint64_t step,
                 uint32_t *ev)
{
    unsigned blocks = (unsigned)(((uint64_t)waves * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL((k_stdp_stream_batch<SK, W32>), dim3(blocks), dim3(BG_BLOCK), 0, s, nets, base, nn, step, ev,
                       ev + 1);
    // event pass: up to 4096 waves looping over the event list
    unsigned eb = std::min(blocks, 1024u);
    hipLaunchKernelGGL((k_stdp_events_batch<SK, W32>), dim3(eb), dim3(BG_BLOCK), 0, s, nets, base, nn, step,
                       (const uint32_t *)ev, (const uint32_t *)(ev + 1));
}
template <int SK>
static void stdp_w(const BatchNet *nets, const uint32_t *base, int nn, uint32_t waves, hipStream_t s, int w32,
                   uint64_t step, uint32_t *ev)
{
    switch (w32) {
    case 1: stdp<SK, 1>(nets, base, nn, waves, s, step, ev); break;
    case 2: stdp<SK, 2>(nets, base, nn, waves, s, step, ev); break;
    case 4: stdp<SK, 4>(nets, base, nn, waves, s, step, ev); break;
    default: stdp<SK, 8>(nets, base, nn, waves, s, step, ev); break;
    }
}

void launch_step_batch(const BatchNet *nets, const uint32_t *base, int nn, uint32_t total, hipStream_t s, int nk,
                       int sk, int w32, uint64_t step, uint32_t *ev)
{
    const bool two_pass = (sk == SK_STDP || sk == SK_DYNSTDP) && ev;
    uint32_t *evc = two_pass ? ev : nullptr;
    unsigned nb = (unsigned)((total + BG_BLOCK - 1) / BG_BLOCK);
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_batch<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step, evc);
    else
        hipLaunchKernelGGL(k_neuron_batch<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, nets, base, nn, step, evc);
    switch (sk) {
    case SK_SPIKING: syn_w<SK_SPIKING>(nets, base, nn, total, s, w32, step); break;
    case SK_DS: syn_w<SK_DS>(nets, base, nn, total, s, w32, step); break;
    case SK_STDP:
        if (two_pass) stdp_w<SK_STDP>(nets, base, nn, total, s, w32, step, ev);
        else syn_w<SK_STDP>(nets, base, nn, total, s, w32, step);
        break;
    default:
        if (two_pass) stdp_w<SK_DYNSTDP>(nets, base, nn, total, s, w32, step, ev);
        else syn_w<SK_DYNSTDP>(nets, base, nn, total, s, w32, step);
        break;
    }
}
