// Multi-network launchers of the fused, per-step Spiking/DS and neuron-only kernels
// (k_fused_batch / k_neuron_batch / k_synapse_batch / k_neuron_only_batch in bghip_kernels.h).
// Every kernel takes the batch's BatchMap by value; block sizes come from the group's BatchShape.
#include "bghip_kernels.h"
#include "bghip_launch.h"
#include <algorithm>

static unsigned wave_blocks(uint32_t waves, int blk)
{
    return (unsigned)(((uint64_t)waves * BG_WAVE + (uint64_t)blk - 1) / (uint64_t)blk);
}

template <int NK, int SK, int W32>
static void fused(const BatchMap &bm, hipStream_t s, int blk, int kmax, int K, int Kfull, int launch)
{
    const dim3 grid(wave_blocks(bm.total, blk)), block(blk);
    if (W32 == 1 && K == BG_K_COMMON)
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, BG_K_COMMON>), grid, block, 0, s, bm, K, Kfull, launch);
    else if (kmax <= 16)
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, 16>), grid, block, 0, s, bm, K, Kfull, launch);
    else
        hipLaunchKernelGGL((k_fused_batch<NK, SK, W32, 32>), grid, block, 0, s, bm, K, Kfull, launch);
}

template <int NK, int SK>
static void fused_w(const BatchMap &bm, hipStream_t s, int blk, int w32, int kmax, int K, int Kfull, int launch)
{
    switch (w32) {
    case 1: fused<NK, SK, 1>(bm, s, blk, kmax, K, Kfull, launch); break;
    case 2: fused<NK, SK, 2>(bm, s, blk, kmax, K, Kfull, launch); break;
    case 4: fused<NK, SK, 4>(bm, s, blk, kmax, K, Kfull, launch); break;
    default: fused<NK, SK, 8>(bm, s, blk, kmax, K, Kfull, launch); break;
    }
}

void launch_fused_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int w32, int kmax,
                        int K, int Kfull, int launch)
{
    if (nk == NK_LIF) {
        if (sk == SK_DS) fused_w<NK_LIF, SK_DS>(bm, s, sh.blk, w32, kmax, K, Kfull, launch);
        else fused_w<NK_LIF, SK_SPIKING>(bm, s, sh.blk, w32, kmax, K, Kfull, launch);
    } else {
        if (sk == SK_DS) fused_w<NK_IZH, SK_DS>(bm, s, sh.blk, w32, kmax, K, Kfull, launch);
        else fused_w<NK_IZH, SK_SPIKING>(bm, s, sh.blk, w32, kmax, K, Kfull, launch);
    }
}

template <int SK, int W32>
static void syn(const BatchMap &bm, hipStream_t s, int blk, uint64_t step)
{
    hipLaunchKernelGGL((k_synapse_batch<SK, W32>), dim3(wave_blocks(bm.total, blk)), dim3(blk), 0, s, bm, step);
}
template <int SK>
static void syn_w(const BatchMap &bm, hipStream_t s, int blk, int w32, uint64_t step)
{
    switch (w32) {
    case 1: syn<SK, 1>(bm, s, blk, step); break;
    case 2: syn<SK, 2>(bm, s, blk, step); break;
    case 4: syn<SK, 4>(bm, s, blk, step); break;
    default: syn<SK, 8>(bm, s, blk, step); break;
    }
}

void launch_neuron_only_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, uint64_t step,
                              int steps)
{
    const dim3 grid(wave_blocks(bm.nnwaves, sh.blk)), block(sh.blk);
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_only_batch<NK_LIF>, grid, block, 0, s, bm, step, steps);
    else
        hipLaunchKernelGGL(k_neuron_only_batch<NK_IZH>, grid, block, 0, s, bm, step, steps);
}

void launch_step_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int w32,
                       uint64_t step, const EvLists &ev)
{
    if (sk == SK_STDP || sk == SK_DYNSTDP) {
        launch_stdp_step_batch(bm, s, sh, nk, sk, w32, step, ev);
        return;
    }
    const dim3 grid(wave_blocks(bm.nnwaves, sh.blk)), block(sh.blk);
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_neuron_batch<NK_LIF>, grid, block, 0, s, bm, step);
    else
        hipLaunchKernelGGL(k_neuron_batch<NK_IZH>, grid, block, 0, s, bm, step);
    if (sk == SK_SPIKING) syn_w<SK_SPIKING>(bm, s, sh.blk, w32, step);
    else syn_w<SK_DS>(bm, s, sh.blk, w32, step);
}
