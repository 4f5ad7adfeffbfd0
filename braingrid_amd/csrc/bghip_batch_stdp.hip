// Multi-network launchers of the four per-step STDP passes (k_stdp_neurons_batch,
// k_stdp_stream_batch, k_stdp_event_syn_batch, k_stdp_event_sum_batch in bghip_kernels.h) and of
// the step-base bump that ends a captured run of batch steps. Block sizes, destinations per
// wavefront and event-pass grids come from the group's BatchShape.
#include "bghip_kernels.h"
#include "bghip_launch.h"
#include <algorithm>

static unsigned wave_blocks(uint32_t waves, int blk)
{
    return (unsigned)(((uint64_t)waves * BG_WAVE + (uint64_t)blk - 1) / (uint64_t)blk);
}

template <int SK, int W32, int D>
static void stream(const BatchMap &bm, hipStream_t s, int blk, uint64_t step, const EvLists &ev)
{
    hipLaunchKernelGGL((k_stdp_stream_batch<SK, W32, D>), dim3(wave_blocks(bm.nchunks, blk)), dim3(blk), 0, s, bm,
                       step, ev);
}

template <int SK, int W32>
static void stdp(const BatchMap &bm, hipStream_t s, const BatchShape &sh, uint64_t step, const EvLists &ev)
{
    switch (sh.dpw) {
    case 1: stream<SK, W32, 1>(bm, s, sh.blk_stream, step, ev); break;
    case 2: stream<SK, W32, 2>(bm, s, sh.blk_stream, step, ev); break;
    default: stream<SK, W32, BG_STREAM_DESTS>(bm, s, sh.blk_stream, step, ev); break;
    }
    hipLaunchKernelGGL(k_stdp_event_syn_batch<SK>, dim3(BG_EV_SHARDS_BATCH * sh.g_syn), dim3(BG_BLOCK), 0, s, bm,
                       step, ev);
    hipLaunchKernelGGL(k_stdp_event_sum_batch, dim3(BG_EV_SHARDS_BATCH * sh.g_sum), dim3(BG_BLOCK), 0, s, bm, ev);
}
template <int SK>
static void stdp_w(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int w32, uint64_t step, const EvLists &ev)
{
    switch (w32) {
    case 1: stdp<SK, 1>(bm, s, sh, step, ev); break;
    case 2: stdp<SK, 2>(bm, s, sh, step, ev); break;
    case 4: stdp<SK, 4>(bm, s, sh, step, ev); break;
    default: stdp<SK, 8>(bm, s, sh, step, ev); break;
    }
}

void launch_stdp_step_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int w32,
                            uint64_t step, const EvLists &ev)
{
    // the neuron pass clears the 2 * BG_EV_SHARDS_BATCH event counters, one per thread
    const unsigned nb =
        std::max(wave_blocks(bm.nnwaves, sh.blk), (unsigned)((2 * BG_EV_SHARDS_BATCH + sh.blk - 1) / sh.blk));
    if (nk == NK_LIF)
        hipLaunchKernelGGL(k_stdp_neurons_batch<NK_LIF>, dim3(nb), dim3(sh.blk), 0, s, bm, step, ev);
    else
        hipLaunchKernelGGL(k_stdp_neurons_batch<NK_IZH>, dim3(nb), dim3(sh.blk), 0, s, bm, step, ev);
    if (sk == SK_STDP) stdp_w<SK_STDP>(bm, s, sh, w32, step, ev);
    else stdp_w<SK_DYNSTDP>(bm, s, sh, w32, step, ev);
}

void launch_batch_bump(hipStream_t s, BatchNet *nets, int nn, uint32_t by)
{
    hipLaunchKernelGGL(k_batch_bump, dim3((unsigned)((nn + 63) / 64)), dim3(64), 0, s, nets, nn, by);
}
