// Per-step neuron and synapse kernels (see k_neuron_step / k_synapse_step / k_stdp_* in
// bghip_kernels.h).
#include "bghip_kernels.h"
#include "bghip_launch.h"
#include <algorithm>

void launch_neuron_step(const DevNet &dn, hipStream_t s, int nk, uint64_t t)
{
    unsigned nb = (unsigned)((dn.N + BG_BLOCK - 1) / BG_BLOCK);
    if (nk == NK_LIF) hipLaunchKernelGGL(k_neuron_step<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, dn, t);
    else hipLaunchKernelGGL(k_neuron_step<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, dn, t);
}

void launch_neuron_only(const DevNet &dn, hipStream_t s, int nk, uint64_t t0, int steps)
{
    unsigned nb = (unsigned)((dn.N + BG_BLOCK - 1) / BG_BLOCK);
    if (nk == NK_LIF) hipLaunchKernelGGL(k_neuron_only<NK_LIF>, dim3(nb), dim3(BG_BLOCK), 0, s, dn, t0, steps);
    else hipLaunchKernelGGL(k_neuron_only<NK_IZH>, dim3(nb), dim3(BG_BLOCK), 0, s, dn, t0, steps);
}

template <int SK, int W32>
static void go(const DevNet &dn, hipStream_t s, uint64_t t)
{
    unsigned blocks = (unsigned)(((uint64_t)dn.N * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL((k_synapse_step<SK, W32>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t);
}
template <int SK>
static void go_w(const DevNet &dn, hipStream_t s, uint64_t t)
{
    switch (dn.W32) {
    case 1: go<SK, 1>(dn, s, t); break;
    case 2: go<SK, 2>(dn, s, t); break;
    case 4: go<SK, 4>(dn, s, t); break;
    default: go<SK, 8>(dn, s, t); break;
    }
}

void launch_synapse_step(const DevNet &dn, hipStream_t s, int sk, uint64_t t)
{
    switch (sk) {
    case SK_SPIKING: go_w<SK_SPIKING>(dn, s, t); break;
    case SK_DS: go_w<SK_DS>(dn, s, t); break;
    case SK_STDP: go_w<SK_STDP>(dn, s, t); break;
    default: go_w<SK_DYNSTDP>(dn, s, t); break;
    }
}

// Event passes: G blocks per event-list shard, more for larger networks (each shard holds about
// N / BG_EV_SHARDS destinations).
static unsigned event_grid(int N, unsigned gmax)
{
    const unsigned g = (unsigned)(((uint64_t)N + 16 * BG_EV_SHARDS - 1) / (16 * BG_EV_SHARDS));
    return BG_EV_SHARDS * std::max(1u, std::min(gmax, g));
}

template <int NK, int SK, int W32>
static void stdp_go(const DevNet &dn, const DevNet *d_self, hipStream_t s, uint32_t off)
{
    const unsigned nb = (unsigned)((dn.N + BG_BLOCK - 1) / BG_BLOCK);
    const uint64_t waves = ((uint64_t)dn.N + BG_STREAM_DESTS - 1) / BG_STREAM_DESTS;
    const unsigned wb = (unsigned)((waves * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    hipLaunchKernelGGL(k_stdp_neurons<NK>, dim3(nb), dim3(BG_BLOCK), 0, s, dn, off);
    hipLaunchKernelGGL((k_stdp_stream<SK, W32>), dim3(wb), dim3(BG_BLOCK), 0, s, dn, off);
    hipLaunchKernelGGL(k_stdp_event_syn<SK>, dim3(event_grid(dn.N, 4)), dim3(BG_BLOCK), 0, s, d_self, off);
    hipLaunchKernelGGL(k_stdp_event_sum, dim3(event_grid(dn.N, 8)), dim3(BG_BLOCK), 0, s, dn);
}
template <int NK, int SK>
static void stdp_go_w(const DevNet &dn, const DevNet *d_self, hipStream_t s, uint32_t t)
{
    switch (dn.W32) {
    case 1: stdp_go<NK, SK, 1>(dn, d_self, s, t); break;
    case 2: stdp_go<NK, SK, 2>(dn, d_self, s, t); break;
    case 4: stdp_go<NK, SK, 4>(dn, d_self, s, t); break;
    default: stdp_go<NK, SK, 8>(dn, d_self, s, t); break;
    }
}

void launch_stdp_step(const DevNet &dn, const DevNet *d_self, hipStream_t s, int nk, int sk, uint32_t t)
{
    if (nk == NK_LIF) {
        if (sk == SK_STDP) stdp_go_w<NK_LIF, SK_STDP>(dn, d_self, s, t);
        else stdp_go_w<NK_LIF, SK_DYNSTDP>(dn, d_self, s, t);
    } else {
        if (sk == SK_STDP) stdp_go_w<NK_IZH, SK_STDP>(dn, d_self, s, t);
        else stdp_go_w<NK_IZH, SK_DYNSTDP>(dn, d_self, s, t);
    }
}

void launch_step_set(hipStream_t s, uint64_t *tbase, uint64_t v)
{
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, s, tbase, v);
}
void launch_step_bump(hipStream_t s, uint64_t *tbase, uint32_t by)
{
    hipLaunchKernelGGL(k_step_bump, dim3(1), dim3(64), 0, s, tbase, by);
}
