// Host launchers of the templated kernels; each family is compiled in its own translation unit so
// the library builds in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include "bghip_device.h"

void launch_fused_lif(const DevNet &dn, hipStream_t s, int sk, int kmax, uint64_t t0, int K, int pull,
                      uint64_t tprev0, int sel);
void launch_fused_izh(const DevNet &dn, hipStream_t s, int sk, int kmax, uint64_t t0, int K, int pull,
                      uint64_t tprev0, int sel);
// implicit-ring fused blocks (Spiking/DS, 32-bit rings): sel = parity of the block's fire words
void launch_fused_impl_lif(const DevNet &dn, hipStream_t s, int sk, int kmax, uint64_t t0, int K, int sel);
void launch_fused_impl_izh(const DevNet &dn, hipStream_t s, int sk, int kmax, uint64_t t0, int K, int sel);
void launch_neuron_step(const DevNet &dn, hipStream_t s, int nk, uint64_t t);
// steps per neuron-only launch
#define BG_QUIET_STEPS 10000
void launch_neuron_only(const DevNet &dn, hipStream_t s, int nk, uint64_t t0, int steps);
void launch_synapse_step(const DevNet &dn, hipStream_t s, int sk, uint64_t t);
// one per-step STDP step at *dn.tbase + off (four launches)
void launch_stdp_step(const DevNet &dn, const DevNet *d_self, hipStream_t s, int nk, int sk, uint32_t off);
void launch_step_set(hipStream_t s, uint64_t *tbase, uint64_t v);
void launch_step_bump(hipStream_t s, uint64_t *tbase, uint32_t by);
// steps per captured graph of the per-step STDP path
#define BG_GRAPH_STEPS 256

struct BatchNet;
// Work items of a batch (destination waves or neuron threads) are numbered across the member
// networks: member i owns items [base[i], base[i+1]) and nof[w] is the member of item w, so a
// kernel finds its network with one load instead of a search.
struct BatchMap {
    const BatchNet *nets;
    const uint32_t *base;   // nn + 1 prefix
    const uint16_t *nof;    // total entries
    uint32_t total;
    int nn;
    const uint2 *chunks;    // per-step STDP streaming pass: {member | shard << 16, first destination} runs
    uint32_t nchunks;
    const uint2 *nwaves;    // thread-per-neuron kernels: {member, first neuron} of each wavefront
    uint32_t nnwaves;
    const uint2 *quads;     // implicit-ring fused kernels: {member, first position} of each workgroup
    uint32_t nquads;        // (four destinations of one member per workgroup)
};
// Launch shape of a batch group's kernels. Every work item computes the same values whatever
// block it lands in and however many destinations share its wavefront, so results do not depend
// on the shape; only the time does.
struct BatchShape {
    int blk;          // threads per block: neuron, synapse, fused and neuron-only kernels (64/128/256)
    int dpw;          // per-step STDP streaming pass: destinations per wavefront (1, 2 or 4)
    int blk_stream;   // threads per block of the streaming pass (64/128/256)
    int g_syn;        // blocks per event-list shard of the learning pass
    int g_sum;        // blocks per event-list shard of the event summation pass
};
void launch_fused_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int w32, int kmax,
                        int K, int Kfull, int launch);
void launch_fused_impl_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int kmax,
                             int K, int Kfull, int launch);
void launch_neuron_only_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, uint64_t step,
                              int steps);
void launch_step_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int w32,
                       uint64_t step, const EvLists &ev);
void launch_stdp_step_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int w32,
                            uint64_t step, const EvLists &ev);
void launch_batch_bump(hipStream_t s, BatchNet *nets, int nn, uint32_t by);
