/*
 * bghip.h -- C ABI of the MI355X spiking-network engine (libbghip.so).
 *
 * The library replaces BrainGrid's GPUSpikingModel (Core/GPUSpikingModel.{h,cu}) behind the
 * reference's model interfaces. A host-side model (the C++ HipSpikingModel in
 * braingrid_amd/host/, or any other caller) keeps BrainGrid's host SoA arrays, uploads them once,
 * advances on the GPU, runs neurite growth on the GPU and downloads state for recorders.
 *
 * Conventions
 *   - every function returns BGHIP_OK (0) or a negative bghip_status;
 *   - handles are opaque; the library owns all device memory, the caller owns host buffers and
 *     only lends them for the duration of a call;
 *   - a bghip_net is externally synchronised (one host thread at a time); distinct nets may be
 *     driven from distinct threads; every net has its own HIP stream;
 *   - advance / update_connections are stream-ordered and return before the GPU finishes; device
 *     errors (delay-ring collision, spike-history overflow, synapse-slot overflow) are reported by
 *     the next bghip_sync or download as BGHIP_ESTATE;
 *   - arrays use the reference's layouts: neuron arrays of length N, synapse "slot" arrays of
 *     length N * max_synapses_per_neuron in destination-major order (iSyn = M * dest + k,
 *     Synapses/AllSynapses.cpp:366-371), spike history N * max_spikes
 *     (Neurons/AllSpikingNeurons.h:185-204), index maps as in Core/SynapseIndexMap.h:38-104.
 *   - noise is the deterministic Philox/Box-Muller stream of include/bg_detmath.h.
 */
#ifndef BGHIP_H
#define BGHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bghip_status {
    BGHIP_OK = 0,
    BGHIP_EINVAL = -1,       /* bad argument or size */
    BGHIP_ENOMEM = -2,       /* device allocation failed */
    BGHIP_EDEVICE = -3,      /* HIP runtime error */
    BGHIP_ESTATE = -4,       /* model state error raised on the device (see bghip_net_error_bits) */
    BGHIP_EUNSUPPORTED = -5  /* configuration outside what the kernels implement */
} bghip_status;

/* device error bits (bghip_net_error_bits) */
#define BGHIP_ERR_RING_COLLISION 1u   /* assert(!(delay_queue & (0x1 << idx))), Synapses/AllSpikingSynapses.cpp:277 */
#define BGHIP_ERR_SPIKE_OVERFLOW 2u   /* assert(spikeCount < max_spikes), Neurons/AllSpikingNeurons.cpp:112 */
#define BGHIP_ERR_SLOT_OVERFLOW 4u    /* no free slot for a new synapse, Synapses/AllSynapses.cpp:359-362 */

enum { BGHIP_NEURON_LIF = 0, BGHIP_NEURON_IZH = 1 };
enum { BGHIP_SYN_SPIKING = 0, BGHIP_SYN_DS = 1, BGHIP_SYN_STDP = 2, BGHIP_SYN_DYNSTDP = 3 };

typedef struct bghip_ctx bghip_ctx;
typedef struct bghip_net bghip_net;

/* Per-synapse-type creation table, index II=0, IE=1, EI=2, EE=3 (Global.h:100). The host fills it
 * from createSynapse (Synapses/AllSpikingSynapses.cpp:191-233, AllDSSynapses.cpp:180-218) so that
 * synapses the GPU creates during growth carry exactly the values the reference would give them.
 * decay must be the host's exp(-dt/tau) (AllSpikingSynapses.cpp:344-354). */
typedef struct bghip_synapse_types {
    float tau[4];
    float decay[4];
    int32_t total_delay[4];
    float U[4], D[4], F[4];
    float u0, r0;                 /* DEFAULT_U, 1.0 (AllDSSynapses.cpp:160-167) */
    /* STDP creation values (AllSTDPSynapses.cpp:252-273) */
    float Apos, Aneg, STDPgap, tauspost, tauspre, taupos, tauneg, Wex, mupos, muneg;
    int32_t useFroemkeDanSTDP;
} bghip_synapse_types;

typedef struct bghip_net_desc {
    int32_t num_neurons;              /* N */
    int32_t max_synapses_per_neuron;  /* M (SimConfig maxSynapsesPerNeuron) */
    int32_t neuron_model;             /* BGHIP_NEURON_* */
    int32_t synapse_model;            /* BGHIP_SYN_* */
    int32_t ring_bits;                /* delay ring length: 32 (reference) .. 256, multiple of 32 */
    int32_t max_spikes;               /* int(epochDuration * maxFiringRate) */
    float dt;                         /* deltaT */
    uint64_t noise_seed;              /* Philox key of the noise stream */
    int32_t block_steps;              /* steps per fused launch: 0 = automatic, 1 = per-step kernels */
    bghip_synapse_types types;
} bghip_net_desc;

/* Host views of AllIFNeurons / AllIZHNeurons / AllSpikingNeurons arrays (length N unless noted).
 * Izhikevich fields may be NULL for LIF networks. */
typedef struct bghip_neurons {
    float *Vm, *Vthresh, *Vreset, *I0, *Inoise, *C1, *C2, *Trefract, *summation_map;
    int32_t *nStepsInRefr;
    float *Aconst, *Bconst, *Cconst, *Dconst, *u, *C3;
    int32_t *spikeCount, *spikeCountOffset;
    uint64_t *spike_history;          /* N * max_spikes */
} bghip_neurons;

/* Host views of the synapse slot arrays (length N*M unless noted). delayQueue holds ring_bits/32
 * words per slot, slot-major; with ring_bits == 32 it is exactly AllSpikingSynapses::delayQueue.
 * Fields of synapse models not in use may be NULL. */
typedef struct bghip_synapses {
    uint8_t *in_use;
    int32_t *sourceNeuronIndex, *destNeuronIndex, *type;
    float *W, *psr, *decay;
    int32_t *total_delay, *delayIdx;
    uint32_t *delayQueue;
    float *r, *u, *U, *D, *F;         /* DS / DynamicSTDP */
    uint64_t *lastSpike;
    uint32_t *delayQueuePost;         /* STDP */
    int32_t *delayIdxPost;
    float *Apos, *Aneg, *STDPgap, *tauspost, *tauspre, *taupos, *tauneg, *Wex, *mupos, *muneg;
    uint8_t *useFroemkeDanSTDP;
    uint32_t *synapse_counts;         /* N */
} bghip_synapses;

/* SynapseIndexMap (Core/SynapseIndexMap.h:38-104); maps have total_synapse_counts entries. */
typedef struct bghip_index_map {
    uint32_t *incomingSynapseIndexMap, *incomingSynapseBegin, *incomingSynapseCount;
    uint32_t *outgoingSynapseIndexMap, *outgoingSynapseBegin, *outgoingSynapseCount;
} bghip_index_map;

/* ConnGrowth parameters (Connections/ConnGrowth.h m_growth) and layout. */
typedef struct bghip_growth {
    float epsilon, beta, rho, targetRate, maxRate, minRadius, startRadius;
    float epochDuration;
} bghip_growth;

typedef struct bghip_growth_stats {
    int64_t total_synapses;   /* after the update */
    int64_t created;
    int64_t erased;
} bghip_growth_stats;

/* Stimulus input of a network (Inputs/ISInput.h; created by FSInput::CreateInstance,
 * Inputs/FSInput.cpp:40-97). Applied on the device at every step before the neuron advance, as
 * Simulator::advanceUntilGrowth calls inputStimulus before IModel::advance (Core/Simulator.cpp:177-182). */
enum { BGHIP_INPUT_NONE = 0, BGHIP_INPUT_POISSON = 1, BGHIP_INPUT_REGULAR = 2 };

typedef struct bghip_input {
    int32_t kind;                 /* BGHIP_INPUT_* */
    /* SInputPoisson (Inputs/SInputPoisson.cpp:24-146): IntParams fr_mean (Hz) and weight; the
     * input synapse of neuron i is a DS synapse of type EI when neuron_type[i] is INH (1), else
     * EE, with W = weight * 1e-8. Interval draws come from the Philox stream of
     * bg_poisson_isi_steps() in include/bg_detmath.h. */
    float fr_mean, weight;
    const uint8_t *mask;          /* N bytes, nonzero = neuron receives input; NULL = all (no <Masks>) */
    const int32_t *neuron_type;   /* N (Layout::neuron_type_map); required for Poisson input */
    /* SInputRegular (Inputs/SInputRegular.cpp:26-110, Inputs/HostSInputRegular.cpp:59-80) */
    float duration, interval;     /* seconds */
    const float *values;          /* N amperes (the 10x10 <Values> table already tiled over the grid) */
    const int32_t *shift;         /* N: nShiftValues (sync "no" / "wave"); NULL = all 0 */
} bghip_input;

/* Mutable input state (for resuming and for checking): the Poisson interval counters and the
 * per-neuron DS input synapse, in the reference's orientation (delayQueue is the 32-bit ring read
 * at delayIdx), plus SInputRegular's nStepsInCycle. Pointers are N long; upload accepts NULL for
 * the state a freshly constructed input has (nISIs 0, psr 0, r 1, u 0.4, no last spike, empty ring). */
typedef struct bghip_input_state {
    int32_t *nISIs;
    float *psr, *r, *u;
    uint64_t *lastSpike;
    uint32_t *delayQueue;
    int32_t *delayIdx;
    int32_t steps_in_cycle;
} bghip_input_state;

/* ---- context ---------------------------------------------------------------------------- */
/* Replaces GPUSpikingModel::setupSim's cudaSetDevice (Core/GPUSpikingModel.cu:109). Per-step runs
 * of the context's networks and batches replay captured HIP graphs of 256 steps; with BGHIP_GRAPHS=0
 * in the environment at this call the same kernels are launched directly (for tools that do not
 * follow graph launches). */
int bghip_ctx_create(int device, bghip_ctx **out);
void bghip_ctx_destroy(bghip_ctx *ctx);
const char *bghip_last_error(const bghip_ctx *ctx);
/* device properties for schedulers: HBM bytes free/total */
int bghip_ctx_mem_info(bghip_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- network lifetime ------------------------------------------------------------------- */
/* allocNeuronDeviceStruct + allocSynapseDeviceStruct (Neurons/IAllNeurons.h:89-142,
 * Synapses/IAllSynapses.h:125-221): one SoA arena per network. */
int bghip_net_create(bghip_ctx *ctx, const bghip_net_desc *desc, bghip_net **out);
/* deleteNeuronDeviceStruct + deleteSynapseDeviceStruct */
void bghip_net_destroy(bghip_net *net);
/* bytes of device memory a network with this description occupies (admission control) */
int64_t bghip_net_footprint(const bghip_net_desc *desc);

/* copyNeuronHostToDevice (Neurons/AllIFNeurons_d.cu, AllIZHNeurons_d.cu) */
int bghip_upload_neurons(bghip_net *net, const bghip_neurons *h);
/* copySynapseHostToDevice + copySynapseIndexMapHostToDevice (Core/GPUSpikingModel.cu:326-358);
 * the index maps are rebuilt on the device from the slot arrays. `step` is g_simulationStep at
 * the time of the upload (fixes the delay-ring phase). */
int bghip_upload_synapses(bghip_net *net, const bghip_synapses *h, uint64_t step);
/* Layout coordinates and neuron types (Layouts/Layout.h xloc, yloc, neuron_type_map: INH=1,
 * EXC=2) plus growth state; needed by bghip_update_connections. */
int bghip_upload_layout(bghip_net *net, const float *xloc, const float *yloc, const int32_t *neuron_type);
int bghip_upload_growth(bghip_net *net, const bghip_growth *g, const float *radii, const float *rates);

/* ISInput::init for SInputPoisson / SInputRegular on the device; `in` == NULL or kind NONE
 * removes the input. The state is taken as of the network's current step. */
int bghip_upload_input(bghip_net *net, const bghip_input *in, const bghip_input_state *state);
int bghip_download_input(bghip_net *net, bghip_input_state *state);

/* AllSynapses::addSynapse (Synapses/AllSynapses.cpp:357-377) for `count` new synapses
 * src[i] -> dest[i], each followed by W[iSyn] = W[i] as ConnStatic::setupConnections does
 * (Connections/ConnStatic.cpp:66-80): every synapse goes to the first slot of its destination left
 * free by the earlier entries, with createSynapse's per-type constants and the synapse type from
 * the uploaded layout; the incoming index map is rebuilt (createSynapseImap). Entries past a full
 * destination are dropped and set BGHIP_ERR_SLOT_OVERFLOW. Needs bghip_upload_layout first;
 * blocks. */
int bghip_add_synapses(bghip_net *net, int64_t count, const int32_t *src, const int32_t *dest, const float *W);
/* The same with per-synapse values a connection rule sets after addSynapse, as the reference's
 * per-synapse arrays allow (total_delay, Synapses/AllSpikingSynapses.h:360; Wex,
 * Synapses/AllSTDPSynapses.h:380): total_delay[i] in steps (1 .. min(255, ring_bits - 1); it also bounds the fused block
 * length) and, for STDP synapses, Wex[i]; either may be NULL for the per-type value. */
int bghip_add_synapses_ex(bghip_net *net, int64_t count, const int32_t *src, const int32_t *dest, const float *W,
                          const int32_t *total_delay, const float *Wex);

/* Checkpoint of the whole network state, device and host side, as a versioned binary file
 * (Model::serialize / deserialize, Core/Model.cpp:62-98, whose round trip the reference breaks:
 * Synapses/AllSynapses.cpp:136-144 vs 170-177). A loaded network continues bit for bit as the saved
 * one would have; the noise and input streams are counter-based, so the step is their state.
 * save blocks; load creates a new network on ctx's device. */
int bghip_net_save(bghip_net *net, const char *path);
int bghip_net_load(bghip_ctx *ctx, const char *path, bghip_net **out);

/* ---- simulation ------------------------------------------------------------------------- */
/* n_steps calls of GPUSpikingModel::advance (Core/GPUSpikingModel.cu:188-225), starting at the
 * network's current step; stream-ordered. */
int bghip_advance(bghip_net *net, uint64_t n_steps);
/* GPUSpikingModel::updateConnections (Core/GPUSpikingModel.cu:246-259): growth rates, radii,
 * overlap, weight update, synapse create/erase, index-map rebuild, all on the device.
 * stats may be NULL; when given the call synchronises. */
int bghip_update_connections(bghip_net *net, bghip_growth_stats *stats);
/* The weight-update half of it with a caller's matrix instead of the overlap areas: the device
 * side of ConnGrowth::updateSynapsesWeights (Connections/ConnGrowth_d.cu:18-61, which uploads
 * W_h[src * N + dest] = W(src, dest) and runs updateSynapsesWeightsDevice,
 * Synapses/AllSynapsesDeviceFuncs_d.cu:858-916). W is host memory, N x N floats, W[src * N + dest]:
 * an existing synapse with W < 0 is erased, otherwise its weight becomes W * sign * 1e-8; a missing
 * one with W > 0 is created in the destination's first free slot (slots freed by this call's erasures
 * included), sources in ascending order; then the incoming map is rebuilt. Radii and rates are not
 * touched. Needs the layout (neuron types); synchronises. */
int bghip_update_weights(bghip_net *net, const float *W, bghip_growth_stats *stats);
/* bghip_update_weights without any device allocation: no counterpart in the reference. It is for
 * callers that cannot allocate device memory (tight serving loops, code that must not call
 * hipMalloc/hipFree). The device copy of W goes into a static scratch array of the library
 * (16 MiB per GPU, held until the process ends), so N is at most 2048; a larger matrix returns
 * BGHIP_EUNSUPPORTED. Same kernels, results identical to bghip_update_weights. Every call on one GPU
 * shares that GPU's scratch: calls on networks of the same GPU must not overlap (from different
 * threads they race); each call synchronises the network's stream before it returns. */
int bghip_update_weights_static(bghip_net *net, const float *W, bghip_growth_stats *stats);
/* The model part of GPUSpikingModel::updateHistory (Core/GPUSpikingModel.cu:266-272): applies the
 * recorder's minRadius clamp (Recorders/XmlGrowthRecorder.cpp:89-90) and clears spike counts
 * (Neurons/AllSpikingNeurons.cpp:79-87). Download spikes and rates/radii before calling it. */
int bghip_update_history(bghip_net *net);
/* On-device growth recorder: once enabled, bghip_update_history also bins the epoch's spikes
 * into burstinessHist (1 s bins) and spikesHistory (10 ms bins) of XmlRecorder::compileHistories
 * (Recorders/XmlRecorder.cpp:71-104) and, with_growth, stores row e of ratesHistory and of the
 * clamped radiiHistory of XmlGrowthRecorder::compileHistories (Recorders/XmlGrowthRecorder.cpp:72-97),
 * all on the device, so no spike history crosses PCIe per epoch. rates0/radii0 are row 0
 * (XmlGrowthRecorder::initDefaultValues, :27-37), N floats each. Histories are (num_epochs+1) x N. */
int bghip_recorder_enable(bghip_net *net, int64_t n_burst, int64_t n_spikes_hist, int32_t num_epochs,
                          int32_t with_growth, const float *rates0, const float *radii0);
/* Copies the recorder's histories to the host (NULL pointers are skipped); blocking. */
int bghip_recorder_download(bghip_net *net, float *burstinessHist, float *spikesHistory, float *ratesHistory,
                            float *radiiHistory);
/* Waits for the network's stream, as the reference's blocking cudaMemcpy calls wait for its kernels
 * (e.g. Neurons/AllSpikingNeurons_d.cu:41-42, checked by HANDLE_ERROR), and reports what its device
 * asserts would have stopped on: BGHIP_ESTATE if a device error bit is set. */
int bghip_sync(bghip_net *net);
/* The same wait without the error report: the error word stays on the device for the next
 * bghip_sync, bghip_net_error_bits or download (for callers that check errors later, off their
 * timed path). No counterpart in the reference, whose waits always come with a copy. */
int bghip_wait(bghip_net *net);
uint32_t bghip_net_error_bits(bghip_net *net);
uint64_t bghip_net_step(const bghip_net *net);
/* 1 when the network's Spiking/DS synapses advance with implicit delay rings (their rings read from
 * the sources' recent fire bits, DS state shared per source, type and creation cohort), 0 when they
 * use the per-synapse ring kernels. No counterpart in the reference (it reports which of this
 * library's kernel families GPUSpikingModel::advance, Core/GPUSpikingModel.cu:188-225, maps to);
 * results are identical either way. BGHIP_IMPLICIT=0 in the environment at bghip_ctx_create keeps
 * a context's networks on the per-synapse ring kernels. */
int bghip_net_implicit(const bghip_net *net);
/* The description the network was created with (or loaded from a checkpoint). */
int bghip_net_get_desc(const bghip_net *net, bghip_net_desc *out);
/* The on-device recorder's sizes as enabled, and `epoch`: the epochs it has stored (history rows
 * after row 0), which is Simulator::simulate's epoch count, currentStep (Core/Simulator.cpp:106-116),
 * for a network loaded from a checkpoint. NULL pointers are skipped; when the recorder is not
 * enabled the sizes are 0 and epoch is -1. */
int bghip_recorder_info(const bghip_net *net, int64_t *n_burst, int64_t *n_spikes_hist, int32_t *num_epochs,
                        int32_t *with_growth, int32_t *epoch);

/* ---- downloads (blocking) --------------------------------------------------------------- */
int bghip_download_neurons(bghip_net *net, bghip_neurons *h);
/* The spike bookkeeping only: AllSpikingNeurons::copyDeviceSpikeCountsToHost and
 * copyDeviceSpikeHistoryToHost (Neurons/AllSpikingNeurons_d.cu:17-44). counts and offsets hold N
 * ints, history N * max_spikes step numbers (row i = neuron i's ring); NULL pointers are skipped. */
int bghip_download_spikes(bghip_net *net, int32_t *counts, int32_t *offsets, uint64_t *history);
/* AllSpikingNeurons::clearSpikeCounts (Neurons/AllSpikingNeurons.cpp:79-87; the device copy
 * Neurons/AllSpikingNeurons_d.cu:53-59): offset = (count + offset) mod max_spikes, count = 0, for a
 * caller with recorders of its own (bghip_update_history does this itself). Stream-ordered. */
int bghip_clear_spike_counts(bghip_net *net);
int bghip_download_synapses(bghip_net *net, bghip_synapses *h);
int64_t bghip_total_synapses(bghip_net *net);
int bghip_download_index_map(bghip_net *net, bghip_index_map *h);
int bghip_download_growth(bghip_net *net, float *radii, float *rates);

/* ---- batched form ------------------------------------------------------------------------ */
/* A batch advances many independent networks with shared launches: Spiking/DS networks through
 * the multi-network fused block kernel, STDP networks through multi-network per-step kernels.
 * Networks keep their own step counters, noise seeds and state; a network's results are the same
 * alone and in any batch. Member networks may be of different sizes and models. No counterpart in
 * the reference, which runs one network per process (Core/GPUSpikingModel.cu). */
typedef struct bghip_batch bghip_batch;
int bghip_batch_create(bghip_ctx *ctx, bghip_batch **out);
void bghip_batch_destroy(bghip_batch *b);
/* Replace the member list (nets stay owned by the caller; a net may be in one batch at a time). */
int bghip_batch_set(bghip_batch *b, bghip_net *const *nets, int n);
/* Every member advances n_steps from its own current step (stream-ordered on the batch's stream;
 * later calls on a member network wait for it). */
int bghip_batch_advance(bghip_batch *b, uint64_t n_steps);
int bghip_batch_sync(bghip_batch *b);

/* Launch shapes. The members are advanced in groups (one per neuron model, synapse model, ring
 * width and mode); each group's kernels run with a launch shape. Per-step groups replay a captured
 * HIP graph of 256 steps for every whole 256 steps of an advance call. Results are bit-identical
 * for every valid shape. The engine's launch-shape autotuner is also without a reference
 * counterpart (the reference fixes its launch shapes at 256 threads per block, e.g.
 * Neurons/AllLIFNeurons_d.cu:25, Synapses/AllSpikingSynapses_d.cu:337, Core/GPUSpikingModel.cu:235). */
typedef struct bghip_batch_shape {
    int32_t mode;                  /* out: 0 neurons only (no synapses), 1 fused blocks, 2 per step */
    int32_t neuron_model, synapse_model, ring_words;   /* out: the group's kind */
    uint32_t networks, work_items; /* out: members, and neurons (= destinations) over all members */
    int32_t block;                 /* threads per block of the neuron, synapse and fused kernels: 64, 128, 256
                                      (implicit-ring fused groups: always 256, four destinations each) */
    int32_t dests_per_wave;        /* per-step STDP streaming pass: destinations per wavefront, 1, 2 or 4 */
    int32_t block_stream;          /* threads per block of that pass: 64, 128, 256 */
    int32_t syn_blocks_per_shard;  /* blocks per event-list shard of the learning pass, 1..64 */
    int32_t sum_blocks_per_shard;  /* blocks per event-list shard of the event summation pass, 1..64 */
    int32_t tuned;                 /* out: 1 if the shape was picked by the autotuner or set */
    float step_us;                 /* out: the autotuner's measured time per step of the shape (0 if unmeasured) */
} bghip_batch_shape;
/* Autotuner: with trial_steps > 0 (at most 1000), the first advance call of a group that is long
 * enough times its candidate shapes on the group's own next steps (trial_steps timed steps each,
 * coordinate descent over the fields above) and keeps the fastest; the choice is remembered per
 * device for groups of the same kind and size class (work items within a factor of two). Trials
 * synchronise the host with the batch stream. 0 turns the tuner off (the default). */
int bghip_batch_autotune(bghip_batch *b, uint32_t trial_steps);
/* Number of groups (re-forms the groups first if a member changed mode). */
int bghip_batch_group_count(bghip_batch *b);
int bghip_batch_get_shape(bghip_batch *b, int group, bghip_batch_shape *out);
/* Sets group `group`'s shape (input fields only) until the groups are re-formed; marks it tuned. */
int bghip_batch_set_shape(bghip_batch *b, int group, const bghip_batch_shape *shape);

/* ---- timing (measurement support) -------------------------------------------------------- */
/* Average duration in milliseconds of the engine's fused step launches recorded with HIP events
 * on the network's stream since the last reset; count receives the number of launches timed. */
int bghip_kernel_timing(bghip_net *net, int enable, double *avg_ms, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif /* BGHIP_H */
