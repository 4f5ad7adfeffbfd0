/*
 * bghost.h -- C ABI of the host-side model (HipSpikingModel) that drives libbghip.so.
 *
 * HipSpikingModel is the MI355X counterpart of GPUSpikingModel (Core/GPUSpikingModel.h): it owns
 * the host state BrainGrid's Model::setupSim creates (layout, neuron parameters drawn from the
 * global MTRand rng(1), growth radii, recorder histories) and runs the IModel life cycle
 *   setupSim -> { advance x steps, updateConnections, updateHistory } per epoch -> saveData
 * with every per-step and per-epoch computation on the GPU. The same class is what a BrainGrid
 * build links behind IModel; this C ABI exposes it to other languages.
 */
#ifndef BGHOST_H
#define BGHOST_H

#include <stdint.h>
#include "bghip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { BGHOST_LAYOUT_FIXED = 0, BGHOST_LAYOUT_DYNAMIC = 1 };
enum { BGHOST_CONN_NONE = 0, BGHOST_CONN_GROWTH = 1, BGHOST_CONN_STATIC = 2 };

/* The simulation parameters of a BrainGrid XML config (Core/SimulationInfo.h:89-155, neuron
 * ranges of Neurons/AllIFNeurons.h and AllIZHNeurons.h, ConnGrowth's m_growth, layout lists). */
typedef struct bghost_params {
    int32_t width, height;
    int32_t neuron_model, synapse_model;     /* BGHIP_NEURON_*, BGHIP_SYN_* */
    int32_t layout_kind, conn_kind;          /* BGHOST_LAYOUT_*, BGHOST_CONN_* */
    int32_t max_synapses_per_neuron, max_firing_rate, num_epochs, ring_bits;
    float dt, epoch_duration;
    uint32_t seed;                           /* XML Seed: key of the deterministic noise */
    const int32_t *inh_list; int32_t n_inh;
    const int32_t *active_list; int32_t n_active;
    float frac_exc, frac_starter;
    float Iinject[2], Inoise[2], Vthresh[2], Vresting[2], Vreset[2], Vinit[2];
    float starter_Vthresh[2], starter_Vreset[2];
    float excA[2], excB[2], excC[2], excD[2], inhA[2], inhB[2], inhC[2], inhD[2];
    float epsilon, beta, rho, target_rate, min_radius, start_radius;
    int32_t block_steps;                     /* 0 = automatic */
    /* ConnStatic (Connections/ConnStatic.cpp:120-210): nConnsPerNeuron, threshConnsRadius and the
     * weight ranges [minExc, maxExc], [minInh, maxInh] */
    int32_t static_nconns;
    float static_thresh, static_exc_w[2], static_inh_w[2];
    /* stimulus input of the -s file (Inputs/FSInput.cpp:40-97): BGHIP_INPUT_NONE / _POISSON /
     * _REGULAR. SInputPoisson: IntParams fr_mean (Hz) and weight (x 1e-8), <Masks> neuron list
     * (n_input_mask < 0: every neuron). SInputRegular: duration and interval (s), sync
     * BGHOST_SYNC_*, the 10x10 <Values><I> table (tiled over the grid). */
    int32_t input_kind;
    float input_fr_mean, input_weight;
    const int32_t *input_mask; int32_t n_input_mask;
    float input_duration, input_interval;
    int32_t input_sync;
    float input_values[100];
} bghost_params;

enum { BGHOST_SYNC_YES = 0, BGHOST_SYNC_NO = 1, BGHOST_SYNC_WAVE = 2 };

typedef struct bghost_model bghost_model;

/* Model::setupSim + HipSpikingModel device setup (allocate, upload), with the stimulus input
 * constructed before the model's draws and initialised after them, as Core/BGDriver.cpp:76-87 and
 * Core/Simulator.cpp:40-43 order them. */
int bghost_model_create(bghip_ctx *ctx, const bghost_params *p, bghost_model **out);
void bghost_model_destroy(bghost_model *m);
bghip_net *bghost_model_net(bghost_model *m);

/* IModel::advance n_steps times (asynchronous). */
int bghost_model_advance(bghost_model *m, uint64_t n_steps);
/* Simulator::advanceUntilGrowth: one epoch of steps. */
int bghost_model_advance_epoch(bghost_model *m);
/* IModel::updateConnections (on the GPU). */
int bghost_model_update_connections(bghost_model *m, bghip_growth_stats *stats);
/* IModel::updateHistory: recorder histories (Recorders/XmlRecorder.cpp:71-104,
 * XmlGrowthRecorder.cpp:72-97), minRadius clamp and spike-count clear, all on the device; the host
 * copies returned by bghost_model_array are refreshed when read. */
int bghost_model_update_history(bghost_model *m);
/* One full epoch: advance_epoch + update_connections + update_history. */
int bghost_model_run_epoch(bghost_model *m);
/* XmlGrowthRecorder::saveSimData (Recorders/XmlGrowthRecorder.cpp:104-154), same bytes. */
int bghost_model_save_xml(bghost_model *m, const char *path);
/* IModel::serialize (Core/Model.cpp:85-98, written by BGDriver -w, Core/BGDriver.cpp:109-116):
 * the device network with its recorder as a bghip_net_save checkpoint. */
int bghost_model_serialize(bghost_model *m, const char *path);
/* IModel::deserialize (Core/Model.cpp:62-76, BGDriver -r after setupSim, Core/BGDriver.cpp:88-95):
 * replaces the model's network by a checkpoint saved from a model of the same configuration and
 * resumes its epoch count; BGHIP_EINVAL for a checkpoint of another configuration. */
int bghost_model_deserialize(bghost_model *m, const char *path);
/* Overwrite all radii on host and device (pre-grown measurement states). */
int bghost_model_set_radii(bghost_model *m, float r);

/* Host arrays owned by the model; count receives the element count. Names: "xloc", "yloc",
 * "neuron_type_map", "starter_map" (uint8), "Vthresh", "Vreset", "Vinit", "Iinject", "Inoise", "I0",
 * "C1", "C2", "Trefract", "Aconst".."Dconst", "C3", "radiiHistory", "ratesHistory",
 * "burstinessHist", "spikesHistory". Returns NULL for unknown names. */
const void *bghost_model_array(bghost_model *m, const char *name, int64_t *count);
int32_t bghost_model_epoch(const bghost_model *m);

#ifdef __cplusplus
}
#endif
#endif
