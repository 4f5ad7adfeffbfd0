/*
 * bg_oracle.h -- CPU oracle for the spiking-network hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A clean-room C restatement of BrainGrid's single-threaded CPU model
 * (Core/SingleThreadedSpikingModel.cpp:25-58): setup draws, LIF / Izhikevich neuron advance,
 * Spiking / DS / STDP / DynamicSTDP synapse advance with the delay bit-ring, the synapse index
 * map, neurite-growth connectivity and the growth recorder's history compilation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as the checker.
 * The product (libbghip.so) never links or calls it.
 *
 * Modes (all default to the reference's own behaviour):
 *   noise_mode 0 = Norm  : MTRand + Marsaglia polar stream in descending neuron order (RNG/Norm.cpp)
 *              1 = Philox: bg_noise_normal(seed, neuron, step) from include/bg_detmath.h
 *   math_mode  0 = libm  : glibc expf/acosf/sinf/powf/exp exactly where the reference calls them
 *              1 = own   : the bg_* functions of include/bg_detmath.h (bit-identical to the GPU)
 *   sum_mode   0 = sequential per-destination summation in incoming-map order (reference)
 *              1 = wave64: the fixed lane-strided 64-lane tree the HIP kernels use
 * The Poisson stimulus input draws from the global MTRand stream in Norm mode, as the reference's
 * HostSInputPoisson does, and from bg_poisson_isi_steps() in Philox mode.
 * Parity pinning of every function is listed in the header of tests/test_oracle_vs_reference.py.
 */
#ifndef BG_ORACLE_H
#define BG_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { BGO_LIF = 0, BGO_IZH = 1 };
enum { BGO_SPIKING = 0, BGO_DS = 1, BGO_STDP = 2, BGO_DYNSTDP = 3 };
enum { BGO_LAYOUT_FIXED = 0, BGO_LAYOUT_DYNAMIC = 1 };
enum { BGO_CONN_NONE = 0, BGO_CONN_GROWTH = 1, BGO_CONN_STATIC = 2 };

typedef struct bgo_params {
    int32_t width, height;
    int32_t neuron_kind, syn_kind, layout_kind, conn_kind;
    int32_t max_synapses_per_neuron;
    int32_t max_firing_rate;
    int32_t num_epochs;
    int32_t ring_bits;            /* 32 = reference; any multiple of 32 up to 1024 */
    float dt, epoch_duration;
    uint32_t seed;                /* XML Seed: Norm seed or Philox key */
    int32_t noise_mode, math_mode, sum_mode;
    /* layout */
    const int32_t *inh_list; int32_t n_inh;
    const int32_t *active_list; int32_t n_active;
    float frac_exc, frac_starter;
    /* neuron parameter ranges [min, max] */
    float Iinject[2], Inoise[2], Vthresh[2], Vresting[2], Vreset[2], Vinit[2];
    float starter_Vthresh[2], starter_Vreset[2];
    float excA[2], excB[2], excC[2], excD[2], inhA[2], inhB[2], inhC[2], inhD[2];
    /* growth */
    float epsilon, beta, rho, target_rate, min_radius, start_radius;
    /* STDP creation-time overrides (reference defaults when stdp_override == 0) */
    int32_t stdp_override;
    float stdp_Apos, stdp_Aneg, stdp_taupos, stdp_tauneg, stdp_tauspre, stdp_tauspost;
    float stdp_gap, stdp_Wex, stdp_mupos, stdp_muneg;
    int32_t stdp_use_froemke_dan;
    /* synapse delay overrides per type II, IE, EI, EE in seconds (0 = reference table) */
    float delay_override[4];
    /* stimulus input (Inputs/FSInput.cpp:40-97): BGO_INPUT_NONE / _POISSON / _REGULAR */
    int32_t input_kind;
    float input_fr_mean, input_weight;    /* SInputPoisson IntParams: fr_mean (Hz), weight (x 1e-8) */
    const int32_t *input_mask;            /* SInputPoisson masked neuron list ... */
    int32_t n_input_mask;                 /* ... its length; < 0: every neuron receives input */
    float input_duration, input_interval; /* SInputRegular IntParams (seconds) */
    int32_t input_sync;                   /* SInputRegular sync: 0 "yes", 1 "no", 2 "wave" */
    float input_values[100];              /* SInputRegular <Values><I>: the 10x10 table */
    /* ConnStatic (Connections/ConnStatic.cpp:120-210): nConnsPerNeuron, threshConnsRadius, weight ranges */
    int32_t static_nconns;
    float static_thresh, static_exc_w[2], static_inh_w[2];
} bgo_params;

enum { BGO_INPUT_NONE = 0, BGO_INPUT_POISSON = 1, BGO_INPUT_REGULAR = 2 };
enum { BGO_SYNC_YES = 0, BGO_SYNC_NO = 1, BGO_SYNC_WAVE = 2 };

typedef struct bgo_net bgo_net;

/* Model::setupSim equivalent. Returns NULL and fills err on failure. */
bgo_net *bgo_create(const bgo_params *p, char *err, int errlen);
void bgo_destroy(bgo_net *n);

/* Advance n_steps whole steps (IModel::advance + g_simulationStep++). 0 = ok, <0 = error. */
int bgo_advance(bgo_net *n, int64_t n_steps);
/* IModel::updateConnections (growth + index map). */
int bgo_update_connections(bgo_net *n);
/* ConnGrowth::updateSynapsesWeights with a caller's dense weight matrix W[src * N + dest] (instead
 * of the overlap areas), then the index map: W < 0 erases an existing synapse, W >= 0 sets its
 * weight, W > 0 creates a missing one. */
int bgo_update_weights(bgo_net *n, const float *W);
/* IModel::updateHistory (recorder compileHistories incl. minRadius clamp, clear spike counts). */
int bgo_update_history(bgo_net *n);
/* One full epoch: advance(epoch steps) + update_connections + update_history. */
int bgo_run_epoch(bgo_net *n);

/* addSynapse (Synapses/AllSynapses.cpp:357-377) with createSynapse defaults, then W := w.
 * Call bgo_rebuild_index_map() after a batch of additions. Returns slot index or <0. */
int64_t bgo_add_synapse(bgo_net *n, int32_t src, int32_t dest, float w);
int bgo_rebuild_index_map(bgo_net *n);
/* Overwrite every radius (pre-grown measurement states). */
void bgo_set_radii(bgo_net *n, float r);

/* Named state arrays. dtype: 'f' float32, 'd' float64, 'i' int32, 'I' uint32, 'q' uint64,
 * 'b' uint8. Returns NULL for an unknown name. */
void *bgo_array(bgo_net *n, const char *name, char *dtype, int64_t *count);
int64_t bgo_step(const bgo_net *n);
int64_t bgo_total_synapses(const bgo_net *n);
const char *bgo_last_error(const bgo_net *n);

/* Direct access to the pieces tests pin individually. */
float bgo_overlap_area(float x1, float y1, float r1, float x2, float y2, float r2, int math_mode);
void bgo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float bgo_noise(uint64_t seed, uint32_t neuron, uint64_t step);
float bgo_norm_stream(uint32_t seed, int n, float *out);   /* first n Norm draws */
double bgo_detmath(int fn, double x, double y);            /* 0 exp 1 log 2 sin 3 cos 4 acos 5 pow 6 cos2pi 7 expf */
/* ConnStatic's candidate order: std::sort of (dist, dest) pairs compared by dist alone
 * (Connections/ConnStatic.cpp:62, ConnStatic.h:138-146). The tie order is the standard library's,
 * so this one function is C++ (oracle/bg_oracle_sort.cpp), compiled against the same libstdc++ as
 * the reference build. */
void bgo_sort_by_dist(float *dist, int32_t *dest, int32_t n);
/* Poisson input inter-spike interval in steps drawn from the counter-based stream. */
int32_t bgo_poisson_isi(uint64_t seed, uint32_t neuron, uint64_t step, float lambda_ms, float dt);

#ifdef __cplusplus
}
#endif
#endif
